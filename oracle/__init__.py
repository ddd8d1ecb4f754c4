"""CPU checkers for the PORTAL hot path (test infrastructure only).

``oracle.bind`` loads the project's C restatement (``liboracle.so``) and, where it
has been built from the reference's own sources, ``_ref/libportal_ref.so``.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import
this package; the product path never does.
"""
