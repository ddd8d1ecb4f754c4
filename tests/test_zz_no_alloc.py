"""portal_inverse_rt_device, the allocation-free inward pass, against portal_inverse_rt
through the C ABI (needs an MI355X).

The outputs are device buffers the caller owns (torch tensors here); the direction
table goes to the library's static scratch array.  Two contexts on the one device are
primed with the allocating entry point, then the allocation-free one is called on
both in a row, the second with the smaller input (fewer nodes, fewer directions), so
the second call runs on scratch that still holds the first call's larger table.  Its
results must equal the allocating entry point's bit for bit, in both precisions.
"""
import numpy as np
import pytest
import torch

from portal_amd import cloud
from portal_amd.api import Portal, PortalError
from portal_amd.lib import PORTAL_FP64, PORTAL_MIXED

pytestmark = pytest.mark.gpu


def _ctx(n_nodes, nth, nph, seed):
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, n_nodes, seed=seed)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    return pt, g, sp, cloud.legendre_all(nth, nph)


@pytest.mark.parametrize("precision", [PORTAL_FP64, PORTAL_MIXED])
def test_no_alloc_inverse_rt_equals_allocating_entry_point(precision):
    big = _ctx(900, 6, 11, seed=21)
    small = _ctx(400, 4, 8, seed=22)
    assert small[3].shape[1] < big[3].shape[1]
    try:
        ref = []
        for pt, g, sp, xth in (big, small):
            J0, J2, _ = pt.inverse_rt(xth, 16, cloud.DX_INWARD, precision=precision)
            ref.append((J0, J2))
        for k, (pt, g, sp, xth) in enumerate((big, small)):
            d0 = torch.full((g.node_num, sp.r_num), float("nan"), dtype=torch.float64, device="cuda")
            d2 = torch.full_like(d0, float("nan"))
            torch.cuda.synchronize()
            pt.inverse_rt_device(xth, 16, cloud.DX_INWARD, d0.data_ptr(), d2.data_ptr(), precision=precision)
            # d0 holds r_num x node_num column-major = node-major rows of r_num
            J0 = d0.cpu().numpy().T
            J2 = d2.cpu().numpy().T
            print(f"no-alloc call {k}: {g.node_num} nodes, {xth.shape[1]} directions, "
                  f"max |J0| {np.abs(J0).max():.3e}", flush=True)
            np.testing.assert_array_equal(J0, ref[k][0])
            np.testing.assert_array_equal(J2, ref[k][1])
            assert np.abs(J0).max() > 0
    finally:
        big[0].close()
        small[0].close()


def test_no_alloc_inverse_rt_refuses_what_it_cannot_do():
    pt, g, sp, xth = _ctx(300, 4, 8, seed=23)
    try:
        d0 = torch.zeros((g.node_num, sp.r_num), dtype=torch.float64, device="cuda")
        d2 = torch.zeros_like(d0)
        # no pass yet: no buffers, paths or launch shape to reuse
        with pytest.raises(PortalError) as e:
            pt.inverse_rt_device(xth, 16, cloud.DX_INWARD, d0.data_ptr(), d2.data_ptr())
        assert e.value.code == -7
        pt.inverse_rt(xth, 16, cloud.DX_INWARD)
        # other directions than the built paths'
        with pytest.raises(PortalError) as e:
            pt.inverse_rt_device(cloud.legendre_all(4, 9), 16, cloud.DX_INWARD, d0.data_ptr(), d2.data_ptr())
        assert e.value.code == -7
        # more directions than the scratch holds: refused before any device work
        many = np.tile(xth, (1, 2000))
        with pytest.raises(PortalError) as e:
            pt.inverse_rt_device(many, 16, cloud.DX_INWARD, d0.data_ptr(), d2.data_ptr())
        assert e.value.code == -1 and "scratch" in str(e.value)
    finally:
        pt.close()
