"""Benchmark: converged grid nodes/s of the mixed-precision + surrogate path against the
plain FP64 GPU path, on the 200,000-node production cloud (10-level CO-like rotor, 200
ray directions, 129 channels), with the Stokes I/Q/U image error (513 x 513 pixels x 129
channels) of the fast path against the FP64 result.

A step is one converged solve: from the resident grid and LTE starting populations,
build the inward paths, then iterate (inward ray integration + SE solve) until the
largest relative level-population change is <= 1e-6 and the largest change of
rho^2_0/rho^0_0 is <= 1e-8.  value = solved nodes / step time of the fast path
(mixed precision, surrogate initial iterate, Ng acceleration).  The plain FP64 path
runs the configuration's 50 iterations (it meets the same tolerance at about the
50th), timed once; the FP64 path with Ng acceleration to the same tolerance (the
like-for-like denominator) and ablations without Ng and without the surrogate are
timed once each.  The CPU baseline is the reference's own inverse_rt compiled from its
Fortran sources (oracle/_ref), on a node sample, on the host's cores.

    python bench.py [--gpus 1] [--steps 1] [--warmup 0] [--nodes 200000]

Prints one JSON line.  Runs on one GPU only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stokes_error(ref, got):
    """Errors of an image against the FP64 image: on the scale of max |I| (the gate), and
    Q and U also on their own scales, with the image's polarisation level beside them."""
    I0, Q0, U0 = ref
    I1, Q1, U1 = got
    s = np.abs(I0).max()
    sq, su = np.abs(Q0).max(), np.abs(U0).max()
    return {"I_max": float(np.abs(I1 - I0).max() / s), "Q_max": float(np.abs(Q1 - Q0).max() / s),
            "U_max": float(np.abs(U1 - U0).max() / s), "I_median": float(np.median(np.abs(I1 - I0)) / s),
            "Q_median": float(np.median(np.abs(Q1 - Q0)) / s), "U_median": float(np.median(np.abs(U1 - U0)) / s),
            "Q_max_own_scale": float(np.abs(Q1 - Q0).max() / sq) if sq > 0 else None,
            "U_max_own_scale": float(np.abs(U1 - U0).max() / su) if su > 0 else None,
            "max_Q_over_max_I": float(sq / s), "max_U_over_max_I": float(su / s)}


def cpu_baseline(g, sp, xth, args, log):
    """The CPU baseline: the reference's own inverse_rt (inverse_rt.f90:1-137, compiled from
    its Fortran sources into oracle/_ref, at its own -O1 or, with --ref-build O3, at -O3), one
    call per node over the host's threads as its OpenMP node loop runs it, on a bounded node
    sample of the same cloud and one FP64 pass; the project's C restatement is timed on the
    same sample beside it.  The reference's SE solve needs a library this image lacks, so the
    sample covers the inward transfer, which is nearly all of the reference's SE pass."""
    if args.cpu_sample <= 0:
        log("cpu baseline skipped: --cpu-sample 0")
        return None
    try:
        from oracle.bind import LIB_REF, Oracle, Reference
        from portal_amd import cloud
        nthr = args.cpu_threads if args.cpu_threads > 0 else min(16, os.cpu_count() or 1)
        rng = np.random.default_rng(1)
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        sub = np.sort(rng.choice(idx, min(args.cpu_sample, len(idx)), replace=False)) + 1
        tp = time.perf_counter()
        Oracle().inverse_nodes(g, sp, xth, args.nvel, cloud.DX_INWARD, sub, nthreads=nthr)
        tp = time.perf_counter() - tp
        lib = LIB_REF if args.ref_build == "O1" else os.path.join(os.path.dirname(LIB_REF), "O3", "libportal_ref.so")
        try:
            ref = Reference(lib)
            kind = "reference"
        except FileNotFoundError:
            ref, kind = None, "port"
        tc = tp
        if ref is not None:
            tc = time.perf_counter()
            ref.inverse_nodes(g, sp, xth, args.nvel, cloud.DX_INWARD, sub, nthreads=nthr)
            tc = time.perf_counter() - tc
        per_pass = len(sub) / tc
        what = (f"the reference's inverse_rt (inverse_rt.f90:1-137, its Fortran built with amdflang -{args.ref_build})"
                if kind == "reference" else "the C restatement of inverse_rt")
        log(f"CPU {kind}: {per_pass:.1f} nodes/s per pass on {nthr} threads (C restatement {len(sub) / tp:.1f})")
        return {"value": per_pass / args.iters, "unit": "converged grid nodes/s", "cores": nthr, "kind": kind,
                "sample": f"{what}: the inward transfer of {len(sub)} seeded nodes of the same cloud, one FP64 "
                          f"pass, {tc:.1f}s on {nthr} threads; converged rate = per-pass rate / {args.iters} "
                          f"iterations (the SE solve, a small part of a pass, is not in the sample)",
                "nodes_per_pass_per_s": per_pass, "port_nodes_per_pass_per_s": len(sub) / tp}
    except Exception as e:  # checkers not built
        log(f"cpu baseline skipped: {e}")
        return None


def image_cpu_port(g, sp, pt, lines, vel_out, args, log):
    """The image pass's CPU baseline: the oracle's restatement of the reference's image
    loops (raytrace.f90, para_finish.f90:382-449; the reference's own image loops need
    em_coefs.f90, which this image cannot build) on 1,024 consecutive pixels about the
    cube's centre, FP64, on the host's threads."""
    if args.cpu_sample <= 0:
        return None
    try:
        from oracle.bind import Oracle
        from portal_amd import cloud, driver
        nthr = args.cpu_threads if args.cpu_threads > 0 else min(16, os.cpu_count() or 1)
        rp = driver.rescale_pol_pops(pt.get_pol_pops(), pt.get_levpop(), sp.lev_g)
        n_pix = (2 * args.im_pix + 1) ** 2
        npx = min(1024, n_pix)
        lo = (n_pix - npx) // 2 + 1
        t = time.perf_counter()
        Oracle().image_pass(g, sp, rp, lines, np.pi / 4, 0.0, 64, vel_out, args.im_pix, 0.7, cloud.DX_IMAGE,
                            n_depth_rt=2048, pix_lo=lo, pix_hi=lo + npx - 1, nthreads=nthr)
        t = time.perf_counter() - t
        rate = npx * 129 * len(lines) / t
        log(f"CPU image (C restatement): {rate:.3g} pixel-channels/s on {nthr} threads")
        return {"pixel_channels_per_s": rate, "cores": nthr, "kind": "port",
                "sample": f"{npx} consecutive pixels about the centre, {len(lines)} line x 129 channels, FP64, "
                          f"{t:.1f}s"}
    except Exception as e:  # checker not built
        log(f"cpu image baseline skipped: {e}")
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--nodes", type=int, default=200000)
    ap.add_argument("--levels", type=int, default=10)
    ap.add_argument("--nth", type=int, default=9)
    ap.add_argument("--nph", type=int, default=32)
    ap.add_argument("--nvel", type=int, default=64)
    ap.add_argument("--iters", type=int, default=50, help="the FP64 configuration's iteration count")
    ap.add_argument("--tol", type=float, default=1e-6, help="relative level-population change at convergence")
    ap.add_argument("--tol-align", type=float, default=1e-8, help="change of rho20/rho00 at convergence")
    ap.add_argument("--no-ablation", action="store_true", help="skip the runs without Ng / without surrogate")
    ap.add_argument("--no-ng", action="store_true", help="fast path without Ng acceleration")
    ap.add_argument("--im-pix", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (default: up to 16)")
    ap.add_argument("--ref-build", choices=("O1", "O3"), default="O1",
                    help="the reference build timed as the CPU baseline: its own -O1 or -O3")
    ap.add_argument("--cpu-only", action="store_true",
                    help="build the cloud, time only the CPU baseline, print its JSON and exit (no GPU)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is "
              "a design only and is not run", file=sys.stderr)
        sys.exit(2)

    from portal_amd import cloud, driver
    from portal_amd.api import Portal
    from portal_amd.lib import PORTAL_FP64, PORTAL_MIXED
    from portal_amd import surrogate

    t0 = time.time()
    sp = cloud.co_like_rotor(args.levels)
    g = cloud.make_cloud(sp, args.nodes, seed=args.seed)
    xth = cloud.legendre_all(args.nth, args.nph)
    N_b = int((g.ab_all[: g.bulk_node_num] > 0).sum())
    log(f"cloud: {g.node_num} nodes ({N_b} solved), {xth.shape[1]} directions, {2 * args.nvel + 1} channels, "
        f"{sp.e_num} levels; built in {time.time() - t0:.1f}s")
    if args.cpu_only:
        cpu = cpu_baseline(g, sp, xth, args, log)
        print(json.dumps({"cpu_baseline": cpu, "nodes": args.nodes, "levels": args.levels,
                          "directions": int(xth.shape[1]), "channels": 2 * args.nvel + 1}), flush=True)
        return

    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    have_surr = surrogate.load_into(pt, sp)
    log(f"surrogate network: {'loaded' if have_surr else 'not available'}")

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731

    lines = driver.image_lines(sp, [1], [2])
    vel_out = driver.channel_velocities(64, 46.875)

    def image(precision=PORTAL_MIXED):
        pp = pt.get_pol_pops()
        lev = pt.get_levpop()
        rp = driver.rescale_pol_pops(pp, lev, sp.lev_g)
        (I, Q, U, T), st = pt.image_pass(rp, lines, np.pi / 4, 0.0, 64, vel_out, args.im_pix, 0.7, cloud.DX_IMAGE,
                                         n_depth_rt=2048, precision=precision)
        return (I, Q, U), st

    out = {}
    img64 = None
    if not args.no_fp64:
        # plain FP64 GPU path: the configuration's fixed iteration count, timed in two parts so that the
        # time at which it met the same tolerance as the fast path is known too
        # in calls of at most 10 iterations (a call continues the previous one's iteration,
        # so the first iteration that meets the tolerance is found exactly), with a line per call
        pt.set_levpop(g.lev_p)
        sync()
        t0 = time.perf_counter()
        done, it_tol, t_tol, walk_ms = 0, None, None, 0.0
        while done < args.iters:
            n = min(10, args.iters - done)
            tol = (args.tol, args.tol_align) if it_tol is None else (0.0, 0.0)
            st = pt.lambda_iterate(xth, args.nvel, cloud.DX_INWARD, n, tol_pop=tol[0], tol_align=tol[1],
                                   precision=PORTAL_FP64, rebuild_paths=1 if done == 0 else 0, outputs=False)
            sync()
            done += st["iterations"]
            walk_ms += st["t_walk_ms"]
            if it_tol is None and st["converged"]:
                it_tol, t_tol = done, time.perf_counter() - t0
            log(f"FP64: {done} iterations, {time.perf_counter() - t0:.1f}s, last change {st['max_rel_change']:.2e}")
        t64 = time.perf_counter() - t0
        if it_tol is None:
            it_tol, t_tol = None, None
        log(f"FP64: tolerance met at iteration {it_tol}; {args.iters} iterations in {t64:.2f}s (walk {walk_ms:.0f} ms)")
        img64, sti = image(PORTAL_FP64)
        log(f"FP64 image: {sti['t_img_ms']:.0f} ms")
        out["fp64"] = {"iterations": args.iters, "seconds": t64, "nodes_per_s": N_b / t64,
                       "iterations_to_tol": it_tol, "seconds_to_tol": t_tol, "image_ms": sti["t_img_ms"]}
        if not args.no_ng:
            # the FP64 path with the same Ng acceleration, to the fast path's tolerance: the
            # like-for-like denominator of the precision + surrogate gain
            pt.set_levpop(g.lev_p)
            sync()
            t0 = time.perf_counter()
            st = pt.lambda_iterate(xth, args.nvel, cloud.DX_INWARD, 4 * args.iters, tol_pop=args.tol,
                                   tol_align=args.tol_align, precision=PORTAL_FP64, rebuild_paths=1, outputs=False,
                                   ng_accel=1)
            sync()
            tng = time.perf_counter() - t0
            log(f"FP64+Ng: {st['iterations']} iterations ({st['ng_steps']} Ng), {tng:.1f}s, "
                f"converged {bool(st['converged'])}")
            out["fp64_ng"] = {"iterations": st["iterations"], "seconds": tng, "nodes_per_s": N_b / tng,
                              "converged": bool(st["converged"])}

    def fast(surr, ng):
        pt.set_levpop(g.lev_p)
        sync()
        t = time.perf_counter()
        st = pt.lambda_iterate(xth, args.nvel, cloud.DX_INWARD, 4 * args.iters, tol_pop=args.tol,
                               tol_align=args.tol_align, precision=PORTAL_MIXED, surrogate=surr, rebuild_paths=1,
                               outputs=False, ng_accel=ng)
        sync()
        return time.perf_counter() - t, st

    surr = 1 if have_surr else 0
    ng = 0 if args.no_ng else 1
    if not args.no_ablation:
        for name, (su, n) in {"mixed": (0, 0), "mixed+surrogate": (surr, 0), "mixed+ng": (0, ng)}.items():
            if (su, n) == (surr, ng) or (name != "mixed" and (su, n) == (0, 0)):
                continue
            t_a, st_a = fast(su, n)
            img_a, _ = image()
            log(f"ablation {name}: {st_a['iterations']} iterations, {t_a:.2f}s")
            out.setdefault("ablation", {})[name] = {
                "iterations": st_a["iterations"], "seconds": t_a, "nodes_per_s": N_b / t_a,
                "speedup_vs_fp64_gpu": out["fp64"]["seconds"] / t_a if "fp64" in out else None,
                "stokes_err_vs_fp64": stokes_error(img64, img_a) if img64 else None}
    times = []
    stm = None
    for k in range(args.warmup + args.steps):
        t, stm = fast(surr, ng)
        log(f"mixed{'+surrogate' if surr else ''}{'+ng' if ng else ''} step {k}: {stm['iterations']} iterations "
            f"({stm['ng_steps']} Ng), {t:.2f}s, fallbacks {stm['fallbacks']}, rt {stm['t_rt_ms']:.0f} ms, "
            f"walk {stm['t_walk_ms']:.0f} ms")
        if k >= args.warmup:
            times.append(t)
    T = max(times)
    img_m, stim = image(PORTAL_MIXED)
    err = stokes_error(img64, img_m) if img64 else None

    # roofline of the dominant kernel (the mixed inward integration): algorithmic bytes per
    # ray-cell crossing, 28 nn (the walk's neighbour ids and positions) + 4 (4 + L) (the
    # FP32 cell record and its L level populations)
    nn = float(np.diff(g.nbr_ptr).mean())
    bytes_per_cross = 28.0 * nn + 4.0 * (4 + sp.e_num)
    launches = max(1, stm["iterations"])
    avg_ms = stm["t_rt_ms"] / launches
    achieved = stm["n_cross"] * bytes_per_cross / (avg_ms * 1e-3) / 1e9
    # the kernel's exponentials: 2 per (crossing, line, channel), against the measured
    # v_exp_f32 issue rate of the device
    exp_rate = 2.0 * sp.r_num * (2 * args.nvel + 1) * stm["n_cross"] / (avg_ms * 1e-3)
    exp_peak = pt.exp_peak()
    # secondary roofline: the SE solve against FP64 peak, with (2/3) c^3 (LU) + 2 c^2
    # (solve) + 2 c^2 (refinement) + 4 c^2 (assembly, dense bound) flops per node
    _, c_dim = pt.species_info()
    se_flops = (2.0 / 3.0) * c_dim ** 3 + 8.0 * c_dim ** 2
    se_ms = stm["t_se_ms"] / max(1, stm["iterations"])
    se_tfs = N_b * se_flops / (se_ms * 1e-3) / 1e12
    secondary = {"se_solve": {"bound": "fp64", "achieved": se_tfs, "peak": 78.6, "unit": "TFLOP/s",
                              "frac": se_tfs / 78.6, "ms_per_pass": se_ms, "c_dim": c_dim,
                              "flops_per_node": se_flops}}
    # the surrogate network's inference (one forward of every solved node, BF16 MFMA with
    # FP32 accumulation): 2 din dout flops per node and layer over the GEMMs, against the
    # dense BF16 MFMA peak; its time is the device time of the normalisation + GEMMs
    if surr and stm.get("t_mlp_ms", 0.0) > 0.0:
        n_in, width, depth, n_out = pt.mlp_dims
        dims = [n_in] + [width] * depth + [n_out]
        mlp_flops = 2.0 * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
        mlp_tfs = N_b * mlp_flops / (stm["t_mlp_ms"] * 1e-3) / 1e12
        secondary["mlp_inference"] = {"bound": "mfma_bf16", "achieved": mlp_tfs, "peak": 2500.0, "unit": "TFLOP/s",
                                      "frac": mlp_tfs / 2500.0, "ms": stm["t_mlp_ms"], "rows": N_b,
                                      "flops_per_row": mlp_flops, "layers": dims}
    # the dominant kernel's HBM traffic and instruction issue, from its rocprofv3 counter
    # runs on this workload (profiles/)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "rt_mixed_traffic.json")
    if os.path.exists(tf):
        prof = json.load(open(tf))
        traffic = prof.get("bytes_per_launch")
        if "valu_instr_per_launch" in prof:
            secondary["rt_valu_issue"] = {
                "bound": "valu", "frac": prof["valu_issue_quad_cycle_frac"],
                "dual_issue_frac": prof["valu_dual_issue_quad_cycle_frac"],
                "valu_instr_per_launch": prof["valu_instr_per_launch"],
                "valu_instr_per_s": prof["valu_instr_per_launch"] / (avg_ms * 1e-3),
                "note": "frac: SIMD quad-cycles that issue a VALU instruction (rocprofv3 counters, "
                        "profiles/" + prof.get("counters", "") + ")"}

    # the image pass (timed separately, as the metric's definition asks): pixel x channel
    # rates, the march's HBM roofline on the algorithmic bytes per visited cell (28 nn for
    # the neighbour ids and positions, 8 (1 + 1 + 3 + 4 lines) for the cell's velocity, b,
    # field and polarisation coefficients) over the start search + march time, and the
    # Stokes kernel's (pixel, channel, cell) update rate
    n_pix, n_ch = (2 * args.im_pix + 1) ** 2, 2 * 64 + 1
    cells = stim["n_cross"]
    img_bytes = cells * (28.0 * nn + 8.0 * (1 + 1 + 3 + 4 * len(lines)))
    march_gbs = img_bytes / (stim["t_walk_ms"] * 1e-3) / 1e9 if stim["t_walk_ms"] > 0 else None
    secondary["image_march"] = {"bound": "hbm", "achieved": march_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": march_gbs / HBM_PEAK_GBS if march_gbs else None, "ms": stim["t_walk_ms"],
                                "visited_cells": cells, "longest_ray_cells": stim["max_path"]}
    secondary["image_stokes_mixed"] = {"updates_per_s": cells * n_ch * len(lines) / (stim["t_rt_ms"] * 1e-3),
                                       "ms": stim["t_rt_ms"],
                                       "note": "(pixel, channel, cell) FP32 Stokes updates, 2 exponentials each"}
    img_cpu = image_cpu_port(g, sp, pt, lines, vel_out, args, log)

    cpu = cpu_baseline(g, sp, xth, args, log)

    res = {
        "metric": "converged grid nodes/s: mixed+surrogate vs FP64 GPU at <=1e-4 rel Stokes err",
        "value": N_b / T,
        "unit": "converged grid nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": T * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic seeded cloud",
        "config": {"workload": f"{args.nodes}-node cloud, {sp.e_num}-level rotor, {xth.shape[1]} directions, "
                               f"{2 * args.nvel + 1} channels; FP64 path {args.iters} lambda iterations, fast path "
                               f"to population change <= {args.tol:g} and alignment change <= {args.tol_align:g}",
                   "nodes": args.nodes, "levels": sp.e_num, "directions": int(xth.shape[1]),
                   "channels": 2 * args.nvel + 1, "surrogate": bool(have_surr),
                   "image": f"{2 * args.im_pix + 1}^2 x 129"},
        "speedup_vs_fp64_gpu": (out["fp64"]["seconds"] / T) if "fp64" in out else None,
        "speedup_vs_fp64_gpu_same_tol": (out["fp64"]["seconds_to_tol"] / T)
        if "fp64" in out and out["fp64"]["seconds_to_tol"] else None,
        # like for like: the same acceleration on both sides
        "speedup_vs_fp64_ng_gpu": (out["fp64_ng"]["seconds"] / T) if "fp64_ng" in out else None,
        "speedup_mixed_surrogate_vs_fp64_gpu": (out["fp64"]["seconds"] / out["ablation"]["mixed+surrogate"]["seconds"])
        if "fp64" in out and "mixed+surrogate" in (out.get("ablation") or {}) else None,
        "fp64_gpu": out.get("fp64"),
        "fp64_ng_gpu": out.get("fp64_ng"),
        "fast_path": {"surrogate": bool(surr), "ng_accel": bool(ng), "iterations": stm["iterations"],
                      "ng_steps": stm["ng_steps"], "surrogate_fallbacks": stm["fallbacks"],
                      "walk_ms": stm["t_walk_ms"], "rt_ms": stm["t_rt_ms"], "se_ms": stm["t_se_ms"],
                      "rt_kernel": f"variant {stm['rt_variant']}, {stm['rt_waves']} waves"},
        "ablation": out.get("ablation"),
        "stokes_err_vs_fp64": err,
        "image": {"pixels": n_pix, "channels": n_ch, "mixed_ms": stim["t_img_ms"],
                  "fp64_ms": out["fp64"]["image_ms"] if "fp64" in out else None,
                  "mixed_pixel_channels_per_s": n_pix * n_ch / (stim["t_img_ms"] * 1e-3),
                  "fp64_pixel_channels_per_s": (n_pix * n_ch / (out["fp64"]["image_ms"] * 1e-3)
                                                if "fp64" in out else None),
                  "cpu_port": img_cpu},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "rt_coop_kernel<float>" if stm["rt_variant"] == 2 else f"variant {stm['rt_variant']}",
                     "avg_launch_ms": avg_ms, "crossings_per_launch": stm["n_cross"],
                     "bytes_per_crossing": bytes_per_cross, "exp_per_s": exp_rate,
                     "exp_peak_per_s": exp_peak, "exp_frac": exp_rate / exp_peak},
        "cpu_baseline": cpu,
        "secondary_rooflines": secondary,
    }
    print(json.dumps(res), flush=True)
    pt.close()


if __name__ == "__main__":
    main()
