#!/usr/bin/env python3
"""Throughput of the SMI shift map on one MI355X: maps/s at 1024² (full-resolution 2048² FFT,
45 correlations), with the roofline of the dominant kernel and a CPU baseline in the same run.

    python bench.py [--steps K] [--warmup W] [--corr 45|28] [--no-cpu-baseline]

A step is one dense shift map of the resident 1024² pair (shift (113, -57), 8-bit intensities):
canvas warp/paste → pack + row FFT → column FFT → cross spectrum → column IFFT → row IFFT →
FP64 per-lag score → argmax. Inputs (the features) are resident in HBM before timing starts.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32; single-issue FMA (no packed FP32 math) tops out at half
# Algorithmic FP32-equivalent operations per destination pixel and pose (counted from the kernels'
# arithmetic, adds and multiplies one each, an FMA two): pose moments = projection 15 + taper 8 +
# 4 bilinear samples 36 + the 45 weighted monomials 89; tangent pass with the Gauss-Newton matrix =
# projection 15 + taper 20 + samples and their gradients 68 + residual 100 + (cx, cy) 20 +
# 8 tangent derivatives 144 + gradient 32 + the 36-entry matrix 288.
POSE_MOMENT_FLOPS = 148
TANGENT_FLOPS = 687


def kernel_bytes(plan, S, canvas_px, nx, ny, ch):
    """Algorithmic HBM bytes per map of each map kernel, summed over its launches (the minimal
    traffic of its own role)."""
    N = nx * ny
    PA, PB, Q, C, T = plan["n_cplx_a"], plan["n_cplx_b"], plan["n_out"], plan["n_corr"], plan["n_sat"]
    TB = plan.get("n_bsat", 0)
    # a translation plan reads both images' own tables (made once per image, S cells each); a
    # rectangle plan builds canvas A's tables per map
    tab_px = S if TB else canvas_px
    out = {
        "warp_stack": 20 * S + 20 * canvas_px,
        "paste_stack": 20 * S + 20 * canvas_px,
        # a translation plan's pack reads the images in place (no canvases)
        "pack_rowfft": 2 * 20 * (S if TB else canvas_px) + (PA + PB) * ch * nx * 8,
        "colfft_fwd": (PA + PB) * (ch * nx * 8 + N * 8),
        "cmul": (PA + PB) * N * 8 + Q * N * 8,
        "colfft_inv": 2 * Q * N * 8,
        "rowifft_extract": Q * N * 8 + C * N * 4,
        "combine": C * N * 4 + N * 4 + (T + TB) * tab_px * 8,
    }
    if T and not TB:
        # band sums and rows each read the canvas stack; the tables are written once
        out["sat_build"] = 2 * 20 * canvas_px + T * canvas_px * 8
    return out


def survey_model_bytes(S, N, L):
    """The minimal-pass traffic of an exact 45-correlation map with every correlation through the
    FFT (8 complex spectra per side, 23 inverse outputs; two passes per 2-D FFT, the pack fused
    into the forward row pass, the cross spectrum into the first inverse pass and the score into
    the last): 40 S + 64·8 N + 16·23 N + 4 L bytes — 3.75 GB for a 1024² pair on a 2048² grid. A
    map's rate against this fixed figure is comparable across plans."""
    return 40 * S + 64 * 8 * N + 16 * 23 * N + 4 * L


def _plan(three, rect, trans=False):
    """Spectra, inverse FFTs and correlations of the exact map plans (rectangle plans: the
    correlations with the fixed weight plane come from summed-area tables; translation plans:
    those with the moving weight plane too, from the fixed side's tables)."""
    if trans:
        # the moving image's weight sums are the rectangles' areas: both images read the tables
        # of their features and feature products only
        nf = 3 if three else 4
        C = nf * nf
        return {"n_cplx_a": (nf + 1) // 2, "n_cplx_b": (nf + 1) // 2, "n_out": (C + 1) // 2, "n_corr": C,
                "n_sat": 9 if three else 14, "n_bsat": 9 if three else 14, "n_real_a": nf, "n_real_b": nf}
    if not rect:
        R = 10 if three else 15
        C = 28 if three else 45
        return {"n_cplx_a": (R + 1) // 2, "n_cplx_b": (R + 1) // 2, "n_out": (C + 1) // 2, "n_corr": C, "n_sat": 0}
    ra, rb, C, T = (4, 9, 18, 10) if three else (5, 14, 30, 15)
    return {"n_cplx_a": (ra + 1) // 2, "n_cplx_b": (rb + 1) // 2, "n_out": (C + 1) // 2, "n_corr": C, "n_sat": T,
            "n_real_a": ra, "n_real_b": rb}


def bench_refine(a, np, torch, zcmir_amd, synth):
    """The refine loop (zc_climb: shift-map hops, Gauss–Newton blocks, roto-scale hops) on the
    1024² map pair from a start pose 6 px and 2° off the truth. Wall time per climb with its
    decisions on the device (climb_graph: one captured graph per iteration, the default) and with
    the host-driven loop, against the GPU time of the climb's launches (HIP events around every
    launch in one more, profiled climb): host_share = 1 − GPU time / wall time."""
    import math
    import time

    mv, fx, Ht = synth.shift_pair((1024, 1024), (113, -57), seed=11, quantize=True)
    t = math.radians(2.0)
    C = np.array([[1, 0, 512.5], [0, 1, 512.5], [0, 0, 1.0]])
    R = np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1]])
    H0 = C @ R @ np.linalg.inv(C) @ Ht @ np.array([[1, 0, 6.0], [0, 1, -4.0], [0, 0, 1]])
    reg = zcmir_amd.Registrar(half=False, symmetric=True, hop=True, hop_fm=True, g_iters=25)
    reg.set_images(mv, fx, preprocess=False)

    def walls_of(graph):
        reg.configure(climb_graph=graph)
        reg.pose = H0
        reg.climb()   # warm-up: tilings, buffers, the captured graph
        walls = []
        for _ in range(a.refine_runs):
            reg.pose = H0
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            c = reg.climb()
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
        return float(np.median(walls)), c

    w_host, c_host = walls_of(False)
    w, c = walls_of(True)
    reg.pose = H0
    reg.profile(True)
    reg.climb()
    rep = reg.profile_report()
    g = sum(tot for _, tot, _ in rep.values()) * 1e-3
    reg.profile(False)
    top = sorted(rep.items(), key=lambda kv: -kv[1][1])[:10]
    corners = np.array([[0.5, 0.5, 1], [1024.5, 0.5, 1], [1024.5, 1024.5, 1], [0.5, 1024.5, 1]]).T
    p, q = c.H @ corners, Ht @ corners
    err = float(np.max(np.hypot(*(p[:2] / p[2] - q[:2] / q[2]))))
    reg.close()
    return {"metric": "refine (climb) wall time", "unit": "ms", "workload": "1024x1024 map pair, start 6 px and "
            "2 deg off the truth, symmetric SMI, shift-map and roto-scale hops, Gauss-Newton blocks",
            "ms_per_climb": round(1e3 * w, 2), "ms_per_climb_host_loop": round(1e3 * w_host, 2),
            "gpu_ms_per_climb": round(1e3 * g, 2), "host_share": round(1 - g / w, 3),
            "host_share_host_loop": round(1 - g / w_host, 3), "iterations": int(c.iterations),
            "iterations_host_loop": int(c_host.iterations), "corner_err_px": round(err, 3), "runs": a.refine_runs,
            "stages_ms_profiled": {k: [v[0], round(v[1], 3)] for k, v in top}}


def bench_map2048(a, np, torch, zcmir_amd, synth):
    """The 2048² pair of the large-map workload (8-bit, shift (113, -57)) at full resolution
    (4096² FFT, 45 correlations, rectangle-sum plan): maps/s and the per-kernel times (HIP events)."""
    import time

    mv, fx, _ = synth.shift_pair((2048, 2048), (113, -57), seed=12, quantize=True)
    reg = zcmir_amd.Registrar(half=False, symmetric=False)
    reg.set_images(mv, fx, preprocess=False)
    sm = reg.shift_map(np.eye(3), exact=True, calibrated=False, resolution=4096, return_map=False)
    if (sm.dx, sm.dy) != (113.0, -57.0) or (sm.n, sm.ny) != (4096, 4096):
        raise SystemExit(f"2048 map workload check failed: peak ({sm.dx}, {sm.dy}), grid {sm.n}x{sm.ny}")
    reg.bench_shift_map(np.eye(3), exact=True, calibrated=False, resolution=4096, iters=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reg.bench_shift_map(np.eye(3), exact=True, calibrated=False, resolution=4096, iters=a.map2048_steps, warm=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    reg.profile(True)
    reg.bench_shift_map(np.eye(3), exact=True, calibrated=False, resolution=4096, iters=a.map2048_steps)
    torch.cuda.synchronize()
    prof = reg.profile_report()
    reg.profile(False)
    reg.close()
    return {"metric": "SMI maps/s @2048^2", "unit": "maps/s", "value": round(a.map2048_steps / wall, 2),
            "ms_per_map": round(1e3 * wall / a.map2048_steps, 3), "fft": [4096, 4096],
            "kernels_avg_ms": {k: round(avg, 4) for k, (cnt, tot, avg) in sorted(prof.items(), key=lambda kv: -kv[1][1])}}


def bench_rs(a, np, torch, zcmir_amd, synth):
    """Roto-scale (log-polar) maps of a 1024² pair (moving = fixed rotated 37°, scaled 1.12),
    symmetric (both directions), exact SMI, the default resolution cap: ms per map and its
    kernels (HIP events)."""
    import time

    mv, fx, _ = synth.rotscale_pair((1024, 1024), theta_deg=37.0, scale=1.12, seed=2)
    reg = zcmir_amd.Registrar(half=False, symmetric=True, exact=True)
    reg.set_images(mv, fx, preprocess=False)
    r = reg.rs_map(return_maps=False)   # warm-up: tilings, tables, buffers
    torch.cuda.synchronize()
    reg.profile(True)
    t0 = time.perf_counter()
    for _ in range(a.rs_runs):
        r = reg.rs_map(return_maps=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = reg.profile_report()
    reg.profile(False)
    reg.close()
    top = sorted(prof.items(), key=lambda kv: -kv[1][1])[:6]
    return {"metric": "roto-scale map time", "unit": "ms", "workload": "1024x1024 pair, moving = fixed rotated "
            "37 deg, scaled 1.12; symmetric (forward and inverse maps), exact SMI, default map_res",
            "ms_per_map": round(1e3 * wall / a.rs_runs, 3), "n": r.n, "n_theta": r.n_th, "n_lam": r.n_lam,
            "peak_dtheta_deg": round(float(np.degrees(r.dth)), 3), "peak_scale": round(float(np.exp(r.dsg)), 4),
            "kernels_ms_per_map": {k: round(tot / a.rs_runs, 4) for k, (cnt, tot, _) in top}, "runs": a.rs_runs}


def bench_stack(a, np, torch, zcmir_amd, synth):
    """Uneven-size stack: slices of 512-1536 px per side (aspect up to 1.5, seeded), each with a
    random homography (rotation ±10°, scale 0.9-1.1, perspective <= 1e-4) and an intensity remap
    (1 - x², x^0.5 or sin(πx)); features of every slice once, then one registration of all slices
    together, once per start-pose correction (the shift-map peak; the best peak of a small
    rotation x scale sweep), each followed by 30 batched Gauss-Newton iterations."""
    import time

    sl = synth.stack_slices(n=a.stack_slices, side=(512, 1536), seed=3)
    reg = zcmir_amd.Registrar(half=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reg.set_stack([q[0] for q in sl], [q[1] for q in sl])
    torch.cuda.synchronize()
    t_set = time.perf_counter() - t0
    out = {"metric": "stack registration time", "unit": "s", "workload": f"{len(sl)} slices 512-1536 px "
           "(aspect <= 1.5), random homographies (rot +-10 deg, scale 0.9-1.1, perspective <= 1e-4), remaps "
           "1-x^2 / x^0.5 / sin(pi x); start-pose correction, then 30 Gauss-Newton iterations",
           "slices": len(sl), "megapixels": round(sum(q[0].size for q in sl) / 1e6, 1), "set_stack_s": round(t_set, 3)}
    reg.register_stack(iters=1, init="map")   # warm-up: tilings, buffers
    for init in ("map", "sweep"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Hs, sc, _, it = reg.register_stack(iters=30, init=init)
        torch.cuda.synchronize()
        t_reg = time.perf_counter() - t0
        errs = []
        for k, (mv, _, Ht, _) in enumerate(sl):
            h, w = mv.shape
            pts = np.array([[0.5, 0.5, 1], [w + 0.5, 0.5, 1], [0.5, h + 0.5, 1], [w + 0.5, h + 0.5, 1]]).T
            p, q = Hs[k] @ pts, Ht @ pts
            errs.append(float(np.max(np.hypot(*(p[:2] / p[2] - q[:2] / q[2])))))
        errs = np.array(errs)
        out["init_" + init] = {"register_s": round(t_reg, 3), "slices_per_s": round(len(sl) / t_reg, 2),
                               "iterations": int(it), "within_1px": round(float(np.mean(errs < 1.0)), 3),
                               "median_corner_err_px": round(float(np.median(errs)), 3)}
    # a profiled registration (HIP events around every launch), apart from the timed ones; the
    # timed ones run each Gauss-Newton iteration on the device as one graph, so this one takes the
    # host-driven loop (the same pose kernels) to time each launch
    reg.configure(gn_graph=False)
    reg.profile(True)
    reg.register_stack(iters=30, init="sweep")
    prof, pbytes, pwork = reg.profile_report(), reg.profile_bytes(), reg.profile_work()
    reg.profile(False)
    reg.configure(gn_graph=True)
    reg.close()
    flops = {"pose_moments": POSE_MOMENT_FLOPS, "tangent_grad": TANGENT_FLOPS}
    kern = {}
    for name, (cnt, tot, avg) in sorted(prof.items(), key=lambda kv: -kv[1][1])[:8]:
        row = {"launches": cnt, "total_ms": round(tot, 3), "avg_ms": round(avg, 4)}
        if name in flops and pwork.get(name):
            row["pixel_poses"] = pwork[name]
            row["tflops"] = round(pwork[name] * flops[name] / (tot * 1e-3) / 1e12, 2)
            row["feature_bytes"] = pbytes.get(name)
            row["gbs_features"] = round(pbytes.get(name, 0) / (tot * 1e-3) / 1e9, 1)
        kern[name] = row
    out["kernels_profiled_registration"] = kern
    d = kern.get("pose_moments")
    if d and "tflops" in d:
        out["roofline"] = {"bound": "valu", "kernel": "pose_moments", "achieved": d["tflops"], "peak": FP32_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(d["tflops"] / FP32_PEAK_TFLOPS, 4),
                           "flops_per_pixel_pose": POSE_MOMENT_FLOPS, "traffic": _traffic("stack", "pose_moments")}
    out["note"] = ("init_map: each start corrected by its shift map's peak (translation only; rotations up to "
                   "10 deg are left to Gauss-Newton); init_sweep: by a +-12 deg, 0.88-1.12 sweep's best peak. "
                   "Timed registrations run each Gauss-Newton iteration on the device (one captured HIP graph per "
                   "iteration, gn_graph); kernel rows come from one more registration (sweep starts, host-driven "
                   "loop, the same pose kernels) with HIP events around every launch.")
    return out


def _traffic(section, kernel):
    """HBM bytes per launch of a kernel from profiles/hbm_traffic.json (PMC, calibrated), or None."""
    tfile = ROOT / "profiles" / "hbm_traffic.json"
    if not tfile.exists():
        return None
    return json.loads(tfile.read_text()).get(section, {}).get(kernel)


def calibrate_copy(mb):
    """A plain device-to-device copy of `mb` MiB (the PMC calibration kernel)."""
    import torch

    n = mb * (1 << 20) // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "reads": n * 4, "writes": n * 4}))


def _kernel_table(prof, pbytes, per):
    """Per-kernel launches, times and achieved algorithmic GB/s (bytes accounted by the engine)."""
    out = {}
    for name, (cnt, tot, avg) in prof.items():
        b = pbytes.get(name)
        out[name] = {"launches_per_unit": cnt / per, "avg_ms": avg, "ms_per_unit": tot / per,
                     "alg_bytes_per_unit": (b / per) if b else None,
                     "gbs": (b / (tot * 1e-3) / 1e9) if b else None}
    return out


def _roofline(kernels, traffic=None, section=None):
    """The dominant kernel's roofline; its PMC traffic per launch from profiles/hbm_traffic.json
    `section` when given."""
    dom = max((k for k in kernels if kernels[k].get("gbs")), key=lambda k: kernels[k]["ms_per_unit"])
    d = kernels[dom]
    if section is not None:
        traffic = _traffic(section, dom)
    return {"bound": "hbm", "kernel": dom, "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
            "alg_bytes_per_launch": d["alg_bytes_per_unit"] / d["launches_per_unit"],
            "avg_launch_ms": round(d["avg_ms"], 5)}


SWEEP_GRID = dict(sw_nth=72, sw_smin=0.8, sw_smax=1.25, sw_ns=9)   # 72 θ (5° steps) × 9 σ, 0.8 – 1.25


def bench_sweep(a, np, torch, zcmir_amd, synth):
    """Sweep poses/s: a 1024² pair whose moving image is the fixed one rotated 37°, scaled 1.12
    (bilinear) and remapped by x^0.5; 648 candidate poses, each a full SMI translation map reduced
    to its top local maxima. Two canvases: 128² cells (the reference's default grid, footprint-
    sized FFTs) and one cell per fixed pixel (full resolution)."""
    mv, fx, _ = synth.rotscale_pair((1024, 1024), theta_deg=37.0, scale=1.12, seed=2)
    out = {}
    for label, kw, steps in (("grid128", dict(sw_grid=128), a.sweep_steps),
                             ("full_resolution", dict(sw_full=True), a.sweep_full_steps)):
        if steps <= 0:
            continue
        reg = zcmir_amd.Registrar(half=False, symmetric=True, **SWEEP_GRID, **kw)
        reg.set_images(mv, fx, preprocess=False)
        info, peaks, best = reg.sweep()
        tuning = [e["text"] for e in reg.events() if e.get("kind") == "log" and "per batch" in e.get("text", "")]
        k = int(np.argmax(best))
        th_deg, sg = -180 + 5 * (k // 9), 1.25 ** ((k % 9 - 4) / 4)
        if th_deg != -35 or abs(sg - 1.25 ** -0.5) > 1e-9:
            raise SystemExit(f"sweep workload check failed ({label}): best candidate θ {th_deg}, σ {sg:.4f}")
        torch.cuda.synchronize()
        # timed: the launches replayed from the captured HIP graph (the sweep_graph setting)
        t0 = time.perf_counter()
        for _ in range(steps):
            reg.sweep()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # per-kernel times: the same launches issued eagerly, one event per launch
        reg.profile(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            reg.sweep()
        torch.cuda.synchronize()
        wall_eager = time.perf_counter() - t0
        kernels = _kernel_table(reg.profile_report(), reg.profile_bytes(), steps)
        reg.profile(False)
        out[label] = {
            "value": round(info["nc"] * steps / wall, 2), "unit": "poses/s", "steps": steps,
            "ms_per_sweep": round(1e3 * wall / steps, 3),
            "ms_per_sweep_eager_profiled": round(1e3 * wall_eager / steps, 3),
            "candidates": info["nc"], "grid": info["grid"], "batch_tuning": tuning,
            "cell_px": round(info["cell"], 4), "largest_fft": info["n"],
            "fft_column_tiling": {str(info["n"]): reg.fft_tiling(info["n"])},
            "roofline": _roofline(kernels, section=f"sweep_{label}"),
            "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                        for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_unit"])},
        }
        reg.close()
    if not a.no_cpu_baseline and "grid128" in out:
        from oracle import cpu_port
        from oracle import smi_ref as O

        pm, pf = O.prepare(mv), O.prepare(fx)
        gv = O.sweep_grid(-180, 180, 72, 0.8, 1.25, 9)
        cores = min(16, os.cpu_count() or 1)
        rate, done, el, _ = cpu_port.time_sweep(pm, pf, np.eye(3), gv, grid=128, budget_s=a.cpu_budget / 2,
                                                workers=cores)
        out["grid128"]["cpu_baseline"] = {
            "value": round(rate, 3), "unit": "poses/s", "cores": cores, "kind": "port",
            "sample": f"the first {done} of the 648 candidates on the 128² canvas in {el:.1f} s: the float32 "
                      f"restatement of the sweep (windows, scipy.fft workers={cores}, score, top-4 peaks)"}
        # the float64 oracle, timed apart (not a baseline: the parity reference)
        t0, n_th, done = time.perf_counter(), 0, 0
        while time.perf_counter() - t0 < a.cpu_budget / 4 and n_th < 72:
            th = -180 + 5 * n_th
            O.sweep(pm, pf, np.eye(3), O.sweep_grid(th, th, 1, 0.8, 1.25, 9), grid=128)
            done += 9
            n_th += 1
        el = time.perf_counter() - t0
        out["grid128"]["oracle_cpu"] = {"value": round(done / el, 3), "unit": "poses/s", "cores": 1,
                                        "kind": "oracle", "sample": f"{done} candidates, float64 NumPy, one thread"}
    return out


def _pose_close(Ha, Hb, w, h):
    """Corner distance (px) and rotation difference (degrees) of two poses."""
    import numpy as np

    pts = np.array([[1, 1, 1], [w, 1, 1], [1, h, 1], [w, h, 1]], float).T
    a, b = Ha @ pts, Hb @ pts
    dist = float(np.abs(a[:2] / a[2] - b[:2] / b[2]).max())
    rot = lambda H: np.degrees(np.arctan2(H[1, 0] - H[0, 1], H[0, 0] + H[1, 1]))
    dth = abs((rot(Ha) - rot(Hb) + 180) % 360 - 180)
    return dist, float(dth)


def bench_prior(a, np, torch, zcmir_amd):
    """Time to pose, learned prior vs the full sweep: 1024² pairs made as the rotation × scale
    workload with a random pose (rotation over the circle, scale 0.8 – 1.25, shift ≤ 16 px) and a
    random remap, held out from training. Full: the sweep search (72 θ × 9 σ, NMS, Gauss–Newton).
    Learned: the network's top-8 (rotation mod 180°, scale) classes → the sweep over the 48 grid
    cells inside their 16 cells → NMS, Gauss–Newton; the full search runs when the proposals' best SMI (128² canvas) falls below
    0.8 × the threshold held out from training. Both at full resolution (one cell per fixed
    pixel)."""
    from zcmir_amd import prior, synth

    weights = Path(a.prior_weights) if a.prior_weights else prior.WEIGHTS
    if not weights.exists():
        return {"error": f"{weights.name} missing"}
    net = prior.Net.load(weights)
    n_eval, n_cal, size = a.prior_pairs, a.prior_cal, 1024

    def pairs(seed, n):
        rng = np.random.default_rng(seed)
        tex = torch.tensor(np.stack([synth._texture(size, size, rng) for _ in range(n)]).astype(np.float32),
                           device="cuda")
        Hs, th, sg, rm = prior.random_poses(rng, n, size)
        mv, fx = prior.make_pairs(tex, np.arange(n), Hs, rm)
        return mv, fx, Hs, th, sg

    # held-out evaluation pairs, and calibration pairs of their own draw
    mv, fx, Hs, th, sg = pairs(20251, n_eval)
    cmv, cfx, cHs, cth, csg = pairs(20252, n_cal)
    torch.cuda.synchronize()
    reg = zcmir_amd.Registrar(half=False, sw_full=True, **SWEEP_GRID)
    # the fallback threshold: the 5th percentile of the calibration pairs' check scores at their true cells
    thr = prior.calibrate(reg, net, cmv, cfx, cHs, cth, csg)
    # warm both paths once (first-use allocations and twiddle tables)
    reg.set_images(mv[0].cpu().numpy(), fx[0].cpu().numpy(), preprocess=False)
    prior.learned_search(reg, net, mv[0], fx[0], threshold=thr, margin=a.prior_margin)
    rows = []
    for i in range(n_eval):
        m_np, f_np = mv[i].cpu().numpy(), fx[i].cpu().numpy()
        reg.set_images(m_np, f_np, preprocess=False)
        reg.pose = np.eye(3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        full = reg.search()
        t_full = time.perf_counter() - t0
        reg.pose = np.eye(3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lrn, info = prior.learned_search(reg, net, mv[i], fx[i], threshold=thr, margin=a.prior_margin)
        torch.cuda.synchronize()
        t_lrn = time.perf_counter() - t0
        d, dth = _pose_close(lrn.H, full.H, size, size)
        # the label-grid cell of the full search's pose (its rotation and scale)
        A = np.asarray(full.H, float)[:2, :2]
        full_cell = prior.truth_cell(float(np.degrees(np.arctan2(A[1, 0] - A[0, 1], A[0, 0] + A[1, 1]))),
                                     float(np.sqrt(abs(np.linalg.det(A)))))
        # ranks of the true and the full search's classes among all of the network's merged classes
        order = prior.propose(net, mv[i][None], fx[i][None], k=prior.N_TH // 2 * prior.N_SG)[0]
        rank = lambda c: order.index(c) // 2 if c in order else -1   # noqa: E731
        rows.append({"t_full_s": round(t_full, 4), "t_learned_s": round(t_lrn, 4), "pose_dist_px": round(d, 4),
                     "rot_diff_deg": round(dth, 4), "fell_back": bool(info["fell_back"]),
                     "check_score": round(info["check_score"], 4),
                     "truth_cell_proposed": prior.truth_cell(th[i], sg[i]) in info["cells"],
                     "full_cell_proposed": full_cell in info["cells"],
                     "truth_class_rank": rank(prior.truth_cell(th[i], sg[i])), "full_class_rank": rank(full_cell),
                     "truth_dist_px": round(_pose_close(full.H, Hs[i], size, size)[0], 3),
                     "learned_truth_dist_px": round(_pose_close(lrn.H, Hs[i], size, size)[0], 3)})
    reg.close()
    tf = float(np.mean([r["t_full_s"] for r in rows]))
    tl = float(np.mean([r["t_learned_s"] for r in rows]))
    same = [r["pose_dist_px"] <= 0.5 and r["rot_diff_deg"] <= 0.5 for r in rows]
    return {"metric": "time-to-pose learned vs full sweep", "value": round(tf / tl, 2), "unit": "x",
            "time_to_pose_full_s": round(tf, 4), "time_to_pose_learned_s": round(tl, 4),
            "same_pose_rate": float(np.mean(same)), "fallbacks": int(sum(r["fell_back"] for r in rows)),
            "threshold": thr, "calibration_pairs": n_cal, "margin_rotations": a.prior_margin, "pairs": rows,
            "network": {"layers": prior.LAYERS, "input": [prior.T, prior.R, prior.C],
                        "trained_pairs": int(net.meta.get("n_pairs", 0))}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--corr", type=int, default=45, choices=(45, 28))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="run a copy kernel of this many MiB and exit")
    ap.add_argument("--fft-cols", type=int, default=0, help="columns per workgroup of the column FFT (0: default)")
    ap.add_argument("--fft-rows", type=int, default=0, help="tiling of the inverse row FFT, 0-3 (0: default)")
    ap.add_argument("--sweep-steps", type=int, default=10, help="timed sweeps on the 128² canvas (0: skip)")
    ap.add_argument("--sweep-full-steps", type=int, default=1, help="timed full-resolution sweeps (0: skip)")
    ap.add_argument("--prior-pairs", type=int, default=24, help="time-to-pose pairs, learned vs full (0: skip)")
    ap.add_argument("--prior-cal", type=int, default=32, help="calibration pairs of the fallback threshold")
    ap.add_argument("--prior-margin", type=int, default=1, help="fine rotations swept beyond each proposed class")
    ap.add_argument("--prior-weights", help="network weights (default: the package's prior_weights.npz)")
    ap.add_argument("--refine-runs", type=int, default=3, help="timed climbs on the 1024² pair (0: skip)")
    ap.add_argument("--rs-runs", type=int, default=10, help="timed roto-scale maps on a 1024² pair (0: skip)")
    ap.add_argument("--sweep-batch", type=int, default=0, help="sweep candidates per batch (0: 256, the default)")
    ap.add_argument("--stack-slices", type=int, default=64, help="slices of the uneven stack workload (0: skip)")
    ap.add_argument("--map2048-steps", type=int, default=20, help="timed maps of the 2048² pair (0: skip)")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path shards by "
              "candidate / slice, never within one map)", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return

    import numpy as np
    import torch

    import zcmir_amd
    from zcmir_amd import synth

    three = a.corr == 28
    if a.fft_cols or a.fft_rows:
        zcmir_amd._load().zc_fft_tune(a.fft_rows, a.fft_cols)
    if a.sweep_batch:
        zcmir_amd._load().zc_sweep_tune(a.sweep_batch)
    size, shift = (1024, 1024), (113, -57)
    mv, fx, _ = synth.shift_pair(size, shift, seed=11, quantize=True)
    reg = zcmir_amd.Registrar(half=False, symmetric=False, smi_features="three" if three else "laplacian")
    reg.set_images(mv, fx, preprocess=False)
    H = np.eye(3)
    sm = reg.shift_map(H, exact=True, calibrated=False, resolution=2048, return_map=False)
    if (sm.dx, sm.dy) != (float(shift[0]), float(shift[1])) or (sm.n, sm.ny) != (2048, 2048):
        raise SystemExit(f"bench workload check failed: peak ({sm.dx}, {sm.dy}), grid {sm.n}x{sm.ny}")

    # warmup, then exactly K maps bracketed by device synchronization (the engine's streams as
    # in use: no per-launch events, and no untimed map inside the window); then the same K maps
    # again with a HIP event around every launch (one stream) for the per-kernel times of the
    # roofline
    reg.bench_shift_map(H, exact=True, calibrated=False, resolution=2048, iters=a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = reg.bench_shift_map(H, exact=True, calibrated=False, resolution=2048, iters=a.steps, warm=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    maps_s = a.steps / wall
    reg.profile(True)
    reg.bench_shift_map(H, exact=True, calibrated=False, resolution=2048, iters=a.steps)
    torch.cuda.synchronize()
    prof = reg.profile_report()
    reg.profile(False)

    rect = (sm.cw, sm.ch) == tuple(sm.canvas)
    # the workload's pose is the identity: an integer translation onto the canvas
    plan = _plan(three, rect, trans=rect)
    kb = kernel_bytes(plan, size[0] * size[1], sm.canvas[0] * sm.canvas[1], sm.n, sm.ny, sm.ch)
    kernels = {}
    for name, (cnt, tot, avg) in prof.items():
        per_map = cnt / a.steps
        b = kb.get(name)
        kernels[name] = {
            "launches_per_map": per_map, "avg_ms": avg, "ms_per_map": tot / a.steps,
            "alg_bytes": (b / per_map) if b else None, "gbs": (b / (tot / a.steps * 1e-3) / 1e9) if b else None,
        }
    dom = max((k for k in kernels if kernels[k]["alg_bytes"]), key=lambda k: kernels[k]["ms_per_map"])
    dk = kernels[dom]
    traffic = None
    tfile = ROOT / "profiles" / "hbm_traffic.json"
    if tfile.exists():
        t = json.loads(tfile.read_text())
        traffic = t.get(f"{a.corr}corr", {}).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(dk["gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(dk["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_ms": round(dk["avg_ms"], 5)}
    total_alg = sum(kb[k] for k in kb if k in kernels)
    # whole-pipeline minimum: features in, spectra written and read back once, inverse spectra,
    # the map, and the tables read once (a rectangle plan also writes canvas A's once)
    model_bytes = 40 * size[0] * size[1] + 32 * (plan["n_cplx_a"] + plan["n_cplx_b"]) * sm.n * sm.ny \
        + 16 * plan["n_out"] * sm.n * sm.ny + 4 * sm.n * sm.ny \
        + (8 * (plan["n_sat"] + plan["n_bsat"]) * size[0] * size[1] if plan.get("n_bsat")
           else 16 * plan["n_sat"] * sm.canvas[0] * sm.canvas[1])
    survey_bytes = survey_model_bytes(size[0] * size[1], sm.n * sm.ny, sm.n * sm.ny)

    cpu = None
    if not a.no_cpu_baseline:
        from oracle import cpu_port

        cores = min(16, os.cpu_count() or 1)
        fa, fb = reg.features(0), reg.features(1)
        sa = np.concatenate([np.ones((1,) + fa.shape[1:], np.float32), fa])
        sb = np.concatenate([np.ones((1,) + fb.shape[1:], np.float32), fb])
        rate, n, el, cmap = cpu_port.time_maps(sa, sb, 2048, 2048, True, three, a.cpu_budget, cores)
        iy, ix = np.unravel_index(int(np.argmax(cmap)), cmap.shape)
        cpu = {"value": round(rate, 4), "unit": "maps/s", "cores": cores, "kind": "port",
               "sample": f"{n} whole 1024^2 maps ({a.corr} correlations, 2048^2 complex64 FFTs, scipy.fft "
                         f"workers={cores}) in {el:.1f} s; peak lag ({ix if ix <= 1024 else ix - 2048}, "
                         f"{iy if iy <= 1024 else iy - 2048})"}

    line = {
        "metric": "SMI maps/s @1024^2", "value": round(maps_s, 2), "unit": "maps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * wall / a.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"1024x1024 pair, shift (113,-57), 8-bit intensities, {a.corr} "
                               "correlations, full-resolution 2048x2048 FFT, exact non-calibrated SMI",
                   "image": list(size), "fft": [sm.n, sm.ny], "correlations": a.corr,
                   "per_lag_whitening": "f64", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "event_ms_per_map": round(ev_ms, 4),
        "fft_column_tiling": {str(L): reg.fft_tiling(L) for L in sorted({sm.n, sm.ny})},
        "alg_bytes_per_map": total_alg,
        "model_bytes_per_map": model_bytes,
        "model_frac": round(model_bytes / (wall / a.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "survey_model_bytes_per_map": survey_bytes,
        "survey_model_frac": round(survey_bytes / (wall / a.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_map"])},
        "device": reg.adapter,
    }
    reg.close()
    if a.sweep_steps > 0 or a.sweep_full_steps > 0:
        line["sweep"] = {"metric": "sweep poses/s", "workload": "1024x1024 pair, moving = fixed rotated 37 deg, "
                         "scaled 1.12 (bilinear), remapped x^0.5; 72 rotations x 9 scales (0.8-1.25) = 648 "
                         "candidates, exact SMI, top-4 local maxima per candidate map",
                         **bench_sweep(a, np, torch, zcmir_amd, synth)}
    if a.prior_pairs > 0:
        line["prior"] = bench_prior(a, np, torch, zcmir_amd)
    if a.refine_runs > 0:
        line["refine"] = bench_refine(a, np, torch, zcmir_amd, synth)
    if a.rs_runs > 0:
        line["rs_map"] = bench_rs(a, np, torch, zcmir_amd, synth)
    if a.stack_slices > 0:
        line["stack"] = bench_stack(a, np, torch, zcmir_amd, synth)
    if a.map2048_steps > 0:
        line["map_2048"] = bench_map2048(a, np, torch, zcmir_amd, synth)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
