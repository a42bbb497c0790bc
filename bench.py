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


def kernel_bytes(plan, S, canvas_px, nx, ny, ch):
    """Algorithmic HBM bytes per launch of each map kernel (minimal traffic of its own role)."""
    N = nx * ny
    P, Q, C = plan["n_cplx"], plan["n_out"], plan["n_corr"]
    return {
        "warp_stack": 20 * S + 20 * canvas_px,
        "paste_stack": 20 * S + 20 * canvas_px,
        "pack_rowfft": 20 * canvas_px + P * ch * nx * 8,
        "colfft_fwd": P * ch * nx * 8 + P * N * 8,
        "cmul": 2 * P * N * 8 + Q * N * 8,
        "colfft_inv": 2 * Q * N * 8,
        "rowifft_extract": Q * N * 8 + C * N * 4,
        "combine": C * N * 4 + N * 4,
    }


def _plan(three):
    R = 10 if three else 15
    C = 28 if three else 45
    return {"n_cplx": (R + 1) // 2, "n_out": (C + 1) // 2, "n_corr": C}


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
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path shards by "
              "candidate / slice, never within one map; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return

    import numpy as np
    import torch

    import zcmir_amd
    from zcmir_amd import synth

    three = a.corr == 28
    if a.fft_cols:
        zcmir_amd._load().zc_fft_tune(0, a.fft_cols)
    size, shift = (1024, 1024), (113, -57)
    mv, fx, _ = synth.shift_pair(size, shift, seed=11, quantize=True)
    reg = zcmir_amd.Registrar(half=False, symmetric=False, smi_features="three" if three else "laplacian")
    reg.set_images(mv, fx, preprocess=False)
    H = np.eye(3)
    sm = reg.shift_map(H, exact=True, calibrated=False, resolution=2048, return_map=False)
    if (sm.dx, sm.dy) != (float(shift[0]), float(shift[1])) or (sm.n, sm.ny) != (2048, 2048):
        raise SystemExit(f"bench workload check failed: peak ({sm.dx}, {sm.dy}), grid {sm.n}x{sm.ny}")

    # warmup, then exactly K maps bracketed by device synchronization, per-kernel HIP events on
    # the engine's stream over the timed region
    reg.bench_shift_map(H, exact=True, calibrated=False, resolution=2048, iters=a.warmup)
    torch.cuda.synchronize()
    reg.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = reg.bench_shift_map(H, exact=True, calibrated=False, resolution=2048, iters=a.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = reg.profile_report()
    reg.profile(False)
    wall = t1 - t0
    maps_s = a.steps / wall

    plan = _plan(three)
    kb = kernel_bytes(plan, size[0] * size[1], sm.canvas[0] * sm.canvas[1], sm.n, sm.ny, sm.ch)
    kernels = {}
    for name, (cnt, tot, avg) in prof.items():
        per_map = cnt / a.steps
        b = kb.get(name)
        kernels[name] = {
            "launches_per_map": per_map, "avg_ms": avg, "ms_per_map": tot / a.steps,
            "alg_bytes": b, "gbs": (b / (avg * 1e-3) / 1e9) if b else None,
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
    total_alg = sum(kb[k] * kernels[k]["launches_per_map"] for k in kb if k in kernels)
    model_bytes = 40 * size[0] * size[1] + 64 * plan["n_cplx"] * sm.n * sm.ny + 16 * plan["n_out"] * sm.n * sm.ny \
        + 4 * sm.n * sm.ny

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
        "config": {"workload": f"config 1: 1024x1024 pair, shift (113,-57), 8-bit intensities, {a.corr} "
                               "correlations, full-resolution 2048x2048 FFT, exact non-calibrated SMI",
                   "image": list(size), "fft": [sm.n, sm.ny], "correlations": a.corr,
                   "per_lag_whitening": "f64", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "event_ms_per_map": round(ev_ms, 4),
        "alg_bytes_per_map": total_alg,
        "model_bytes_per_map": model_bytes,
        "model_frac": round(model_bytes / (wall / a.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_map"])},
        "device": reg.adapter,
    }
    reg.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
