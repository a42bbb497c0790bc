"""Build libzcsmi.so (the HIP kernels and the C ABI) in-tree for gfx950 with hipcc.

    python -m zcmir_amd.build            # compile what changed, link zcmir_amd/libzcsmi.so
    python -m zcmir_amd.build --force    # recompile everything
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libzcsmi.so"
ARCH = os.environ.get("ZCSMI_ARCH", "gfx950")

SOURCES = ["k_prep.hip", "k_pack.hip", "k_ifft.hip", "k_score.hip", "k_fft.hip", "k_pose.hip", "k_sweep.hip", "k_gn.hip", "noalloc.hip", "cnn.hip", "pre.hip", "engine.cpp", "sweep.cpp", "gn.cpp", "stack.cpp", "rs.cpp", "climb.cpp", "abi.cpp"]
HEADERS = ["kdev.hpp", "gn_dev.hpp", "events.hpp", "fft.hpp", "fft_regs.hpp", "plan.hpp", "kernels.hpp", "engine.hpp", "settings.hpp", "geometry.hpp",
           "search.hpp", "lie.hpp", "pre.hpp"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (ordinary float code only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


# the preprocessing kernels round every float32 operation as the reference writes it (no fused
# multiply-add), so that they are reproducible bit for bit
PER_SOURCE = {"pre.hip": ["-ffp-contract=off"]}


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (Path(c).exists() or c == "hipcc"):
            return c
    raise FileNotFoundError("hipcc not found")


def _deps_mtime(dfile: Path, fallback: float) -> float:
    """Newest mtime of the headers a source included when it was last compiled (its -MMD file);
    `fallback` (every header) when there is no dependency file."""
    try:
        text = dfile.read_text().replace("\\\n", " ")
    except OSError:
        return fallback
    newest = 0.0
    for tok in text.split(":", 1)[-1].split():
        try:
            newest = max(newest, Path(tok).stat().st_mtime)
        except OSError:
            return fallback   # a dependency went away: rebuild
    return newest


def build(force: bool = False, verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    hipcc = _hipcc()
    hdr_mtime = max((CSRC / h).stat().st_mtime for h in HEADERS)
    procs = []
    objs = []
    for src in SOURCES:
        s = CSRC / src
        o = OBJ / (src + ".o")
        d = OBJ / (src + ".d")
        objs.append(o)
        if not force and o.exists() and o.stat().st_mtime > max(s.stat().st_mtime, _deps_mtime(d, hdr_mtime)):
            continue
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [hipcc, *FLAGS, *PER_SOURCE.get(src, []), *lang, "-MMD", "-MF", str(d), "-c", str(s), "-o", str(o)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if p.returncode != 0:
            failed.append((src, out))
        elif out and verbose:
            print(out)
    if failed:
        for src, out in failed:
            print(f"--- {src} ---\n{out}", file=sys.stderr)
        raise RuntimeError("hipcc failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.force, a.verbose))
