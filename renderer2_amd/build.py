"""Build the native library (HIP for gfx950 + host C++) in-tree.

    python -m renderer2_amd.build            # librenderer2.so
    python -m renderer2_amd.build --oracle   # also the CPU oracle (test infrastructure)

Output: renderer2_amd/lib/librenderer2.so (git-ignored, travels with the tree).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = Path(os.environ.get("R2_LIB_DIR", PKG / "lib"))

SOURCES = [
    "r2_api.cpp",
    "r2_synth.cpp",
    "r2_grid.hip",
    "r2_render.hip",
    "r2_render_packed.hip",
    "r2_filter.hip",
    "r2_server.cpp",
    "r2_post.hip",
    "r2_frontend.cpp",
    "r2_nvdb.cpp",
    "r2_image.cpp",
]
# the command-line program (links the library)
MAIN = "r2_main.cpp"

BIN = Path(__file__).resolve().parent / "bin"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("R2_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", str(ROOT / "include"),
    *os.environ.get("R2_EXTRA_FLAGS", "").split(),
]


# The render kernels are compiled for throughput, and their results are held to
# statistical bars against the oracle:
#   * fused multiply-adds (nvcc's default for the reference's kernels too);
#   * FP32 division and square root by the hardware reciprocal / square-root
#     instructions (within 2 ulp) instead of the correctly rounded expansions;
#   * FP32 denormals flushed to zero (no render value relies on them: the images
#     are identical with and without the flush).
# Together these give about 27 % more paths per second.  Every other source keeps
# each operation separately and correctly rounded, because the filter, grid
# build and post-process are bit-exact with the oracle.
# Without sinking of common instructions out of branch arms: LLVM's SimplifyCFG
# otherwise merges the stores of a lane's state variables from the two arms of a
# branch into one store through a selected address, which keeps that state in
# scratch memory, reloaded at every scheduling step of the path program (a
# code-layout choice only: no arithmetic changes).
# With the AMDGPU register-pressure trackers in the instruction scheduler: at the
# 128-register budget of four wavefronts per SIMD the integrator then spills 12 B
# per lane instead of 52 B, and runs 3 % faster (scheduling only: the same
# instructions, so the same results).
RENDER_SOURCES = {"r2_render.hip", "r2_render_packed.hip"}
RENDER_FLAGS = ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero",
                "-mllvm", "-simplifycfg-sink-common=false", "-mllvm", "-amdgpu-use-amdgpu-trackers"]


def _compile(src: Path, obj: Path) -> None:
    extra = RENDER_FLAGS if src.name in RENDER_SOURCES else []
    cmd = [HIPCC, *COMMON_FLAGS, *extra, "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr[-6000:]}")


def build_library(jobs: int | None = None, verbose: bool = True) -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    missing = [s.name for s in srcs if not s.exists()]
    if missing:
        raise FileNotFoundError(f"missing sources: {missing}")
    objs = [LIB / (s.stem + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, o): s for s, o in zip(srcs, objs)}
        for f in cf.as_completed(futs):
            f.result()
            if verbose:
                print(f"[build] compiled {futs[f].name}", flush=True)
    out = LIB / "librenderer2.so"
    tmp = LIB / "librenderer2.so.tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, out)
    for o in objs:
        o.unlink(missing_ok=True)
    if verbose:
        print(f"[build] linked {out}", flush=True)
    build_cli(verbose)
    return out


def build_cli(verbose: bool = True) -> Path:
    """renderer2 <config.json> [out_dir]: the front end's command-line program."""
    BIN.mkdir(parents=True, exist_ok=True)
    exe = BIN / "renderer2"
    cmd = [HIPCC, "-O2", "-std=c++17", str(CSRC / MAIN), "-o", str(exe), f"-L{LIB}", "-lrenderer2",
           "-Wl,-rpath,$ORIGIN/../lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"building the command-line program failed:\n{res.stderr[-4000:]}")
    if verbose:
        print(f"[build] linked {exe}", flush=True)
    build_valu_peak(verbose)
    return exe


def build_valu_peak(verbose: bool = True) -> Path:
    """valu_peak [iters]: the device's FP32 VALU issue rate (the ceiling the
    kernels' VALU issue rates are read against)."""
    BIN.mkdir(parents=True, exist_ok=True)
    exe = BIN / "valu_peak"
    cmd = [HIPCC, "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-Wno-unused-command-line-argument", str(CSRC / "r2_valu_peak.hip"), "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"building valu_peak failed:\n{res.stderr[-4000:]}")
    if verbose:
        print(f"[build] linked {exe}", flush=True)
    return exe


def build_oracle(verbose: bool = True) -> Path:
    """Compile the CPU oracle (test infrastructure) with its own Makefile."""
    odir = ROOT / "oracle"
    res = subprocess.run(["make", "-C", str(odir), "-j8"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout[-3000:]}\n{res.stderr[-3000:]}")
    if verbose:
        print(f"[build] oracle {odir / 'lib' / 'libr2oracle.so'}", flush=True)
    return odir / "lib" / "libr2oracle.so"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        print("hipcc not found", file=sys.stderr)
        return 2
    build_library(a.j)
    if a.oracle:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
