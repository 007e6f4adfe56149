"""Builds libvlasov_hip.so (gfx950) in-tree with hipcc.

    python vlasov.jl_amd/build.py            # compile every source, link
    python vlasov.jl_amd/build.py --check    # only report whether the .so is current

Device code is compiled with -ffp-contract=off (no multiply-add contraction
beyond the explicit fma calls, as the reference's Julia kernels) and with the
packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libvlasov_hip.so"
ARCH = os.environ.get("VLASOV_HIP_ARCH", "gfx950")

SOURCES = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
HEADERS = sorted(list(CSRC.glob("*.h")) + [HERE.parent / "include" / "vlasov_hip.h"])

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
# -amdgpu-mfma-vgpr-form: matrix-core results go to ordinary VGPRs, which the
# epilogues read directly, instead of accumulation registers copied out one
# v_accvgpr_read at a time.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-mfma-vgpr-form", "-fno-gpu-rdc"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the HIP extension cannot be built")
    return exe


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    cmd = [hipcc(), "-c", str(src), "-o", str(out)] + COMMON
    if src.suffix == ".hip":
        cmd += ["-x", "hip"] + DEVICE
    else:
        cmd += ["-x", "c++"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def is_current() -> bool:
    if not LIB.exists():
        return False
    t = LIB.stat().st_mtime
    return all(p.stat().st_mtime <= t for p in SOURCES + HEADERS)


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")), 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fno-gpu-rdc", "-o", str(tmp)] + \
        [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB, {ARCH})")
    return LIB


if __name__ == "__main__":
    if "--check" in sys.argv:
        print("current" if is_current() else "stale")
        sys.exit(0 if is_current() else 1)
    build()
