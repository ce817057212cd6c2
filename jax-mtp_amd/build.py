"""Build libmtp_hip.so (HIP kernels + C ABI) for gfx950, in-tree.

Usage:  python jax-mtp_amd/build.py [--force]

hipcc cross-compiles for gfx950 without a GPU.  Every HIP translation unit
is compiled with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmtp_hip.so"
MD = PKG / "md" / "md_nve"
# the pair style's own source compiled against the LAMMPS stand-ins of tests/lammps_stub (its test driver)
PAIR_CHECK = ROOT / "tests" / "lammps_stub" / "run_pair"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MTP_OFFLOAD_ARCH", "gfx950")
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
SOURCES = ["mtp_kernels.hip", "mtp_train.hip", "mtp_head.hip", "mtp_pack.hip", "mtp_capi.cpp", "mtp_plan.cpp",
           "mtp_format.cpp"]
HEADERS = ["mtp_pack.hpp", "mtp_kernels.hpp", "mtp_train.hpp", "mtp_head.hpp", "mtp_device.hpp", "mtp_plan.hpp", "mtp_format.hpp", "nbr_pack.hpp"]


def _compile(src: str) -> pathlib.Path:
    out = OBJ / (src + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *lang, "-c", str(CSRC / src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def build(force: bool = True) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    stub = ROOT / "tests" / "lammps_stub"
    pair_src = [PKG / "lammps" / "pair_jax_mtp.cpp", PKG / "lammps" / "pair_jax_mtp.h", stub / "run_pair.cpp",
                stub / "lammps_stub.h"]
    deps = ([CSRC / s for s in SOURCES + HEADERS] + [ROOT / "include" / "mtp_hip.h", PKG / "md" / "md_nve.cpp"]
            + [p for p in pair_src if p.exists()])
    outs = [LIB, MD] + ([PAIR_CHECK] if stub.exists() else [])
    if (not force and all(o.exists() for o in outs)
            and min(o.stat().st_mtime for o in outs) >= max(d.stat().st_mtime for d in deps)):
        return LIB
    try:
        with cf.ThreadPoolExecutor(max_workers=4) as ex:
            objs = list(ex.map(_compile, SOURCES))
        tmp = LIB.with_suffix(".so.tmp")
        subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)],
                       check=True, capture_output=True, text=True)
        os.replace(tmp, LIB)
        # the standalone NVE driver over the pair style's host path (plain C++, links the library)
        subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{CSRC}",
                        str(PKG / "md" / "md_nve.cpp"), f"-L{PKG}", "-lmtp_hip", "-Wl,-rpath,$ORIGIN/..",
                        "-o", str(MD)], check=True, capture_output=True, text=True)
        if stub.exists():
            subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", f"-I{stub}", f"-I{PKG / 'lammps'}",
                            f"-I{ROOT / 'include'}", f"-I{CSRC}", str(stub / "run_pair.cpp"),
                            str(PKG / "lammps" / "pair_jax_mtp.cpp"), f"-L{PKG}", "-lmtp_hip",
                            "-Wl,-rpath,$ORIGIN/../../jax-mtp_amd", "-o", str(PAIR_CHECK)],
                           check=True, capture_output=True, text=True)
    except subprocess.CalledProcessError as e:
        sys.stderr.write(e.stdout or "")
        sys.stderr.write(e.stderr or "")
        raise
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv or True))
