"""Build the native libraries in-tree (no JIT cache, no setuptools).

* ``libtorchlpc_amd_hip.so``: the HIP kernels and the C ABI of
  ``include/torchlpc_amd.h``, compiled by ``hipcc --offload-arch=gfx950``.
* ``_C.so``: the PyTorch operator layer (``torchlpc_amd/csrc/torch_ops.cpp``),
  plain C++ compiled by ``g++`` against the installed torch headers and linked
  to the HIP library with an ``$ORIGIN`` rpath.

Run ``python -m torchlpc_amd._build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
HIP_LIB = PKG / "libtorchlpc_amd_hip.so"
OPS_LIB = PKG / "_C.so"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("TLPC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    # packed FP32 (v_pk_*_f32) is kept out of generated code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _hip_sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build_hip(jobs: int = 8) -> Path:
    hipcc = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
    BUILD.mkdir(exist_ok=True)
    objs = []

    def compile_one(src: Path) -> Path:
        obj = BUILD / (src.stem + ".o")
        _run([hipcc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, _hip_sources()))
    tmp = HIP_LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_ops() -> Path:
    import torch
    from torch.utils import cpp_extension as ce

    tlib = Path(torch.__file__).resolve().parent / "lib"
    incs = ce.include_paths(device_type="cuda")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [
        os.environ.get("CXX", "g++"),
        "-O2",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-D__HIP_PLATFORM_AMD__=1",
        "-DUSE_ROCM=1",
        f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        "-DTORCH_EXTENSION_NAME=_C",
        *[f"-I{p}" for p in incs],
        f"-I{sysconfig.get_paths()['include']}",
        str(CSRC / "torch_ops.cpp"),
        "-o",
        str(OPS_LIB.with_suffix(".so.tmp")),
        f"-L{tlib}",
        f"-L{PKG}",
        "-ltorchlpc_amd_hip",
        "-lc10_hip",
        "-ltorch_hip",
        "-ltorch_cpu",
        "-lc10",
        "-Wl,-rpath,$ORIGIN",
        f"-Wl,-rpath,{tlib}",
    ]
    _run(cmd)
    os.replace(OPS_LIB.with_suffix(".so.tmp"), OPS_LIB)
    return OPS_LIB


def build(jobs: int = 8) -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f_hip = ex.submit(build_hip, jobs)
        f_hip.result()
        # the op layer links against the HIP library, so it goes second
        ex.submit(build_ops).result()


if __name__ == "__main__":
    build()
    print(f"built {HIP_LIB.name} and {OPS_LIB.name}")
