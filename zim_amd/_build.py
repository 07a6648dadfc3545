"""Builds zim_amd/libpsgc.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m zim_amd._build          # or __graft_entry__.build()
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libpsgc.so")

SOURCES = ["capi.hip", "plain.hip", "tiled.hip", "pipeline.hip", "adjoint.hip", "compact.hip",
           "roi.hip", "cloud.hip", "probe.hip", "plane.hip", "hoststage.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PSGC_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no mul+add fusion anywhere: integer outputs must match the CPU restatement
    "-ffp-contract=off",
    # packed FP32 instructions stay off (they have produced wrong results on
    # this hardware generation in some forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libpsgc.so", *objs,
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def build_torch(verbose: bool = False) -> str:
    """Builds the libtorch extension (csrc/torch_ext.cpp) into _build/torch,
    linked against libpsgc.so (which must be built first)."""
    from torch.utils import cpp_extension
    out_dir = os.path.join(BUILD, "torch")
    os.makedirs(out_dir, exist_ok=True)
    cpp_extension.load(
        name="psgc_torch",
        sources=[os.path.join(CSRC, "torch_ext.cpp")],
        extra_include_paths=[os.path.join(ROOT, "include")],
        extra_cflags=["-O2"],
        # /root/repo resolves on the GPU hosts as well (the tree is linked there)
        extra_ldflags=[f"-L{HERE}", "-l:libpsgc.so", f"-Wl,-rpath,{HERE}"],
        build_directory=out_dir,
        verbose=verbose,
        is_python_module=True,
    )
    path = os.path.join(out_dir, "psgc_torch.so")
    if verbose:
        print(f"built {path}")
    return path


if __name__ == "__main__":
    build()
    if len(sys.argv) > 1 and sys.argv[1] == "torch":
        build_torch(verbose=True)
    sys.exit(0)
