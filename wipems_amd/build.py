"""Build the native libraries in-tree.

* ``wipems_amd/libwipems_fdtd.so``   -- HIP engine for gfx950 (hipcc)
* ``wipems_amd/libwipems_operator.so`` -- host operator builder (g++)
* ``oracle/liboracle.so`` and ``oracle/libcpubaseline.so`` -- CPU checker and
  CPU baseline (gcc), built only when ``with_oracle`` is set.

Every HIP source is compiled with contraction off (the reference compiles for
baseline x86-64, so no FMA), FP32 denormals flushed (the SSE engines run with
FTZ|DAZ) and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wipems_amd")
INC = os.path.join(ROOT, "include")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fgpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC,
]

HIP_SOURCES = ["csrc/fdtd_engine.hip", "csrc/fdtd_batch.hip", "csrc/nf2ff.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_engine(verbose=False):
    """Each HIP source is compiled to its own object in parallel (no
    relocatable device code: every translation unit holds its own kernels),
    then linked into one shared library."""
    from concurrent.futures import ThreadPoolExecutor
    import tempfile
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    srcs = [os.path.join(PKG, s) for s in HIP_SOURCES if os.path.exists(os.path.join(PKG, s))]
    out = os.path.join(PKG, "libwipems_fdtd.so")
    tmp = out + ".tmp"
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory() as td:
        objs = [os.path.join(td, os.path.basename(src) + ".o") for src in srcs]
        with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
            list(ex.map(lambda so: _run([hipcc] + cflags + ["-c", "-o", so[1], so[0]]), zip(srcs, objs)))
        _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_operator(verbose=False):
    out = os.path.join(PKG, "libwipems_operator.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-pthread", "-I" + INC,
          "-o", tmp, os.path.join(PKG, "host", "operator.cpp")])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


HDF5_ROOT = "/opt/conda"


def build_dump(verbose=False):
    """HDF5 field-dump writer (host only), linked against the image's HDF5."""
    out = os.path.join(PKG, "libwipems_dump.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + INC, "-I" + os.path.join(HDF5_ROOT, "include"),
          "-o", tmp, os.path.join(PKG, "host", "h5dump.cpp"), "-L" + os.path.join(HDF5_ROOT, "lib"), "-lhdf5",
          "-Wl,-rpath," + os.path.join(HDF5_ROOT, "lib")])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_oracle(verbose=False):
    d = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-j4", "all"], cwd=d)
    if verbose:
        print("built oracle in", d)


def build_all(with_oracle=True, verbose=False):
    build_operator(verbose)
    build_dump(verbose)
    build_engine(verbose)
    if with_oracle:
        build_oracle(verbose)


if __name__ == "__main__":
    build_all(with_oracle="--no-oracle" not in sys.argv, verbose=True)
