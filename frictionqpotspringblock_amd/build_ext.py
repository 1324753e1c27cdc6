"""Build the native parts of frictionqpotspringblock_amd in-tree.

* ``libfqsb.so``: the HIP kernels for gfx950 and the C ABI (include/fqsb.h),
  compiled with hipcc, one translation unit per process; the resident-kernel
  instantiation units are generated first (csrc/gen_instances.py).
* ``_fqsb.<ext>.so``: the pybind11 module over the C++ host classes
  (include/fqsb.hpp), compiled with g++ and linked against libfqsb.so.

Run ``python -m frictionqpotspringblock_amd.build_ext`` (or call ``build()``).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FQSB_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # evaluate every expression as written: no fused multiply-add, so device
    # results are bit-identical to a CPU evaluation of the same expressions
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
]

HIP_SOURCES = [
    "fq_capi.hip",
    "fq_service.hip",
    "fq_stream.hip",
    "fq_longrange.hip",
    "fq_adjoint.hip",
    "fq_tile.hip",
]


def _generate_units():
    """The resident-kernel instantiation units (fq_res_*.hip, fq_res_units.hpp)
    are generated from the table in csrc/gen_instances.py."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("_fq_gen_instances", os.path.join(CSRC, "gen_instances.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.generate(CSRC)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _hip_object(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    sources = HIP_SOURCES + _generate_units()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_hip_object, sources))
    out = os.path.join(HERE, "libfqsb.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_module() -> str:
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_fqsb" + ext)
    tmp = out + ".tmp"
    _run(
        [
            "g++",
            "-O2",
            "-std=c++17",
            "-shared",
            "-fPIC",
            "-fvisibility=hidden",
            f"-I{pybind11.get_include()}",
            f"-I{sysconfig.get_paths()['include']}",
            f"-I{INCLUDE}",
            os.path.join(CSRC, "pybind_module.cpp"),
            "-o",
            tmp,
            f"-L{HERE}",
            "-lfqsb",
            "-Wl,-rpath,$ORIGIN",
        ]
    )
    os.replace(tmp, out)
    return out


def build(jobs: int = 8) -> None:
    build_lib(jobs)
    build_module()


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print("built", os.path.join(HERE, "libfqsb.so"))
