"""Build the song_los shared library (HIP for gfx950 + C host) and the CPU oracle.

Usage: python -m song_amd.build [--no-oracle]

Products (in-tree and git-ignored, so they travel with a copy of the tree):
  song_amd/libsong_los.so   the LOS stage: transfer2_* C-ABI + HIP kernels
  oracle/liblos_oracle.so   CPU oracle (test infrastructure only)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

C_SOURCES = ["los_special.c", "los_lm.c", "los_bessel2.c", "los_scenario.c", "los_transfer2.c"]
# (source, extra flags); the exact kernels must not contract a*b+c into FMA
HIP_SOURCES = [
    ("los_engine.hip", []),
    ("los_kernels_exact.hip", ["-ffp-contract=off"]),
    ("los_kernels_fast.hip", []),
    ("los_serve.hip", []),
    ("los_bessel2_gpu.hip", ["-ffp-contract=off"]),
    ("los_k3int.hip", ["-ffp-contract=off"]),
    ("los_cl2.hip", ["-ffp-contract=off"]),
]

C_FLAGS = ["-O2", "-std=c99", "-D_GNU_SOURCE", "-fPIC", "-fopenmp", "-Wall",
           "-Wno-format-truncation", "-Wno-unused-function"]
HIP_FLAGS = ["-O3", "-fPIC", f"--offload-arch={ARCH}", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def _jobs() -> int:
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(n or (os.cpu_count() or 1), 16))


def build_library(out=os.path.join(HERE, "libsong_los.so")):
    from concurrent.futures import ThreadPoolExecutor

    with tempfile.TemporaryDirectory(prefix="songlos_") as tmp:
        cmds, objs = [], []
        for src in C_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            cmds.append(["gcc", *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        for src, extra in HIP_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            cmds.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        # the translation units are independent: compile them side by side
        with ThreadPoolExecutor(max_workers=_jobs()) as pool:
            for _ in pool.map(_run, cmds):
                pass
        tmp_out = out + ".tmp"
        _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp_out, *objs, "-lgomp", "-lm"])
        os.replace(tmp_out, out)
    return out


def build_oracle(out=os.path.join(ROOT, "oracle", "liblos_oracle.so")):
    """The oracle twice, with the reference's flags (gcc -O3 -std=c99, Makefile:15-25, plus
    the -fopenmp it leaves commented out, Makefile:36-37): with -DDEBUG (the reference's
    default, Makefile:19: its runtime checks on; the tests load this one) and without
    (liblos_oracle_nodebug.so, timed beside it as the CPU baseline)."""
    src = os.path.join(ROOT, "oracle", "los_ref.c")
    nodebug = out.replace(".so", "_nodebug.so")
    for dst, flags in ((out, ["-DDEBUG"]), (nodebug, [])):
        tmp_out = dst + ".tmp"
        _run(["gcc", "-O3", "-std=c99", "-D_GNU_SOURCE", "-fopenmp", "-fPIC", "-shared",
              *flags, "-Wno-format-truncation", "-o", tmp_out, src, "-lm"])
        os.replace(tmp_out, dst)
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    lib = build_library()
    print("built", lib)
    if "--no-oracle" not in argv:
        print("built", build_oracle())


if __name__ == "__main__":
    main()
