"""Builds the HIP library libmatfun_hip.so (gfx950) in-tree from matfun_amd/csrc with hipcc.

    python -m matfun_amd.build            (or build() from Python)

Flags: -O3, C++17, position independent, packed-FP32 instruction selection disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops), FP contraction as hipcc's default (FMA).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmatfun_hip.so")
SOURCES = ["matfun_hip.hip", "smf_kernels.hip", "gmf_kernels.hip", "gmf_sorted.hip", "batch.hip", "batch_fused.hip",
           "batch_smf.hip", "batch_gmf.hip", "hencky.hip", "store_probe.hip"]
# the heterogeneous-batch kernels: MachineLICM off, so that no per-tile constant is hoisted out of the persistent loop
# over tiles and held in registers across it (batch_kernels.hpp)
PER_FILE = {"batch_fused.hip": ["-mllvm", "-disable-machine-licm"], "batch_smf.hip": ["-mllvm", "-disable-machine-licm"],
            "batch_gmf.hip": ["-mllvm", "-disable-machine-licm"]}
ARCH = os.environ.get("MATFUN_HIP_ARCH", "gfx950")


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags(extra=()):
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-I" + CSRC, *extra]


def build(verbose=False, extra=()):
    """compile the sources in parallel (the batch kernels take minutes each), link libmatfun_hip.so"""
    from concurrent.futures import ThreadPoolExecutor
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmds.append([hipcc(), *flags(list(extra) + PER_FILE.get(src, [])), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
        if verbose:
            print(" ".join(cmds[-1]), flush=True)
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        for r in list(ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds)):
            if r.returncode != 0:
                sys.stderr.write(r.stderr)
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, extra=sys.argv[1:]))
