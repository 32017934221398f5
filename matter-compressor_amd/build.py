"""Build the MI355X codec library in-tree.

Produces, next to this file:
  lib/libmc_amd.so   HIP kernels (gfx950) + the mchip C ABI + the plain-C
                     matter-compressor drop-in API (include/*.h)
  lib/libmc_synth.so synthetic workload generator (benchmarks and tests)
and, through oracle/Makefile, the CPU oracle used by the tests.

Usage: python matter-compressor_amd/build.py [--no-oracle]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
SYNTH = os.path.join(HERE, "synth")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: every f32 multiply and add stays a separate IEEE
# operation (bit-exact with the reference's deterministic build); f32
# division and square root are correctly rounded, as on the host.
# -packed-fp32-ops is disabled: no v_pk_*_f32 instructions are generated.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function",
]
C_FLAGS = ["-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           "-Wno-unused-function", "-I", INCLUDE, "-D_GNU_SOURCE"]

HIP_SOURCES = ["mchip_kernels.hip", "mchip_api.hip", "mchip_serve.hip", "mchip_render.hip"]
HOST_SOURCES = ["mc_host.c", "mc_host_region.c", "mc_host_serve.c", "mc_host_stream.c"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "argument unused during compilation" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(oracle=True, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", o])
    for s in HOST_SOURCES:
        src = os.path.join(HOST, s)
        if not os.path.exists(src):
            continue
        o = os.path.join(OBJ, s.replace(".c", ".o"))
        objs.append(o)
        jobs.append(["gcc", *C_FLAGS, "-c", src, "-o", o])
    with ThreadPoolExecutor(max_workers=4) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libmc_amd.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", so + ".tmp", *objs, "-lpthread", "-lm"])
    os.replace(so + ".tmp", so)
    synth = os.path.join(LIB, "libmc_synth.so")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-o", synth + ".tmp", os.path.join(SYNTH, "mc_synth.c"), "-lm", "-lpthread"])
    os.replace(synth + ".tmp", synth)
    if oracle:
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return so


if __name__ == "__main__":
    print(build(oracle="--no-oracle" not in sys.argv, verbose=True))
