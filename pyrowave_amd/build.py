"""Builds pyrowave_amd/libpyrowave-shared.so (HIP kernels for gfx950 + C ABI).

Usage: python -m pyrowave_amd.build   (or pyrowave_amd.build.build()).
Every source is compiled by hipcc for --offload-arch=gfx950 into objects under
pyrowave_amd/_build, then linked into one shared library that exports only the
pyrowave_* symbols (version script csrc/pyrowave.map).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpyrowave-shared.so")
SOURCES = ["pw_dwt.hip", "pw_codec.hip", "pw_packets.hip", "pw_psnr.hip", "pw_host.cpp", "pw_capi.cpp", "pw_server.cpp", "pw_external.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYROWAVE_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden",
    "-DPYROWAVE_EXPORT_SYMBOLS", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    # No packed FP32 instructions: hipcc must not form v_pk_*_f32.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, warn in results:
        if warn and verbose:
            print(warn, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        "-Wl,--version-script=" + os.path.join(CSRC, "pyrowave.map"),
        "-Wl,-soname,libpyrowave-shared.so.0",
    ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
