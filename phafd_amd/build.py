"""Build libphafd_hip.so (gfx950 HIP kernels + C ABI) in-tree.

Each csrc/*.hip translation unit is compiled with hipcc for gfx950, then all
objects are linked into phafd_amd/libphafd_hip.so.  Flags:
  -ffp-contract=off   no implicit FMA, so list/bin arithmetic follows the
                      reference's rounding exactly (explicit fma() is used where
                      wanted, e.g. FFT butterflies)
  -packed-fp32-ops    disabled target feature (packed FP32 VALU ops are not
                      emitted)
The host script driver (host/*.cpp) is linked into the same library and into
the phafd_hip executable.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libphafd_hip.so")
EXE = os.path.join(HERE, "phafd_hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + " ".join(cmd[-3:]))
    return r


def sources():
    hip = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    host = []
    if os.path.isdir(HOST):
        host = sorted(os.path.join(HOST, f) for f in os.listdir(HOST)
                      if f.endswith(".cpp") and f != "main.cpp")
    return hip, host


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hip, host = sources()
    jobs_list = []
    for src in hip:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs_list.append([HIPCC] + COMMON + DEVICE + ["-c", "-o", obj, src])
    for src in host:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs_list.append(["g++"] + COMMON + ["-c", "-o", obj, src])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs_list))
    objs = [c[c.index("-o") + 1] for c in jobs_list]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        tmpx = EXE + ".tmp"
        _run(["g++"] + COMMON + ["-o", tmpx, main, "-L" + HERE, "-lphafd_hip",
                                 "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"], verbose)
        os.replace(tmpx, EXE)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
