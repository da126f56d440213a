"""Builds the native parts of fastrecon_amd in-tree.

  libfastrecon_hip.so   HIP kernels + C ABI (include/fastrecon_hip.h), gfx950
  libfastrecon.so       C++ API (ReconData / GridLut / ImageFilter / ImageData)
  fastrecon             command-line program (reference Program/Main.cpp)

HIP code is compiled for gfx950 only, with FP contraction off (the gridding
and SOS kernels reproduce the reference's FP32 operations one by one) and
packed FP32 instructions disabled.
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CPP = os.path.join(PKG, "cpp")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_LIB = os.path.join(PKG, "libfastrecon_hip.so")
API_LIB = os.path.join(PKG, "libfastrecon.so")
CLI = os.path.join(PKG, "fastrecon")

HIPFLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-value",
    "-I", INCLUDE, "-I", CSRC,
]
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off",
            "-I", INCLUDE, "-I", os.path.join(CPP, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = r.stdout
    # hipcc's host pass does not know the packed-fp32 target feature: harmless
    out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in srcs]

    def comp(pair):
        s, o = pair
        return _run([HIPCC, *HIPFLAGS, "-c", s, "-o", o])

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(comp, zip(srcs, objs)):
            if verbose and out:
                print(out)
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L", os.path.join(ROCM, "lib"), "-lrocfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_cpp(verbose=False):
    """C++ API library (links libfastrecon_hip.so) and the CLI program."""
    srcs = sorted(glob.glob(os.path.join(CPP, "src", "*.cpp")))
    if not srcs:
        return None
    lib_srcs = [s for s in srcs if not s.endswith("fastrecon_main.cpp")]
    tmp = API_LIB + ".tmp"
    _run(["g++", *CXXFLAGS, "-shared", "-o", tmp, *lib_srcs,
          "-L", PKG, "-lfastrecon_hip", "-lpthread", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, API_LIB)
    main = os.path.join(CPP, "src", "fastrecon_main.cpp")
    if os.path.exists(main):
        tmp = CLI + ".tmp"
        _run(["g++", *CXXFLAGS, "-o", tmp, main, "-L", PKG, "-lfastrecon", "-lfastrecon_hip",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, CLI)
    return API_LIB


def build_test_programs(verbose=False):
    """C++ programs the GPU tests run against the C++ API (tests/cpp/*.cpp ->
    tests/cpp/<name>, linked to libfastrecon.so)."""
    tdir = os.path.join(ROOT, "tests", "cpp")
    out = []
    for src in sorted(glob.glob(os.path.join(tdir, "*.cpp"))):
        exe = os.path.splitext(src)[0]
        tmp = exe + ".tmp"
        _run(["g++", *CXXFLAGS, "-o", tmp, src, "-L", PKG, "-lfastrecon", "-lfastrecon_hip", "-lpthread",
              "-Wl,-rpath,$ORIGIN/../../fastrecon_amd"])
        os.replace(tmp, exe)
        out.append(exe)
    return out


def build_all(verbose=False):
    build_hip(verbose=verbose)
    build_cpp(verbose=verbose)
    build_test_programs(verbose=verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", HIP_LIB)
