"""Build the gfx950 library in-tree.

    python tensornetworks_amd/build.py

produces tensornetworks_amd/lib/libtn_hip.so (HIP kernels + C ABI) and
tensornetworks_amd/lib/libtensornetworks.so (C++ API over the C ABI).
Object files are cached under tensornetworks_amd/_build/ and rebuilt when
their sources or headers change.
"""
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed fp32 instructions stay off (some forms have given wrong results on
# this hardware); the host pass warns that it does not know the feature.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", INCLUDE,
]

HIP_SOURCES = ["hip/tn_kernels.hip", "hip/tn_engine.hip", "hip/tn_chain.hip", "hip/tn_batch.hip", "hip/tn_precond.hip"]
HIP_HEADERS = ["hip/tn_kernels.h", "hip/tn_engine.h", "hip/tn_chain.h"]
CXX_SOURCES = ["host/tensornetworks.cpp"]
CXX_HEADERS = []
# host-only C++ linked into the ABI library (no device code)
ABI_HOST_SOURCES = ["host/tn_mpo.cpp"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE]


def _digest(paths, flags):
    h = hashlib.sha256()
    for p in paths:
        with open(p, "rb") as f:
            h.update(f.read())
    h.update(" ".join(flags).encode())
    return h.hexdigest()[:16]


def _compile_hip(src):
    path = os.path.join(CSRC, src)
    deps = [path] + [os.path.join(CSRC, x) for x in HIP_HEADERS] + [os.path.join(INCLUDE, "tn_abi.h")]
    obj = os.path.join(BUILD, os.path.basename(src) + "." + _digest(deps, HIP_FLAGS) + ".o")
    if not os.path.exists(obj):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        # drop stale objects of this source
        base = os.path.basename(src) + "."
        for f in os.listdir(BUILD):
            if f.startswith(base) and os.path.join(BUILD, f) != obj:
                os.remove(os.path.join(BUILD, f))
    return obj


def _compile_host(src):
    path = os.path.join(CSRC, src)
    deps = [path, os.path.join(INCLUDE, "tn_abi.h")]
    obj = os.path.join(BUILD, os.path.basename(src) + "." + _digest(deps, HOST_FLAGS) + ".o")
    if not os.path.exists(obj):
        r = subprocess.run(["g++"] + HOST_FLAGS + ["-c", path, "-o", obj], capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed on {src}")
        base = os.path.basename(src) + "."
        for f in os.listdir(BUILD):
            if f.startswith(base) and os.path.join(BUILD, f) != obj:
                os.remove(os.path.join(BUILD, f))
    return obj


def build(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile_hip, HIP_SOURCES))
    objs += [_compile_host(s) for s in ABI_HOST_SOURCES]
    lib = os.path.join(LIBDIR, "libtn_hip.so")
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    cxx = [os.path.join(CSRC, s) for s in CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cxx:
        api = os.path.join(LIBDIR, "libtensornetworks.so")
        tmp = api + ".tmp"
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, "-o", tmp] + cxx + [
            "-L", LIBDIR, "-ltn_hip", "-Wl,-rpath,$ORIGIN"]
        subprocess.run(cmd, check=True)
        os.replace(tmp, api)
        if verbose:
            print("built", api)
        test_src = os.path.join(ROOT, "tests", "cpp", "test_variational.cpp")
        if os.path.exists(test_src):
            exe = os.path.join(LIBDIR, "test_variational")
            subprocess.run(["g++", "-O2", "-std=c++17", "-I", INCLUDE, "-o", exe, test_src, "-L", LIBDIR,
                            "-ltensornetworks", "-ltn_hip", "-Wl,-rpath,$ORIGIN"], check=True)
            if verbose:
                print("built", exe)
    return lib


if __name__ == "__main__":
    build()
