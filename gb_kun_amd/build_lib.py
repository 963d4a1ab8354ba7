"""Builds libgbk.so (HIP kernels for gfx950 + host C++) in-tree.

Host C++ that uses OpenMP is compiled with g++; every .hip file is compiled by
hipcc for gfx950 with packed-FP32 instructions disabled; the objects are
linked into gb_kun_amd/libgbk.so.  Objects are cached under gb_kun_amd/_build
by content hash of the source and headers, so a rebuild recompiles only what
changed.
"""
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgbk.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GBK_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-fopenmp", "-I", os.path.join(ROOT, "include")]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _headers_digest():
    h = hashlib.sha256()
    for d in (CSRC, os.path.join(ROOT, "include")):
        for f in sorted(os.listdir(d)):
            if f.endswith((".h", ".hpp")):
                with open(os.path.join(d, f), "rb") as fh:
                    h.update(fh.read())
    return h.hexdigest()


def _obj_path(src, flags, hdr):
    h = hashlib.sha256()
    with open(os.path.join(CSRC, src), "rb") as fh:
        h.update(fh.read())
    h.update(" ".join(flags).encode())
    h.update(hdr.encode())
    return os.path.join(BUILD, f"{src}.{h.hexdigest()[:16]}.o")


_HOST_NOISE = "'-packed-fp32-ops' is not a recognized feature for this target"


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    # the host compilation pass does not know the device-only feature flag
    return "\n".join(l for l in r.stdout.splitlines() if _HOST_NOISE not in l and l.strip())


def build(verbose=True, jobs=None, force=False):
    """force=True recompiles every source even when an object of the same
    content is cached in _build/."""
    os.makedirs(BUILD, exist_ok=True)
    hip, cpp = _sources()
    hdr = _headers_digest()
    jobs_list = []
    objs = []
    for s in hip:
        o = _obj_path(s, HIP_FLAGS, hdr)
        objs.append(o)
        if force or not os.path.exists(o):
            jobs_list.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in cpp:
        o = _obj_path(s, CXX_FLAGS, hdr)
        objs.append(o)
        if force or not os.path.exists(o):
            jobs_list.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    if jobs_list:
        n = jobs or min(8, os.cpu_count() or 4)
        with ThreadPoolExecutor(max_workers=n) as ex:
            for cmd, out in zip(jobs_list, ex.map(_run, jobs_list)):
                if verbose:
                    print("compiled", os.path.basename(cmd[-3]), flush=True)
                if out.strip() and verbose:
                    print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-fopenmp=libgomp", "-lgomp"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, flush=True)
    build_cpp_drivers(verbose)
    return LIB


def build_cpp_drivers(verbose=True):
    """Compiles the C++ front-end drivers under tests/cpp against include/gbk/lablas.hpp and libgbk.so."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    if not os.path.isdir(src_dir):
        return
    os.makedirs(out_dir, exist_ok=True)
    for f in sorted(os.listdir(src_dir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, f[:-4])
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(src_dir, f),
              "-L", HERE, "-lgbk", "-Wl,-rpath,$ORIGIN/../../../gb_kun_amd", "-o", exe])
        if verbose:
            print("built", exe, flush=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
