"""Build the package's native libraries in-tree.

    python -m ftlm_extended_hubbard_amd.build

- libftlm_hip.so: every HIP source under csrc/, compiled with hipcc for gfx950 only, with
  packed-FP32 VALU instructions disabled (-packed-fp32-ops); exports the C ABI of
  include/ftlm_hip.h.
- libftlm.so: the reference's C++ API (csrc/host/include/ftlm/*.hpp) over that C ABI.
- _ftlm_core*.so: the pybind11 module over libftlm.so.
- tests/cpp/_bin/*: the C++ API tests (tests/cpp/*.cpp).
"""

from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libftlm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")

ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    # a kernel whose host-side instantiation fails leaves its launch stub undefined without any
    # diagnostic; refuse such a library here rather than at load time on the GPU
    r = subprocess.run(["nm", "-D", "--undefined-only", tmp], capture_output=True, text=True)
    undef = [ln.split()[-1] for ln in r.stdout.splitlines() if "ftlm_hip" in ln]
    if undef:
        raise RuntimeError("libftlm_hip.so has undefined library symbols:\n" + "\n".join(undef[:10]))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


HOST = os.path.join(CSRC, "host")
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}",
              f"-I{os.path.join(HOST, 'include')}"]
CXX = os.environ.get("CXX", "g++")


def _run(cmd: list[str], what: str) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_host(verbose: bool = False) -> None:
    """libftlm.so, the _ftlm_core extension and the C++ test programs (needs libftlm_hip.so)."""
    lib = os.path.join(PKG, "libftlm.so")
    _run([CXX, *HOST_FLAGS, "-shared", os.path.join(HOST, "ftlm_host.cpp"), "-o", lib + ".tmp", f"-L{PKG}",
          "-lftlm_hip", "-Wl,-rpath,$ORIGIN"], "libftlm.so")
    os.replace(lib + ".tmp", lib)
    import sysconfig

    try:
        import pybind11

        suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
        mod = os.path.join(PKG, "_ftlm_core" + suffix)
        inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
        _run([CXX, *HOST_FLAGS, *inc, "-shared", os.path.join(HOST, "bindings.cpp"), "-o", mod + ".tmp", f"-L{PKG}",
              "-lftlm", "-lftlm_hip", "-Wl,-rpath,$ORIGIN"], "_ftlm_core")
        os.replace(mod + ".tmp", mod)
    except ImportError:
        if verbose:
            print("pybind11 not importable: _ftlm_core not built")
    tdir = os.path.join(ROOT, "tests", "cpp")
    bdir = os.path.join(tdir, "_bin")
    os.makedirs(bdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(tdir, "test_*.cpp")))

    def one(src: str) -> None:
        out = os.path.join(bdir, os.path.splitext(os.path.basename(src))[0])
        _run([CXX, *HOST_FLAGS, src, "-o", out, f"-L{PKG}", "-lftlm", "-lftlm_hip",
              "-Wl,-rpath,$ORIGIN/../../../ftlm_extended_hubbard_amd"], os.path.basename(src))

    with cf.ThreadPoolExecutor(max_workers=max(1, min(4, len(srcs)))) as ex:
        list(ex.map(one, srcs))
    # the grand-canonical n(mu) drivers (C ABI only)
    bin_dir = os.path.join(PKG, "bin")
    os.makedirs(bin_dir, exist_ok=True)
    drv = os.path.join(HOST, "drivers", "ftlm_nmu.cpp")
    for name, k in (("ftlm_nmu_rect", 0), ("ftlm_nmu_rect_k", 1)):
        out = os.path.join(bin_dir, name)
        _run([CXX, *HOST_FLAGS, f"-DFTLM_NMU_K={k}", drv, "-o", out + ".tmp", f"-L{PKG}", "-lftlm_hip",
              "-Wl,-rpath,$ORIGIN/.."], name)
        os.replace(out + ".tmp", out)
    # the small benchmark drivers (C ABI plus HIP runtime device buffers)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    bench = os.path.join(HOST, "drivers", "ftlm_bench.cpp")
    for name, k in (("ftlm_bench_2x2", 1), ("ftlm_bench_4x2_ftlm", 2), ("ftlm_bench_2x2_nmu", 3)):
        out = os.path.join(bin_dir, name)
        _run([CXX, *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include", f"-DFTLM_BENCH={k}", bench, "-o",
              out + ".tmp", f"-L{PKG}", "-lftlm_hip", f"-L{rocm}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/..",
              f"-Wl,-rpath,{rocm}/lib"], name)
        os.replace(out + ".tmp", out)
    if verbose:
        print("built", lib, "and", len(srcs), "C++ tests, drivers in", bin_dir)


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    build_host(verbose=True)
    sys.exit(0)
