"""Builds libprism_amd.so (the HIP library) in-tree with hipcc for gfx950.

Usage: python -m prism_amd.build [--force]

Flags that matter for results: -ffp-contract=off (no FMA contraction, so
the device evaluates the predictor exactly as written and in the same
rounding steps as any IEEE host), correctly rounded f32 division, no
denormal flushing; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(OUT, "libprism_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PRISM_AMD_ARCH", "gfx950")

SOURCES = ["layout.cpp", "extrema.hip", "predictor.hip", "bitplane.hip", "codec.hip", "strategy.hip", "quality.hip",
           "batch.hip"]
HEADERS = ["common.hpp", "batch.hpp", "extrema.hpp"]

# C++ API (prism::Compressor & co.) and the `prism` command line, host code
# over the C ABI
CPP = os.path.join(PKG, "cpp")
CPP_SOURCES = ["parameters.cpp", "compressor.cpp", "prism.cpp"]
CPP_LIB = os.path.join(OUT, "libprism_cpp.so")
CLI = os.path.join(OUT, "prism")
# a driver written against the reference's header names only (include path
# = the directory of the API headers), linked like any external caller
EXAMPLE_SRC = os.path.join(CPP, "examples", "reference_driver.cpp")
EXAMPLE = os.path.join(OUT, "reference_driver")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include"), "-I", SRC]
DEVICE = [f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _obj(src: str) -> str:
    return os.path.join(OUT, os.path.splitext(src)[0] + ".o")


def _stale(target: str, deps) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, force: bool) -> str:
    path = os.path.join(SRC, src)
    obj = _obj(src)
    deps = [path] + [os.path.join(SRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "prism_amd.h")]
    if not force and not _stale(obj, deps):
        return obj
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _cpp_headers():
    inc = os.path.join(CPP, "include", "prism")
    return [os.path.join(inc, f) for f in sorted(os.listdir(inc))] + [os.path.join(ROOT, "include", "prism_amd.h")]


def _compile_cpp(src: str, force: bool) -> str:
    path = os.path.join(CPP, "src", src)
    obj = os.path.join(OUT, "cpp_" + os.path.splitext(src)[0] + ".o")
    if not force and not _stale(obj, [path] + _cpp_headers()):
        return obj
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(CPP, "include"), "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _link(target: str, inputs, extra, force: bool):
    if not force and not _stale(target, inputs):
        return
    cmd = [HIPCC, "-o", target + ".tmp"] + extra + list(inputs)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed for {target}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(target + ".tmp", target)


def _build_example(force: bool):
    if not force and not _stale(EXAMPLE, [EXAMPLE_SRC, CPP_LIB] + _cpp_headers()):
        return
    obj = EXAMPLE + ".o"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(CPP, "include", "prism"), "-I",
           os.path.join(ROOT, "include"), "-c", EXAMPLE_SRC, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for the reference-API driver:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    _link(EXAMPLE, [obj, CPP_LIB, LIB], ["-Wl,-rpath,$ORIGIN"], True)


def build(force: bool = False, verbose: bool = False) -> str:
    """Builds libprism_amd.so (kernels + C ABI), libprism_cpp.so (C++ API),
    the `prism` executable and the reference-API driver into
    prism_amd/_build; returns the first."""
    os.makedirs(OUT, exist_ok=True)
    jobs = min(len(SOURCES) + len(CPP_SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
        cpp_objs = list(ex.map(lambda s: _compile_cpp(s, force), CPP_SOURCES + ["main.cpp"]))
    _link(LIB, objs, [f"--offload-arch={ARCH}", "-shared", "-fPIC"], force)
    rpath = "-Wl,-rpath,$ORIGIN"
    _link(CPP_LIB, cpp_objs[:-1] + [LIB], ["-shared", "-fPIC", rpath], force)
    _link(CLI, cpp_objs[-1:] + [CPP_LIB, LIB], [rpath], force)
    _build_example(force)
    if verbose:
        print(LIB)
        print(CPP_LIB)
        print(CLI)
        print(EXAMPLE)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
