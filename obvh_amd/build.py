"""Build the HIP shared library obvh_amd/libobvh.so for gfx950.

Every HIP source is compiled with hipcc in parallel and linked into one
shared library that exports the C ABI of include/obvh/obvh.h.  The two
intersector plugins of include/obvh/korangar_ads.h (liblbvh_gpu.so,
libatrbvh_gpu.so: one source, two builder settings) link against it and find
it next to themselves (RPATH $ORIGIN).

Floating point flags are part of the parity contract with the CPU oracle:
  -ffp-contract=off                          no FMA contraction
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                             no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
# OBVH_LIB_DIR: where the libraries are built and loaded from (default: the
# package directory; the sanitizer run uses its own)
OUT = os.environ.get("OBVH_LIB_DIR") or HERE
BUILD = os.path.join(OUT, "_build")
LIB = os.path.join(OUT, "libobvh.so")
ARCH = os.environ.get("OBVH_ARCH", "gfx950")

# plugin library name -> value of OBVH_ADS_ATRBVH for ads_plugin.hip
PLUGINS = {"lbvh_gpu": 0, "atrbvh_gpu": 1}
# host-only C++ linked into libobvh.so (g++, no FMA contraction)
HOST_SOURCES = ["cpu_dito.cpp"]
# the bench integrator plugin (host C++ over libobvh.so)
BENCH_PLUGIN = "bench"
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Werror",
              f"-I{os.path.join(ROOT, 'include')}"]

SOURCES = ["abi.hip", "dito.hip", "transform.hip", "lbvh.hip", "collapse.hip", "atrbvh.hip", "traverse.hip", "upload.hip"]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


# extra flags for every HIP / host compile (set by the sanitizer build)
HIP_EXTRA: list = []
HOST_EXTRA: list = []
HOST_CXX = None


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *FLAGS, *HIP_EXTRA, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def plugin_path(name: str) -> str:
    return os.path.join(OUT, f"lib{name}.so")


def _compile_plugin(name: str) -> str:
    obj = os.path.join(BUILD, f"ads_plugin_{name}.o")
    cmd = [hipcc(), *FLAGS, *HIP_EXTRA, f"-DOBVH_ADS_ATRBVH={PLUGINS[name]}", "-c",
           os.path.join(CSRC, "ads_plugin.hip"), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on ads_plugin.hip ({name}):\n{r.stderr}")
    return obj


def gxx() -> str:
    for cand in (os.environ.get("CXX"), shutil.which("g++")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("g++ not found: the host C++ sources cannot be built")


def _compile_host(src: str, extra=()) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".host.o")
    cmd = [HOST_CXX or gxx(), *HOST_FLAGS, *HOST_EXTRA, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on {src}:\n{r.stderr}")
    return obj


def _link(objs, out, extra=()) -> None:
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, *extra, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of {os.path.basename(out)} failed:\n{r.stderr}")
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    """Compile every HIP source, link libobvh.so, then the two intersector
    plugins against it (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES) + len(PLUGINS) + 2, os.cpu_count() or 4)) as ex:
        lib_objs = ex.map(_compile, SOURCES)
        host_objs = ex.map(_compile_host, HOST_SOURCES)
        bench_obj = ex.submit(_compile_host, "bench_plugin.cpp", ("-fvisibility=hidden",))
        plugin_objs = {name: ex.submit(_compile_plugin, name) for name in PLUGINS}
        objs = list(lib_objs) + list(host_objs)
        plugin_objs = {name: f.result() for name, f in plugin_objs.items()}
        bench_obj = bench_obj.result()
    _link(objs, LIB)
    for name, obj in plugin_objs.items():
        _link([obj], plugin_path(name), ("-L" + OUT, "-lobvh", "-Wl,-rpath,$ORIGIN"))
    _link([bench_obj], plugin_path(BENCH_PLUGIN), ("-L" + OUT, "-lobvh", "-Wl,-rpath,$ORIGIN"))
    if verbose:
        print(f"built {LIB} and {', '.join(os.path.basename(plugin_path(n)) for n in [*PLUGINS, BENCH_PLUGIN])}",
              file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
