"""Host sanitizer run: AddressSanitizer + UndefinedBehaviorSanitizer over the
host code of libobvh.so (entry points, argument checks, workspace carving,
tree validation, SAH, the host DiTO), the plugins, and the CPU oracle, under
the CPU test suite.

    python -m obvh_amd.sanitize [pytest args...]

builds instrumented copies into obvh_amd/_asan/ (device code is built as
usual; only the host side is instrumented) and oracle/_asan/, then runs
`pytest -m "not gpu"` with the clang ASan runtime preloaded and the copies
selected by OBVH_LIB_DIR / OBVH_ORACLE_LIB.  Exit status: pytest's (a
sanitizer report aborts the run).  The regular libraries are not touched.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_asan")
ORACLE_OUT = os.path.join(ROOT, "oracle", "_asan")
LLVM = "/opt/rocm/llvm"
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


def asan_runtime() -> str:
    hits = sorted(glob.glob(os.path.join(LLVM, "lib", "clang", "*", "lib", "linux", "libclang_rt.asan-x86_64.so")))
    if not hits:
        raise RuntimeError("clang's ASan runtime not found under /opt/rocm/llvm")
    return hits[-1]


def build_library() -> None:
    os.environ["OBVH_LIB_DIR"] = OUT
    os.makedirs(OUT, exist_ok=True)
    import importlib

    from . import build

    importlib.reload(build)  # pick up OBVH_LIB_DIR
    host = []
    for f in SAN:
        host += ["-Xarch_host", f]
    build.HIP_EXTRA[:] = host
    build.HOST_EXTRA[:] = SAN
    build.HOST_CXX = os.path.join(LLVM, "bin", "clang++")
    build.build(verbose=True)


def build_oracle() -> str:
    os.makedirs(ORACLE_OUT, exist_ok=True)
    srcs = ["dito.c", "transform.c", "lbvh.c", "cpu_transform.c", "collapse.c", "atrbvh.c", "traverse.c", "par.c"]
    out = os.path.join(ORACLE_OUT, "liboracle.so")
    cmd = [os.path.join(LLVM, "bin", "clang"), "-O1", "-g", "-std=c11", "-pthread", "-fPIC", "-ffp-contract=off",
           *SAN, "-shared", "-o", out, *[os.path.join(ROOT, "oracle", s) for s in srcs], "-lm"]
    subprocess.run(cmd, check=True)
    return out


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    build_library()
    oracle_lib = build_oracle()
    env = dict(os.environ)
    env.update({
        "OBVH_LIB_DIR": OUT,
        "OBVH_ORACLE_LIB": oracle_lib,
        "LD_PRELOAD": asan_runtime(),
        # Python and torch keep their allocations to the end: no leak report
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1:detect_odr_violation=0",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
    })
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-m", "not gpu", "-q", "-p", "no:xdist",
           *argv]
    print("sanitized run:", " ".join(cmd), flush=True)
    return subprocess.run(cmd, env=env, cwd=ROOT).returncode


if __name__ == "__main__":
    raise SystemExit(main())
