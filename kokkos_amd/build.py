"""Build the MI355X execution space's shared libraries in-tree with hipcc.

Libraries (under kokkos_amd/lib/):
  libkkmx_rt.so     C-ABI runtime (include/kkmx_runtime.h)
  libkkmx_paths.so  hot-path C-ABI entry points (include/kkmx_paths.h)
  further libraries listed in LIBS below.
The restated reference unit tests are test code: their library is built from
tests/cpp/selftest.hip into tests/cpp/lib/ (TEST_LIBS below), next to the
tests that load it.

Usage: python -m kokkos_amd.build [--jobs N] [names...]
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIBDIR = PKG / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KKMX_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # Packed FP32 instructions stay off: hipcc only forms them when allowed.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
    "-I", str(PKG / "include"),
    "-Wno-unused-command-line-argument",
    # Let the memory-dependence analysis look far enough back for the Range
    # scan's final pass to reuse its first pass's loads of invariant inputs
    # (the two are hundreds of instructions apart).
    "-mllvm", "-memdep-block-scan-limit=100000",
    "-mllvm", "-memdep-block-number-limit=100000",
    "-mllvm", "-gvn-max-num-deps=100000",
]

# name -> (sources relative to the package, extra flags, link against runtime)
LIBS = {
    "kkmx_rt": (["csrc/runtime.cpp"], [], False),
    "kkmx_paths": (["csrc/paths.hip"], [], True),
    "kkmx_peak": (["csrc/peak.hip"], [], True),
    "kkmx_tuner": (["csrc/tuner.hip"], [], True),
    "kkmx_sweep": (["csrc/sweep.hip"], [], True),
}
# Test libraries: name -> (sources relative to the repository root, output dir).
TEST_LIBDIR = ROOT / "tests" / "cpp" / "lib"
TEST_LIBS = {
    "kkmx_selftest": (["tests/cpp/selftest.hip"], TEST_LIBDIR),
}


def lib_path(name: str) -> Path:
    if name in TEST_LIBS:
        return TEST_LIBS[name][1] / f"lib{name}.so"
    return LIBDIR / f"lib{name}.so"


def build_one(name: str, verbose: bool = False) -> Path:
    if name in TEST_LIBS:
        rel, outdir = TEST_LIBS[name]
        srcs, extra, link_rt = [str(ROOT / s) for s in rel], [], True
        rpath = "-Wl,-rpath," + os.path.join("$ORIGIN", os.path.relpath(LIBDIR, outdir))
    else:
        rel, extra, link_rt = LIBS[name]
        srcs, outdir, rpath = [str(PKG / s) for s in rel], LIBDIR, "-Wl,-rpath,$ORIGIN"
    outdir.mkdir(parents=True, exist_ok=True)
    out = lib_path(name)
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, *COMMON, *extra, *srcs, "-o", str(tmp)]
    if link_rt:
        cmd += ["-L", str(LIBDIR), "-lkkmx_rt", rpath]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {name}")
    os.replace(tmp, out)
    return out


def build(names=None, jobs: int = 4, verbose: bool = False) -> list[Path]:
    names = list(names or [*LIBS.keys(), *TEST_LIBS.keys()])
    outs = []
    # The runtime first: everything else links against it.
    if "kkmx_rt" in names:
        outs.append(build_one("kkmx_rt", verbose))
        names.remove("kkmx_rt")
    elif not lib_path("kkmx_rt").exists():
        outs.append(build_one("kkmx_rt", verbose))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for p in ex.map(lambda n: build_one(n, verbose), names):
            outs.append(p)
    return outs


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    jobs = 4
    for a in sys.argv[1:]:
        if a.startswith("--jobs="):
            jobs = int(a.split("=", 1)[1])
    for p in build(args or None, jobs=jobs, verbose="--verbose" in sys.argv):
        print(p)
