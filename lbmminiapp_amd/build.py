"""Builds the gfx950 HIP library ``liblbm_hip.so`` (C ABI: include/lbm_capi.h)
and the C++ host layer (reference-compatible classes and the deck-driven
driver) in-tree.

    python -m lbmminiapp_amd.build            # incremental
    python -m lbmminiapp_amd.build --force    # rebuild everything

Compiler flags:
  --offload-arch=gfx950                  MI355X only, no other targets
  -ffp-contract=off                      no implicit FMA: the FP64 path must be the
                                         reference's IEEE operation sequence; the
                                         reduced-precision collision re-enables
                                         contraction locally with a pragma
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                         no packed FP32 instructions
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
BIN = os.path.join(PKG, "bin")
LIB = os.path.join(PKG, "liblbm_hip.so")
HOSTLIB = os.path.join(PKG, "liblbm_host.so")
DRIVER = os.path.join(BIN, "lbm-proxy-app")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"),
             "-I" + HOST]


def _sources(d, ext):
    if not os.path.isdir(d):
        return []
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(ext))


def _headers(d):
    return [p for p in _sources(d, ".hpp") + _sources(d, ".h")]


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _run(cmd):
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + out.stdout + out.stderr)
    return out


def build(force: bool = False, verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    hdrs = _headers(CSRC) + [os.path.join(ROOT, "include", "lbm_capi.h")]
    srcs = _sources(CSRC, ".hip")
    objs = []
    todo = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s)[:-4] + ".o")
        objs.append(o)
        if force or _stale(o, [s] + hdrs):
            todo.append((s, o))
    jobs = jobs or min(8, max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", s, "-o", o]) for s, o in todo]
        for f in futs:
            r = f.result()
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr)
    if force or todo or _stale(LIB, objs):
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB, *objs])
    # host layer: reference-compatible classes over the C ABI, and the driver
    host_srcs = [s for s in _sources(HOST, ".cpp") if not s.endswith("main.cpp")]
    host_hdrs = _headers(HOST) + _sources(HOST, ".hh") + [os.path.join(ROOT, "include", "lbm_capi.h")]
    if host_srcs and (force or _stale(HOSTLIB, host_srcs + host_hdrs + [LIB])):
        _run(["g++", *CXX_FLAGS, "-shared", "-o", HOSTLIB, *host_srcs, "-L" + PKG, "-llbm_hip",
              "-Wl,-rpath,$ORIGIN"])
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main) and (force or _stale(DRIVER, [main, HOSTLIB] + host_hdrs)):
        _run(["g++", *CXX_FLAGS, "-o", DRIVER, main, "-L" + PKG, "-llbm_host", "-llbm_hip",
              "-Wl,-rpath,$ORIGIN/.."])
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    lib = build(force=args.force, verbose=args.verbose)
    print(lib)


if __name__ == "__main__":
    sys.exit(main())
