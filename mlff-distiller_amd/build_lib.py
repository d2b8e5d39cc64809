"""Build libmlff_amd.so for gfx950 with hipcc (in-tree, no JIT cache).

Each .hip translation unit is compiled to an object, then linked into one
shared library under ``_lib/``.  Flags:

* ``--offload-arch=gfx950``: MI355X only;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: keep the compiler from
  emitting packed FP32 arithmetic;
* ``-ffp-contract=fast`` for the model kernels; the neighbour-list unit is
  built with ``-ffp-contract=off`` (its distance must round exactly like
  torch.norm on the CPU; the device code also pins the order with explicit
  fmaf and a contract(off) pragma).
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "libmlff_amd.so"
SOURCES = ("capi.hip", "nl.hip", "painn.hip", "md.hip", "train.hip", "train_update.hip", "train_filter.hip",
           "pack.hip", "probe.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def lib_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        flags = list(COMMON)
        flags.append("-ffp-contract=off" if src == "nl.hip" else "-ffp-contract=fast")
        cmd = [HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, out)
    for o in objs:
        try:
            os.remove(o)
        except OSError:
            pass
    return out


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out.decode(errors='replace')}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
