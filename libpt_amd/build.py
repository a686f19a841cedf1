"""Build libpt_amd.so (HIP for gfx950) in-tree.

    python -m libpt_amd.build          # compile and link
    python -m libpt_amd.build --check  # also print kernel register usage

Every HIP source is compiled with ``--offload-arch=gfx950`` and with packed
FP32 instructions disabled (``-target-feature -packed-fp32-ops``); the
host-only sources go through the same driver so they see the HIP headers.
No hidden code-generation options: the tile kernels keep their MFMA
accumulators in ordinary VGPRs (no accumulator copies) with the default
code generation.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
LIB = os.path.join(HERE, "libpt_amd.so")

SOURCES = ["pt_kernel.hip", "pt_noalloc.hip", "pt_host.cpp", "pt_terms.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def _compile(src: str, extra=()) -> subprocess.Popen:
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", out]
    return subprocess.Popen(cmd, stdout=subprocess.PIPE,
                            stderr=subprocess.STDOUT, text=True)


def _noise(line: str) -> bool:
    return "packed-fp32-ops" in line


def build(check: bool = False, verbose: bool = False) -> str:
    procs = []
    for s in SOURCES:
        extra = ["-Rpass-analysis=kernel-resource-usage"] if (
            check and s.endswith(".hip")) else []
        procs.append((s, _compile(s, extra)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if not _noise(ln)]
        if p.returncode != 0:
            failed.append(s)
            print("\n".join(lines), file=sys.stderr)
        elif verbose or check:
            if lines:
                print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    objs = [os.path.join(OBJ, s + ".o") for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs,
           "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        print(r.stdout, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(check="--check" in sys.argv, verbose="-v" in sys.argv))
