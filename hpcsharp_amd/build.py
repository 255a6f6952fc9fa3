"""Build libhpcsort.so (the gfx950 HIP library) in-tree with hipcc.

Usage: python -m hpcsharp_amd.build  (or hpcsharp_amd.build.build()).
       python -m hpcsharp_amd.build --comparator  also builds rocprim_compare,
       which times libhpcsort against rocPRIM's radix sort (a yardstick only:
       the library never uses rocPRIM).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libhpcsort.so")
SOURCES = [os.path.join(HERE, "csrc", s) for s in ("hpcs_api.hip", "hpcs_serve.hip")]
HEADERS = [os.path.join(HERE, "csrc", h)
           for h in ("hpcs_kernels.hpp", "hpcs_shapes.hpp", "hpcs_merge.hpp", "hpcs_digit11.hpp", "hpcs_msd.hpp", "hpcs_segments.hpp",
                     "hpcs_topk.hpp", "hpcs_arena.hpp", "hpcs_serve_core.hpp", "hpcs_tuning.hpp")]
DEPS = SOURCES + HEADERS + [os.path.join(REPO, "include", "hpcsort.h")]
ARCH = os.environ.get("HPCS_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every HIP build of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    # a part leaves the helpers of the other parts' entry points unused
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]
# hpcs_api.hip is compiled once per part (-DHPCS_PART=k, in parallel); each
# part holds one group of entry points and the kernels they instantiate
API_PARTS = 7
OBJ_DIR = os.path.join(HERE, "_build")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    inc = ["-I", os.path.join(REPO, "include")]
    api, serve = SOURCES
    jobs = [([hipcc(), *FLAGS, *inc, f"-DHPCS_PART={k}", "-c", api, "-o", os.path.join(OBJ_DIR, f"hpcs_api_{k}.o")])
            for k in range(API_PARTS)]
    jobs.append([hipcc(), *FLAGS, *inc, "-c", serve, "-o", os.path.join(OBJ_DIR, "hpcs_serve.o")])
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    failed = False
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed compiling libhpcsort.so")
    tmp = LIB + ".tmp"
    link = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *[j[-1] for j in jobs], "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libhpcsort.so")
    os.replace(tmp, LIB)
    return LIB


COMPARATOR = os.path.join(HERE, "rocprim_compare")


def build_comparator(verbose: bool = False) -> str:
    """hpcsharp_amd/rocprim_compare: links libhpcsort.so (found next to the
    binary) and rocPRIM's headers; not part of build()."""
    src = os.path.join(HERE, "csrc", "rocprim_compare.hip")
    cmd = [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(REPO, "include"), src, "-L", HERE, "-lhpcsort", "-Wl,-rpath,$ORIGIN",
           "-o", COMPARATOR + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building rocprim_compare")
    os.replace(COMPARATOR + ".tmp", COMPARATOR)
    return COMPARATOR


if __name__ == "__main__":
    print(build(verbose=True))
    if "--comparator" in sys.argv[1:]:
        print(build_comparator(verbose=True))
