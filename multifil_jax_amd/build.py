"""Compile the gfx950 engine library in-tree with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_HERE)
SRC = os.path.join(_HERE, "csrc", "mf_engine.hip")
OUT = os.path.join(_HERE, "libmultifil_amd.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-shared", "-fPIC",
    # keep hipcc from emitting packed FP32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_engine(verbose: bool = False) -> str:
    """Compile the library from source (always: no up-to-date check), and
    print one line saying that it did, with the time taken and a digest of
    the sources."""
    import hashlib
    import time
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), SRC, "-o", tmp]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-8000:])
    if verbose and r.stderr:
        print(r.stderr[-4000:])
    os.replace(tmp, OUT)
    h = hashlib.sha256()
    srcs = [os.path.join(_HERE, "csrc", f) for f in sorted(os.listdir(os.path.join(_HERE, "csrc")))]
    for f in srcs + [os.path.join(ROOT, "include", "multifil_amd.h")]:
        with open(f, "rb") as fh:
            h.update(fh.read())
    print(f"build: compiled {os.path.relpath(OUT, ROOT)} from source in {time.time() - t0:.1f} s "
          f"(gfx950, sources sha256 {h.hexdigest()[:16]})", flush=True)
    return OUT


if __name__ == "__main__":
    print(build_engine(verbose=True))
