"""Compile the gfx950 engine library in-tree with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_HERE)
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


def _sources():
    csrc = os.path.join(_HERE, "csrc")
    return sorted(os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".hip"))


def build_engine(verbose: bool = False, jobs: int = None) -> str:
    """Compile the library from source (always: no up-to-date check), one
    object per translation unit in parallel, then link; print one line
    saying that it did, with the time taken and a digest of the sources."""
    import hashlib
    import tempfile
    import time
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    srcs = _sources()
    jobs = jobs or max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1), 16))
    inc = ["-I", os.path.join(ROOT, "include")]
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="mf_build_") as tmpd:
        def compile_one(src):
            obj = os.path.join(tmpd, os.path.basename(src) + ".o")
            r = subprocess.run([hipcc(), *compile_flags, *inc, "-c", src, "-o", obj],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n" + r.stderr[-8000:])
            if verbose and r.stderr:
                print(r.stderr[-4000:])
            return obj
        with ThreadPoolExecutor(jobs) as ex:
            objs = list(ex.map(compile_one, srcs))
        tmp = OUT + ".tmp"
        r = subprocess.run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr[-8000:])
    os.replace(tmp, OUT)
    h = hashlib.sha256()
    hdrs = [os.path.join(_HERE, "csrc", f) for f in sorted(os.listdir(os.path.join(_HERE, "csrc"))) if f.endswith(".h")]
    for f in srcs + hdrs + [os.path.join(ROOT, "include", "multifil_amd.h")]:
        with open(f, "rb") as fh:
            h.update(fh.read())
    print(f"build: compiled {os.path.relpath(OUT, ROOT)} from source in {time.time() - t0:.1f} s "
          f"({len(srcs)} translation units, {jobs} jobs; gfx950, sources sha256 {h.hexdigest()[:16]})", flush=True)
    return OUT


if __name__ == "__main__":
    print(build_engine(verbose=True))
