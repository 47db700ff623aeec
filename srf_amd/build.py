"""Build libsrfx.so (all HIP sources under csrc/) for gfx950 with hipcc.

Usage: python -m srf_amd.build [--force]

Each .hip file is compiled to an object under csrc/_build/ and the objects
are linked into srf_amd/libsrfx.so, which the package loads with ctypes.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libsrfx.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops off: no v_pk_*_f32 generation.  No fast-math: float
# division and sqrt stay correctly rounded where the kernels rely on it.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    "-I", INCLUDE,
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build(force=False, verbose=False, csrc=CSRC, lib=LIB, jobs=None):
    """Compile every .hip file of ``csrc`` (in parallel, at most ``jobs``
    compilers at once) and link them into ``lib``."""
    from concurrent.futures import ThreadPoolExecutor
    build_dir = os.path.join(csrc, "_build")
    os.makedirs(build_dir, exist_ok=True)
    headers = [os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith(".h")]
    headers.append(os.path.join(INCLUDE, "srfx.h"))
    objs, cmds = [], []
    for src in sorted(f for f in os.listdir(csrc) if f.endswith(".hip")):
        s = os.path.join(csrc, src)
        o = os.path.join(build_dir, src[:-4] + ".o")
        objs.append(o)
        if force or _newer(o, [s] + headers):
            cmds.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    if cmds:
        jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            list(pool.map(run, cmds))
    if force or _newer(lib, objs):
        tmp = lib + ".tmp"
        run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
        os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
