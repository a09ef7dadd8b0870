"""Build the gfx950 shared library libpamr_mg.so in the package directory.

Every .hip source under csrc/ is compiled with hipcc for gfx950 and linked
into one shared library.  Device code is compiled with -ffp-contract=off
(no FMA contraction: the FP64 stencils then round exactly like the
reference's Fortran) and without packed-FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpamr_mg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(HERE, '..', 'include')}",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return out


APPS = os.path.join(HERE, "apps")
INC = os.path.join(HERE, "..", "include")


def build_apps() -> list:
    """The example application hooks (plain C, built with gcc) as shared
    libraries that link libpamr_mg.so from the package directory."""
    out = []
    for src in sorted(f for f in os.listdir(APPS) if f.endswith(".c")):
        so = os.path.join(APPS, "lib" + src[:-2] + ".so")
        cmd = ["gcc", "-std=gnu99", "-O2", "-fPIC", "-shared", "-Wall", "-D__HIP_PLATFORM_AMD__",
               f"-I{INC}", "-I/opt/rocm/include", os.path.join(APPS, src), "-o", so + ".tmp",
               f"-L{HERE}", "-lpamr_mg", "-L/opt/rocm/lib", "-lamdhip64",
               "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"gcc failed for {src}:\n{r.stderr[-6000:]}")
        os.replace(so + ".tmp", so)
        out.append(so)
    return out


def _stale(src) -> bool:
    """True when src's object is missing or older than src or any header."""
    out = os.path.join(OBJ, src + ".o")
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    deps = [os.path.join(CSRC, src)]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps += [os.path.join(INC, f) for f in os.listdir(INC) if f.endswith(".h")]
    return any(os.path.getmtime(d) > t for d in deps)


def build(jobs: int = 8, incremental: bool = False) -> str:
    """Compile every source (incremental: only those whose object is stale)
    and link libpamr_mg.so and the example applications."""
    srcs = sources()
    todo = [s for s in srcs if _stale(s)] if incremental else srcs
    if todo:
        with ThreadPoolExecutor(max_workers=min(jobs, len(todo))) as ex:
            list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ, s + ".o") for s in srcs]
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(LIB + ".tmp", LIB)
    build_apps()
    return LIB


if __name__ == "__main__":
    print(build(8, incremental="-i" in sys.argv[1:]))
