"""Build libatjup_hip.so in-tree for gfx950 with hipcc (no JIT, no torch extension)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libatjup_hip.so")
SOURCES = ["atjup_kernels.hip", "atjup_physics.hip", "atjup_adjoint.hip", "atjup_report.hip", "atjup_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the FP64 kernels keep the reference's expression trees (no FMA contraction),
# which is what makes them reproduce the CPU oracle bit for bit.
# -packed-fp32-ops off: packed FP32 instructions are not used on this hardware pool.
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compile every source and link `out` (default: the in-tree library). `defines` are extra
    -D flags for tuning variants built elsewhere."""
    objs = []
    procs = []
    tag = "" if out == LIB else "." + os.path.basename(out).replace(".so", "")
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", tag + ".o"))
        cmd = [HIPCC, *FLAGS, *["-D" + d for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
        objs.append(obj)
    for p, src in procs:
        log, _ = p.communicate()
        text = log.decode(errors="replace")
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{text}")
        if verbose and text.strip():
            print(text)
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, out)
    return out


def build_pyatjup(verbose: bool = False) -> str:
    """The Cython binding atjup_amd/pyatjup*.so: pyatjup.pyx -> C++ (Cython) -> g++, linked against the
    in-tree libatjup_hip.so (found next to it at run time through an $ORIGIN rpath)."""
    import sysconfig
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    cpp = os.path.join(build_dir, "pyatjup.cpp")
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = [sys.executable, "-m", "cython", "--cplus", "-3", "-I", inc, os.path.join(HERE, "pyatjup.pyx"), "-o", cpp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("cython failed:\n" + r.stdout.decode(errors="replace"))
    out = os.path.join(HERE, "pyatjup" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-w", "-I", sysconfig.get_paths()["include"], "-I", inc,
           cpp, "-L", HERE, "-latjup_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("g++ failed on pyatjup:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_pyatjup(verbose="-v" in sys.argv))
