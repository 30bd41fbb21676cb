"""Build of the HIP library (patchmatch_amd/libpmp.so) for gfx950.

Every HIP source is compiled with
  -ffp-contract=off     no implicit fma: the fma/mul/add sequence written in
                        the sources is the one executed (bit reproducibility);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                        no packed FP32 instructions.
Objects are compiled in parallel and linked into one shared library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_objs")
LIB = os.path.join(HERE, "libpmp.so")
SOURCES = ["pmp_context.hip", "pmp_estimator.hip", "pmp_queue.hip", "pmp_preinit.hip", "pmp_server.hip", "pmp_sched.cpp", "pmp_trace.cpp", "pmp_fusion.hip", "pmp_io.hip", "pmp_thumb.hip", "pmp_frontend.hip", "pmp_debug.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PMP_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_hip(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


SYNTH_LIB = os.path.join(HERE, "libpmpsynth.so")


def build_synth(verbose: bool = False) -> str:
    """Benchmark-input ray caster (csrc/synth_render.hip -> libpmpsynth.so)."""
    tmp = SYNTH_LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", os.path.join(CSRC, "synth_render.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for synth_render.hip:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SYNTH_LIB)
    if verbose:
        print(f"built {SYNTH_LIB}")
    return SYNTH_LIB


EXAMPLES = ["facade_window", "frontend_stream"]


def build_examples(verbose: bool = False) -> list:
    """C++ programs over the C-ABI and the header-only facade (linked to libpmp.so)."""
    out = []
    for name in EXAMPLES:
        src = os.path.join(HERE, "examples", name + ".cpp")
        exe = os.path.join(HERE, "examples", name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}", src, f"-L{HERE}", "-lpmp",
               "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example {name} failed to build:\n{r.stdout}\n{r.stderr}")
        out.append(exe)
        if verbose:
            print(f"built {exe}")
    return out


def build_oracle(verbose: bool = False) -> str:
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle/build/liboracle.so")
    return os.path.join(ROOT, "oracle", "build", "liboracle.so")


if __name__ == "__main__":
    build_hip(verbose=True)
    build_synth(verbose=True)
    build_examples(verbose=True)
    build_oracle(verbose=True)
    sys.exit(0)
