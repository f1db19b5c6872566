"""Builds anns_fhe_amd/libfhea.so (gfx950) from anns_fhe_amd/csrc with hipcc.

Run as `python -m anns_fhe_amd.build` (or through __graft_entry__.build()).
Objects go to anns_fhe_amd/_build/; the shared library sits in-tree next to
this file so that it travels with the repository snapshot.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfhea.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FHEA_ARCH", "gfx950")

SOURCES = ["fhea_context.cpp", "fhea_client.cpp", "fhea_ntt.hip", "fhea_ntt_f64.hip", "fhea_conv_cols.hip", "fhea_kernels.hip",
           "fhea_engine.hip", "fhea_probe.hip"]

FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Wall", "-Wno-unused-function",
    # packed-FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ host programs over the C-ABI (the reference's main_sim flow: seeded keys,
    # and the context and keys read from a test-vector directory)
    exes = []
    for name, src in (("c1_sequence", "c1_sequence_main.cpp"), ("dir_sequence", "dir_sequence_main.cpp")):
        exe = os.path.join(HERE, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"),
               os.path.join(CSRC, "host", src), "-o", exe + ".tmp", "-L", HERE, "-lfhea", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host program build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
        exes.append(exe)
    if verbose:
        print(f"built {LIB} and {', '.join(exes)}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
