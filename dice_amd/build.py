"""Build the HIP library and the drop-in executable in-tree for gfx950.

    python -m dice_amd.build          # -> dice_amd/lib/libshci_gpu.so, dice_amd/bin/shci_var

Every source is compiled by hipcc with FP contraction off (the Hamiltonian
elements reproduce the reference's FP64 summation order term by term) and
with packed-FP32 instructions disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["prims.hip", "kernels.hip", "mlp.hip", "host_io.cpp", "engine.cpp", "learned.cpp", "capi.cpp"]
ARCH = os.environ.get("SHCI_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + CSRC, "-I" + INCLUDE,
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(jobs=4, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libshci_gpu.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", so] + objs)
    exe = os.path.join(BIN, "shci_var")
    _run([HIPCC] + FLAGS + [os.path.join(CSRC, "shci_var_main.cpp"), "-o", exe, "-L" + LIB,
                            "-lshci_gpu", "-Wl,-rpath,$ORIGIN/../lib"])
    return so


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
