"""Build the MI355X-native libcfd2lcs libraries in-tree.

Produces libcfd2lcs_amd/lib/libcfd2lcs_dp.so (lcsdata_t = double) and
libcfd2lcs_amd/lib/libcfd2lcs_sp.so (lcsdata_t = float) with hipcc for
gfx950.  Usage: python -m libcfd2lcs_amd.build  (or build() from Python).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")

ARCH = os.environ.get("LCS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    ok = True
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines() if "unknown target feature" not in l and "packed-fp32" not in l]
        if p.returncode != 0:
            ok = False
            sys.stderr.write("FAILED: " + " ".join(c) + "\n" + "\n".join(lines) + "\n")
    if not ok:
        raise RuntimeError("hipcc build failed")


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I", CSRC]
    objs = {
        "kernels": os.path.join(BUILD, "kernels.o"),
        "particles": os.path.join(BUILD, "particles.o"),
        "adjoint": os.path.join(BUILD, "adjoint.o"),
        "scratch": os.path.join(BUILD, "scratch.o"),
        "grid": os.path.join(BUILD, "grid_host.o"),
        "io": os.path.join(BUILD, "lcs_io.o"),
        "api_dp": os.path.join(BUILD, "api_dp.o"),
        "api_sp": os.path.join(BUILD, "api_sp.o"),
    }
    cmds = [
        [HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, "kernels.hip"), "-o", objs["kernels"]],
        [HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, "particles.hip"), "-o", objs["particles"]],
        # hardware floating-point atomics for the adjoint scatters
        [HIPCC, *FLAGS, "-munsafe-fp-atomics", *inc, "-c", os.path.join(CSRC, "adjoint.hip"), "-o",
         objs["adjoint"]],
        [HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, "scratch.hip"), "-o", objs["scratch"]],
        [HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, "grid_host.cpp"), "-o", objs["grid"]],
        [HIPCC, *FLAGS, *inc, "-pthread", "-c", os.path.join(CSRC, "lcs_io.cpp"), "-o", objs["io"]],
        [HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, "api.cpp"), "-o", objs["api_dp"]],
        [HIPCC, *FLAGS, *inc, "-DLCS_SINGLE", "-c", os.path.join(CSRC, "api.cpp"), "-o", objs["api_sp"]],
    ]
    if verbose:
        for c in cmds:
            print(" ".join(c))
    _run(cmds)
    links = []
    for tag in ("dp", "sp"):
        out = os.path.join(LIB, f"libcfd2lcs_{tag}.so")
        tmp = out + ".tmp"
        links.append([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
                      objs[f"api_{tag}"], objs["kernels"], objs["particles"], objs["adjoint"], objs["scratch"],
                      objs["grid"], objs["io"], "-pthread"])
    _run(links)
    for tag in ("dp", "sp"):
        out = os.path.join(LIB, f"libcfd2lcs_{tag}.so")
        os.replace(out + ".tmp", out)
    return [os.path.join(LIB, f"libcfd2lcs_{t}.so") for t in ("dp", "sp")]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
