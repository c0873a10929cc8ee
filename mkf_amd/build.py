"""Build the gfx950 HIP library (mkf_amd/libmkhip.so) in-tree with hipcc.

No JIT, no torch extension machinery: one hipcc invocation over the .hip
sources, so the built .so travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmkhip.so")
SOURCES = ["kernels.hip", "kernels_mkf.hip", "mkhip.hip", "mkhip_plan.hip", "mkhip_kklss.hip", "mkhip_mkf.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # exact integer path; the one float64 step (ModUp overflow estimate) must not
    # be contracted or reassociated
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wno-unused-value",
    # packed fp32 ops are disabled for all HIP code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("hipcc failed building %s" % LIB)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
