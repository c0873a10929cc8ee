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
SOURCES = ["kernels.hip", "kernels_ext.hip", "kernels_mkf.hip", "kernels_bfv.hip", "kernels_fused.hip", "kernels_galois.hip",
           "kernels_keygen.hip", "kernels_poly.hip", "mkhip.hip", "mkhip_plan.hip", "mkhip_kklss.hip", "mkhip_mkf.hip", "mkhip_bfv.hip",
           "mkhip_rotate.hip", "mkhip_keygen.hip", "mkhip_poly.hip", "mkhip_batch.hip",
           "mkhip_server.hip", "mkhip_peaks.hip", "mkhip_noalloc.hip"]
OBJDIR = os.path.join(HERE, "_build")
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


# per-file compiler flags (none at present: LLVM's max-ILP scheduling strategy
# measured 4-5 % faster for the external products of kernels_ext.hip with
# per-stage reduced inverse transforms, but with the unreduced inverse
# transforms it makes them spill, and the default then measures faster)
FILE_FLAGS = {}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    return res.returncode, out


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel, then link the shared library."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    cmds = [[HIPCC] + compile_flags + FILE_FLAGS.get(s, []) + ["-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 2) - 1), 8)
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    for (rc, out), src in zip(results, SOURCES):
        if rc != 0:
            sys.stderr.write(out + "\n")
            raise RuntimeError("hipcc failed compiling %s" % src)
    tmp = LIB + ".tmp"
    rc, out = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    if rc != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("hipcc failed linking %s" % LIB)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
