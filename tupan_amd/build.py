"""In-tree build of the HIP library and the libtupan.h drop-in shims.

    python -m tupan_amd.build          # or __graft_entry__.build()

Produces, under tupan_amd/lib/:
  libtupan_hip.so   all gfx950 kernels + the tupan_* C ABI (include/tupan_hip.h)
  libtupan_f64.so   drop-in for the reference's FP64 library: exports the
  libtupan_f32.so   libtupan.h symbols (phi_kernel, acc_jerk_kernel, ...)
                    with REAL = double / float, forwarding to libtupan_hip.so

Every HIP translation unit is compiled with packed FP32 disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["runtime.hip", "gravity.hip", "pnacc.hip", "sakura.hip",
               "integrate.hip", "adjoint.hip", "ensemble.hip", "nreg.hip",
               "ieee.hip", "reduce.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TUPAN_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       universal_newlines=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def _compile_hip(src, verbose):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj],
         verbose)
    return obj


def build_all(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), os.cpu_count() or 1, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile_hip(s, verbose), srcs))
    core = os.path.join(LIBDIR, "libtupan_hip.so")
    tmp = core + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
         + objs, verbose)
    os.replace(tmp, core)
    # libtupan.h drop-in shims: plain C, no HIP headers; resolve the core
    # library next to themselves.
    shim = os.path.join(CSRC, "libtupan_abi.c")
    if os.path.exists(shim):
        for bits, defs in (("64", ["-DTUPAN_ABI_F64"]), ("32", [])):
            out = os.path.join(LIBDIR, "libtupan_f%s.so" % bits)
            _run(["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-Wall"]
                 + defs + ["-I" + os.path.join(ROOT, "include"), shim,
                           "-L" + LIBDIR, "-ltupan_hip",
                           "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"],
                 verbose)
            os.replace(out + ".tmp", out)
    return core


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", os.path.join(LIBDIR, "libtupan_hip.so"))
