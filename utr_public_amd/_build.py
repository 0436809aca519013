"""In-tree build of the native parts of utr_public_amd.

Outputs (git-ignored, shipped with the tree):
  utr_public_amd/_lib/libutr_hip.so    HIP kernels + C ABI (gfx950)
  utr_public_amd/_lib/libutr_synth.so  seeded synthetic inputs (host C++)
  utr_public_amd/_lib/libutr_io.so     file formats of the command-line program (host C++)
  utr_public_amd/bin/utr               command-line drop-in for UTR.exe
"""
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
BINDIR = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-munsafe-fp-atomics",
    # keep packed FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    hip_lib = os.path.join(LIBDIR, "libutr_hip.so")
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", hip_lib + ".tmp",
                                os.path.join(CSRC, "utr_api.hip"),
                                os.path.join(CSRC, "utr_plan.cpp")], verbose)
    os.replace(hip_lib + ".tmp", hip_lib)
    synth = os.path.join(LIBDIR, "libutr_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", synth + ".tmp",
          os.path.join(CSRC, "synth.cpp")], verbose)
    os.replace(synth + ".tmp", synth)
    inc = ["-I", os.path.join(ROOT, "include")]
    io_lib = os.path.join(LIBDIR, "libutr_io.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", io_lib + ".tmp",
          os.path.join(CSRC, "utr_io.cpp")] + inc, verbose)
    os.replace(io_lib + ".tmp", io_lib)
    exe = os.path.join(BINDIR, "utr")
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-Wall", "-o", exe + ".tmp",
          os.path.join(CSRC, "utr_cli.cpp"), os.path.join(CSRC, "utr_io.cpp")] + inc +
         ["-L", LIBDIR, "-lutr_hip", "-Wl,-rpath,$ORIGIN/../_lib",
          "-Wl,-rpath-link,/opt/rocm/lib"], verbose)
    os.replace(exe + ".tmp", exe)
    return hip_lib


if __name__ == "__main__":
    build(verbose=True)
