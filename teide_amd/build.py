"""Build the native libraries in-tree.

    teide_amd/lib/libteide_amd.so   the engine: C17 host (gcc) + HIP device
                                     layer for gfx950 (hipcc), C ABI = td.h
    teide_amd/lib/libtdgen.so        synthetic workload generators
    oracle/_build/liborc.so          test-only CPU oracle (oracle/Makefile)
    oracle/_ref/*                    test-only reference build, when the
                                     reference sources are present

Run `python -m teide_amd.build` (or __graft_entry__.build()).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(REPO, "_build", "obj")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("TEIDE_OFFLOAD_ARCH", "gfx950")

C_FLAGS = ["-std=c17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(CSRC, "host")]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
             # packed-FP32 instructions off for all device code
             # no a*b+c -> fma contraction: aggregate finalisation (variance)
             # must round like the reference's separate multiply and add
             "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument",
             "-I" + os.path.join(REPO, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    return out


def _compile(src):
    base = os.path.relpath(src, CSRC).replace(os.sep, "_")
    obj = os.path.join(OBJ, base + ".o")
    if src.endswith(".hip"):
        _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    else:
        _run(["gcc"] + C_FLAGS + ["-c", src, "-o", obj])
    return obj


def _sources():
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
                  if f.endswith(".c"))
    dev = sorted(os.path.join(CSRC, "dev", f) for f in os.listdir(os.path.join(CSRC, "dev"))
                 if f.endswith(".hip"))
    return host, dev


def build_engine(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    host, dev = _sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, dev + host))
    out = os.path.join(LIB, "libteide_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
         ["-Wl,--no-undefined", "-lpthread", "-lm"])
    os.replace(tmp, out)
    return out


def build_datagen():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libtdgen.so")
    _run(["gcc", "-std=c17", "-O2", "-fPIC", "-shared", "-o", out,
          os.path.join(CSRC, "datagen.c"), "-lm", "-lpthread"])
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "-j8", "all"])


def build_all():
    build_datagen()
    build_engine()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIB, "libteide_amd.so"), file=sys.stderr)
