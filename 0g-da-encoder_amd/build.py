"""Build lib/libzgr.so (the gfx950 encoder behind include/zgr.h) with hipcc.

    python 0g-da-encoder_amd/build.py [-q] [--incremental]

Compiles in-tree so the shared library travels with the repository snapshot.  The
translation units compile in parallel, then link into one library.  build() recompiles
every unit; --incremental (build(incremental=True)) recompiles only the units whose
sources or included headers changed since their object was written.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libzgr.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=" + ARCH,
    "-fPIC",
    # packed FP32 forms have produced silent wrong results on these boxes; the
    # encoder is integer-only, but keep the compiler from emitting them anywhere.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

# zgr.hip: the encoder kernels and every host entry point; zgr_params.hip: the
# parameter-file kernels (the slowest unit to compile, rarely changed)
SOURCES = ["zgr.hip", "zgr_params.hip"]


def _filter(out):
    return "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)


def _deps_newer(obj, dep):
    """True when the object is missing or older than any file its dependency list names."""
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return True
    t = os.path.getmtime(obj)
    text = open(dep).read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else []
    return any(not os.path.exists(f) or os.path.getmtime(f) > t for f in files)


def build(verbose=False, incremental=False):
    os.makedirs(OBJDIR, exist_ok=True)
    procs = []
    for s in SOURCES:
        base = os.path.splitext(s)[0]
        obj = os.path.join(OBJDIR, base + ".o")
        dep = os.path.join(OBJDIR, base + ".d")
        if incremental and not _deps_newer(obj, dep):
            continue
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-MMD", "-MF", dep, "-o", obj + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors = []
    for s, obj, p in procs:
        out = _filter(p.communicate()[0])
        if p.returncode != 0:
            errors.append("hipcc failed on %s:\n%s" % (s, out))
        else:
            os.replace(obj + ".tmp", obj)
            if verbose and out.strip():
                print(out)
    if errors:
        raise RuntimeError("\n".join(errors))
    tmp = LIB + ".tmp"
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    # roctx: the host-side ranges of the entry points (rocprofv3 --marker-trace)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-fPIC", "-shared", "-Wl,-z,defs"] + objs + \
        ["-lrocprofiler-sdk-roctx", "-o", tmp]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + _filter(res.stdout))
    os.replace(tmp, LIB)
    build_host_tools(verbose)
    return LIB


HOST = os.path.join(HERE, "host")
ENCODE_FILE = os.path.join(LIBDIR, "zg_encode_file")


def build_host_tools(verbose=False):
    """The C++ host mirror's example program (host code only, linked to libzgr.so)."""
    tmp = ENCODE_FILE + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall",
           os.path.join(HOST, "zg_encode_file.cpp"), "-L" + LIBDIR, "-lzgr", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed:\n" + res.stdout)
    os.replace(tmp, ENCODE_FILE)
    return ENCODE_FILE


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv, incremental="--incremental" in sys.argv))
