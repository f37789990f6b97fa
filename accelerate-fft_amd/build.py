"""Builds libafft.so (the gfx950 FFT engine) in-tree with hipcc.

    python accelerate-fft_amd/build.py [-j N] [--sanitize]

Every HIP translation unit is compiled for gfx950 only, with packed-FP32 instruction
selection disabled (-Xclang -target-feature -Xclang -packed-fp32-ops), then linked into
accelerate-fft_amd/lib/libafft.so.  Translation units compile in parallel.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libafft.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
]
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature: harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    return src, r.returncode, err, obj


def build(jobs=None, verbose=True):
    stems = [os.path.splitext(f)[0] for f in SOURCES]
    if len(set(stems)) != len(stems):
        raise RuntimeError("two sources share an object name: " + ", ".join(SOURCES))
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = jobs or min(8, len(SOURCES))
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, err, obj in ex.map(_compile, SOURCES):
            if err.strip() and verbose:
                print(err, file=sys.stderr)
            if rc != 0:
                raise RuntimeError(f"hipcc failed on {src}")
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


# The host-sanitised variant: the host-side C++ sources (planner, plan cache, ragged planning,
# autotuner, VJP front-end) compiled with AddressSanitizer and UndefinedBehaviorSanitizer on the
# host pass only (-Xarch_host before every -fsanitize), linked with the kernels of the regular
# build into lib_san/libafft.so.  Load it with AFFT_LIB=<that path> and the ASan runtime preloaded
# (san_runtime() names it).
SAN_OBJ = os.path.join(HERE, "_build_san")
SAN_LIB = os.path.join(HERE, "lib_san", "libafft.so")
HOST_SOURCES = sorted(f for f in SOURCES if f.endswith(".cpp"))
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]


def san_runtime():
    """The clang ASan runtime (it carries the UBSan handlers too), or None."""
    import glob
    hits = sorted(glob.glob(os.path.join(os.path.dirname(HIPCC), "..", "llvm", "lib", "clang", "*", "lib",
                                         "linux", "libclang_rt.asan-x86_64.so")))
    return os.path.realpath(hits[-1]) if hits else None


def sanitized_stale():
    """True when lib_san/libafft.so is missing or older than a source, a header or the regular
    library it links the kernels of."""
    if not os.path.exists(SAN_LIB):
        return True
    deps = [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(HERE, "..", "include", "afft.h")]
    if os.path.exists(LIB):
        deps.append(LIB)
    newest = max(os.path.getmtime(d) for d in deps if os.path.isfile(d))
    return os.path.getmtime(SAN_LIB) < newest


def build_sanitized(verbose=True):
    rt = san_runtime()
    if rt is None:
        raise RuntimeError("no clang ASan runtime next to hipcc")
    missing = [f for f in SOURCES if f.endswith(".hip") and
               not os.path.exists(os.path.join(OBJ, os.path.splitext(f)[0] + ".o"))]
    if missing:
        build(verbose=verbose)
    os.makedirs(SAN_OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(SAN_LIB), exist_ok=True)
    objs = []
    for src in SOURCES:
        stem = os.path.splitext(src)[0]
        if src not in HOST_SOURCES:
            objs.append(os.path.join(OBJ, stem + ".o"))
            continue
        obj = os.path.join(SAN_OBJ, stem + ".o")
        cmd = [HIPCC] + COMMON + SAN_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n" + r.stderr)
        objs.append(obj)
    rtdir = os.path.dirname(rt)
    tmp = SAN_LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + \
          ["-L" + rtdir, "-lclang_rt.asan-x86_64", "-Wl,-rpath," + rtdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, SAN_LIB)
    if verbose:
        print("built", SAN_LIB)
    return SAN_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--sanitize", action="store_true",
                    help="also build lib_san/libafft.so (host code with ASan + UBSan)")
    a = ap.parse_args()
    build(a.j)
    if a.sanitize:
        build_sanitized()
