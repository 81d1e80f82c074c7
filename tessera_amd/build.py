"""Builds libtessera_mi355x.so in-tree: every .hip source under csrc/ compiled for gfx950 by hipcc.

Usage: python -m tessera_amd.build   (or tessera_amd.build.build() from Python)
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtessera_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # No packed FP32 instructions anywhere (the library does integer work; this keeps the
    # compiler from forming v_pk_* FP32 ops in any float code).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(os.path.dirname(HERE), "include"),
]


# Host code only under AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize.sh): each
# sanitizer flag applies to the host pass, the device code is built as always.
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer",
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-shared-libasan"]


def _compile(src, out_dir=OUT_DIR, extra=()):
    obj = os.path.join(out_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, *extra, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False, sanitize=False):
    """The library, or with sanitize=True its ASan/UBSan host build at _san/libtessera_mi355x.so."""
    out_dir = os.path.join(HERE, "_san", "obj") if sanitize else OUT_DIR
    lib_path = os.path.join(HERE, "_san", "libtessera_mi355x.so") if sanitize else LIB
    extra = SAN_FLAGS if sanitize else ()
    os.makedirs(out_dir, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(lambda s: _compile(s, out_dir, extra), sources))
    tmp = lib_path + ".tmp"
    link_extra = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-shared-libasan"] if sanitize else []
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *link_extra, "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib_path)
    if verbose:
        print("built", lib_path)
    return lib_path


if __name__ == "__main__":
    build(verbose=True, sanitize="--sanitize" in sys.argv)
    sys.exit(0)
