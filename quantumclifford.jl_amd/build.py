"""Build libqcf.so (HIP, gfx950) in-tree.

Every HIP/C++ source under csrc/ is compiled with hipcc for gfx950 into
_build/ and linked into qcf/libqcf.so, next to the Python host package that
loads it.  Run `python quantumclifford.jl_amd/build.py` or call build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "qcf", "libqcf.so")
ARCH = os.environ.get("QCF_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 instructions are switched off for every device compile.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]


def _sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


GEN = os.path.join(BUILD, "gen")


def _embed_device_header() -> None:
    """qcf_device.h as a C++ raw string literal, for run-time compilation."""
    os.makedirs(GEN, exist_ok=True)
    text = open(os.path.join(CSRC, "qcf_device.h")).read()
    assert ")QCFDEV\"" not in text
    out = 'R"QCFDEV(' + text + ')QCFDEV"\n'
    path = os.path.join(GEN, "qcf_device_src.inc")
    if not os.path.exists(path) or open(path).read() != out:
        with open(path, "w") as f:
            f.write(out)


# Host-only sanitizer build (build(sanitize=True)): AddressSanitizer and
# UndefinedBehaviorSanitizer on the host code of the C++ sources (the
# circuit compiler, the matching decoder, the kernel generator, the runtime),
# each -fsanitize= after -Xarch_host so that no device code is built with it.
SAN_OUT = os.path.join(HERE, "qcf", "libqcf_san.so")
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all",
             "-g"]


def _compile(src: str, sanitize: bool = False) -> str:
    san = sanitize and src.endswith(".cpp")
    obj = os.path.join(BUILD, os.path.basename(src) + (".san.o" if san else ".o"))
    cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE_FLAGS, *(SAN_FLAGS if san else []), f"-I{GEN}",
           "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False, sanitize: bool = False) -> str:
    """libqcf.so, or with `sanitize` qcf/libqcf_san.so: the same library with
    the host C++ sanitized, loaded instead of libqcf.so when QCF_LIBRARY
    names it and the sanitizer runtime is preloaded (CPU test runs only)."""
    os.makedirs(BUILD, exist_ok=True)
    _embed_device_header()
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, sanitize), srcs))
    out = SAN_OUT if sanitize else OUT
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           *(SAN_FLAGS + ["-Xarch_host", "-shared-libsan"] if sanitize else []),
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_oracle(verbose: bool = False) -> str:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    root = os.path.dirname(HERE)
    odir = os.path.join(root, "oracle")
    r = subprocess.run(["make", "-C", odir, "-B", "liboracle.so"], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {os.path.join(odir, 'liboracle.so')}")
    return os.path.join(odir, "liboracle.so")


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        build(verbose=True, sanitize=True)
        sys.exit(0)
    build(verbose=True)
    if "--no-oracle" not in sys.argv:
        build_oracle(verbose=True)
