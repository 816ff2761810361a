"""Build the STZ HIP library (gfx950) in-tree.

    python sc25-stz2_amd/build.py          # -> sc25-stz2_amd/lib/libstz_hip.so
                                           #    sc25-stz2_amd/lib/stz_cli
    python sc25-stz2_amd/build.py --sanitize-host
        # -> sc25-stz2_amd/_build/asan/libstz_hip.so: host code (C ABI, codebooks,
        #    planner, header parsing) with AddressSanitizer + UBSan, device code as
        #    usual; for the CPU tests, e.g.
        #    STZ_HIP_LIB=sc25-stz2_amd/_build/asan/libstz_hip.so \
        #    LD_PRELOAD=$(hipcc --print-file-name=libclang_rt.asan-x86_64.so) \
        #    ASAN_OPTIONS=detect_leaks=0 python -m pytest tests/test_abi.py

Device code is built with -ffp-contract=off (no FMA contraction, matching the
reference's x86 build without -march) and -fgpu-flush-denormals-to-zero
(the reference runs with FTZ/DAZ on).  Packed-FP32 instruction selection is
disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
SOURCES = ["k_huff.hip", "k_field.hip", "k_residual.hip", "k_interp.hip", "k_lorenzo.hip", "k_tuner.hip", "k_stage.hip", "k_staged.hip",
           "host_codec.cpp", "stz_api.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("STZ_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fgpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src):
    out = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    so = os.path.join(LIB, "libstz_hip.so")
    tmp = so + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-l:libzstd.so.1", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, so)
    cli_src = os.path.join(CSRC, "stz_cli.cpp")
    if os.path.exists(cli_src):
        cli = os.path.join(LIB, "stz_cli")
        cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), cli_src, "-o", cli + ".tmp",
               "-L" + LIB, "-lstz_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("cli build failed: %s\n%s" % (r.stdout, r.stderr))
        os.replace(cli + ".tmp", cli)
    if verbose:
        print("built", so)
    return so


# host code only: each sanitizer flag applies to the host compilation
HOST_SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-omit-frame-pointer"]


def build_sanitized_host(verbose=False):
    out_dir = os.path.join(OBJ, "asan")
    os.makedirs(out_dir, exist_ok=True)

    def comp(src):
        out = os.path.join(out_dir, src + ".o")
        extra = ["-g", "-O1"] + HOST_SANITIZE if src.endswith(".cpp") else []
        cmd = [HIPCC] + CFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        return out

    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(comp, SOURCES))
    so = os.path.join(out_dir, "libstz_hip.so")
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + HOST_SANITIZE + ["-o", so] + objs + \
        ["-l:libzstd.so.1", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", so)
    return so


if __name__ == "__main__":
    if "--sanitize-host" in sys.argv[1:]:
        build_sanitized_host(verbose=True)
    else:
        build(verbose=True)
    sys.exit(0)
