"""Builds libdgamd.so (HIP kernels + C ABI, gfx950 only) in this directory.

    python build.py            # compile every csrc/*.hip and link
    python build.py --clean    # remove the build products first

Every HIP source is compiled by hipcc for --offload-arch=gfx950 with the
packed-FP32 target feature disabled; objects go to _build/ and the shared
library to ./libdgamd.so.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdgamd.so")
HOSTLIB = os.path.join(HERE, "libdghost.so")
APPS = os.path.join(HERE, "apps")
BIN = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROCM, "include")]
HOST_LIBS = ["-L" + HERE, "-ldgamd", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
             "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless.
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


# Kernels that feed LDS straight from global memory count their own
# outstanding loads and place their waits by hand; a register spill would put
# scratch traffic between those loads and the waits (and has gone with wrong
# results), so any spill in them fails the build.
NO_SPILL = ("_dma_", "crt_gemm_kernel", "oz_plane_kernel")


def _check_spills(src, remarks):
    name = None
    for line in remarks.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:")[1].split("[")[0].strip()
        elif name and ("VGPRs Spill:" in line or "SGPRs Spill:" in line or "ScratchSize [bytes/lane]:" in line):
            n = int(line.split(":")[-1].split("[")[0].strip())
            if n and any(t in name for t in NO_SPILL):
                raise RuntimeError("build failed: %s spills (%s) in %s" % (name, line.split("remark:")[-1].strip(), src))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    out = _run([HIPCC] + HIP_FLAGS + ["-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj])
    _check_spills(src, out)
    return obj


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def host_sources():
    if not os.path.isdir(HOST):
        return []
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))


def build(verbose=False, jobs=None):
    """Compiles every HIP source and links libdgamd.so (and the C++ host API
    library libdghost.so on top of it).  Always recompiles."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    hsrcs = host_sources()
    if hsrcs:
        # The host API is plain C++ over the HIP runtime: no device code.
        tmp = HOSTLIB + ".tmp"
        _run(["g++"] + HOST_FLAGS + ["-shared", "-o", tmp] + hsrcs + HOST_LIBS)
        os.replace(tmp, HOSTLIB)
        # Command-line drivers and the host-API self-test.
        os.makedirs(BIN, exist_ok=True)
        for app in sorted(f for f in os.listdir(APPS) if f.endswith(".cpp")):
            exe = os.path.join(BIN, app[:-4])
            _run(["g++"] + HOST_FLAGS + ["-o", exe + ".tmp", os.path.join(APPS, app), "-L" + HERE, "-ldghost"]
                 + HOST_LIBS + ["-Wl,-rpath,$ORIGIN/.."])
            os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    if "--clean" in sys.argv:
        shutil.rmtree(OBJ, ignore_errors=True)
        for f in (LIB, HOSTLIB):
            if os.path.exists(f):
                os.remove(f)
    build(verbose=True)
