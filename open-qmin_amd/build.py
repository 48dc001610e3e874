"""Build libqmin.so (gfx950) in-tree: regenerate the stencil code, compile each
HIP translation unit with hipcc in parallel, link one shared library."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build", "obj")
LIB = os.path.join(HERE, "lib", "libqmin.so")
DRIVER = os.path.join(HERE, "lib", "openQmin-hip")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
SOURCES = ["ldg_kernels.hip", "step_kernels.hip", "fire_stream.hip", "qmin_capi.hip", "minimizers.hip", "ensemble.hip", "warmstart.hip"]
# host code that must match the reference's g++ build bit for bit
HOST_SOURCES = ["host_init.cpp"]
GXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + CSRC]


def run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    subprocess.run([sys.executable, os.path.join(HERE, "stencils", "gen_stencils.py")], check=True,
                   stdout=subprocess.DEVNULL)
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        out = run([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    def compile_host(src):
        obj = os.path.join(OBJ, src.replace(".cpp", ".o"))
        run([GXX, "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-c", os.path.join(CSRC, src),
             "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES) + len(HOST_SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES)) + list(ex.map(compile_host, HOST_SOURCES))
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the openQmin-style C++ driver over the drop-in classes
    run([GXX, "-std=c++17", "-O2", "-Wall", "-I" + INCLUDE, "-I" + os.path.join(HERE, "host"),
         os.path.join(HERE, "host", "openqmin_hip.cpp"), "-o", DRIVER + ".tmp", "-L" + os.path.dirname(LIB),
         "-lqmin", "-Wl,-rpath,$ORIGIN"])
    os.replace(DRIVER + ".tmp", DRIVER)
    # reference-style GPUArray / ArrayHandle use of the drop-in classes
    example = os.path.join(os.path.dirname(LIB), "array-views-example")
    run([GXX, "-std=c++17", "-O2", "-Wall", "-I" + INCLUDE, "-I" + os.path.join(HERE, "host"),
         os.path.join(HERE, "host", "array_views_example.cpp"), "-o", example + ".tmp", "-L" + os.path.dirname(LIB),
         "-lqmin", "-Wl,-rpath,$ORIGIN"])
    os.replace(example + ".tmp", example)
    # streaming-ceiling measurement for the fused pass's access pattern
    ceiling = os.path.join(os.path.dirname(LIB), "qmin-stream-ceiling")
    run([HIPCC] + FLAGS + [os.path.join(CSRC, "stream_ceiling.hip"), "-o", ceiling + ".tmp"])
    os.replace(ceiling + ".tmp", ceiling)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
