"""Build the MI355X library in-tree: nalu-wind_amd/lib/libnaluwind_amd.so.

HIP sources (csrc/*.hip) are compiled for gfx950 with hipcc; host C++
(host/*.cpp) with g++ -fopenmp; both are linked into one shared library
whose C ABI is declared in include/nalu_amd.h and include/nalu_amd_mesh.h.

    python nalu-wind_amd/build.py [-j N]
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "lib", "libnaluwind_amd.so")
STREAM_LIB = os.path.join(HERE, "lib", "libnw_stream.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-I" + INC, "-I" + os.path.join(HERE, "csrc"),
    "-I" + os.path.join(HERE, "host"),
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-march=x86-64-v2",
             "-I" + INC, "-I/opt/rocm/include", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    hips = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    cpps = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
    cmds = []
    objs = []
    for src in hips:
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", o])
        objs.append(o)
    for src in cpps:
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] +
         objs + ["-lgomp"])
    os.replace(tmp, LIB)
    # the streaming-copy probe the benchmark measures the HBM rate with
    stmp = STREAM_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", stmp,
                                os.path.join(HERE, "stream", "nw_stream.hip")])
    os.replace(stmp, STREAM_LIB)
    # C++ adapter test driver (mirrors the reference unit tests)
    test_src = os.path.join(ROOT, "tests", "cpp", "unit_edge_kernels.cpp")
    if os.path.exists(test_src):
        _run(["g++", "-O2", "-std=c++17", "-I" + INC,
              "-I" + os.path.join(HERE, "host"), test_src, "-o",
              os.path.join(OBJ, "unit_edge_kernels"), "-L" + os.path.dirname(LIB),
              "-lnaluwind_amd", "-Wl,-rpath," + os.path.dirname(LIB)])
    return LIB


if __name__ == "__main__":
    j = None
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    print(build(j, verbose=True))
