"""Build libsmg.so (gfx950 kernels + host engine + C ABI) in-tree with hipcc."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib", "libsmg.so")

SOURCES = ["smg_kernels.hip", "smg_vor3d.hip", "smg_lset_dev.hip", "smg_engine.cpp", "smg_api.cpp", "smg_lset.cpp", "smg_server.cpp",
           "smg_tri.cpp", "smg_tet.cpp", "smg_rst.cpp", "smg_io.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    """Compile every source and link the shared library; raises on failure."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    obj_dir = os.path.join(HERE, "lib", "obj")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(src)[0] + ".o") for src in SOURCES]

    def compile_one(k):
        cmd = [hipcc] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC,
                                       "-c", os.path.join(CSRC, SOURCES[k]), "-o", objs[k]]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    # the translation units are independent: compile them side by side
    workers = max(1, min(len(SOURCES), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(compile_one, k) for k in range(len(SOURCES))]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


DRIVER = os.path.join(HERE, "lib", "smg_run")


def build_driver(verbose=False):
    """The C++ host driver (host/smg_run.cpp over the C ABI), plain g++, linked to
    lib/libsmg.so with an rpath so that it runs from the tree."""
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host"),
           os.path.join(HERE, "host", "smg_run.cpp"), "-o", DRIVER + ".tmp",
           "-L", os.path.dirname(LIB), "-lsmg", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
