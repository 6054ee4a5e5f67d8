"""Build libmesh_hip.so (gfx950) in-tree with hipcc.

Run as ``python -m mesh_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmesh_hip.so")
SOURCES = ["runtime.hip", "su3_hmc.hip", "landau.hip", "u1.hip", "flow.hip", "pflow.hip", "pcflow.hip", "su2.hip"]
HEADERS = ["mesh_internal.h", "su3_dev.h", "rng_dev.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every multiply-add is an explicit fma(): keep the operation order fixed
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def build(verbose=False, extra=(), out=LIB, tag="default"):
    """Compile every source and link ``out``; ``extra`` are additional hipcc
    flags (e.g. -DMESH_SITE_LB=5 for another site-block length)."""
    from concurrent.futures import ThreadPoolExecutor

    bdir = os.path.join(HERE, "_build", tag)
    os.makedirs(bdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stderr=subprocess.PIPE, text=True)
        # the host pass does not know the GPU-only target feature: drop that note
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        if err.strip():
            print(err, file=sys.stderr)
        if r.returncode != 0:
            raise subprocess.CalledProcessError(r.returncode, cmd)
        return obj

    # the sources are independent translation units: compile them side by side
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    tmp = out + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], check=True)
    os.replace(tmp, out)
    if out == LIB:
        build_apps(verbose)
    return out


APPS = {"mesh_hmc": "hmc.cpp", "mesh_landau": "landau.cpp", "mesh_sweep_beta": "sweep_beta.cpp",
        "test_lattice_gauge": "test_lattice_gauge.cpp", "test_io": "test_io.cpp"}
# programs with device code of their own (the generic lattice operations of
# mesh/lattice_apply.h), compiled by hipcc
HIP_APPS = {"test_lattice_apply": "test_lattice_apply.hip"}


def build_apps(verbose=False):
    """Host programs over the C ABI (g++, no HIP headers): the C++ call
    surface of mesh/gauge.h exercised by a driver and a test program."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    root = os.path.dirname(HERE)
    for exe, src in APPS.items():
        cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{root}/include", f"-I{HERE}/host",
               os.path.join(HERE, "apps", src), f"-L{HERE}", "-lmesh_hip", "-Wl,-rpath,$ORIGIN/..",
               "-Wl,-rpath-link,/opt/rocm/lib", "-o", os.path.join(bindir, exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    for exe, src in HIP_APPS.items():
        cmd = [HIPCC, *FLAGS, "-std=c++20", f"-I{root}/include", f"-I{HERE}/host", os.path.join(HERE, "apps", src),
               f"-L{HERE}", "-lmesh_hip", "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(bindir, exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stderr=subprocess.PIPE, text=True)
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        if err.strip():
            print(err, file=sys.stderr)
        if r.returncode != 0:
            raise subprocess.CalledProcessError(r.returncode, cmd)


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
