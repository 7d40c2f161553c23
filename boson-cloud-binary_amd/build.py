"""Build libgm.so (gfx950 HIP kernels + host solver + C ABI) in-tree.

    python boson-cloud-binary_amd/build.py            # build
    python boson-cloud-binary_amd/build.py --clean    # remove objects first

Objects go to boson-cloud-binary_amd/_build/, the library to
boson-cloud-binary_amd/gm/libgm.so, next to the Python binding that loads it.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "gm", "libgm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["gm_kernels.hip", "gm_theta_f64.hip", "gm_math.cpp", "gm_theta_tables.cpp", "gm_ilut.cpp", "gm_solver.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", CSRC, "-I", os.path.join(REPO, "include")]
# Packed FP32 VALU ops stay off in every device pass (v_pk_* have returned wrong
# results on this hardware); the host pass warns that it does not know the feature.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# matrix-core results in ordinary VGPRs (no accumulator copies in the Gaunt-form
# coupling loop); the polar-quadrature apply keeps its accumulators in AGPRs
DEVICE_EXTRA = {"gm_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def compile_one(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE + DEVICE_EXTRA.get(src, []) + ["-x", "hip"]
    else:
        cmd += ["--offload-host-only"]  # host code: no device pass
    cmd += ["-c", os.path.join(CSRC, src), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in ln]
    if res.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), "\n".join(msgs)))
    return out


def build(clean=False, verbose=True):
    if clean and os.path.isdir(OBJ):
        shutil.rmtree(OBJ)
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), res.stdout + res.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_driver(verbose)
    return LIB


CPP_TESTS = os.path.join(REPO, "tests", "cpp")
# C++ programs over the headers in include/: the gm.hpp class driver, the
# reference's main driver over the drop-in class, the config reader probe and
# the reference's operator checks over the drop-in class
DRIVERS = (("cnm_driver.cpp", True), ("gm_main.cpp", True), ("config_probe.cpp", False), ("physics_ops.cpp", True))


# host-only programs built from the library's own host sources (no GPU)
HOST_DRIVERS = (("theta_tables.cpp", ("gm_theta_tables.cpp", "gm_math.cpp")),)


def build_driver(verbose=True):
    """The C++ test programs of tests/cpp, linked against libgm.so, and the
    host-only ones built from the library's host sources."""
    out_dir = os.path.join(CPP_TESTS, "_bin")
    os.makedirs(out_dir, exist_ok=True)
    for src, deps in HOST_DRIVERS:
        out = os.path.join(out_dir, os.path.splitext(src)[0])
        cmd = [HIPCC, "--offload-host-only", "-O2", "-std=c++17", "-I", CSRC, os.path.join(CPP_TESTS, src)]
        cmd += [os.path.join(CSRC, d) for d in deps] + ["-o", out]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("driver build failed: %s\n%s" % (" ".join(cmd), res.stdout + res.stderr))
        if verbose:
            print("built", out)
    for src, link in DRIVERS:
        out = os.path.join(out_dir, os.path.splitext(src)[0])
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"),
               os.path.join(CPP_TESTS, src), "-o", out]
        if link:
            cmd += ["-L", os.path.dirname(LIB), "-lgm", "-Wl,-rpath,$ORIGIN/../../../boson-cloud-binary_amd/gm"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("driver build failed: %s\n%s" % (" ".join(cmd), res.stdout + res.stderr))
        if verbose:
            print("built", out)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    args = ap.parse_args()
    try:
        build(clean=args.clean)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
