"""Build the engine library (libhawaii_gpu.so, gfx950) and the CPU oracle in-tree.

    python -m hawaiihybrid_amd.build          # or __graft_entry__.build()

Objects go to hawaiihybrid_amd/_build/; the shared library lands next to this file.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhawaii_gpu.so")
REFLIB = os.path.join(HERE, "libhawaii.so")   # the reference-named C API (include/hawaii_ref.h)
DRIVER = os.path.join(HERE, "hawaii_gpu")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + CSRC]
# packed FP32 instructions stay off for the device code
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = [
    # the integrators' arithmetic is what their expressions say (a * b + c inside one expression
    # becomes one FMA, nothing across statements), so kernels of different shape (one item per
    # thread, the lane pool, stage vectors in LDS or in the buffer arena) compute the same bits
    ("hh_kernels.hip", DEVICE + ["-ffp-contract=on"]),
    ("hh_aux.hip", DEVICE),
    ("hh_pipnn.hip", DEVICE),
    ("hh_zimm.hip", DEVICE),
    ("hh_vegas.hip", DEVICE),
    ("hh_prmu.hip", DEVICE),
    ("hh_ch4co2.hip", DEVICE),
    ("hh_moments.cpp", ["-x", "hip"] + DEVICE),
    ("hh_host.cpp", ["-x", "hip"] + DEVICE),
    ("hh_pool.cpp", ["-x", "hip"] + DEVICE),
    ("hh_session.cpp", ["-x", "hip"] + DEVICE),
    ("hh_conf.cpp", ["-x", "hip"] + DEVICE),
    # bit-exact initial conditions: no floating-point contraction in the sampler
    ("hh_sampler.cpp", ["-ffp-contract=off"]),
    # PROCESSING numerics (baseline fit, LOESS): the window choice of the smoothing compares
    # distances computed as the reference computes them, so no contraction there either
    ("hh_process.cpp", ["-ffp-contract=off"]),
]


def _compile(src, extra):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build_engine(verbose=True):
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda a: _compile(*a), SOURCES))
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(LIB + ".tmp", LIB)
    # the .conf driver executable, linked against the library next to it
    main_obj = _compile("hh_driver_main.cpp", [])
    cmd = [HIPCC, "-o", DRIVER + ".tmp", main_obj, "-L" + HERE, "-lhawaii_gpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(DRIVER + ".tmp", DRIVER)
    # the reference-named C API, linked against the library next to it
    ref_obj = _compile("hh_refapi.cpp", ["-x", "hip"] + DEVICE)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", REFLIB + ".tmp", ref_obj, "-L" + HERE,
           "-lhawaii_gpu", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(REFLIB + ".tmp", REFLIB)
    if verbose:
        print("built", LIB)
        print("built", DRIVER)
        print("built", REFLIB)
    return LIB


def build_oracle(verbose=True):
    r = subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if os.path.isdir("/root/reference"):
        r = subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "-s", "ref"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("reference checker build failed:\n" + r.stdout + r.stderr)
    if verbose:
        print("built oracle")


def build(verbose=True):
    build_oracle(verbose)
    return build_engine(verbose)


if __name__ == "__main__":
    build()
    sys.exit(0)
