"""Build recipes for libhtx.so (HIP, gfx950) and the CPU oracle libraries.

Everything builds in-tree with plain hipcc / gcc command lines:

  hermite-teno_amd/htx/libhtx.so     the product: HIP kernels + C ABI
  oracle/liboracle.so                the FP64 checker (-O2, no FP contraction)
  oracle/liboracle_omp.so            the same source, OpenMP, -O3, for the
                                     CPU baseline timing
  oracle/liboracle_count.so          the same source as C++ with an
                                     op-counting double (oracle/flop_count.h),
                                     for the flop count of the path
  tests/host/libfv2d_host.so         CPU build of the stage kernel's tiling
                                     (test infrastructure)
  tests/host/libbatch_host.so        CPU walk of the batched launches'
                                     bookkeeping (test infrastructure)
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(PKG_DIR))
CSRC = os.path.join(os.path.dirname(PKG_DIR), "csrc")
ORACLE = os.path.join(ROOT, "oracle")

LIBHTX = os.path.join(PKG_DIR, "libhtx.so")
LIBORACLE = os.path.join(ORACLE, "liboracle.so")
LIBORACLE_OMP = os.path.join(ORACLE, "liboracle_omp.so")
LIBORACLE_COUNT = os.path.join(ORACLE, "liboracle_count.so")

HIP_SOURCES = ["htx_fv2d.hip", "htx_fvh2d.hip", "htx_strict.hip", "htx_strict_fvh.hip", "htx_ic.hip", "htx_learn.hip",
               "htx_batch.hip", "htx_api.hip", "htx_noalloc.hip", "htx_classes.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # contract a*b+c into FMA, except where a function asks not to (the
    # reconstruction chain, see HTX_NO_CONTRACT in htx_numerics.h)
    "-ffp-contract=fast-honor-pragmas",
    # packed FP32 instructions stay off: some of their forms have produced
    # silently wrong results on this hardware pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass does not know the packed-fp32 target feature; that warning is harmless
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip() != "1 warning generated.")
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


# per-file flags: the strict build of the kernels turns contraction off
# (the classifier's matrix-core products write ordinary VGPRs, which spares
# the accumulator-register reads before the GELU)
PER_FILE_FLAGS = {"htx_strict.hip": ["-ffp-contract=off"], "htx_ic.hip": ["-ffp-contract=off"],
                  "htx_strict_fvh.hip": ["-ffp-contract=off", "-mllvm", "-disable-machine-licm"],
                  "htx_classes.hip": ["-ffp-contract=off"],
                  "htx_learn.hip": ["-ffp-contract=off", "-mllvm", "-amdgpu-mfma-vgpr-form"]}


def build_htx(extra=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        flags = [f for f in HIPCC_FLAGS if f != "-shared"] + PER_FILE_FLAGS.get(src, [])
        cmds.append([hipcc, *flags, *extra, "-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    # the translation units are independent: compile them side by side
    from concurrent.futures import ThreadPoolExecutor

    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIBHTX + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIBHTX)
    return LIBHTX


def build_oracle():
    src = os.path.join(ORACLE, "htx_oracle.c")
    _run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", "-o", LIBORACLE, src, "-lm"])
    _run(["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-o", LIBORACLE_OMP, src,
          "-lm"])
    _run(["g++", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-std=c++17", "-include",
          os.path.join(ORACLE, "flop_count.h"), "-x", "c++", "-o", LIBORACLE_COUNT, src])
    return LIBORACLE


HOST_HARNESS = os.path.join(ROOT, "tests", "host", "libfv2d_host.so")
HOST_HARNESS_FAST = os.path.join(ROOT, "tests", "host", "libfv2d_host_fast.so")
HOST_HARNESS_HERMITE = os.path.join(ROOT, "tests", "host", "libfvh2d_host.so")
HOST_HARNESS_HERMITE_FAST = os.path.join(ROOT, "tests", "host", "libfvh2d_host_fast.so")
HOST_HARNESS_1D = os.path.join(ROOT, "tests", "host", "libfv1d_host.so")
HOST_HARNESS_BATCH = os.path.join(ROOT, "tests", "host", "libbatch_host.so")
HOST_HARNESS_CONTRACT = os.path.join(ROOT, "tests", "host", "libfv2d_host_contract.so")


def build_host_harness():
    """CPU builds of the stage kernel's tiling (tests/host/fv2d_host.hip), with
    contraction off: one with the reference's division-by-division evaluation
    (compared bit for bit with the oracle) and one with the product's
    strength-reduced divisions (compared within tolerance)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    src = os.path.join(ROOT, "tests", "host", "fv2d_host.hip")
    base = [hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off"]
    _run(base + ["-DHTX_EXACT_DIVISION=1", "-o", HOST_HARNESS, src])
    _run(base + ["-o", HOST_HARNESS_FAST, src])
    srch = os.path.join(ROOT, "tests", "host", "fvh2d_host.hip")
    _run(base + ["-DHTX_EXACT_DIVISION=1", "-o", HOST_HARNESS_HERMITE, srch])
    _run(base + ["-o", HOST_HARNESS_HERMITE_FAST, srch])
    _run(base + ["-DHTX_EXACT_DIVISION=1", "-o", HOST_HARNESS_1D, os.path.join(ROOT, "tests", "host", "fv1d_host.hip")])
    # the product's contraction setting, with hardware FMA on the host: the
    # learned path's features must not depend on it
    _run([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=fast-honor-pragmas", "-mfma",
          "-o", HOST_HARNESS_CONTRACT, src])
    srcb = os.path.join(ROOT, "tests", "host", "batch_host.hip")
    _run(base + ["-o", HOST_HARNESS_BATCH, srcb])
    return HOST_HARNESS


EXAMPLES = os.path.join(os.path.dirname(PKG_DIR), "examples")
EXAMPLE_BIN = os.path.join(EXAMPLES, "build", "main_fv2d")


CLASSES_CHECK = os.path.join(ROOT, "tests", "host", "classes_check")


def build_examples():
    """The main.cpp-style C++ driver over the facade headers (C++11, g++),
    and the test driver of the physics and scheme classes."""
    os.makedirs(os.path.dirname(EXAMPLE_BIN), exist_ok=True)
    _run(["g++", "-std=c++11", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", CLASSES_CHECK,
          os.path.join(ROOT, "tests", "host", "classes_check.cpp"), "-L", PKG_DIR, "-lhtx", "-Wl,-rpath," + PKG_DIR])
    _run(["g++", "-std=c++11", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", EXAMPLE_BIN,
          os.path.join(EXAMPLES, "main_fv2d.cpp"), "-L", PKG_DIR, "-lhtx", "-Wl,-rpath," + PKG_DIR,
          "-Wl,-rpath,$ORIGIN/../../htx"])
    return EXAMPLE_BIN


GPU_HARNESS_DIV = os.path.join(ROOT, "tests", "gpu_harness", "libdivcheck.so")


def build_gpu_harness():
    """GPU test harnesses (tests/gpu_harness): the shared-reciprocal division
    check, built like the product's strict kernels."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    flags = [f for f in HIPCC_FLAGS if not f.startswith("-ffp-contract")] + ["-ffp-contract=off"]
    _run([hipcc, *flags, "-o", GPU_HARNESS_DIV, os.path.join(ROOT, "tests", "gpu_harness", "divcheck.hip")])
    return GPU_HARNESS_DIV


def build_all():
    build_oracle()
    build_host_harness()
    build_gpu_harness()
    build_htx()
    build_examples()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "htx": build_htx, "oracle": build_oracle, "harness": build_host_harness,
         "gpu_harness": build_gpu_harness,
         "examples": build_examples}[w]()
