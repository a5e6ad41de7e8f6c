"""Build libesf.so (HIP kernels for gfx950 + the C ABI + host scene code)
and the headless driver esf_run, in-tree, with hipcc.

    python -m eulersf_amd.build        # or: python eulersf_amd/build.py

Every translation unit is compiled for gfx950 only.  FP64 kernels mirror the
reference's left-to-right arithmetic, so floating-point contraction is off;
packed-FP32 instruction selection is disabled as well.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libesf.so")
RUNNER = os.path.join(HERE, "esf_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
SOURCES = [
    os.path.join(CSRC, "esf_pressure.hip"),
    os.path.join(CSRC, "esf_solid.hip"),
    os.path.join(CSRC, "esf_grid.hip"),
    os.path.join(CSRC, "esf_mesh.hip"),
    os.path.join(CSRC, "esf_sweep.hip"),
    os.path.join(CSRC, "esf_flip.hip"),
    os.path.join(CSRC, "esf_api.hip"),
    os.path.join(CSRC, "esf_batch.hip"),
    os.path.join(HOST, "esf_scene.cpp"),
    os.path.join(HOST, "esf_frames.cpp"),
]


# host-only sources: the fiber switch is x86-64 assembly, kept out of the
# offload compiler's device pass
HOST_ONLY = [os.path.join(HOST, "esf_fiber.cpp")]
GXX = os.environ.get("CXX", "g++")


def _compile_host(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [GXX, "-O2", "-std=c++17", "-fPIC", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _compile(src):
    if src in HOST_ONLY:
        return _compile_host(src)
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC, *FLAGS, *lang, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES + HOST_ONLY))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-lz", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cmd = [HIPCC, *FLAGS, "-x", "hip", os.path.join(HOST, "esf_run.cpp"), "-o", RUNNER,
           f"-L{HERE}", "-lesf", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} and {RUNNER}")
    return LIB


if __name__ == "__main__":
    try:
        build(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
