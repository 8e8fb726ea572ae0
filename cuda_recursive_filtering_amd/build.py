"""Build libiirx.so (HIP kernels for gfx950 + host C ABI) in-tree.

    python -m cuda_recursive_filtering_amd.build      (or build() from Python)

hipcc compiles every .hip / .cpp under csrc/ for --offload-arch=gfx950 in
parallel, then links one shared library next to this file.  The compiler's
kernel resource report (VGPR/SGPR/spill/LDS/occupancy) is kept in
build/resource_usage.txt.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "libiirx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IIRX_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed fp32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-Wall", "-Wno-unused-result",
]


# bs_setup.hip reproduces the host FP64 precompute bit for bit: no fused
# multiply-add may be formed there (a later -ffp-contract wins)
FILE_FLAGS = {"bs_setup.hip": ["-ffp-contract=off"]}


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    extra = FILE_FLAGS.get(os.path.basename(src), [])
    cmd = [HIPCC] + COMMON + extra + ["-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return obj, p.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs, reports = [], []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for obj, rep in ex.map(_compile, srcs):
            objs.append(obj)
            reports.append(rep)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    with open(os.path.join(OBJ, "resource_usage.txt"), "w") as f:
        f.write("\n".join(reports))
    # the C++ driver of the reference's test_accuracy / test_performance, linked to the library
    drv = os.path.join(HERE, "iirx_driver")
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-Wno-unused-result", "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(HERE, "harness", "iirx_driver.cpp"), "-L", HERE, "-liirx", "-Wl,-rpath,$ORIGIN",
           "-o", drv + ".tmp"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"driver build failed\n{' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    os.replace(drv + ".tmp", drv)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
