"""Build the native parts of mrtrix3_amd in-tree.

Products (all under mrtrix3_amd/):
  lib/libmrcsd.so   HIP kernels for gfx950 + the C ABI of include/*.h + the C++ host set-up code
  bin/dwi2fod       the command-line drop-in (C++), linked against lib/libmrcsd.so
The CPU oracle under oracle/ is test infrastructure and is built by its own Makefile.

Every HIP translation unit is compiled with `-Xclang -target-feature -Xclang -packed-fp32-ops`
disabled (see HIPFLAGS): packed FP32 instructions are avoided on this target.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + CSRC, "-I" + os.path.join(REPO, "include")]

LIB_SOURCES = [
    "kernels/csd_solve.hip",
    "kernels/icls_solve.hip",
    "kernels/mlp.hip",
    "kernels/peaks.hip",
    "kernels/resp.hip",
    "kernels/fmls.hip",
    "kernels/gt.hip",
    "host/sh.cpp",
    "host/dense.cpp",
    "host/gradient.cpp",
    "host/csd.cpp",
    "host/icls.cpp",
    "host/msmt.cpp",
    "host/fmls.cpp",
    "host/voxel.cpp",
    "host/npz.cpp",
    "host/gt.cpp",
    "capi_common.cpp",
    "capi_csd.cpp",
    "capi_icls.cpp",
    "capi_warm.cpp",
    "capi_peaks.cpp",
    "capi_resp.cpp",
    "capi_fmls.cpp",
    "capi_tune.cpp",
    "capi_gt.cpp",
    "icls_noalloc.hip",
]


def _gen_inc(data, inc, npairs, origin, suffix=""):
    src = os.path.join(PKG, "data", data)
    dst = os.path.join(CSRC, "host", inc)
    vals = []
    for line in open(src):
        line = line.split("#")[0].strip()
        if line:
            vals.extend(line.split())
    if len(vals) != 2 * npairs:
        raise RuntimeError("expected %d [az el] pairs in %s" % (npairs, src))
    text = ("// Generated from mrtrix3_amd/data/%s by mrtrix3_amd/build.py.\n"
            "// [azimuth elevation] pairs, MRtrix3 %s (MPL-2.0).\n" % (data, origin))
    text += "".join("%s%s, %s%s,\n" % (vals[i], suffix, vals[i + 1], suffix) for i in range(0, 2 * npairs, 2))
    if not os.path.exists(dst) or open(dst).read() != text:
        open(dst, "w").write(text)


def _gen_dirs_inc():
    _gen_inc("electrostatic_repulsion_300.txt", "dirs300.inc", 300, "src/dwi/directions/predefined.cpp:92-393")
    # double literals converted to float where the table is declared, as the reference's float array is
    _gen_inc("sh2peaks_seeds_60.txt", "seeds60.inc", 60, "cmd/sh2peaks.cpp:363-424")
    _gen_inc("tesselation_1281.txt", "dirs1281.inc", 1281, "src/dwi/directions/predefined.cpp")


def _compile(rel):
    src = os.path.join(CSRC, rel)
    obj = os.path.join(OBJ, rel.replace("/", "_") + ".o")
    lang = ["-x", "hip"] if rel.endswith(".hip") else []
    cmd = [HIPCC] + HIPFLAGS + lang + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    _gen_dirs_inc()
    jobs = min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    so = os.path.join(LIB, "libmrcsd.so")
    tmp = so + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, so)
    for name, extra in (("dwi2fod", []), ("sh2peaks", []), ("fod2fixel", []),
                        ("amp2response", ["gradient.cpp", "dense.cpp"]),
                        ("dwi2response", ["gradient.cpp", "dense.cpp"])):
        cli = os.path.join(CSRC, name + ".cpp")
        exe = os.path.join(BIN, name)
        host = [os.path.join(CSRC, "host", f) for f in ["mif.cpp", "nifti.cpp"] + extra]
        cmd = [HIPCC] + HIPFLAGS + [cli] + host + ["-o", exe + ".tmp",
                                    "-L" + LIB, "-lmrcsd",
                                    "-Wl,-rpath,$ORIGIN/../lib", "-lpthread", "-lz"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
        os.replace(exe + ".tmp", exe)
    # the test driver of the per-voxel classes (runs on the GPU box, which does not build)
    tdir = os.path.join(REPO, "tests", "native")
    tbin = os.path.join(tdir, "_bin")
    os.makedirs(tbin, exist_ok=True)
    exe = os.path.join(tbin, "voxel_classes")
    cmd = [HIPCC] + HIPFLAGS + [os.path.join(tdir, "voxel_classes.cpp"), "-o", exe + ".tmp", "-L" + LIB,
                                "-lmrcsd", "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIB, tbin), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", so)
    return so


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed\n%s%s" % (r.stdout, r.stderr))


if __name__ == "__main__":
    build(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle()
