"""Build the MI355X library and the FP64 oracle in-tree.

    python -m blade_amd.build            # everything
    python -m blade_amd.build --lib      # libblade_mi355x.so only

The library is compiled with hipcc for gfx950 only; the oracle (test
infrastructure) with g++ and OpenMP.  Outputs:
    blade_amd/libblade_mi355x.so
    oracle/liboracle.so
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libblade_mi355x.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIP_SOURCES = ["domdec.hip", "nbdirect.hip", "pme.hip", "fft_mfma.hip", "bonded.hip", "update.hip", "runtime.hip"]
CXX_SOURCES = ["setup.cpp", "api.cpp", "files.cpp", "autotune.cpp", "xtc.cpp", "script.cpp"]

# gfx950 only; the packed FP32 instructions (v_pk_*_f32) are disabled: some
# of their forms have given wrong results on this hardware without an error
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmds.append([cc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        cmds.append([cc] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-C", ORACLE_DIR, "-B", "liboracle.so"])
    # the reference's xdrfile library, the checker of the XTC writer, where
    # the reference tree is present
    if os.path.isdir("/root/reference/src/xdr"):
        _run(["make", "-C", ORACLE_DIR, "-B", "_ref/libxdrfile.so"])
    return os.path.join(ORACLE_DIR, "liboracle.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    both = not (a.lib or a.oracle)
    if a.lib or both:
        print("built", build_lib(a.j))
    if a.oracle or both:
        print("built", build_oracle())


if __name__ == "__main__":
    sys.exit(main())
