"""Build libedgeto_amd.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python edge-to_amd/build.py [-j N]
The library is written next to the Python package (edgeto_amd/libedgeto_amd.so).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "edgeto_amd")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libedgeto_amd.so")

SOURCES = ["stiffness.hip", "filter.hip", "mg.hip", "coarse.hip", "blas.hip", "oc.hip",
           "solver.hip", "op_tiled.hip", "galerkin_elem.hip", "op_mfma.hip", "op_matrix.hip", "op_wht.hip", "level1.hip", "k0_structure.cpp", "fea.hip", "krylov.hip", "batch.cpp", "multimat.hip", "unstructured.hip", "capi.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          # packed FP32 instructions stay off: their results are not trusted on this hardware
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
          "-I" + os.path.join(HERE, "..", "include")]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-x", "hip", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I" + os.path.join(HERE, "..", "include"), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, r.stdout + r.stderr, obj


def build(jobs=8, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    failed = False
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, out, obj in ex.map(_compile, srcs):
            # hipcc's host pass does not know the packed-fp32 target feature; that warning is benign
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if rc != 0:
                failed = True
                sys.stderr.write(f"[build] {src} FAILED\n" + "\n".join(lines) + "\n")
            elif verbose and lines:
                sys.stderr.write(f"[build] {src}\n" + "\n".join(lines) + "\n")
            objs.append(obj)
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lrocsolver", "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
