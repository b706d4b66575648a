"""Builds libdgv_amd.so (all HIP sources of csrc/) in-tree for gfx950."""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdgv_amd.so")
ARCH = os.environ.get("DGV_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep packed FP32 instructions out of every kernel.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
            "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-lambda-capture",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value", "-munsafe-fp-atomics"]


# Per-file additions.  The matrix-core kernels keep their sums in ordinary VGPRs
# (the batched kernel folds them on the VALU every few steps and reads them
# without v_accvgpr_read moves).
FILE_FLAGS = {"dgv_batch.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
              "dgv_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    procs = []
    for s in srcs:
        o = os.path.join(bdir, os.path.basename(s) + ".o")
        cmd = [hipcc(), *HIPFLAGS, *FILE_FLAGS.get(os.path.basename(s), []), "-x", "hip", "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(o)
    failed = []
    for s, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((s, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for s, out in failed:
            sys.stderr.write(f"--- {s}\n{out}\n")
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    subprocess.run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


SANITIZE_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


def build_sanitized(outdir: str, verbose: bool = False) -> str:
    """A copy of libdgv_amd.so whose host code is built with AddressSanitizer and
    UndefinedBehaviorSanitizer (device code unchanged), into outdir.  Loaded by
    the host-side tests through DGV_LIB_PATH with the sanitizer runtime preloaded."""
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    os.makedirs(outdir, exist_ok=True)
    procs, objs = [], []
    for s in srcs:
        o = os.path.join(outdir, os.path.basename(s) + ".o")
        cmd = [hipcc(), *HIPFLAGS, *FILE_FLAGS.get(os.path.basename(s), []), *SANITIZE_HOST, "-g", "-x", "hip", "-c",
               s, "-o", o]
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(o)
    failed = [(s, p.communicate()[0].decode(errors="replace")) for s, p in procs]
    failed = [(s, out) for (s, out), (_, p) in zip(failed, procs) if p.returncode != 0]
    if failed:
        for s, out in failed:
            sys.stderr.write(f"--- {s}\n{out}\n")
        raise RuntimeError("sanitized HIP compilation failed")
    lib = os.path.join(outdir, "libdgv_amd.so")
    subprocess.run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *SANITIZE_HOST, "-o", lib, *objs],
                   check=True)
    return lib


def amdflang():
    for c in (os.environ.get("FC_AMD"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose: bool = False):
    """The iso_c_binding shim (fortran/dgv_gpu.f90) and its example driver."""
    fc = amdflang()
    if fc is None:
        return None
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    obj = os.path.join(bdir, "dgv_gpu.o")
    exe = os.path.join(HERE, "dgv_driver")
    cmds = [[fc, "-O2", f"-J{bdir}", "-c", os.path.join(fdir, "dgv_gpu.f90"), "-o", obj],
            [fc, "-O2", f"-I{bdir}", f"-J{bdir}", os.path.join(fdir, "dgv_driver.f90"), obj,
             f"-L{HERE}", "-ldgv_amd", "-Wl,-rpath,$ORIGIN", "-o", exe]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
