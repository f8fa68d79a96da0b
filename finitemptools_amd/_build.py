"""Compile the HIP sources into ``libmpsheff.so`` next to this file (gfx950)."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libmpsheff.so"
SOURCES = [
    "mh_runtime.hip",
    "mh_gemm.hip",
    "mh_gemm_mixed.hip",
    "mh_heff.hip",
    "mh_krylov.hip",
    "mh_blocksparse.hip",
    "mh_adjoint.hip",
    "mh_eigen.hip",
    "mh_tridc.hip",
    "mh_batched.hip",
    "mh_heev.hip",
    "mh_svd.hip",
    "mh_bs_planner.cpp",
    "mh_schur.cpp",
]
# rocSOLVER's divide-and-conquer tridiagonal eigensolver (mh_stedc_batched)
LIBS = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=" + ARCH,
    "-fPIC",
    "-shared",
    # Packed FP32 instructions stay off for every kernel of this library.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# MFMA kernels keep their accumulators in VGPRs: with the default heuristic the
# compiler may pick AGPRs and then copy every accumulator AGPR<->VGPR on each
# K-loop iteration.
EXTRA_FLAGS = {
    "mh_gemm.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
    "mh_gemm_mixed.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
    "mh_blocksparse.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
    "mh_batched.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def lib_path():
    return os.path.join(HERE, LIB_NAME)


def build(verbose=False, jobs=None):
    """Compile every source (separately, in parallel) and link the library."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objdir = os.path.join(HERE, "_objs")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc] + [f for f in HIPCC_FLAGS if f != "-shared"] + EXTRA_FLAGS.get(src, []) + [
            "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; harmless.
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
        if p.returncode != 0:
            errors.append(f"{src}:\n" + "\n".join(lines))
        elif verbose and lines:
            print("\n".join(lines))
    if errors:
        raise RuntimeError("HIP compilation failed:\n" + "\n".join(errors))
    tmp = lib_path() + ".tmp"
    cmd = [hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + LIBS
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, lib_path())
    return lib_path()


if __name__ == "__main__":
    print(build(verbose=True))
