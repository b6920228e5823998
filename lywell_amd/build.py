"""Build recipe for the lywell shared libraries (in-tree, gfx950).

Produces
  lywell_amd/liblywell.so      HIP kernels + C ABI (include/lywell.h)
  lywell_amd/liblywell_f.so    Fortran modules modMLFMA / modIterativo / modRCS /
                               modMonostatica
                               (ISO_C_BINDING shims over the C ABI)
  lywell_amd/lw_solve          Fortran host program driving the shims (testMLFMA)
  lywell_amd/lw_rcsmono        Fortran host program of the monostatic sweep
and, as test infrastructure, oracle/liboracle.so (the C oracle) and, when the
reference sources are present, oracle/_ref/ref_driver.

Run:  python -m lywell_amd.build   (or python lywell_amd/build.py)
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lywell_amd")
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(ROOT, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "lib", "llvm", "bin", "flang")
ARCH = "gfx950"

# hipcc flags: gfx950 only; packed FP32 instructions disabled (they are not
# trusted on this platform) - the host pass warns that it does not know the
# feature, which is harmless.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-result",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip():
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    units = [
        ("lw_device.hip", []),
        ("lw_batch.hip", []),
        ("lw_mlp.hip", []),
        ("lw_rcs_static.hip", []),
        # the two triangular-solve kernels sum rows in the same order: no
        # contraction, so that they give the same bits
        ("lw_lusol.hip", ["-ffp-contract=off"]),
        # bit-exact elimination order and roundings of ilu0C: no contraction
        ("lw_ilu0.hip", ["-ffp-contract=off"]),
        # bit-exact Legendre sums: no contraction into FMAs in this unit
        ("lw_setup_kernels.hip", ["-ffp-contract=off"]),
        ("lw_plan.cpp", ["-x", "hip"]),
        ("lw_mesh.cpp", ["-x", "hip"]),
        ("lw_dat.cpp", ["-x", "hip"]),
        ("lw_sweep.cpp", ["-x", "hip"]),
    ]
    for src, extra in units:
        obj = os.path.join(BUILD, src + ".o")
        run([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(PKG, "liblywell.so")
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out] + objs)
    return out


def build_fortran():
    """Fortran modules with the reference's public names over the C ABI."""
    if not os.path.exists(FLANG):
        raise RuntimeError("flang not found at " + FLANG)
    fdir = os.path.join(BUILD, "fortran")
    os.makedirs(fdir, exist_ok=True)
    srcs = ["modGlobalParam.f90", "lw_capi.f90", "modMLFMA.f90", "modIterativo.f90", "modRCS.f90",
            "modMonostatica.f90"]
    objs = []
    for s in srcs:
        obj = os.path.join(fdir, s + ".o")
        run([FLANG, "-O2", "-fPIC", "-J", fdir, "-c", os.path.join(FSRC, s), "-o", obj])
        objs.append(obj)
    lib = os.path.join(PKG, "liblywell_f.so")
    run([FLANG, "-shared", "-fPIC", "-o", lib] + objs + ["-L" + PKG, "-llywell", "-Wl,-rpath,$ORIGIN"])
    for prog in ("lw_solve", "lw_rcsmono"):
        drv_obj = os.path.join(fdir, prog + ".f90.o")
        run([FLANG, "-O2", "-J", fdir, "-c", os.path.join(FSRC, prog + ".f90"), "-o", drv_obj])
        exe = os.path.join(PKG, prog)
        run([FLANG, "-o", exe, drv_obj, "-L" + PKG, "-llywell_f", "-llywell", "-Wl,-rpath,$ORIGIN"])
    return lib


def build_oracle():
    """C oracle (test infrastructure) and, if the sources exist, the reference."""
    odir = os.path.join(ROOT, "oracle")
    run(["make", "-C", odir, "liboracle.so"])
    ref = os.environ.get("LYWELL_REFERENCE", "/root/reference")
    if os.path.exists(os.path.join(ref, "modMLFMA.f90")):
        run(["bash", os.path.join(odir, "build_ref.sh")])


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_hip()
    if "--no-fortran" not in argv:
        build_fortran()
    if "--no-oracle" not in argv:
        build_oracle()


if __name__ == "__main__":
    main()
