"""Build the gfx950 HIP library and the Fortran host program in-tree.

libses3d_hip.so   hipcc --offload-arch=gfx950, one translation unit
                  (csrc/s3d_api.hip includes csrc/s3d_kernels.hip)
bin/copy_probe, bin/contraction_probe
                  stand-alone measurement programs (HBM streams; the
                  derivative contraction on the VALU and on the matrix cores)
bin/ses3d_nt      flang: the Fortran host (fortran/*.f90) linked against the
                  library through ISO_C_BINDING
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]

FORTRAN_SOURCES = ["s3d_c_binding.f90", "gll_mod.f90", "ses3d_config.f90", "earth_models.f90",
                   "sac_io.f90", "channels.f90", "evolution_mod.f90",
                   "snapshots.f90", "grad.f90", "ses3d_nt.f90"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s" % " ".join(cmd))


def build_hip():
    src = os.path.join(PKG, "csrc", "s3d_api.hip")
    out = os.path.join(PKG, "libses3d_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + [src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_probe(name):
    """A stand-alone measurement program csrc/<name>.hip -> bin/<name>."""
    src = os.path.join(PKG, "csrc", name + ".hip")
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, name)
    flags = [f for f in HIP_FLAGS if f not in ("-shared", "-fPIC")]
    _run([HIPCC] + flags + [src, "-o", exe])
    return exe


def build_copy_probe():
    return build_probe("copy_probe")


def build_fortran():
    fdir = os.path.join(PKG, "fortran")
    srcs = [os.path.join(fdir, f) for f in FORTRAN_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    if not os.path.exists(FLANG):
        print("flang not found; Fortran host not built", file=sys.stderr)
        return None
    bdir = os.path.join(PKG, "_build", "mod")
    os.makedirs(bdir, exist_ok=True)
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, "ses3d_nt")
    objs = []
    for s in srcs:
        o = os.path.join(bdir, os.path.basename(s).replace(".f90", ".o"))
        # the built-in models keep default-kind literals on purpose
        _run([FLANG, "-O2", "-Wno-real-constant-widening", "-module-dir", bdir, "-I", bdir,
              "-c", s, "-o", o])
        objs.append(o)
    _run([FLANG, "-O2"] + objs + ["-L" + PKG, "-lses3d_hip", "-Wl,-rpath," + PKG,
                                  "-o", exe])
    return exe


def build_all():
    lib = build_hip()
    build_copy_probe()
    build_probe("contraction_probe")
    exe = build_fortran()
    return lib, exe


if __name__ == "__main__":
    build_all()
