"""Build the gfx950 HIP library and the Fortran host program in-tree.

libses3d_hip.so   hipcc --offload-arch=gfx950: csrc/s3d_api.hip (the C ABI and
                  every kernel but the element kernel, csrc/s3d_kernels.hip)
                  and csrc/s3d_elem.hip once per GLL order 2..7 (the element
                  kernel's instantiations), compiled in parallel and linked
bin/copy_probe, bin/contraction_probe
                  stand-alone measurement programs (HBM streams; the
                  derivative contraction on the VALU and on the matrix cores)
bin/ses3d_nt      flang: the Fortran host (fortran/*.f90) linked against the
                  library through ISO_C_BINDING
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

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


ELEM_ORDERS = range(2, 8)   # GLL orders with element-kernel instantiations


def _run_parallel(cmds):
    """Run the commands concurrently (at most MAX_JOBS or the CPU count at
    once); raise if any fails, after all have ended."""
    def one(cmd):
        print("+", " ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        return cmd, r
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    failed = []
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for cmd, r in ex.map(one, cmds):
            out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
            if out.strip():
                print(out)
            if r.returncode != 0:
                failed.append(" ".join(cmd))
    if failed:
        raise RuntimeError("commands failed: %s" % "; ".join(failed))


def build_hip():
    csrc = os.path.join(PKG, "csrc")
    odir = os.path.join(PKG, "_build", "hip")
    os.makedirs(odir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(odir, "s3d_api.o")]
    cmds = [[HIPCC] + cflags + ["-c", os.path.join(csrc, "s3d_api.hip"), "-o", objs[0]]]
    for p in ELEM_ORDERS:
        o = os.path.join(odir, "s3d_elem_p%d.o" % p)
        cmds.append([HIPCC] + cflags + ["-DS3D_ELEM_P=%d" % p, "-c", os.path.join(csrc, "s3d_elem.hip"),
                                        "-o", o])
        objs.append(o)
    for o in objs:   # no stale object can be linked if a compile fails
        if os.path.exists(o):
            os.remove(o)
    _run_parallel(cmds)
    out = os.path.join(PKG, "libses3d_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp])
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
