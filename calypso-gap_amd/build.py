"""Build the gapx native libraries in-tree.

    python calypso-gap_amd/build.py            # HIP library + Fortran facade + f2py module

Outputs (git-ignored, they travel with the tree):
    calypso-gap_amd/lib/libgapx.so              HIP kernels + C ABI (gfx950)
    calypso-gap_amd/lib/libgapx_fortran.so      Fortran host facade (FGAP_CALC, ...)
    calypso-gap_amd/libgap/libgap*.so           f2py module libgap.libgap

Every HIP file is compiled for gfx950 only, with packed-FP32 instruction
selection disabled.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"

HIP_SOURCES = ["api.hip", "neighbors.hip", "descriptors.hip", "gpr.hip", "gpr_mixed.hip", "bond.hip", "mlp.hip", "relax.hip",
               "variance.hip", "md.hip", "ceilings.hip"]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wno-unused-result", "-Wno-unused-value",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, **kw)


def build_hip(jobs=8):
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    procs = []
    objs = []
    for s in srcs:
        obj = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stderr=subprocess.PIPE, text=True))
    failed = False
    for p, s in zip(procs, srcs):
        _, err = p.communicate()
        err = "\n".join(ln for ln in err.splitlines() if "packed-fp32-ops" not in ln)
        if p.returncode != 0:
            failed = True
            sys.stderr.write(err + "\n")
    if failed:
        raise SystemExit("HIP compilation failed")
    out = os.path.join(LIB, "libgapx.so")
    run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", out] + objs)
    return out


def build_fortran():
    """Fortran facade (FGAP_CALC/FGAP_READ/CAR2ACSF/FGET_BOND) over the C ABI,
    and the f2py extension module libgap.libgap."""
    srcs = [os.path.join(FSRC, f) for f in ("gapx_iface.f90", "gap_facade.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    objdir = os.path.join(HERE, "_build")
    moddir = os.path.join(objdir, "mod")
    os.makedirs(moddir, exist_ok=True)
    objs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s).replace(".f90", ".o"))
        run([FLANG, "-O2", "-fPIC", "-module-dir", moddir, "-c", s, "-o", o])
        objs.append(o)
    flib = os.path.join(LIB, "libgapx_fortran.so")
    run([FLANG, "-shared", "-fPIC", "-o", flib] + objs + ["-L" + LIB, "-lgapx", "-Wl,-rpath,$ORIGIN"])
    # f2py wrapper: generated C from our own signature file, compiled with gcc
    pyf = os.path.join(FSRC, "libgap.pyf")
    gen = os.path.join(objdir, "f2py")
    os.makedirs(gen, exist_ok=True)
    run([sys.executable, "-m", "numpy.f2py", pyf, "--build-dir", gen, "--quiet"], cwd=gen)
    import numpy
    import numpy.f2py
    f2src = os.path.join(os.path.dirname(numpy.f2py.__file__), "src")
    inc = [sysconfig.get_paths()["include"], numpy.get_include(), f2src]
    cobjs = []
    for c in ("libgapmodule.c", os.path.join(f2src, "fortranobject.c")):
        src = c if os.path.isabs(c) else os.path.join(gen, c)
        o = os.path.join(gen, os.path.basename(c).replace(".c", ".o"))
        run(["gcc", "-O2", "-fPIC", "-c", src, "-o", o] + ["-I" + p for p in inc])
        cobjs.append(o)
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    mod = os.path.join(HERE, "libgap", "libgap" + ext)
    run(["gcc", "-shared", "-o", mod] + cobjs +
        ["-L" + LIB, "-lgapx_fortran", "-lgapx", "-Wl,-rpath,$ORIGIN/../lib"])
    return mod


def build_oracle():
    """The oracle: its C restatement always; the reference Fortran when it is mounted."""
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/gappy/libgap"):
        run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def main():
    build_hip()
    build_fortran()
    build_oracle()


if __name__ == "__main__":
    main()
