"""Build the native library libsfx.so (HIP kernels for gfx950 + C ABI + host model builder).

The library is built in-tree so that it travels with the repository snapshot; nothing is
installed. `python -m sfincs_amd.build` compiles it; tests and the bench only load it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsfx.so")

SOURCES = ["sfx_model.cpp", "sfx_meteo.cpp", "sfx_netcdf.cpp", "sfx_device.hip", "sfx_libm.hip"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-level parity with the reference: no FMA contraction, IEEE div/sqrt
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = ["hipcc"] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs + ["-lpthread"]
    _run(cmd)
    os.replace(tmp, LIB)
    tables()
    return LIB


def tables():
    """Enumerate (or load and re-verify) the C library's powf, cosf and sinf tables next to
    the library."""
    import ctypes
    lib = ctypes.CDLL(LIB)
    counts = (ctypes.c_int64 * 6)()
    lib.sfx_libm_tables.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
    rc = lib.sfx_libm_tables(None, counts)
    if rc != 0:
        raise RuntimeError("powf table generation failed")
    print("libm tables (powf 1/3, 5/3, 7/3, 1.5; cosf; sinf): " + " + ".join(str(c) for c in counts) + " entries",
          flush=True)


def build_fortran_host():
    """Fortran host (sfincs_lib entry points over the C ABI), built with amdflang."""
    fdir = os.path.join(HERE, "fortran")
    out = os.path.join(fdir, "libsfincs_host.so")
    exe = os.path.join(fdir, "sfincs_mi355x")
    mod_dir = os.path.join(fdir, "_mod")
    os.makedirs(mod_dir, exist_ok=True)
    src = os.path.join(fdir, "sfincs_lib.f90")
    main = os.path.join(fdir, "sfincs_main.f90")
    if not os.path.exists(src):
        return None
    _run(["amdflang", "-O2", "-fPIC", "-shared", "-module-dir", mod_dir, "-o", out, src,
          "-L" + HERE, "-lsfx", "-Wl,-rpath,$ORIGIN/.."])
    if os.path.exists(main):
        _run(["amdflang", "-O2", "-I", mod_dir, "-o", exe, main, "-L" + fdir, "-lsfincs_host",
              "-L" + HERE, "-lsfx", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."])
    return out


if __name__ == "__main__":
    build()
    if "--fortran" in sys.argv:
        build_fortran_host()
