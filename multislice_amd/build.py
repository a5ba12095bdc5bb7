"""Build libmsx.so (the HIP engine behind include/msx.h) for gfx950, in-tree.

    python -m multislice_amd.build [-j JOBS]

Each supported line length gets its own object from msx_launch.hip
(-DMSX_N=<n>) so the template instantiations compile in parallel.  Packed
FP32 instructions are disabled for all device code.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmsx.so")
SIZES = [16, 32, 64, 128, 256, 384, 512, 768, 1024, 2048]
ARCH = os.environ.get("MSX_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-fvisibility=hidden",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + CSRC, "-I" + os.path.join(HERE, "..", "include")]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines()
                    if "unknown target feature" not in l and "packed-fp32" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + err)
    return err


def build(jobs=None, verbose=True, defines=(), out=None, bdir=None):
    """Compile every unit and link libmsx.so.  `defines` (e.g. MSX_COL_TW=8)
    and `out` build a shape variant elsewhere (used by the tuner)."""
    out = out or OUT
    bdir = bdir or os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    cc = hipcc()
    units = []
    for n in SIZES:
        units.append((os.path.join(CSRC, "msx_launch.hip"), os.path.join(bdir, f"launch_{n}.o"),
                      [f"-DMSX_N={n}"]))
    for name in ("msx_potential", "msx_api", "msx_unet", "msx_unet_kernels", "msx_specimen", "msx_formats",
                 "msx_binning", "msx_lms", "msx_series"):
        units.append((os.path.join(CSRC, name + ".hip"), os.path.join(bdir, name + ".o"), []))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        dflags = ["-D" + d for d in defines]
        futs = [ex.submit(_run, [cc] + flags() + dflags + extra + ["-c", src, "-o", obj])
                for src, obj, extra in units]
        for f in futs:
            f.result()
    tmp = out + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + [u[1] for u in units])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs)
    sys.exit(0)
