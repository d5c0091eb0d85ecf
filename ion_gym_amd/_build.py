"""Build the engine's shared library, libion_gym_amd.so, for gfx950.

Each HIP source is compiled by hipcc into an object in ``_build/`` and the
objects are linked in-tree into ``ion_gym_amd/libion_gym_amd.so``, which the
ctypes binding (``ion_gym_amd._lib``) loads.  The benchmark's gather-rate
probe (csrc/probe_gather.hip, measurement only) is linked on its own into
``ion_gym_amd/libig_probe.so``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libion_gym_amd.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ("ig_engine.hip", "ig_rz_f64.hip", "ig_rz_f32s.hip", "ig_rz_f32.hip",
           "ig_xyz_f64.hip", "ig_xyz_f32s.hip", "ig_xyz_f32.hip",
           "ig_planar_f64.hip", "ig_planar_f32s.hip", "ig_planar_f32.hip",
           "ig_births.hip", "ig_solve.hip")
PROBE_SOURCES = ("probe_gather.hip", "probe_hw_math.hip")
PROBE_OUT = os.path.join(HERE, "libig_probe.so")
ARCH = os.environ.get("ION_GYM_AMD_ARCH", "gfx950")


def hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin",
                           "hipcc"), "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    raise FileNotFoundError("hipcc not found (set ROCM_PATH)")


def flags():
    return [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
            # packed FP32 VALU forms are kept out of the generated code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-result",
            "-I", os.path.join(HERE, "..", "include")]


def build(verbose=False, jobs=None, incremental=False):
    """Compile every source and link the library; raises on failure.
    ``incremental`` skips objects newer than every source and header."""
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    deps = [os.path.join(CSRC, f) for f in os.listdir(CSRC)]
    deps.append(os.path.join(HERE, "..", "include", "ion_gym_amd.h"))
    newest = max(os.path.getmtime(d) for d in deps)

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        if (incremental and os.path.exists(obj)
                and os.path.getmtime(obj) >= newest):
            return obj
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs) + 1)) as ex:
        objs = list(ex.map(one, srcs + list(PROBE_SOURCES)))
    n_probe = len(PROBE_SOURCES)
    for out, group in ((OUT, objs[:-n_probe]), (PROBE_OUT, objs[-n_probe:])):
        tmp = out + ".tmp"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
               *group]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
        os.replace(tmp, out)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
