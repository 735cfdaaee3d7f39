"""Builds the HIP library libomeinsum_mi355x.so (gfx950) in-tree with hipcc via make, and the
calibration program calib/mfma_loop (the packed GEMM's inner loop without global memory)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "omeinsum_mi355x", "libomeinsum_mi355x.so")
CALIB_SRC = os.path.join(HERE, "calib", "mfma_loop.hip")
CALIB = os.path.join(HERE, "calib", "mfma_loop")
HIPFLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(jobs=8, force=False, verbose=False):
    if force:
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, capture_output=not verbose)
    r = subprocess.run(["make", "-C", CSRC, "-j%d" % jobs], capture_output=not verbose, text=True)
    if r.returncode != 0:
        sys.stderr.write((r.stdout or "") + (r.stderr or ""))
        raise RuntimeError("building libomeinsum_mi355x.so failed")
    if force or not os.path.exists(CALIB) or os.path.getmtime(CALIB) < os.path.getmtime(CALIB_SRC):
        r = subprocess.run(["/opt/rocm/bin/hipcc"] + HIPFLAGS + ["-o", CALIB + ".tmp", CALIB_SRC],
                           capture_output=not verbose, text=True)
        if r.returncode != 0:
            sys.stderr.write((r.stdout or "") + (r.stderr or ""))
            raise RuntimeError("building calib/mfma_loop failed")
        os.replace(CALIB + ".tmp", CALIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))
