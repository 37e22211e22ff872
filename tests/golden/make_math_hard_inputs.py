"""Writes math_hard_inputs.npz: for exp, log, sin, cos, tan, exp10, exp2, log10 and log2, every
float input whose glibc double result lies within 256 double ulps of a float rounding boundary
(math_hard_inputs.c scans all 2^32 inputs).  tests/test_gpu_math_glibc.py evaluates the
operators on them and compares with the oracle's glibc results.  Run in the build container:
    python3 tests/golden/make_math_hard_inputs.py"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FNS = ["exp", "log", "sin", "cos", "tan", "exp10", "exp2", "log10", "log2"]


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "scan")
        subprocess.run(["gcc", "-O2", "-fopenmp", "-o", exe, os.path.join(HERE, "math_hard_inputs.c"), "-lm"],
                       check=True)
        subprocess.run([exe, tmp], check=True)
        arrays = {fn: np.fromfile(os.path.join(tmp, f"hard_{fn}.u32"), dtype=np.uint32) for fn in FNS}
    np.savez_compressed(os.path.join(HERE, "math_hard_inputs.npz"), **arrays)


if __name__ == "__main__":
    main()
