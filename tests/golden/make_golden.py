"""Write the DST-I golden vectors used to pin the oracle and the GPU DST.

scipy.fft.dst(type=1, norm='ortho') is the published orthonormal DST-I that
the reference's FFTW-based transform equals (2D_Program/solver_functions/
DST.c:6-44 with coef sqrt(2/(N+1)), solver.c:25).  Inputs are seeded complex
rows; the fixture stores inputs and outputs only.

    python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = [1, 2, 3, 7, 16, 31, 63, 100, 127, 255]


def main():
    rng = np.random.default_rng(20240601)
    out = {"scipy_version": np.array(scipy.__version__)}
    for n in SIZES:
        x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
        out[f"x_{n}"] = x
        out[f"y_{n}"] = sf.dst(x, type=1, norm="ortho", axis=-1)
    np.savez_compressed(os.path.join(HERE, "dst1_golden.npz"), **out)


if __name__ == "__main__":
    main()
