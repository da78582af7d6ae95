"""Writes tests/golden/fft_golden.npz: numpy (pocketfft) FP64 DFTs of seeded
random complex inputs, the third-party reference the oracle's and the GPU's
FFTs are checked against.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")


def main():
    rng = np.random.default_rng(20200604)
    data = {}
    for n in (8, 12, 32, 96, 128, 384):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        data[f"in1d_{n}"] = x
        data[f"fwd1d_{n}"] = np.fft.fft(x)
        data[f"bwd1d_{n}"] = np.fft.ifft(x) * n
    for shape in ((8, 6, 4), (16, 32, 1), (12, 8, 2)):
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        key = "x".join(map(str, shape))
        data[f"in3d_{key}"] = x
        data[f"fwd3d_{key}"] = np.fft.fftn(x)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
