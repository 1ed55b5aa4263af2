"""Write the FFT golden vectors used to pin the oracle's r2c / c2r.

numpy.fft.rfftn / irfftn (pocketfft) are the library routine the restated
transforms must agree with.  The c2r inputs include deliberately non-Hermitian
spectra: irfftn then keeps only Re() of the kx = 0 and kx = Nx/2 bins after
the inverse transforms along z and y, which is the synthesis rule the hot
path's c2r of the gradient spectra relies on.

Run:  python tests/golden/gen_fft_golden.py   (writes tests/golden/fft_golden.npz)
"""
import os

import numpy as np

SHAPES = [(8, 8, 8), (4, 4, 4), (16, 8, 32), (2, 4, 8), (8, 16, 4)]   # (Nz, Ny, Nx)


def main():
    rng = np.random.default_rng(20241202)
    out = {}
    for s in SHAPES:
        tag = "x".join(map(str, s))
        a = rng.standard_normal(s)
        out[f"r2c_in_{tag}"] = a
        out[f"r2c_out_{tag}"] = np.fft.rfftn(a, axes=(0, 1, 2))
        h = (s[0], s[1], s[2] // 2 + 1)
        c = rng.standard_normal(h) + 1j * rng.standard_normal(h)   # non-Hermitian on purpose
        out[f"c2r_in_{tag}"] = c
        # unnormalised c2r = N * irfftn
        out[f"c2r_out_{tag}"] = np.fft.irfftn(c, s=s, axes=(0, 1, 2)) * np.prod(s)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
