"""Writes tests/golden/golden.npz: small input/output vectors from numpy and
scipy (independent FFT/DCT/DST/spline implementations) that pin the oracle.

Run: python tests/golden/make_golden.py   (numpy 2.2, scipy 1.15)
"""
import os

import numpy as np
import scipy.fft as sf
from scipy.interpolate import CubicSpline

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20240601)
    g = {}
    # complex 3-D DFT (sign -1, unscaled) including a non power-of-two axis
    a = rng.standard_normal((8, 6, 5)) + 1j * rng.standard_normal((8, 6, 5))
    g["fft_in_re"], g["fft_in_im"] = a.real.copy(), a.imag.copy()
    f = np.fft.fftn(a)
    g["fft_out_re"], g["fft_out_im"] = f.real.copy(), f.imag.copy()
    # DCT-II (scipy's is 2x the unnormalised sum) and DST-II for lengths 12 and 7
    for n in (12, 7):
        x = rng.standard_normal(n)
        g[f"dct_in_{n}"] = x
        g[f"dct2_{n}"] = sf.dct(x, type=2) / 2.0
        g[f"dst2_{n}"] = sf.dst(x, type=2) / 2.0
    # periodic Poisson solution with numpy's FFT, zero-mean gauge
    n = (8, 6, 4)
    d = (0.3, 0.4, 0.5)
    r = rng.standard_normal(n)
    r -= r.mean()
    lam = np.zeros(n)
    for ax in range(3):
        k = np.arange(n[ax])
        e = -4.0 * np.sin(np.pi * k / n[ax]) ** 2 / d[ax] ** 2
        sh = [1, 1, 1]
        sh[ax] = n[ax]
        lam = lam + e.reshape(sh)
    R = np.fft.fftn(r)
    lam[0, 0, 0] = 1.0
    S = R / lam
    S[0, 0, 0] = 0.0
    g["pois_rhs"], g["pois_sol"] = r, np.fft.ifftn(S).real
    # not-a-knot cubic spline (MATLAB interp1 'spline') on a 6-snapshot window
    t0 = np.linspace(0.0, 1.0, 6)
    y = rng.standard_normal((6, 5))
    tt = np.linspace(0.2, 0.8, 7)
    g["spl_t0"], g["spl_y"], g["spl_tt"] = t0, y, tt
    g["spl_out"] = CubicSpline(t0, y, bc_type="not-a-knot", axis=0)(tt)
    np.savez(os.path.join(HERE, "golden.npz"), **g)


if __name__ == "__main__":
    main()
