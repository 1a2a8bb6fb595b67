#!/usr/bin/env python3
"""Golden Boys-function values F_m(T) = int_0^1 t^(2m) exp(-T t^2) dt.

Computed with mpmath 1.3.0 at 50 significant digits through the confluent
hypergeometric form F_m(T) = 1F1(m+1/2; m+3/2; -T) / (2m+1).  mpmath is a
third-party arbitrary-precision library, unrelated to ChronusQ; the table
pins the oracle's and the device kernels' Boys / Rys evaluations.
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))

T = [0.0, 1e-12, 1e-8, 1e-5, 1e-3, 0.01, 0.1, 0.3, 0.5, 1.0, 1.5, 2.0, 3.0,
     5.0, 7.5, 10.0, 12.5, 15.0, 17.5, 20.0, 25.0, 30.0, 33.0, 35.0, 35.99,
     36.01, 38.0, 40.0, 45.0, 50.0, 60.0, 75.0, 100.0, 150.0, 300.0, 1000.0]
rng = np.random.Generator(np.random.PCG64(7))
T += list(np.round(rng.uniform(0.0, 60.0, size=40), 6))
MMAX = 24


def boys(m, t):
    return mp.hyp1f1(m + mp.mpf(1) / 2, m + mp.mpf(3) / 2, -mp.mpf(t)) / (2 * m + 1)


def main():
    Ts = np.array(T, dtype=np.float64)
    F = np.array([[float(boys(m, t)) for m in range(MMAX + 1)] for t in Ts])
    np.savez_compressed(os.path.join(HERE, "boys_mpmath.npz"), T=Ts, F=F)


if __name__ == "__main__":
    main()
