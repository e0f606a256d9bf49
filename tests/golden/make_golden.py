"""Writes the golden fixtures in this directory (run: python tests/golden/make_golden.py).

The reference ships no tests or golden vectors and cannot be built here, so
the fixtures restate its algorithms with independent tools:
  periodic_poisson_128x64.npz : FFTW_periodic.h:77-92 spectral solve done with
                                numpy.fft (pocketfft) on a seeded RHS;
  freespace_direct_64x32.npz  : FFTW_freespace.cpp:100-137 Green's function,
                                evaluated as a direct O(N^2) sum (no FFT);
  advdiff_96x64.npz           : advDiff.cpp:14-116 on seeded fields, written
                                from the reference formulas in numpy.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import numpy_ref as R  # noqa: E402


def main():
    rng = np.random.default_rng(20261202)
    r = rng.standard_normal((64, 128))
    np.savez(os.path.join(HERE, "periodic_poisson_128x64.npz"), rhs=r, p=R.poisson_periodic(r))

    h = 1.0 / 64
    r = rng.standard_normal((32, 64))
    np.savez(os.path.join(HERE, "freespace_direct_64x32.npz"), rhs=r, h=h,
             p=R.poisson_freespace_direct(r, h))

    u = rng.standard_normal((64, 96)) * 0.3
    v = rng.standard_normal((64, 96)) * 0.3
    dt, h, nu, uinf = 1e-3, 1.0 / 96, 1e-3, (0.2, -0.05)
    u1, v1, corr = R.advdiff(u, v, dt, h, nu, uinf)
    np.savez(os.path.join(HERE, "advdiff_96x64.npz"), u=u, v=v, dt=dt, h=h, nu=nu,
             uinf=np.array(uinf), u1=u1, v1=v1, corr=corr)


if __name__ == "__main__":
    main()
