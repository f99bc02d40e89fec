#!/usr/bin/env python3
"""Pack the analytic seismograms that ship with the reference's two regression examples
into tests/golden/analytic_series.npz (data only):

  testsh_t, testsh_u   EXAMPLES/TestSH/uyref.mat (tref, uref): SH displacement of a
                       point force, at the example's fifth receiver (analytic_sh2d of
                       SEMLAB, per the example's analyze_test.m)
  lamb_ux, lamb_uz     EXAMPLES/LambsProblem/{Ux,Uz}_file_ascii: Lamb's problem at the
                       example's two receivers, [sample][receiver] (EX2DDIR, per the
                       example's analyze_test.m), without the t = 0 row

The .mat file is read with scipy.io.loadmat, which parses the MAT-file and runs nothing
from it.

Usage: python tests/golden/make_analytic.py   (needs /root/reference)
"""
import os

import numpy as np
import scipy.io

EX = "/root/reference/EXAMPLES"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "analytic_series.npz")


def main():
    m = scipy.io.loadmat(os.path.join(EX, "TestSH", "uyref.mat"))
    ux = np.loadtxt(os.path.join(EX, "LambsProblem", "Ux_file_ascii"))
    uz = np.loadtxt(os.path.join(EX, "LambsProblem", "Uz_file_ascii"))
    nsta = 2
    np.savez_compressed(OUT, testsh_t=np.asarray(m["tref"], dtype=np.float64).ravel(),
                        testsh_u=np.asarray(m["uref"], dtype=np.float64).ravel(),
                        lamb_ux=ux.reshape(nsta, -1).T, lamb_uz=uz.reshape(nsta, -1).T)
    print(OUT)


if __name__ == "__main__":
    main()
