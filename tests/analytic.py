"""The reference's analytic regression tests (EXAMPLES/TestSH/analyze_test.m and
EXAMPLES/LambsProblem/analyze_test.m), restated: the float32 seismograms of a run
against the analytic seismograms packed in tests/golden/analytic_series.npz, error
max|u - u_analytic| / max|u_analytic| per trace, and the examples' own pass marks (2%
for TestSH, 0.5% for Lamb's problem)."""
import os

import numpy as np

from conftest import GOLDEN

BARS = {"testsh": 0.02, "lamb": 0.005}


def errors(name, seismo):
    """{trace: error} for case ``name`` from FP64 seismograms [nsamp][ndof][nrec]."""
    g = np.load(os.path.join(GOLDEN, "analytic_series.npz"))
    u = seismo.astype(np.float32).astype(np.float64)  # as read back from U?_sem2d.dat
    if name == "testsh":
        ref = g["testsh_u"]
        ours = u[:, 0, 4]  # u(:,5)
        assert ours.shape == ref.shape
        return {"uy5": float(np.abs(ours - ref).max() / np.abs(ref).max())}
    if name == "lamb":
        out = {}
        # analytic series start at t = dt: a zero row for t = 0
        ux = np.vstack([np.zeros((1, 2)), g["lamb_ux"]])
        uz = np.vstack([np.zeros((1, 2)), g["lamb_uz"]])
        assert u.shape[0] == ux.shape[0]
        # max(abs([ux-uxa, uz-uza])) ./ max(abs([uxa, uza])): each trace's largest misfit
        # over that trace's largest value
        for c, ref in ((0, ux), (1, uz)):
            for r in range(2):
                out[f"u{'xz'[c]}{r + 1}"] = float(np.abs(u[:, c, r] - ref[:, r]).max() / np.abs(ref[:, r]).max())
        return out
    raise KeyError(name)
