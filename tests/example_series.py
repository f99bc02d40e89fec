"""The fault time series shipped with the reference's RateState example
(EXAMPLES/RateState/{Ux,Vx,Tau}_{0,3,6,9}km_ascii, packed into
tests/golden/ratestate_series.npz by tests/golden/make_ratestate_series.py).

Sample k of a shipped series is the state after time step k+1, i.e. row k+1 of the
solver's fault records (row 0 is the initial state).  The shipped series are not
the output of the reference source as it stands: the oracle, which reproduces that
source's fault records bit for bit on the same input (tests/test_oracle_golden.py,
case ratestate), differs from them by up to 3% relative L2 in slip rate and shear
traction at 9 km, with the same peak step everywhere.  The bars below hold that
difference with a margin; the peak steps must be equal.
"""
import os

import numpy as np

from conftest import GOLDEN

# relative L2 over the 803 samples, per quantity (slip, slip rate, shear traction)
BARS = {"Ux": 1.5e-2, "Vx": 5e-2, "Tau": 3e-2}
COLUMN = {"Ux": 0, "Vx": 1, "Tau": 2}


def compare(pb, records):
    """[(quantity, x, relative L2, shipped peak step, our peak step)] for the four points."""
    g = np.load(os.path.join(GOLDEN, "ratestate_series.npz"))
    n1 = np.asarray(pb.fault["node1"])
    xf = pb.coord[n1 - 1, 0]
    out = []
    for qi, q in enumerate(g["quantities"].tolist()):
        for p, x in enumerate(g["x"]):
            j = int(np.argmin(np.abs(xf - x)))
            assert abs(xf[j] - x) < 1e-6
            shipped = g["series"][qi, p]
            ours = records[1:1 + shipped.size, COLUMN[q], j]
            assert ours.shape == shipped.shape
            err = float(np.linalg.norm(ours - shipped) / np.linalg.norm(shipped))
            out.append((q, float(x), err, int(np.abs(shipped).argmax()), int(np.abs(ours).argmax())))
    return out


def check(rows):
    lines = [f"  {q:3s} x={x / 1e3:.0f} km  rel L2 {e:.2e} (bar {BARS[q]:.1e})  peak step {a}/{b}"
             for q, x, e, a, b in rows]
    print("\n" + "\n".join(lines), flush=True)
    for q, x, e, a, b in rows:
        assert e < BARS[q], (q, x, e)
        assert a == b, (q, x, a, b)
