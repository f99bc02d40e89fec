#!/usr/bin/env python3
"""Pack the fault time series that ship with the reference's RateState example
(EXAMPLES/RateState/{Ux,Vx,Tau}_{0,3,6,9}km_ascii: slip, slip rate and shear traction
change at four points of the fault, one sample per time step, 803 steps) into
tests/golden/ratestate_series.npz as data: series[quantity][point][step], with the
point abscissae in metres.

Usage: python tests/golden/make_ratestate_series.py   (needs /root/reference)
"""
import os

import numpy as np

SRC = "/root/reference/EXAMPLES/RateState"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ratestate_series.npz")
QUANTITIES = ("Ux", "Vx", "Tau")
KM = (0, 3, 6, 9)


def main():
    series = np.stack([np.stack([np.loadtxt(os.path.join(SRC, f"{q}_{km}km_ascii")) for km in KM])
                       for q in QUANTITIES])
    np.savez_compressed(OUT, series=series, quantities=np.array(QUANTITIES),
                        x=np.array([1e3 * km for km in KM]))
    print(OUT, series.shape)


if __name__ == "__main__":
    main()
