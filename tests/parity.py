"""Shared comparison helpers of the GPU parity tests (GPU run vs the oracle).

Bars: FP64 relative L2 error <= 1e-10 on displacement, velocity, acceleration,
seismograms, fault slip, slip rate, shear traction, friction and potency.  For
rate-and-state faults (and any other case whose error exceeds 1e-10) the bar of each
quantity is max(1e-10, spread), where spread is
the largest error against the oracle among the oracle's perturbation ensemble
(oracle/build_oracle.py: libm results moved by +1 ulp, by -1 ulp, and by -1/0/+1 ulp
drawn per call with four seeds; the element forces assembled in reverse and in colour
order; FMA contraction; FMA with colour order).  The ensemble
measures how far the reference arithmetic moves under round-off that a correct
implementation may legitimately change: on a creeping rate-and-state fault the slip
rate comes from the difference of two tractions of ~1e8 Pa
(SRC/bc_dynflt.f90:685-687), so one ulp in the friction coefficient moves it by ~1e-7;
with a Kelvin-Voigt layer along a slip-weakening fault the acceleration field moves by
~7e-10 under a change of summation order alone.  No multiplier is applied to the spread.
"""
import numpy as np

TOL = 1e-10


def rel(x, y):
    n = np.linalg.norm(y)
    return float(np.linalg.norm(x - y) / (n if n > 0 else 1.0))


def quantities(pb, s):
    """{name: (array, error function)} of everything a run is compared on."""
    rows = {}
    for nm, a in zip("dva", s.fields()):
        rows[nm] = (a, rel)
    if pb.receivers is not None:
        rows["seismo"] = (s.seismograms(), rel)
    if pb.fault is not None:
        r = s.fault_records()
        rsf = pb.fault["law"] == "RSF"
        for q, nm in enumerate(("slip", "slip_rate", "shear", "normal", "mu", "tstick")):
            if rsf and nm == "tstick":
                continue  # unset in the reference's RSF branch
            if nm == "mu" and pb.fault.get("twf") is not None:
                # time weakening sets huge(1d0) ahead of the nucleation front: equal there,
                # relative L2 elsewhere
                def err(x, y):
                    big = np.abs(y) > 1e300
                    return rel(x[~big], y[~big]) if np.array_equal(x[big], y[big]) else np.inf
            elif nm == "normal":
                # zero up to round-off in these symmetric set-ups: absolute, against |T0n|
                scale = np.abs(pb.fault["T0"][:, 1]).max() * np.sqrt(r[:, q].size)
                err = (lambda x, y, sc=scale: float(np.linalg.norm(x - y) / sc))
            else:
                err = rel
            rows[nm] = (r[:, q], err)
        if pb.fault.get("osides"):  # per-side records: d, v of each side
            rows["sides"] = (s.fault_sides(), rel)
        p = s.potency()
        # against the largest potency of the same field (displacement, velocity columns):
        # single columns can be pure round-off (no opening on a planar fault)
        h = p.shape[1] // 2
        scale = np.repeat([max(np.abs(p[:, :h]).max(), 1e-300), max(np.abs(p[:, h:]).max(), 1e-300)], h)
        rows["potency"] = (p, lambda x, y, sc=scale: float((np.abs(x - y) / sc).max()))
    return rows


def compare(pb, gpu, ref, members=()):
    """errors of gpu against ref, bars, and the ensemble spread per quantity.
    The error functions (scales) are those of the oracle run."""
    qr = quantities(pb, ref)
    qg = quantities(pb, gpu)
    errs = {k: qr[k][1](qg[k][0], qr[k][0]) for k in qr}
    spread = {k: 0.0 for k in qr}
    for m in members:
        qm = quantities(pb, m)
        for k in qr:
            spread[k] = max(spread[k], qr[k][1](qm[k][0], qr[k][0]))
    bars = {k: max(TOL, spread[k]) for k in qr}
    return errs, bars, spread
