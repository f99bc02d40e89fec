import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

CASES = os.path.join(HERE, "golden", "cases")
GOLDEN = os.path.join(HERE, "golden")

# every golden case: a Par.inp under golden/cases/<name>/ and the reference's outputs
# in golden/<name>.npz (tests/golden/make_golden.py)
GOLDEN_CASES = sorted(n for n in os.listdir(CASES)
                      if os.path.exists(os.path.join(CASES, n, "Par.inp"))
                      and os.path.exists(os.path.join(GOLDEN, n + ".npz")))
# the friction branches each pinned by their own golden case
FRICTION_CASES = ["swf_exp", "swf_pow", "swf_alpha", "swf_heal", "rsf_k1", "rsf_slip", "rsf_vshape"]

# (case, quantity, error, bar) of every GPU parity comparison, printed at the end
REPORT = []


def report(case, errs, bars):
    for q in errs:
        REPORT.append((case, q, float(errs[q]), float(bars[q])))


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if not REPORT:
        return
    tr = terminalreporter
    tr.section("GPU parity: relative error / bar per case (worst ratio first)")
    by_case = {}
    for case, q, e, b in REPORT:
        by_case.setdefault(case, []).append((q, e, b))
    for case, rows in by_case.items():
        worst = max(rows, key=lambda r: r[1] / r[2])
        keys = [r for r in rows if r[0] in ("d", "v", "a", "seismo", "slip", "slip_rate", "shear", "potency", "sides")]
        body = " ".join(f"{q}={e:.1e}" + ("" if b == 1e-10 else f"/{b:.0e}") for q, e, b in keys)
        tr.write_line(f"{case[:30]:30s} worst {worst[0]} {worst[1]:.2e}/{worst[2]:.1e} | {body}")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def case_par(name):
    return os.path.join(CASES, name, "Par.inp")


@pytest.fixture(scope="session")
def oracle_api():
    import oracle
    return oracle.load()


@pytest.fixture(scope="session")
def oracle_ensemble():
    """[(name, api)]: the perturbed builds of the oracle (oracle/build_oracle.py), each
    changing one thing a correct implementation may do differently (libm results by
    +-1 ulp, the element assembly order, FMA contraction)"""
    import oracle
    return oracle.ensemble()
