"""The oracle on the reference's two analytic regression examples (TestSH: SH point
force, NGLL=6; LambsProblem: P-SV point force under a free surface, NGLL=9; both with
the leapfrog scheme), against the analytic seismograms they ship with, at the examples'
own pass marks (tests/analytic.py)."""
import pytest

from analytic import BARS, errors
from conftest import case_par
from sem2dpack_amd import builder, solver


@pytest.mark.parametrize("name", ["testsh", "lamb"])
def test_oracle_matches_the_analytic_solution(name, oracle_api):
    pb = builder.build(case_par(name))
    s = solver.Solver(pb, api=oracle_api, compact=False)
    s.step(pb.nt)
    err = errors(name, s.seismograms())
    print(f"\n  {name}: " + "  ".join(f"{k}={v:.2e}" for k, v in err.items()) + f"  (pass mark {BARS[name]})")
    assert all(v < BARS[name] for v in err.values()), err
