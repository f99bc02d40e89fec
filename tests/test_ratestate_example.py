"""The oracle over the reference's full RateState example (one-sided SH fault, slip
law, 270x90 elements, 803 steps) against the fault series the example ships with
(slip, slip rate, shear traction at x = 0, 3, 6, 9 km)."""
from conftest import case_par
from example_series import check, compare
from sem2dpack_amd import builder, solver


def test_oracle_reproduces_the_shipped_series(oracle_api):
    pb = builder.build(case_par("ratestate"))
    assert pb.nt == 803 and pb.fault["oitd"] == 1
    s = solver.Solver(pb, api=oracle_api, compact=False)
    s.step(pb.nt)
    check(compare(pb, s.fault_records()))
