"""Progress rows and the non-finite watchdog of the time loop, in the oracle (the rows the
HIP library is checked against): the row of the last step holds max|v| and max|d| of
the fields at that step (SRC/main.f90:73-76), and a non-finite value planted in the
state stops the next step that ends on a row with error 21."""
import numpy as np
import pytest

from conftest import case_par
from sem2dpack_amd import builder, solver


def test_oracle_progress_rows(oracle_api):
    pb = builder.build(case_par("small_psv"), nsteps=150)
    s = solver.Solver(pb, api=oracle_api, compact=False, itinfo=50)
    s.step(150)
    rows = s.info_rows()
    assert rows.shape == (3, 4)
    assert rows[:, 0].tolist() == [50.0, 100.0, 150.0]
    assert rows[:, 1].tolist() == [k * pb.dt for k in (50, 100, 150)]
    d, v, a = s.fields()
    assert rows[-1, 2] == np.abs(v).max() and rows[-1, 3] == np.abs(d).max()
    assert s.field_info() == (np.abs(v).max(), np.abs(d).max(), 0)
    v[1, 7] = np.inf
    d[0, 3] = np.nan
    s.set_fields(d, v, a)
    assert s.field_info()[2] == 2
    s.step(49)  # no row in these steps
    with pytest.raises(solver.SolverError, match=r"\(21\).*step 200"):
        s.step(1)


def test_itinfo_from_the_input_file():
    pb = builder.build(case_par("c1_psv"), nsteps=1)
    assert pb.itinfo == 500
