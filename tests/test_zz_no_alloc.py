"""sem2d_compute_fint_noalloc (staging buffers in the library's static device scratch
array, no device allocation) against sem2d_compute_fint (the same kernels with
allocated staging buffers): bitwise equal, on two calls in a row with the second on a
smaller input, for FP64 and FP32 fields; an input beyond the scratch limit is refused
with an error and leaves the next call correct."""
import numpy as np
import pytest

from conftest import case_par
from sem2dpack_amd import builder, solver, workloads

pytestmark = pytest.mark.gpu


def _check(pb, fp32, seed):
    s = solver.Solver(pb, bulk_fp32=fp32)
    d = np.random.default_rng(seed).standard_normal((pb.ndof, pb.npoin))
    f_ref = s.fint(d)
    f_new = s.fint_noalloc(d)
    s.close()
    assert np.isfinite(f_ref).all() and np.abs(f_ref).max() > 0
    assert np.array_equal(f_new, f_ref)


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_matches_allocating_entry_point(fp32):
    big = builder.build(case_par("c1_psv"), nsteps=1)     # 160,801 nodes x 2
    small = builder.build(case_par("sh_swf"), nsteps=1)   # a smaller SH grid
    print(f"\n[noalloc fp32={fp32} npoin {big.npoin} then {small.npoin}]", flush=True)
    assert small.ndof * small.npoin < big.ndof * big.npoin
    _check(big, fp32, 1)
    _check(small, fp32, 2)


def test_noalloc_rejects_inputs_beyond_the_limit():
    pb = workloads.build_throughput(nx=420, nz=400, lx=42e3, lz=40e3, nt=1)
    print(f"\n[noalloc limit: ndof*npoin = {pb.ndof * pb.npoin}]", flush=True)
    assert pb.ndof * pb.npoin > 4_194_304
    s = solver.Solver(pb)
    with pytest.raises(solver.SolverError, match="scratch limit"):
        s.fint_noalloc(np.zeros((pb.ndof, pb.npoin)))
    s.close()
    _check(builder.build(case_par("small_psv"), nsteps=1), False, 3)
