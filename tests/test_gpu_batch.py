"""Batched ensemble (sem2d_batch_*) vs the same scenarios stepped one by one.

Every member of a batch runs exactly its own solver's arithmetic, so the bar is
bitwise equality with the member's single-scenario run: fields, fault state, fault
records and seismograms.  The batch loop is a hipGraph replay; the checks come after
each of several replays, with the one-by-one runs stepping eagerly in between, and
again after the autotuner has switched the tile-kernel launch configuration.
Member 0 is also compared with the oracle at the single-scenario tolerance.
"""
import numpy as np
import pytest

import conftest
from conftest import case_par
from parity import compare
from sem2dpack_amd import builder, parfile, solver, workloads

pytestmark = pytest.mark.gpu


def rupture_members(sizes, nt, fp32):
    pbs, members = [], []
    for k, nx in enumerate(sizes):
        par = parfile.parse_par(workloads.par_rupture(nx, nx // 2, 0.009 + 0.002 * k, 9100 + k, nt=nt))
        pb = builder.build(par)
        pbs.append(pb)
        members.append(solver.Solver(pb, bulk_fp32=fp32))
    return pbs, members


def source_members(shapes, nt):
    """small P-SV boxes of different sizes with an explosion and a receiver line, fixed dt"""
    txt = open(case_par("small_psv")).read()
    pbs, members = [], []
    for nx, nz in shapes:
        t = txt.replace("nelem=20,16", f"nelem={nx},{nz}").replace("Courant=0.3d0", "Dt=2.0d-3")
        t = t.replace("NbSteps=300", f"NbSteps={nt}")
        pb = builder.build(parfile.parse_par(t), fname_dir=None)
        pbs.append(pb)
        members.append(solver.Solver(pb))
    return pbs, members


def solo_copies(pbs, fp32):
    return [solver.Solver(pb, bulk_fp32=fp32) for pb in pbs]


def assert_same(batch, solos, m, fault):
    for x, y in zip(batch.fields(m), solos[m].fields()):
        assert np.array_equal(x, y)
    if fault:
        fb, fs = batch.fault_state(m), solos[m].fault_state()
        for k in fs:
            assert np.array_equal(fb[k], fs[k]), k
        assert np.array_equal(batch.fault_records(m), solos[m].fault_records())
        assert np.array_equal(batch.potency(m), solos[m].potency())
    if getattr(solos[m], "nrec", 0):
        assert np.array_equal(batch.seismograms(m), solos[m].seismograms())


@pytest.mark.parametrize("fp32", [False, True])
def test_rupture_batch_bitwise(fp32):
    print(f"\n[batch rupture fp32={fp32}]", flush=True)
    sizes = (24, 33, 41, 17)
    pbs, members = rupture_members(sizes, 300, fp32)
    solos = solo_copies(pbs, fp32)
    b = solver.Batch(members)
    for rep in range(3):  # three replays of the 64-step graph
        b.step(64)
        for s in solos:
            s.step(64)
        for m in range(len(sizes)):
            assert_same(b, solos, m, True)
    b.step(22)  # a second graph of another length
    for s in solos:
        s.step(22)
    for m in range(len(sizes)):
        assert_same(b, solos, m, True)


def test_rupture_batch_matches_oracle(oracle_api, oracle_ensemble):
    """a batch member vs the oracle, with the bars of the single-scenario parity tests
    (tests/parity.py: 1e-10, or the oracle's perturbation-ensemble spread for the
    rate-and-state quantities)"""
    pbs, members = rupture_members((24, 30), 200, False)
    b = solver.Batch(members)
    b.step(200)
    ref = solver.Solver(pbs[1], api=oracle_api, compact=False)
    ens = [solver.Solver(pbs[1], api=api, compact=False) for _, api in oracle_ensemble]
    for s in [ref] + ens:
        s.step(200)
    solo = solver.Solver(pbs[1])
    solo.step(200)
    # the batch member is bitwise its own solver (test_rupture_batch_bitwise); compare
    # that run with the oracle at the parity bars
    assert all(np.array_equal(x, y) for x, y in zip(b.fields(1), solo.fields()))
    errs, bars, _ = compare(pbs[1], solo, ref, ens)
    conftest.report("batch member 1 (30x15 RSF)", errs, bars)
    bad = {k: (v, bars[k]) for k, v in errs.items() if not v < bars[k]}
    assert not bad, bad


def test_source_batch_bitwise_and_tuned():
    print("\n[batch sources + receivers, tuner]", flush=True)
    shapes = ((20, 16), (27, 11), (12, 24))
    pbs, members = source_members(shapes, 260)
    solos = solo_copies(pbs, False)
    b = solver.Batch(members)
    b.step(100)
    for s in solos:
        s.step(100)
    for m in range(len(shapes)):
        assert_same(b, solos, m, False)
    choice = b.tune(trial_steps=3)
    print("  tuned:", choice, b.config(), flush=True)
    b.step(160)
    for s in solos:
        s.step(160)
    for m in range(len(shapes)):
        assert_same(b, solos, m, False)


def test_batch_rejects_mismatched_members():
    txt = open(case_par("small_psv")).read()
    pa = builder.build(parfile.parse_par(txt.replace("Courant=0.3d0", "Dt=2.5d-3")))
    pb = builder.build(parfile.parse_par(txt.replace("Courant=0.3d0", "Dt=2.0d-3")))
    with pytest.raises(solver.SolverError, match="time step"):
        solver.Batch([solver.Solver(pa), solver.Solver(pb)])
