"""The configurations at their own scale (the parity cases elsewhere are smaller):

  C2  SH rupture, 600 x 300 elements (2,886,002 nodes, 2,401 fault pairs): 1000 of its
      4000 steps, FP64 vs the oracle at the parity bar, and the FP32 bulk variant
      reported against the FP64 oracle;
  C4  the 256-scenario rate-and-state ensemble (nx ~ U{100..400}): 300 steps batched,
      bitwise equal to the same scenarios stepped one by one, and two members against
      the oracle at the parity bars;
  C5  the adjoint workload, 400 x 400 elements, 16 sources, 64 receivers, 3000 steps:
      gradients of both storage plans equal to round-off, and central finite differences
      along 4 random directions within 1e-3.
"""
import math
import time

import numpy as np
import pytest

import conftest
from parity import TOL, compare, rel
from sem2dpack_amd import solver, workloads

pytestmark = pytest.mark.gpu


def test_c2_sh_rupture_600x300(oracle_api):
    nt = 1000
    pb = workloads.build_sh_rupture(600, 300, nt=nt)
    print(f"\n[C2 600x300 SH rupture: npoin {pb.npoin}, {len(pb.fault['node1'])} pairs, {nt} steps]", flush=True)
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    gpu = solver.Solver(pb)
    g32 = solver.Solver(pb, bulk_fp32=True)
    for s in (gpu, g32, ref):
        s.step(nt)
    errs, bars, _ = compare(pb, gpu, ref)
    conftest.report("C2 600x300 FP64", errs, bars)
    rr, r32 = ref.fault_records(), g32.fault_records()
    e32 = {"slip": rel(r32[:, 0], rr[:, 0]), "slip_rate": rel(r32[:, 1], rr[:, 1]),
           "shear": rel(r32[:, 2], rr[:, 2])}
    conftest.report("C2 600x300 FP32 bulk", e32, {k: 1e-3 for k in e32})
    print("  FP64 " + "  ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    print("  FP32 bulk vs FP64 oracle " + "  ".join(f"{k}={v:.2e}" for k, v in e32.items()), flush=True)
    assert np.abs(rr[-1, 0]).max() > 1e-3  # the rupture has started by then
    assert all(v < bars[k] for k, v in errs.items()), errs
    assert all(v < 1e-3 for v in e32.values()), e32


@pytest.mark.timeout(600)
def test_c4_ensemble_256_bitwise_and_oracle(oracle_api, oracle_ensemble):
    nt = 300
    scen = workloads.rupture_ensemble(256)
    t0 = time.time()
    pbs = [workloads.build_rupture(sc, nt=nt) for sc in scen]
    print(f"\n[C4 256 scenarios built in {time.time() - t0:.0f} s; npoin total "
          f"{sum(p.npoin for p in pbs)}]", flush=True)
    members = [solver.Solver(pb, bulk_fp32=True) for pb in pbs]
    solos = [solver.Solver(pb, bulk_fp32=True) for pb in pbs]
    b = solver.Batch(members)
    b.step(nt)
    for s in solos:
        s.step(nt)
    for m in range(len(pbs)):
        fb, fs = b.fault_state(m), solos[m].fault_state()
        for q in fs:
            assert np.array_equal(fb[q], fs[q]), (m, q)
        assert np.array_equal(b.fault_records(m), solos[m].fault_records()), m
        assert np.array_equal(b.potency(m), solos[m].potency()), m
        if m % 8 == 0:
            for x, y in zip(b.fields(m), solos[m].fields()):
                assert np.array_equal(x, y), m
    b.close()
    for s in members + solos:
        s.close()
    # two members (the smallest and the largest) against the oracle, FP64 bulk
    sizes = [p.npoin for p in pbs]
    for m in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        pb = pbs[m]
        gpu = solver.Solver(pb)
        ref = solver.Solver(pb, api=oracle_api, compact=False)
        ens = [solver.Solver(pb, api=api, compact=False) for _, api in oracle_ensemble]
        for s in [gpu, ref] + ens:
            s.step(nt)
        errs, bars, _ = compare(pb, gpu, ref, ens)
        conftest.report(f"C4 member {m} ({scen[m]['nx']}x{scen[m]['nz']})", errs, bars)
        print(f"  member {m} nx {scen[m]['nx']}: " + "  ".join(f"{k}={v:.1e}/{bars[k]:.0e}" for k, v in errs.items()),
              flush=True)
        assert all(v < bars[k] for k, v in errs.items()), (m, errs)


def test_c5_adjoint_400x400_both_plans(tmp_path):
    from sem2dpack_amd import adjoint
    nt = 3000
    path, back, true = workloads.adjoint_setup(str(tmp_path), nx=400, nz=400, nt=nt, nsrc=16, nrec=64, f0=3.0)
    m = adjoint.SeismogramModel(path)
    obs = m.forward(*true)
    m.close()
    grads = {}
    for seg in (0, int(math.ceil(math.sqrt(nt)))):
        mm = adjoint.SeismogramModel(path, seg=seg)
        seis = mm.forward(*back)
        hist, segl = mm._state.solver.adjoint_plan()
        g = mm.backward(seis - obs)
        print(f"\n[C5 plan history={bool(hist)} segment={segl}] device ms {mm.last_timing}", flush=True)
        grads[bool(hist)] = g
        mm.close()
    adjoint.release()
    assert set(grads) == {True, False}
    e = [rel(a, b) for a, b in zip(grads[True], grads[False])]
    print(f"  history vs checkpoints rel-L2 (rho, lam, mu): {e}", flush=True)
    assert max(e) < 1e-9
    rng = np.random.Generator(np.random.PCG64(20260503))
    fdm = adjoint.SeismogramModel(path)
    eps = 1e-2
    worst = 0.0
    for k in range(4):
        dm = [rng.standard_normal(p.size) * p * 1e-2 for p in back]
        Jp = 0.5 * float(np.sum((fdm.forward(*[p + eps * q for p, q in zip(back, dm)]) - obs) ** 2))
        Jm = 0.5 * float(np.sum((fdm.forward(*[p - eps * q for p, q in zip(back, dm)]) - obs) ** 2))
        fd = (Jp - Jm) / (2 * eps)
        an = sum(np.dot(a, b) for a, b in zip(grads[True], dm))
        worst = max(worst, abs(an - fd) / abs(fd))
        print(f"  direction {k}: FD {fd:.6e} adjoint {an:.6e} rel {abs(an - fd) / abs(fd):.2e}", flush=True)
    fdm.close()
    conftest.REPORT.append(("C5 400x400 adjoint", "fd_max_rel", worst, 1e-3))
    assert worst < 1e-3
