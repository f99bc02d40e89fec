"""HIP time step vs the oracle (C restatement of the reference) on the same inputs.

Bars (tests/parity.py): FP64 relative L2 error <= 1e-10 on displacement, velocity,
acceleration, seismograms, fault slip, slip rate, shear traction, friction and
potency; element forces <= 1e-13 (one application of -K).  Rate-and-state cases, and
any other case with an error above 1e-10, take max(1e-10, the spread of the oracle's
perturbation ensemble) per quantity.  The normal
traction change, zero up to round-off in these symmetric set-ups, is compared in
absolute terms against |T0_n|.  Every case's errors and bars are printed in the
terminal summary.
"""
import os

import numpy as np
import pytest

import conftest
from conftest import GOLDEN, case_par
from parity import TOL, compare, rel
from sem2dpack_amd import builder, solver

pytestmark = pytest.mark.gpu


def _pair(name, nsteps, oracle_api, **kw):
    pb = builder.build(case_par(name), nsteps=nsteps)
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    gpu = solver.Solver(pb, **kw)
    return pb, ref, gpu


@pytest.mark.parametrize("name", ["small_psv", "sh_swf", "psv_rsf", "c1_psv", "mesh2d_psv", "ng3_psv", "ng4_kv",
                                  "testsh", "ng7_swf", "lamb", "ng10_psv", "tpv3_anti"])
def test_fint(name, oracle_api):
    pb, ref, gpu = _pair(name, 1, oracle_api)
    rng = np.random.default_rng(7)
    d = rng.standard_normal((pb.ndof, pb.npoin))
    f_ref = ref.fint(d)
    f_gpu = gpu.fint(d)
    assert rel(f_gpu, f_ref) < 1e-13


# (case, steps): every golden case over its full length.  c1_psv is the plain
# CPU-runnable configuration (100x100 elements, 2000 steps: the P and S arrivals and
# the absorbing-boundary reflections cross the receiver line); hete_psv runs 2000 steps
# (1.28 s) of its 0.3 s-onset explosion, past the arrivals and boundary reflections.
CASES = [("small_psv", 300), ("c1_psv", 2000), ("sh_swf", 600), ("multi_src", 300), ("interp_rec", 300),
         ("sh_kv", 600), ("psv_kv", 500), ("psv_rsf", 500), ("hete_psv", 2000), ("rsf_ens", 400),
         ("psv_stacey", 400), ("mesh2d_psv", 500), ("mesh2d_curved", 500), ("mesh2d_sh", 500),
         ("mesh2d_flt", 500), ("ratestate", 803), ("tpv3_kv", 800), ("tpv3_dirfirst", 600),
         ("vshape_25d", 600), ("psv_25d", 600), ("lf_psv", 500), ("lf_swf_kv", 600), ("lf_25d", 600),
         ("testsh", 1987), ("lamb", 3000), ("tpv3_anti", 1500), ("ng3_psv", 400), ("ng4_kv", 500),
         ("ng7_swf", 600), ("ng10_psv", 400), ("sides_psv", 500), ("sides_sh1", 500), ("domains", 500),
         ("twf_inplane", 600), ("twf_rsf", 500), ("twf_only", 600), ("perio_psv", 700), ("perio_sh", 800),
         ("bimat", 156), ("nor3_rsf", 400), ("nor0_swf", 500), ("hht_flt1d", 455), ("hht_psv", 600),
         ("hht_twf", 400), ("wave_psv", 825), ("wave_sh", 251), ("lsf_psv", 886), ("aniso_psv", 559), ("aniso_sh", 403),
         ("layer_psv", 869), ("layer_sh", 532)] + \
    [(n, 200 if n == "rsf_slip" else 500) for n in conftest.FRICTION_CASES]


@pytest.mark.parametrize("name,nsteps", CASES)
def test_time_series(name, nsteps, oracle_api, oracle_ensemble):
    print(f"\n[case {name} nsteps {nsteps}]", flush=True)
    pb, ref, gpu = _pair(name, nsteps, oracle_api)
    rsf = pb.fault is not None and pb.fault["law"] == "RSF"
    members = [solver.Solver(pb, api=api, compact=False) for _, api in oracle_ensemble] if rsf else []
    for s in [ref, gpu] + members:
        s.step(nsteps)
    errs, bars, spread = compare(pb, gpu, ref, members)
    if not members and any(not v < bars[k] for k, v in errs.items()):
        # above 1e-10: the bar becomes the oracle's own spread under the ensemble's
        # reorderings, measured on this case
        members = [solver.Solver(pb, api=api, compact=False) for _, api in oracle_ensemble]
        for s in members:
            s.step(nsteps)
        errs, bars, spread = compare(pb, gpu, ref, members)
    conftest.report(name, errs, bars)
    print("  error  " + "  ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    if members:
        print("  spread " + "  ".join(f"{k}={v:.1e}" for k, v in spread.items()), flush=True)
    bad = {k: (v, bars[k]) for k, v in errs.items() if not v < bars[k]}
    assert not bad, bad


def test_rsf_ens_history_and_float32_records(oracle_api, oracle_ensemble):
    """The single-scenario template of the rate-and-state ensemble: slip and slip-rate
    error against the oracle at every record (every 10 steps), the ensemble's spread at
    the same records, and how many of the GPU's float32-rounded fault records differ
    from the reference's own FltXX_sem2d.dat (slip, slip rate, shear, friction)."""
    name, nt = "rsf_ens", 400
    pb, ref, gpu = _pair(name, nt, oracle_api)
    members = [solver.Solver(pb, api=api, compact=False) for _, api in oracle_ensemble]
    for s in [ref, gpu] + members:
        s.step(nt)
    rg, rr = gpu.fault_records(), ref.fault_records()
    rm = [m.fault_records() for m in members]
    oitd = pb.fault["oitd"]
    lines = []
    worst = {}
    for q, nm in ((0, "slip"), (1, "slip_rate")):
        e = np.array([rel(rg[k, q], rr[k, q]) for k in range(1, rr.shape[0])])
        sp = np.array([max(rel(m[k, q], rr[k, q]) for m in rm) for k in range(1, rr.shape[0])])
        worst[nm] = (e.max(), sp.max())
        lines.append(f"  {nm:9s} error by step " + " ".join(f"{(k + 1) * oitd}:{e[k]:.1e}" for k in range(0, len(e), 8)))
        lines.append(f"  {nm:9s} spread       " + " ".join(f"{(k + 1) * oitd}:{sp[k]:.1e}" for k in range(0, len(sp), 8)))
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    cols = [0, 1, 2, 4]
    ours = rg[:, cols].astype(np.float32)
    theirs = g["flt_f32_all"][:, cols]
    assert ours.shape == theirs.shape
    diff = ours != theirs
    ulps = np.abs(ours.view(np.int32).astype(np.int64) - theirs.view(np.int32).astype(np.int64))
    lines.append(f"  float32 records: {int(diff.sum())} of {diff.size} values differ from the reference's "
                 f"Flt05_sem2d.dat (per column {diff.sum(axis=(0, 2)).tolist()}), max {int(ulps.max())} float32 ulps")
    print("\n" + "\n".join(lines), flush=True)
    conftest.REPORT.append(("rsf_ens f32 records", "values_differing", float(diff.sum()), float(diff.size)))
    for nm, (e, sp) in worst.items():
        assert e <= max(TOL, sp), (nm, e, sp)


@pytest.mark.parametrize("name", ["sh_swf", "psv_rsf", "rsf_ens", "psv_kv", "twf_inplane"])
def test_fp32_bulk_rupture_vs_fp64_oracle(name, oracle_api):
    """FP32 bulk update with FP64 friction, state and fault records (the reduced-
    precision form of the rupture configurations) against the FP64 oracle: slip, slip
    rate and seismogram error reported, bar 1e-3 relative L2 (as the FP32 wave cases)."""
    nt = {"sh_swf": 600, "psv_rsf": 500, "rsf_ens": 400, "psv_kv": 500, "twf_inplane": 600}[name]
    pb, ref, g32 = _pair(name, nt, oracle_api, bulk_fp32=True)
    ref.step(nt)
    g32.step(nt)
    rr, rg = ref.fault_records(), g32.fault_records()
    errs = {"slip": rel(rg[:, 0], rr[:, 0]), "slip_rate": rel(rg[:, 1], rr[:, 1])}
    if pb.receivers is not None:
        errs["seismo"] = rel(g32.seismograms(), ref.seismograms())
    print(f"\n  FP32 bulk {name}: " + "  ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    conftest.report(name + " fp32", errs, {k: 1e-3 for k in errs})
    assert all(v < 1e-3 for v in errs.values()), errs


def test_stacey_general_path_and_fp32(oracle_api):
    """Stacey boundaries (BC_ABSORB stacey=T) through the general path (colour-sorted
    element kernel, k_abso then k_stacey per boundary) at the parity bar, and with FP32
    fields (structured path) against the FP64 oracle."""
    nt = 400
    pb = builder.build(case_par("psv_stacey"), nsteps=nt)
    assert all(ab["stacey"] is not None for ab in pb.abso)
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    gen = solver.Solver(pb, structured=False)
    g32 = solver.Solver(pb, bulk_fp32=True)
    for s in (ref, gen, g32):
        s.step(nt)
    e_gen = rel(gen.seismograms(), ref.seismograms())
    e_32 = rel(g32.seismograms(), ref.seismograms())
    print(f"\n  Stacey general path rel-L2 {e_gen:.2e}, FP32 structured {e_32:.2e}", flush=True)
    assert e_gen < TOL
    for x, y in zip(gen.fields(), ref.fields()):
        assert rel(x, y) < TOL
    assert e_32 < 1e-3


def test_mesh2d_fp32(oracle_api):
    """Unstructured MESH2D mesh (general-geometry elements, 10 coefficients, general path)
    with FP32 fields against the FP64 oracle."""
    nt = 500
    pb = builder.build(case_par("mesh2d_psv"), nsteps=nt)
    assert pb.nelast == 10
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    g32 = solver.Solver(pb, bulk_fp32=True)
    ref.step(nt)
    g32.step(nt)
    err = rel(g32.seismograms(), ref.seismograms())
    print(f"\n  MESH2D FP32 seismogram rel-L2 error vs FP64 oracle: {err:.3e}", flush=True)
    assert err < 1e-3


def test_deterministic(oracle_api):
    pb = builder.build(case_par("sh_swf"), nsteps=200)
    outs = []
    for _ in range(2):
        s = solver.Solver(pb)
        s.step(200)
        outs.append(s.fields())
        s.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_fp32_bulk_error(oracle_api):
    pb = builder.build(case_par("c1_psv"), nsteps=600)
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    g32 = solver.Solver(pb, bulk_fp32=True)
    ref.step(600)
    g32.step(600)
    err = rel(g32.seismograms(), ref.seismograms())
    print(f"\nFP32 bulk seismogram rel-L2 error vs FP64 oracle: {err:.3e}")
    assert err < 1e-3


def test_pipelined_tile_kernel(oracle_api):
    """The persistent pipelined tile kernel (runs of several tiles per workgroup, with
    the next tile's loads in flight) on a 400x200-element window of the heterogeneous
    throughput medium with its explosion source: FP64 vs the oracle at the parity bar,
    and bitwise equal to the one-workgroup-per-tile kernel for both launch widths and
    both bulk precisions (every variant sums each node's forces in the same order)."""
    from sem2dpack_amd import workloads
    nt = 40
    pb = workloads.build_throughput(nx=400, nz=200, lx=40e3, lz=20e3, nt=nt, per_element_coefs=True)
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    ref.step(nt)
    fr = ref.fields()
    for fp32 in (False, True):
        outs = {}
        for threads, per_cu in ((256, 0), (256, 1), (512, 1), (512, 2)):
            print(f"\n[pipe fp32={fp32} threads={threads} wg_per_cu={per_cu}]", flush=True)
            g = solver.Solver(pb, bulk_fp32=fp32)
            g.set_tile_config(threads, per_cu)
            g.step(nt)
            outs[(threads, per_cu)] = g.fields()
            g.close()
        base = outs[(256, 0)]
        for key, f in outs.items():
            for x, y in zip(f, base):
                assert np.array_equal(x, y), key
        if not fp32:
            for nm, x, y in zip("dva", base, fr):
                assert rel(x, y) < TOL, nm


@pytest.mark.parametrize("name,structured", [("c1_psv", True), ("mesh2d_psv", False), ("sh_swf", True)])
def test_progress_rows_and_watchdog(name, structured, oracle_api):
    """The progress rows of the time loop (it, it*dt, max|v|, max|d| every ItInfo steps,
    SRC/main.f90:73-76) against the oracle's, at the parity bar; then a non-finite value
    planted in the state from the host: the next row finds it and sem2d_step returns 21."""
    nt, itinfo = 300, 50
    pb = builder.build(case_par(name), nsteps=nt)
    ref = solver.Solver(pb, api=oracle_api, compact=False, itinfo=itinfo)
    gpu = solver.Solver(pb, structured=structured, itinfo=itinfo)
    ref.step(nt)
    gpu.step(nt)
    rr, rg = ref.info_rows(), gpu.info_rows()
    print(f"\n  {name}: {rg.shape[0]} progress rows, last {rg[-1].tolist()}", flush=True)
    assert rr.shape == rg.shape == (nt // itinfo, 4)
    assert np.array_equal(rg[:, :2], rr[:, :2])
    assert rel(rg[:, 2:], rr[:, 2:]) < TOL
    vm, dm, nb = gpu.field_info()
    assert nb == 0 and vm == rg[-1, 2] and dm == rg[-1, 3]
    d, v, a = gpu.fields()
    v[0, pb.npoin // 2] = np.nan
    gpu.set_fields(d, v, a)
    assert gpu.field_info()[2] == 1
    with pytest.raises(solver.SolverError, match=r"\(21\).*non-finite field values at step 350"):
        gpu.step(itinfo)


def test_dirichlet_after_fault_on_the_tile_path(oracle_api):
    """The TPV3-type case without Kelvin-Voigt damping runs on the fused tile kernels: its
    Dirichlet side, listed after the fault, zeroes the vertical force of the corner node
    after the fault solve (inside k_fault) and that of the other side nodes in
    k_deferred; FP64 vs the oracle at the parity bar."""
    from sem2dpack_amd import parfile
    txt = open(case_par("tpv3_kv")).read()
    txt = txt.replace("kind='ELAST','KV'", "kind='ELAST'").replace("&MAT_KV etaH='GAUSSIAN' /\n", "")
    txt = txt.replace("&DIST_GAUSSIAN length=1d6,282.094, ampli=0.1d0 /\n", "")
    pb = builder.build(parfile.parse_par(txt), nsteps=500)
    assert pb.kv_eta is None and pb.dirichlet and pb.dirichlet[0]["after_fault"]
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    gpu = solver.Solver(pb)
    ref.step(500)
    gpu.step(500)
    errs, bars, _ = compare(pb, gpu, ref)
    conftest.report("tpv3 no KV, tile path", errs, bars)
    print("\n  error  " + "  ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < bars[k] for k, v in errs.items()), errs


def test_ratestate_example_shipped_series():
    """The reference's full RateState example (803 steps) on the fused tile path
    against the fault series it ships with (tests/example_series.py): relative L2
    within the oracle's own distance from them, same peak steps."""
    from example_series import BARS, check, compare as compare_series
    pb = builder.build(case_par("ratestate"))
    gpu = solver.Solver(pb)
    gpu.step(pb.nt)
    rows = compare_series(pb, gpu.fault_records())
    for q, x, e, _, _ in rows:
        conftest.REPORT.append(("ratestate vs shipped series", f"{q}@{x / 1e3:.0f}km", e, BARS[q]))
    check(rows)


@pytest.mark.parametrize("name", ["testsh", "lamb"])
def test_analytic_examples(name):
    """The reference's analytic regression examples (TestSH, NGLL=6; LambsProblem,
    NGLL=9; leapfrog) on the GPU: float32 seismograms against the analytic ones at the
    examples' own pass marks (tests/analytic.py)."""
    from analytic import BARS, errors
    pb = builder.build(case_par(name))
    gpu = solver.Solver(pb)
    gpu.step(pb.nt)
    err = errors(name, gpu.seismograms())
    print(f"\n  {name}: " + "  ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
    for k, v in err.items():
        conftest.REPORT.append((f"{name} vs analytic", k, v, BARS[name]))
    assert all(v < BARS[name] for v in err.values()), err
