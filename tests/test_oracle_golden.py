"""The oracle (C restatement, oracle/sem2d_oracle.c) vs the reference's own FP64 outputs.

The oracle follows the reference statement by statement in the same floating-point
order, so the comparison is bit for bit: element forces, final fields, receiver
series, fault jumps; the reference's float32 fault file is matched to float32.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_CASES, case_par
from sem2dpack_amd import builder, solver
import refdump  # noqa: F401  (fixture readers live beside the tests)

CASES = GOLDEN_CASES


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def harness_d(pb):
    """The deterministic displacement of oracle/ref_harness.f90."""
    x, z = pb.coord[:, 0], pb.coord[:, 1]
    k = np.arange(1, pb.npoin + 1, dtype=np.float64)
    out = np.empty((pb.ndof, pb.npoin))
    for c in range(1, pb.ndof + 1):
        out[c - 1] = 1e-3 * np.sin(1e-3 * x * float(c) + 7e-4 * z) + 1e-4 * np.cos(0.37 * k + 1.1 * float(c))
    return out


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name, oracle_api):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    nt = int(g["nt"])
    pb = builder.build(case_par(name), nsteps=nt)
    s = solver.Solver(pb, api=oracle_api, compact=False)
    f = s.fint(harness_d(pb))
    idx = g["sample_idx"]
    # sin/cos of numpy vs the Fortran runtime may differ by an ulp: compare forces to 1e-13
    assert np.abs(f[:, idx] - g["fint_sample"]).max() <= 1e-13 * np.abs(g["fint_sample"]).max()
    s.step(nt)
    d, v, a = s.fields()
    assert np.array_equal(d[:, idx], g["final_d_sample"])
    assert np.array_equal(v[:, idx], g["final_v_sample"])
    assert np.array_equal(a[:, idx], g["final_a_sample"])
    if pb.receivers is not None and pb.receivers["atnode"]:
        assert np.array_equal(s.seismograms(), g["seismo"])
    elif pb.receivers is not None:
        # receivers inside elements: the reference keeps only its float32 samples of the
        # interpolated series; FP64 series rounded to float32 must match them exactly
        assert np.array_equal(s.seismograms().astype(np.float32), g["seis_f32"])
    if "flt_f32_all" in g:
        # decimated fault output: every oitd steps at nodes oix1:oixn:oixd
        rec = s.fault_records()
        assert rec.shape[0] == g["flt_f32_all"].shape[0] and rec.shape[2] == g["flt_f32_all"].shape[2]
        assert np.array_equal(rec[:, [0, 1, 2, 4]].astype(np.float32), g["flt_f32_all"][:, [0, 1, 2, 4]])
        if pb.fault.get("osides"):  # per-side rows after the six of each record
            assert np.array_equal(s.fault_sides().astype(np.float32), g["flt_f32_all"][:, 6:])
        oitd, oix1, oixd = int(g["flt_oitd"]), int(g["flt_oix1"]), int(g["flt_oixd"])
        st_every = int(g["fault_every"])
        assert st_every % oitd == 0
        if "fault_dd" in g:  # split nodes: slip from the reference's FP64 jumps
            sel = slice(oix1 - 1, pb.fault["oixn"], oixd)
            n1 = pb.fault["n1"][sel]
            dd = g["fault_dd"][:, :, sel]
            slip = n1[:, 1] * dd[:, 0] - n1[:, 0] * dd[:, 1] if pb.ndof == 2 else dd[:, 0]
            assert np.array_equal(rec[::st_every // oitd, 0], slip)
    elif "fault_node1" in g:
        st_every = int(g["fault_every"])
        rec = s.fault_records()  # every step (otd = dt in these cases)
        if "fault_dd" in g:  # split nodes: slip from the reference's FP64 jumps
            n1 = pb.fault["n1"]
            dd = g["fault_dd"]
            if pb.ndof == 2:
                slip = n1[:, 1] * dd[:, 0] - n1[:, 0] * dd[:, 1]
            else:
                slip = dd[:, 0]
            if pb.beta == 0.0:
                assert np.array_equal(rec[::st_every, 0], slip)
            else:
                # beta /= 0 (HHT-alpha): the fault's slip is dD + beta dt^2 dA, equal to the
                # jump of the corrected fields up to rounding (the float32 records below
                # are compared exactly)
                assert np.abs(rec[::st_every, 0] - slip).max() <= 1e-13 * np.abs(slip).max()
        if "flt_f32" in g:
            ref32 = g["flt_f32"]
            with np.errstate(over="ignore"):  # huge(1d0) friction (time weakening) -> +Inf, as real()
                ours = rec[::st_every].astype(np.float32)
            cols = [0, 1, 2, 4] + ([5] if pb.fault["law"] == "SWF" else [])
            assert np.array_equal(ours[:, cols], ref32[:, cols])
            if pb.fault.get("osides"):  # per-side rows after the six of each record
                assert np.array_equal(s.fault_sides()[::st_every].astype(np.float32), ref32[:, 6:])


@pytest.mark.parametrize("name", ["sh_swf", "psv_rsf", "rsf_ens", "sh_kv", "ratestate", "tpv3_kv", "tpv3_dirfirst",
                                  "vshape_25d", "psv_25d"])
def test_potency_file_matches_reference(name, oracle_api, tmp_path):
    """FltXX_potency_sem2d.tab written from the oracle's potency rows is byte-identical to
    the reference's file (BC_DYNFLT_potency sums, (6D24.16) formatting)."""
    from sem2dpack_amd import outputs
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    nt = int(g["nt"])
    pb = builder.build(case_par(name), nsteps=nt)
    s = solver.Solver(pb, api=oracle_api, compact=False)
    s.step(nt)
    pot = s.potency()
    assert pot.shape == (nt + 1, 2 * (pb.ndof + 1))
    tag = int(pb.fault["tags"][0])
    outputs.write_potency(str(tmp_path), tag, pot)
    ours = open(tmp_path / ("Flt%02d_potency_sem2d.tab" % tag)).read()
    assert ours == str(g["potency_txt"])


@pytest.mark.parametrize("name", ["small_psv", "sh_swf", "sides_sh1"])
def test_output_files_match_reference(name, oracle_api, tmp_path):
    """Seismogram and fault files written from the oracle's FP64 series are byte-compatible
    with the reference's float32 files."""
    from sem2dpack_amd import outputs
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    nt = int(g["nt"])
    pb = builder.build(case_par(name), nsteps=nt)
    s = solver.Solver(pb, api=oracle_api, compact=False)
    s.step(nt)
    r = pb.receivers
    outputs.write_seismograms(str(tmp_path), s.seismograms(), pb.dt, r["isamp"], r["coord"])
    back = outputs.read_seismograms(str(tmp_path), r["nsamp"], len(r["iglob"]), pb.ndof)
    assert np.array_equal(back, g["seis_f32"])
    if pb.fault is not None:
        sides = s.fault_sides() if pb.fault["osides"] else None
        outputs.write_fault(str(tmp_path), pb.fault, s.fault_records(), pb.dt, nt, sides=sides)
        nout = (pb.fault["oixn"] - pb.fault["oix1"]) // pb.fault["oixd"] + 1
        tag = int(pb.fault["tags"][0])
        f32 = outputs.read_fault(str(tmp_path / ("Flt%02d_sem2d.dat" % tag)), nout, g["flt_f32"].shape[1])
        assert np.array_equal(f32[::int(g["fault_every"])], g["flt_f32"])


def _header_cases():
    g = np.load(os.path.join(GOLDEN, "headers.npz"))
    return sorted({k.split(":")[0] for k in g.files})


@pytest.mark.parametrize("name", _header_cases())
def test_text_headers_match_reference(name, tmp_path):
    """SeisHeader_sem2d.hdr, FltXX_sem2d.hdr and FltXX_init_sem2d.tab written by
    sem2dpack_amd.outputs from the builder's problem are byte-identical to the files the
    reference writes (list-directed records: shortest digits, F or 1P E form)."""
    from sem2dpack_amd import outputs
    g = np.load(os.path.join(GOLDEN, "headers.npz"))
    pb = builder.build(case_par(name))
    if pb.receivers is not None:
        r = pb.receivers
        outputs.write_seismogram_header(str(tmp_path), pb.dt, r["isamp"], r["nsamp"], r["coord"])
    if pb.fault is not None:
        outputs.write_fault_headers(str(tmp_path), pb.fault, pb.nt)
    keys = [k for k in g.files if k.startswith(name + ":")]
    assert keys
    for k in keys:
        ours = open(tmp_path / k.split(":")[1]).read()
        assert ours == str(g[k]), k
