"""The drop-in boundary from Fortran: tests/native/fortran_driver (a host program that
uses the ISO_C_BINDING module sem2dpack_amd/fortran/sem2d_hip.f90, as a SEM2DPACK main
program would) creates the solver from the builder's arrays, registers the absorbing
boundaries, the point force and the receivers, steps and fetches.  Its fields and
seismograms must equal those of the same calls made through ctypes (bitwise: the same
library on the same inputs) and match the oracle at the parity bar."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, case_par
from parity import TOL, rel
from sem2dpack_amd import builder, solver

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "native", "fortran_driver")


def write_case(path, pb, nsteps):
    r = pb.receivers
    m = pb.mesh
    cells = None
    if m.get("nx", 0) > 0:
        cells = np.empty(m["nx"] * m["nz"], dtype=np.int32)
        cells[pb.elem_iz * m["nx"] + pb.elem_ix] = np.arange(1, pb.nelem + 1, dtype=np.int32)
    srcs = [s for s in pb.sources if s["kind"] == "force"]
    assert len(srcs) == len(pb.sources), "the driver registers point forces only"
    coef_mode = 1 if pb.coef_mode == "shared" else 0
    hdr = np.array([pb.ndof, pb.nelem, pb.npoin, pb.nelast, coef_mode, m.get("nx", 0), m.get("nz", 0),
                    max(m.get("ezflt", 0), 0), nsteps, len(pb.abso), len(srcs),
                    len(r["iglob"]) if r is not None else 0, r["isamp"] if r is not None else 1,
                    solver.Solver.FIELD_CODE[r["field"]] if r is not None else 1,
                    r["nsamp"] if r is not None else 0], dtype=np.int32)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(np.array([pb.dt, pb.beta, pb.gamma]).tobytes())
        fh.write(np.ascontiguousarray(pb.ibool, dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(pb.hprime.T).tobytes())
        fh.write(np.ascontiguousarray(pb.rmass).tobytes())
        fh.write(np.ascontiguousarray(pb.coefs, dtype=np.float64).tobytes())
        if cells is not None:
            fh.write(cells.tobytes())
        for ab in pb.abso:
            node = np.ascontiguousarray(ab["node"], dtype=np.int32)
            fh.write(np.int32(len(node)).tobytes())
            fh.write(node.tobytes())
            fh.write(np.ascontiguousarray(ab["C"].T).tobytes())
        for s in srcs:
            tab = np.ascontiguousarray(s["stf_table"], dtype=np.float64)
            fh.write(np.int32(s["iglob"]).tobytes())
            fh.write(np.ascontiguousarray(s["dir"], dtype=np.float64).tobytes())
            fh.write(np.int32(len(tab)).tobytes())
            fh.write(tab.tobytes())
        if r is not None:
            fh.write(np.ascontiguousarray(r["iglob"], dtype=np.int32).tobytes())


@pytest.mark.skipif(not os.path.exists(EXE), reason="fortran_driver not built (flang missing)")
@pytest.mark.parametrize("name,nsteps", [("c1_psv", 2000), ("multi_src", 300)])
def test_fortran_host_program(name, nsteps, oracle_api, tmp_path):
    pb = builder.build(case_par(name), nsteps=nsteps)
    if pb.coef_mode != "shared":
        pb = builder.build(case_par(name), nsteps=nsteps, per_element_coefs=True)
    write_case(str(tmp_path / "case.bin"), pb, nsteps)
    print(f"\n[fortran driver {name} {nsteps} steps]", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0 and "fortran_driver: done" in r.stdout, r.stdout + r.stderr
    n = pb.ndof * pb.npoin
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    d, v, a = (out[k * n:(k + 1) * n].reshape(pb.ndof, pb.npoin) for k in range(3))
    rr = pb.receivers
    seis = out[3 * n:].reshape(-1, pb.ndof, len(rr["iglob"]))
    # the same calls through ctypes (compact coefficients off: the driver passes the
    # reference's per-element arrays)
    g = solver.Solver(pb, compact=False)
    g.step(nsteps)
    for x, y in zip((d, v, a), g.fields()):
        assert np.array_equal(x, y)
    assert np.array_equal(seis, g.seismograms())
    ref = solver.Solver(pb, api=oracle_api, compact=False)
    ref.step(nsteps)
    errs = {nm: rel(x, y) for nm, x, y in zip("dva", (d, v, a), ref.fields())}
    errs["seismo"] = rel(seis, ref.seismograms())
    print("  error vs oracle  " + "  ".join(f"{k}={e:.2e}" for k, e in errs.items()), flush=True)
    assert all(e < TOL for e in errs.values()), errs
