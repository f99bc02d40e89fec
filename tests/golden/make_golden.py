#!/usr/bin/env python3
"""Generate the golden fixtures of tests/golden/*.npz from the reference solver.

Runs oracle/_ref/ref_harness (the reference SEM2DPACK modules built from
/root/reference by oracle/build_ref.py, plus an FP64 dump driver) on every case
directory under tests/golden/cases/ and stores compact FP64 vectors:

  ibool_sha256, coord_sha256, rmass_sha256, npoin, dt   numbering/mass fingerprints
  rec_nodes, seismo[nsamp][ndof][nrec]                  FP64 receiver time series
                                                        (the field the case records)
  fint_d_sum, fint_sha256, fint_sample                  f = -K d for the harness's d
  final_{d,v,a}_sample, sample_idx                      final fields at sampled nodes
  fault_* (fault cases)                                 node pairing and FP64 jumps
                                                        d2-d1, v2-v1 of the fault-node
                                                        fields, every fault_every steps
                                                        (a fault with one side: its
                                                        nodes, fault_one_sided = 1)
  flt_f32 (fault cases)                                 the reference's own float32
                                                        FltXX_sem2d.dat records
  potency_txt (fault cases)                             the reference's own
                                                        FltXX_potency_sem2d.tab text
  seis_f32                                              the reference's own float32
                                                        U?_sem2d.dat traces

Usage: python tests/golden/make_golden.py [case ...]   (needs /root/reference)
"""
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import refdump  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def read_flt(path, nout, ndat=None):
    """Sequential unformatted float32 records with 4-byte markers (POST/python/sem2d_read_fault.py);
    ndat (rows per output time) from the FltXX_sem2d.hdr beside the file unless given."""
    if ndat is None:
        with open(path[:-4] + ".hdr") as fh:
            fh.readline()
            ndat = int(fh.readline().split()[1])
    raw = np.fromfile(path, dtype=np.float32)
    rec = nout + 2
    n = raw.size // rec
    return raw[:n * rec].reshape(n, rec)[:, 1:-1].reshape(-1, ndat, nout)


def run_case(name):
    src = os.path.join(HERE, "cases", name)
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        for f in os.listdir(src):
            if not f.endswith(".py"):
                shutil.copy(os.path.join(src, f), tmp)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {exe}"], cwd=tmp, capture_output=True, text=True)
        if "FATAL" in r.stdout or "ref_harness: done" not in r.stdout:
            raise RuntimeError(r.stdout[-2000:])
        g = refdump.read_grid(tmp)
        nd = refdump.read_nodes(tmp, g["ndof"])
        dd, ff = refdump.read_fint(tmp, g["npoin"], g["ndof"])
        fd, fv, fa = refdump.read_final(tmp, g["npoin"], g["ndof"])
        out = dict(npoin=g["npoin"], nelem=g["nelem"], ndof=g["ndof"], nt=g["nt"], dt=g["dt"],
                   ibool_sha256=sha(g["ibool"]), coord_sha256=sha(g["coord"]), rmass_sha256=sha(g["rmass"]),
                   rmass_sum=float(np.sum(1.0 / g["rmass"][0])))
        if g["nelem"] <= 4000:
            out["ibool"] = g["ibool"]
        idx = np.unique(np.linspace(0, g["npoin"] - 1, 997).astype(np.int64))
        out["sample_idx"] = idx
        out["fint_sha256"] = sha(ff)
        out["fint_sample"] = ff[:, idx]
        out["final_d_sample"], out["final_v_sample"], out["final_a_sample"] = fd[:, idx], fv[:, idx], fa[:, idx]
        nrec = nd["nrec"]
        out["rec_nodes"] = nd["nodes"][:nrec]
        # the harness records every step; keep the field the case's REC_LINE asks for
        import sem2dpack_amd.parfile as pf
        rl = pf.read_par(os.path.join(src, "Par.inp")).first("REC_LINE")
        field = str(rl.get("field", "V")).strip() if rl is not None else "V"
        isamp = int(rl.get("isamp", 1)) if rl is not None else 1
        series = {"D": nd["d"], "V": nd["v"], "A": nd["a"]}[field]
        out["seismo"] = series[::isamp, :, :nrec]
        # the reference's own float32 seismogram files (REC_write, direct access)
        names = ["Uy_sem2d.dat"] if g["ndof"] == 1 else ["Ux_sem2d.dat", "Uz_sem2d.dat"]
        if all(os.path.exists(os.path.join(tmp, n)) for n in names):
            nsamp = out["seismo"].shape[0]
            out["seis_f32"] = np.stack([np.fromfile(os.path.join(tmp, n), np.float32).reshape(nrec, nsamp).T
                                        for n in names], axis=1)
        has_dynflt = pf.read_par(os.path.join(src, "Par.inp")).first("BC_DYNFLT") is not None
        if 5 in g["bounds"] and 6 in g["bounds"] and has_dynflt:
            n1, n2 = g["bounds"][5]["node"], g["bounds"][6]["node"]
            keep = n1 != n2
            npf = int(keep.sum())
            out["fault_node1"], out["fault_node2"] = n1[keep], n2[keep]
            bd = pf.read_par(os.path.join(src, "Par.inp")).first("BC_DYNFLT")
            otd = float(bd.get("otd", 0.0))
            oitd = max(1, int(np.floor(otd / g["dt"] + 0.5)))
            # jumps every `every` steps, a multiple of the fault output interval
            every = oitd * max(1, g["nt"] // 40 // oitd)
            out["fault_every"] = every
            sel = slice(0, None, every)
            k1 = nrec + np.nonzero(keep)[0]
            k2 = nrec + len(n1) + np.nonzero(keep)[0]
            # displacement and velocity jumps d(node2) - d(node1), xz frame
            out["fault_dd"] = nd["d"][sel][:, :, k2] - nd["d"][sel][:, :, k1]
            out["fault_dv"] = nd["v"][sel][:, :, k2] - nd["v"][sel][:, :, k1]
            pot = os.path.join(tmp, "Flt05_potency_sem2d.tab")
            if os.path.exists(pot):
                with open(pot) as fh:
                    out["potency_txt"] = np.array(fh.read())
            flt = os.path.join(tmp, "Flt05_sem2d.dat")
            oxi = [int(x) for x in bd.get("oxi", [1, 2 ** 31 - 1, 1])]
            if os.path.exists(flt) and oitd == 1 and oxi[2] == 1:
                out["flt_f32"] = read_flt(flt, npf)[::every]
            elif os.path.exists(flt):
                # decimated output (every oitd steps at nodes oix1:oixn:oixd): keep all records
                nout = (min(oxi[1], npf) - oxi[0]) // oxi[2] + 1
                out["flt_f32_all"] = read_flt(flt, nout)
                out["flt_oitd"], out["flt_oix1"], out["flt_oixd"] = oitd, oxi[0], oxi[2]
        par = pf.read_par(os.path.join(src, "Par.inp"))
        for bd in par.all("BC_DEF"):
            # a fault with one side (BC_DEF tag = t, on a domain boundary): the harness dumps
            # no jumps for it; keep its node list, float32 records and potency text
            t = int(bd.get("tag", 0))
            if str(bd.get("kind", "")).upper() != "DYNFLT" or t <= 0 or t not in g["bounds"]:
                continue
            n1 = g["bounds"][t]["node"]
            out["fault_node1"] = n1
            out["fault_one_sided"] = 1
            every = max(1, g["nt"] // 40)
            out["fault_every"] = every
            pot = os.path.join(tmp, "Flt%02d_potency_sem2d.tab" % t)
            if os.path.exists(pot):
                with open(pot) as fh:
                    out["potency_txt"] = np.array(fh.read())
            flt = os.path.join(tmp, "Flt%02d_sem2d.dat" % t)
            fb = par.first("BC_DYNFLT")
            oxi = [int(x) for x in fb.get("oxi", [1, 2 ** 31 - 1, 1])]
            oitd = max(1, int(np.floor(float(fb.get("otd", 0.0)) / g["dt"] + 0.5)))
            if os.path.exists(flt) and oitd == 1 and oxi[2] == 1 and oxi[0] == 1:
                out["flt_f32"] = read_flt(flt, len(n1))[::every]
            elif os.path.exists(flt):
                nout = (min(oxi[1], len(n1)) - oxi[0]) // oxi[2] + 1
                out["flt_f32_all"] = read_flt(flt, nout)
                out["flt_oitd"], out["flt_oix1"], out["flt_oixd"] = oitd, oxi[0], oxi[2]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, "ok", os.path.getsize(os.path.join(HERE, name + ".npz")), "bytes")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    names = sys.argv[1:] or sorted(os.listdir(os.path.join(HERE, "cases")))
    for n in names:
        run_case(n)
