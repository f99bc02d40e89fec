"""Output files in the formats of the reference solver (what its post-processing reads).

* Seismograms (``REC_write``, ``SRC/receivers.f90:351-391``; header from ``REC_init``,
  ``SRC/receivers.f90:196-205``): ``Ux_sem2d.dat`` / ``Uz_sem2d.dat`` (P-SV) or
  ``Uy_sem2d.dat`` (SH), direct-access files with one record per station holding
  its ``nsamp`` single-precision samples, and ``SeisHeader_sem2d.hdr``.
* Fault records (``BC_DYNFLT_init`` / ``BC_DYNFLT_write``,
  ``SRC/bc_dynflt.f90:419-520, 751-778``): ``FltXX_sem2d.dat``, a sequential
  unformatted file whose records (4-byte length markers around each) are, per output
  time, slip, slip rate, shear traction, normal traction, friction coefficient and
  stick traction of nodes ``oix1:oixn:oixd`` in single precision
  (``POST/python/sem2d_read_fault.py`` reads it); ``FltXX_sem2d.hdr`` and
  ``FltXX_init_sem2d.tab``.
* Potency (``BC_DYNFLT_write`` / ``BC_DYNFLT_potency``, ``SRC/bc_dynflt.f90:757,
  832-855``): ``FltXX_potency_sem2d.tab``, one formatted ``(6D24.16)`` line per step.

Numbers in the text headers use Python formatting; the binary files are byte
compatible with the reference's.
"""
import os

import numpy as np


def write_seismograms(outdir, seis, dt, isamp, coords):
    """seis: [nsamp][ndof][nrec] (FP64 samples); coords: [nrec][2]."""
    nsamp, ndof, nrec = seis.shape
    s32 = seis.astype(np.float32)
    names = ["Uy_sem2d.dat"] if ndof == 1 else ["Ux_sem2d.dat", "Uz_sem2d.dat"]
    for c, nme in enumerate(names):
        with open(os.path.join(outdir, nme), "wb") as fh:
            for r in range(nrec):
                fh.write(np.ascontiguousarray(s32[:, c, r]).tobytes())
    with open(os.path.join(outdir, "SeisHeader_sem2d.hdr"), "w") as fh:
        fh.write(" DT NSAMP NSTA\n")
        fh.write(f" {float(np.float32(dt * isamp))!r} {nsamp} {nrec}\n")
        fh.write(" XSTA ZSTA\n")
        for x, z in coords:
            fh.write(f" {x!r} {z!r}\n")


def read_seismograms(outdir, nsamp, nrec, ndof):
    names = ["Uy_sem2d.dat"] if ndof == 1 else ["Ux_sem2d.dat", "Uz_sem2d.dat"]
    out = np.empty((nsamp, ndof, nrec), dtype=np.float32)
    for c, nme in enumerate(names):
        a = np.fromfile(os.path.join(outdir, nme), dtype=np.float32).reshape(nrec, nsamp)
        out[:, c, :] = a.T
    return out


def write_fault(outdir, fault, records, dt, nt):
    """records: [nsamp][6][nout] FP64 fault records (sem2d_fetch_fault)."""
    tag = int(fault["tags"][0])
    nout = records.shape[2]
    with open(os.path.join(outdir, f"Flt{tag:02d}_sem2d.dat"), "wb") as fh:
        marker = np.array([4 * nout], dtype=np.int32).tobytes()
        for rec in records.astype(np.float32):
            for q in range(6):
                fh.write(marker)
                fh.write(np.ascontiguousarray(rec[q]).tobytes())
                fh.write(marker)
    sel = slice(fault["oix1"] - 1, fault["oixn"], fault["oixd"])
    nsamp = (nt - fault["oit"]) // fault["oitd"] + 1
    with open(os.path.join(outdir, f"Flt{tag:02d}_sem2d.hdr"), "w") as fh:
        fh.write(" NPTS NDAT NSAMP DELT\n")
        fh.write(f" {nout} 6 {nsamp} {fault['odt']!r}\n")
        fh.write(" Slip:Slip_Rate:Shear_Stress:Normal_Stress:Friction:T_stick\n")
        fh.write(" XPTS ZPTS\n")
        for x, z in fault["coord"][sel]:
            fh.write(f" {x!r} {z!r}\n")
    with open(os.path.join(outdir, f"Flt{tag:02d}_init_sem2d.tab"), "w") as fh:
        for t0, tn, mu, b in zip(fault["T0"][sel, 0], fault["T0"][sel, 1], fault["MU0"][sel], fault["B"][sel, 0]):
            fh.write(f" {t0!r} {tn!r} {mu!r} {b!r}\n")


def fortran_d(x, w=24, d=16):
    """Fortran Dw.d edit descriptor: 0.ddd...D+ee (sign, leading zero, d digits)."""
    if x == 0.0:
        body = "0." + "0" * d + "D+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        exp = f"D{ex:+03d}" if abs(ex) < 100 else f"{ex:+04d}"
        body = ("-" if x < 0 else "") + "0." + digits + exp
    return body.rjust(w)


def write_potency(outdir, tag, pot):
    """FltXX_potency_sem2d.tab: one line per row, format (6D24.16) as BC_DYNFLT_write
    (SRC/bc_dynflt.f90:757); a row of 4 values (SH) fills one line of 4 fields."""
    with open(os.path.join(outdir, f"Flt{tag:02d}_potency_sem2d.tab"), "w") as fh:
        for row in pot:
            for k in range(0, len(row), 6):
                fh.write("".join(fortran_d(float(x)) for x in row[k:k + 6]) + "\n")


def read_potency(path):
    """Values of a potency file (D exponents), one row per line."""
    rows = []
    with open(path) as fh:
        for line in fh:
            rows.append([float(line[i:i + 24].replace("D", "E")) for i in range(0, len(line.rstrip("\n")), 24)])
    return rows


def read_fault(path, nout, ndat=6):
    """FltXX_sem2d.dat -> [nsamp][ndat][nout] float32 (cf. POST/python/sem2d_read_fault.py)."""
    raw = np.fromfile(path, dtype=np.float32)
    rec = nout + 2
    n = raw.size // rec
    return raw[:n * rec].reshape(n, rec)[:, 1:-1].reshape(-1, ndat, nout)
