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

The binary files (with the per-side rows of osides = T) and the text headers (list-directed ``WRITE(*,*)`` records, see
:func:`list_directed`) are byte compatible with the reference's.
"""
import decimal
import os

import numpy as np


def _shortest_digits(v, kind):
    """(digits, X) with |v| read back from 0.digits x 10**X, as the reference's Fortran
    runtime chooses them: the fewest significant digits n for which n-digit decimals lie
    strictly between the midpoints to |v|'s two neighbours, and among those the middle
    one ((lowest + highest) // 2 in units of the n-th digit)."""
    ftype = np.float32 if kind == 4 else np.float64
    a = ftype(abs(v))
    x = decimal.Decimal(float(a))
    below = decimal.Decimal(float(np.nextafter(a, ftype(0))))
    if a < np.finfo(ftype).max:
        above = decimal.Decimal(float(np.nextafter(a, ftype(np.inf))))
    else:  # the largest finite value: its rounding interval is as wide above as below
        above = 2 * x - below
    less, more = (x + below) / 2, (x + above) / 2
    e10 = x.adjusted()  # x = d.ddd x 10**e10
    with decimal.localcontext() as ctx:
        ctx.prec = 80
        for n in range(1, 40):
            unit = decimal.Decimal(1).scaleb(e10 - n + 1)
            lo = (less / unit).to_integral_value(rounding=decimal.ROUND_FLOOR) + 1
            hi = (more / unit).to_integral_value(rounding=decimal.ROUND_CEILING) - 1
            if lo <= hi:
                d = (lo + hi) // 2
                digits = str(int(d))
                X = e10 + 1 + (len(digits) - n)
                return digits.rstrip("0") or "0", X
    raise ValueError(v)


def _real_item(x, kind):
    """One REAL of kind 4 or 8 as the reference's Fortran runtime writes it in
    list-directed output: the digits of :func:`_shortest_digits`, in F form (no leading
    zero: ``.677``, ``25.``, ``81600000.``) while the decimal exponent X of 0.ddd x 10**X
    lies in 0 .. max(6, precision + 1) (7 for kind 4, 16 for kind 8), else in 1P E form
    (``7.914186712941471E-03``, ``5.E-04``)."""
    v = np.float32(x) if kind == 4 else np.float64(x)
    if v == 0:
        return "-0." if np.signbit(v) else "0."
    digits, X = _shortest_digits(v, kind)
    sign = "-" if v < 0 else ""
    if X < 0 or X > (7 if kind == 4 else 16):
        e = X - 1
        return f"{sign}{digits[0]}.{digits[1:]}E{'-' if e < 0 else '+'}{abs(e):02d}"
    if X >= len(digits):
        return f"{sign}{digits}{'0' * (X - len(digits))}."
    return f"{sign}{digits[:X]}.{digits[X:]}"


def list_directed(*items):
    """One record of list-directed output (``WRITE(unit,*) items``): every item after
    one blank; ints as integers, Python floats and numpy float64 as REAL(8), numpy
    float32 as REAL(4), strings as they are."""
    out = []
    for it in items:
        if isinstance(it, str):
            out.append(it)
        elif isinstance(it, (bool, np.bool_)):
            out.append("T" if it else "F")
        elif isinstance(it, (int, np.integer)):
            out.append(str(int(it)))
        elif isinstance(it, np.float32):
            out.append(_real_item(it, 4))
        else:
            out.append(_real_item(float(it), 8))
    return "".join(" " + o for o in out) + "\n"


def write_seismograms(outdir, seis, dt, isamp, coords):
    """seis: [nsamp][ndof][nrec] (FP64 samples); coords: [nrec][2]."""
    nsamp, ndof, nrec = seis.shape
    s32 = seis.astype(np.float32)
    names = ["Uy_sem2d.dat"] if ndof == 1 else ["Ux_sem2d.dat", "Uz_sem2d.dat"]
    for c, nme in enumerate(names):
        with open(os.path.join(outdir, nme), "wb") as fh:
            for r in range(nrec):
                fh.write(np.ascontiguousarray(s32[:, c, r]).tobytes())
    write_seismogram_header(outdir, dt, isamp, nsamp, coords)


def write_seismogram_header(outdir, dt, isamp, nsamp, coords):
    """SeisHeader_sem2d.hdr of REC_init (SRC/receivers.f90:214-221): real(tsamp), the
    number of samples and stations, then the station coordinates."""
    nrec = len(coords)
    with open(os.path.join(outdir, "SeisHeader_sem2d.hdr"), "w") as fh:
        fh.write(list_directed("DT NSAMP NSTA"))
        fh.write(list_directed(np.float32(dt * isamp), nsamp, nrec))
        fh.write(list_directed("XSTA ZSTA"))
        for x, z in coords:
            fh.write(list_directed(float(x), float(z)))


def read_seismograms(outdir, nsamp, nrec, ndof):
    names = ["Uy_sem2d.dat"] if ndof == 1 else ["Ux_sem2d.dat", "Uz_sem2d.dat"]
    out = np.empty((nsamp, ndof, nrec), dtype=np.float32)
    for c, nme in enumerate(names):
        a = np.fromfile(os.path.join(outdir, nme), dtype=np.float32).reshape(nrec, nsamp)
        out[:, c, :] = a.T
    return out


def write_fault(outdir, fault, records, dt, nt, sides=None):
    """records: [nsamp][6][nout] FP64 fault records (sem2d_fetch_fault); sides, with
    osides = T: [nsamp][4*ndof][nout] per-side records (sem2d_fetch_fault_sides), written
    after the six rows of each output time (BC_DYNFLT_write, SRC/bc_dynflt.f90:769-774)."""
    tag = int(fault["tags"][0])
    nout = records.shape[2]
    rows = records if sides is None else np.concatenate([records, sides], axis=1)
    with np.errstate(over="ignore"):  # real() of huge(1d0) (time-weakening friction) is +Inf
        rows32 = rows.astype(np.float32)
    with open(os.path.join(outdir, f"Flt{tag:02d}_sem2d.dat"), "wb") as fh:
        marker = np.array([4 * nout], dtype=np.int32).tobytes()
        for rec in rows32:
            for q in range(rec.shape[0]):
                fh.write(marker)
                fh.write(np.ascontiguousarray(rec[q]).tobytes())
                fh.write(marker)
    write_fault_headers(outdir, fault, nt)


def write_fault_headers(outdir, fault, nt):
    """FltXX_sem2d.hdr and FltXX_init_sem2d.tab of BC_DYNFLT_init
    (SRC/bc_dynflt.f90:446-452,479-500) for a run of nt steps."""
    tag = int(fault["tags"][0])
    sel = slice(fault["oix1"] - 1, fault["oixn"], fault["oixd"])
    nout = (fault["oixn"] - fault["oix1"]) // fault["oixd"] + 1
    nsamp = (nt - fault["oit"]) // fault["oitd"] + 1
    ndof = fault["B"].shape[1]
    names = "Slip:Slip_Rate:Shear_Stress:Normal_Stress:Friction:T_stick"
    ndat = 6
    if fault.get("osides"):
        ndat += 4 * ndof
        names += ":D1t:D2t:V1t:V2t" if ndof == 1 else ":D1t:D1n:D2t:D2n:V1t:V1n:V2t:V2n"
    with open(os.path.join(outdir, f"Flt{tag:02d}_sem2d.hdr"), "w") as fh:
        fh.write(list_directed("NPTS NDAT NSAMP DELT"))
        fh.write(list_directed(nout, ndat, nsamp, float(fault["odt"])))
        fh.write(" " + names + "\n")
        fh.write(list_directed("XPTS ZPTS"))
        for x, z in fault["coord"][sel]:
            fh.write(list_directed(float(x), float(z)))
    with open(os.path.join(outdir, f"Flt{tag:02d}_init_sem2d.tab"), "w") as fh:
        for t0, tn, mu, b in zip(fault["T0"][sel, 0], fault["T0"][sel, 1], fault["MU0"][sel], fault["B"][sel, 0]):
            fh.write(list_directed(float(t0), float(tn), float(mu), float(b)))


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
