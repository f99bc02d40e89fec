"""Host-side problem set-up for structured (``MESH_CART``) SEM2DPACK inputs.

This module restates, in numpy plus one native routine (RCM), the part of the
reference initialisation ``init_main`` (``SRC/init.f90:16-131``) that produces
the arrays consumed by the explicit time step.  Every array keeps the
reference's meaning, units and (1-based) node numbering so that outputs can be
compared node by node with the reference:

=====================  =====================================================
reference routine      here
=====================  =====================================================
``CART_build``         :func:`_cart_mesh`  (``SRC/mesh_cartesian.f90:219-314``)
``MESH_STRUCTURED_*``  :func:`_cart_mesh`, :func:`_rcm_perm`
                       (``SRC/mesh_structured.f90:12-269``, ``SRC/rcm.f90``)
``SE_init_numbering``  :func:`_numbering` (``SRC/spec_grid.f90:198-314``)
``SE_init_coord``      :func:`_numbering` (``SRC/spec_grid.f90:321-374``)
``get_GLL_info``       :func:`gll_info` (``SRC/gll.f90:19-686``)
``SE_Jacobian`` etc.   :func:`_jacobians` (``SRC/spec_grid.f90:560-660``)
``CHECK_grid``         :func:`_grid_cfl` (``SRC/init.f90:145-289``)
``TIME_init``          :func:`build` (``SRC/time.f90:323-395``)
``MAT_ELAST_init_a``   :func:`_elastic_coefs` (``SRC/mat_elastic.f90:290-360``)
``MAT_MASS_init``      :func:`build` (``SRC/mat_mass.f90:29-61``)
``BC_ABSO_init``       :func:`_abso_init` (``SRC/bc_abso.f90:115-266``)
``BC_DYNFLT_init``     :func:`_dynflt_init` (``SRC/bc_dynflt.f90:231-520``)
``SO_init``            :func:`_sources` (``SRC/src_gen.f90:216-262``)
``REC_init``           :func:`_receivers` (``SRC/receivers.f90:168-303``)
=====================  =====================================================

Floating-point expressions follow the reference's operation order where the
order is visible in the source (shape functions, Jacobians, coefficient
products, mass assembly order, GLL iterations), so that numbering, coordinates,
time step, mass and coefficients agree bit for bit with the reference.
"""
import ctypes
import math
import sys
import os

import numpy as np

from . import parfile as _pf
from ._lib import load_host_lib

PI = 3.141592653589793  # SRC/constants.f90:44
TINY_XABS = 1e-3        # SRC/constants.f90:36

# side and edge identifiers of SRC/fem_grid.f90:74-87
EDGE_D, EDGE_R, EDGE_U, EDGE_L = 1, 2, 3, 4
SIDE_D, SIDE_R, SIDE_U, SIDE_L = 1, 2, 3, 4
FAULT_D, FAULT_U = 5, 6


class InputError(ValueError):
    """Raised where the reference would stop with IO_abort (SRC/stdio.f90:205-214)."""


# ---------------------------------------------------------------------------
# GLL quadrature (Gauss-Lobatto-Jacobi points of Fischer & Patera's spectral
# library, as used by SRC/gll.f90), restated so that the nodes, weights and
# derivative matrix agree bit for bit with the reference.
# ---------------------------------------------------------------------------
def _gammaf(x):
    """gammaf (SRC/gll.f90:205-226): table of the values the quadrature needs."""
    sp = math.sqrt(PI)
    table = {-0.5: -2.0 * sp, 0.5: sp, 1.0: 1.0, 2.0: 1.0, 1.5: sp / 2.0, 2.5: 1.5 * sp / 2.0,
             3.5: 2.5 * 1.5 * sp / 2.0, 3.0: 2.0, 4.0: 6.0, 5.0: 24.0, 6.0: 120.0}
    return table.get(x, 1.0)


def _jacobf(n, alp, bet, x):
    """Jacobi polynomial P_n and derivative with the two previous orders (SRC/gll.f90:377-418)."""
    apb = alp + bet
    poly, pder = 1.0, 0.0
    psave = pdsave = 0.0
    if n == 0:
        return poly, pder, 0.0, 0.0, 0.0, 0.0
    polyl, pderl = poly, pder
    poly = (alp - bet + (apb + 2.0) * x) / 2.0
    pder = (apb + 2.0) / 2.0
    if n == 1:
        return poly, pder, 0.0, 0.0, 0.0, 0.0
    for k in range(2, n + 1):
        dk = float(k)
        a1 = 2.0 * dk * (dk + apb) * (2.0 * dk + apb - 2.0)
        a2 = (2.0 * dk + apb - 1.0) * (alp * alp - bet * bet)
        b3 = (2.0 * dk + apb - 2.0)
        a3 = b3 * (b3 + 1.0) * (b3 + 2.0)
        a4 = 2.0 * (dk + alp - 1.0) * (dk + bet - 1.0) * (2.0 * dk + apb)
        polyn = ((a2 + a3 * x) * poly - a4 * polyl) / a1
        pdern = ((a2 + a3 * x) * pder - a4 * pderl + a3 * poly) / a1
        psave, pdsave = polyl, pderl
        polyl, poly = poly, polyn
        pderl, pder = pder, pdern
    return poly, pder, polyl, pderl, psave, pdsave


def _jacg(np_, alpha, beta):
    """Gauss-Jacobi points by Newton iteration with deflation (SRC/gll.f90:300-375)."""
    xjac = [0.0] * np_
    n = np_ - 1
    dth = 4.0 * math.atan(1.0) / (2.0 * float(n) + 2.0)
    xlast = 0.0
    for j in range(1, np_ + 1):
        if j == 1:
            x = math.cos((2.0 * (float(j) - 1.0) + 1.0) * dth)
        else:
            x1 = math.cos((2.0 * (float(j) - 1.0) + 1.0) * dth)
            x = (x1 + xlast) / 2.0
        for _ in range(10):
            p, pd = _jacobf(np_, alpha, beta, x)[:2]
            recsum = 0.0
            for i in range(1, j):
                recsum = recsum + 1.0 / (x - xjac[np_ - i])
            delx = -p / (pd - recsum * p)
            x = x + delx
            if abs(delx) < 1.0e-12:
                break
        xjac[np_ - j] = x
        xlast = x
    for i in range(np_):  # selection sort, ascending
        xmin, jmin = 2.0, i
        for j in range(i, np_):
            if xjac[j] < xmin:
                xmin, jmin = xjac[j], j
        if jmin != i:
            xjac[i], xjac[jmin] = xjac[jmin], xjac[i]
    return xjac


def _pnormj(n, alpha, beta):
    """SRC/gll.f90:476-507."""
    dn = float(n)
    const = alpha + beta + 1.0
    if n <= 1:
        prod = _gammaf(dn + alpha) * _gammaf(dn + beta)
        prod = prod / (_gammaf(dn) * _gammaf(dn + alpha + beta))
        return prod * 2.0 ** const / (2.0 * dn + const)
    prod = _gammaf(alpha + 1.0) * _gammaf(beta + 1.0)
    prod = prod / (2.0 * (1.0 + const) * _gammaf(const + 1.0))
    prod = prod * (1.0 + alpha) * (2.0 + alpha)
    prod = prod * (1.0 + beta) * (2.0 + beta)
    for i in range(3, n + 1):
        di = float(i)
        frac = (di + alpha) * (di + beta) / (di * (di + alpha + beta))
        prod = prod * frac
    return prod * 2.0 ** const / (2.0 * dn + const)


def _endw(n, alpha, beta, which):
    """endw1 / endw2 (SRC/gll.f90:79-203)."""
    apb = alpha + beta
    if n == 0:
        return 0.0
    if which == 1:
        f1 = _gammaf(alpha + 2.0) * _gammaf(beta + 1.0) / _gammaf(apb + 3.0)
    else:
        f1 = _gammaf(alpha + 1.0) * _gammaf(beta + 2.0) / _gammaf(apb + 3.0)
    f1 = f1 * (apb + 2.0) * 2.0 ** (apb + 2.0) / 2.0
    if n == 1:
        return f1
    if which == 1:
        fint1 = _gammaf(alpha + 2.0) * _gammaf(beta + 1.0) / _gammaf(apb + 3.0)
    else:
        fint1 = _gammaf(alpha + 1.0) * _gammaf(beta + 2.0) / _gammaf(apb + 3.0)
    fint1 = fint1 * 2.0 ** (apb + 2.0)
    fint2 = _gammaf(alpha + 2.0) * _gammaf(beta + 2.0) / _gammaf(apb + 4.0)
    fint2 = fint2 * 2.0 ** (apb + 3.0)
    if which == 1:
        f2 = (-2.0 * (beta + 2.0) * fint1 + (apb + 4.0) * fint2) * (apb + 3.0) / 4.0
    else:
        f2 = (2.0 * (alpha + 2.0) * fint1 - (apb + 4.0) * fint2) * (apb + 3.0) / 4.0
    if n == 2:
        return f2
    f3 = 0.0
    for i in range(3, n + 1):
        di = float(i - 1)
        abn = alpha + beta + di
        abnn = abn + di
        a1 = -(2.0 * (di + alpha) * (di + beta)) / (abn * abnn * (abnn + 1.0))
        a2 = (2.0 * (alpha - beta)) / (abnn * (abnn + 2.0))
        a3 = (2.0 * (abn + 1.0)) / ((abnn + 2.0) * (abnn + 1.0))
        f3 = -(a2 * f2 + a1 * f1) / a3
        f1, f2 = f2, f3
    return f3


def _pnleg(z, n):
    p1, p2, p3 = 1.0, z, z
    for k in range(1, n):
        fk = float(k)
        p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0)
        p1, p2 = p2, p3
    return p3


def _pndleg(z, n):
    p1, p2, p1d, p2d, p3d = 1.0, z, 0.0, 1.0, 1.0
    for k in range(1, n):
        fk = float(k)
        p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0)
        p3d = ((2.0 * fk + 1.0) * p2 + (2.0 * fk + 1.0) * z * p2d - fk * p1d) / (fk + 1.0)
        p1, p2, p1d, p2d = p2, p3, p2d, p3d
    return p3d


def gll_info(n=5):
    """get_GLL_info (SRC/gll.f90:19-36): nodes x, weights w, H[i, j] = h'_i(x_j)."""
    z = [0.0] * n
    w = [0.0] * n
    nm1 = n - 2
    if nm1 == 1:  # zwgjd with one point (SRC/gll.f90:594-598), alpha = beta = 1
        z[1] = (1.0 - 1.0) / (2.0 + 2.0)
        w[1] = _gammaf(2.0) * _gammaf(2.0) / _gammaf(4.0) * 2.0 ** (2.0 + 1.0)
    elif nm1 > 0:  # zwgjd on the interior with alpha = beta = 1
        zi = _jacg(nm1, 1.0, 1.0)
        np1, np2 = nm1, nm1 + 1
        dnp1, dnp2 = float(np1), float(np2)
        fac1 = dnp1 + 1.0 + 1.0 + 1.0
        fac2 = fac1 + dnp1
        fac3 = fac2 + 1.0
        fnorm = _pnormj(np1, 1.0, 1.0)
        rcoef = (fnorm * fac2 * fac3) / (2.0 * fac1 * dnp2)
        for i in range(nm1):
            p, pd, pm1, pdm1, _, _ = _jacobf(np2, 1.0, 1.0, zi[i])
            z[i + 1] = zi[i]
            w[i + 1] = -rcoef / (p * pdm1)
    z[0], z[n - 1] = -1.0, 1.0
    for i in range(1, n - 1):
        w[i] = w[i] / (1.0 - z[i] * z[i])
    N = n - 1
    pd = _jacobf(N, 0.0, 0.0, z[0])[1]
    w[0] = _endw(N, 0.0, 0.0, 1) / (2.0 * pd)
    pd = _jacobf(N, 0.0, 0.0, z[n - 1])[1]
    w[n - 1] = _endw(N, 0.0, 0.0, 2) / (2.0 * pd)
    if n % 2 != 0:
        z[(n - 1) // 2] = 0.0
    H = np.zeros((n, n))
    for ix in range(n):
        for ip in range(n):
            i, j = ip, ix
            if i == 0 and j == 0:
                v = -float(N) * (float(N) + 1.0) / 4.0
            elif i == N and j == N:
                v = float(N) * (float(N) + 1.0) / 4.0
            elif i == j:
                v = 0.0
            else:
                r1, r2, r3 = _pnleg(z[j], N), _pndleg(z[j], N), _pnleg(z[i], N)
                v = r1 / (r3 * (z[j] - z[i])) + (1.0 - z[j] * z[j]) * r2 / (
                    float(N) * (float(N) + 1.0) * r3 * (z[j] - z[i]) * (z[j] - z[i]))
            H[ip, ix] = v
    return np.array(z), np.array(w), H


def gll5():
    return gll_info(5)


# ---------------------------------------------------------------------------
# Problem container
# ---------------------------------------------------------------------------
class Problem:
    """Arrays of one SEM problem, reference numbering and meaning.

    Layouts (C order, equal to the reference's Fortran column-major memory):
      ibool[nelem, ngll, ngll]   ibool[e, j, i] = reference ibool(i, j, e), 1-based
      coord[npoin, 2]            reference coord(:, k)
      rmass[ndof, npoin]         reference rmass(k, c) (inverse of the assembled mass)
      coefs                      shared:      [nelast, ngll, ngll]      (a(i,j,k) -> [k, j, i])
                                 per element: [nelem, nelast, ngll, ngll]
      hprime[ngll, ngll]         hprime[i, j] = h'_i(x_j)
    """

    def __init__(self):
        self.sources = []
        self.abso = []
        self.fault = None
        self.receivers = None


# ---------------------------------------------------------------------------
# mesh + numbering
# ---------------------------------------------------------------------------
def _cart_mesh(g, domains=()):
    """CART_read (SRC/mesh_cartesian.f90:86-191): the box, its fault row and fault zone,
    and the MESH_CART_DOMAIN subdomains (tag, element ranges ex, ez, 1-based)."""
    xlim = g.get("xlim")
    zlim = g.get("zlim")
    nel = g.get("nelem")
    if xlim is None or zlim is None:
        raise InputError("CART_read: you must set xlim and zlim")
    nx, nz = int(nel[0]), int(nel[1])
    if nx <= 0 or nz <= 0:
        raise InputError("CART_read: nelem must be positive")
    ezflt = int(g.get("ezflt", 0))
    if g.get("faultx", False):
        ezflt = -1
    if ezflt >= nz:
        raise InputError("CART_read: ezflt must be < nelem(2)")
    if ezflt == -1:
        ezflt = nz // 2
    if ezflt < -1:
        raise InputError("CART_read: ezflt must be >= -1")
    if g.get("split", False):
        raise InputError("MESH_CART split is not supported by this build")
    doms = []
    for d in domains:
        tag = int(d.get("tag", 0))
        ex = [int(v) for v in d.get("ex", [0, 0])]
        ez = [int(v) for v in d.get("ez", [0, 0])]
        if tag < 1:
            raise InputError("CART_read: tag null, negative or missing")
        if not all(1 <= v <= nx for v in ex) or not all(1 <= v <= nz for v in ez):
            raise InputError("CART_read: ex or ez out of bounds or missing")
        doms.append((tag, ex, ez))
    if not doms:  # a single domain spans the whole box
        doms = [(1, [1, nx], [1, nz])]
    return dict(xmin=float(xlim[0]), xmax=float(xlim[1]), zmin=float(zlim[0]), zmax=float(zlim[1]),
                nx=nx, nz=nz, ezflt=ezflt, fztag=int(g.get("fztag", 0)), fznz=int(g.get("fznz", 1)),
                domains=doms)


def _layered_mesh(par, g, fname_dir=None):
    """MESH_LAYERS_read / MESH_LAYERS_build (SRC/mesh_layers.f90:39-113,161-230,266-330,
    340-470) for four-node elements: layer tops that are constant (ztop) or DIST_LINEAR
    profiles (ztopH='LINEAR'), from MESH_LAYER blocks (the first block is the top layer,
    default tag = its position) or from a file of 'ztop nz tag' lines (top layer first).
    Control rows: row 0 at zmin, then in each layer (bottom up) the rows
    bot + j/nz (top - bot), j = 1..nz; the row on top of element row ezflt is doubled."""
    xlim, zmin = g.get("xlim"), g.get("zmin")
    if xlim is None:
        raise InputError("MESH_LAYERS_read: you must set xlim")
    if zmin is None:
        raise InputError("MESH_LAYERS_read: you must set zmin")
    nx, nlayer = int(g.get("nx", 1)), int(g.get("nlayer", 0))
    fname = str(g.get("file", "")).strip()
    ezflt, fztag, fznz = int(g.get("ezflt", 0)), int(g.get("fztag", 0)), int(g.get("fznz", 1))
    if nx <= 0:
        raise InputError("MESH_LAYERS_read: nx must be positive")
    if nlayer <= 0 and not fname:
        raise InputError("MESH_LAYERS_read: nlayer or file must be set")
    if ezflt < -1:
        raise InputError("MESH_LAYERS_read: ezflt must be >= -1")
    if ezflt == 0:
        fztag = 0
    if fztag < 0:
        raise InputError("MESH_LAYERS_read: fztag must be positive")
    if fznz < 1:
        raise InputError("MESH_LAYERS_read: fznz must be strictly positive")
    layers = []  # top layer first, as read
    if fname:
        path = os.path.join(fname_dir, fname) if fname_dir else fname
        with open(path) as fh:
            rows = [ln.replace("d", "e").replace("D", "e").split() for ln in fh if ln.strip()]
        for r in rows:
            layers.append(dict(top=("const", float(r[0])), nez=int(float(r[1])), tag=int(float(r[2]))))
    else:
        blocks = par.all("MESH_LAYER")
        if len(blocks) < nlayer:
            raise InputError("MESH_LAYER input block not found")
        for pos, b in enumerate(blocks[:nlayer]):
            zh = str(b.get("ztoph", "")).strip()
            if zh and zh.upper() != "LINEAR":
                # QSPLINE, CSPLINE and the other distributions need 8- or 9-node elements
                raise InputError(f"MESH_LAYER ztopH={zh!r} (curved layer interfaces) is not supported "
                                 "by this build")
            if not zh and b.get("ztop") is None:
                raise InputError("MESH_LAYERS_read: ztop or ztopH must be set")
            top = _read_cd(par, iter(par.next_groups(b)), b.get("ztop", 0.0), zh)
            layers.append(dict(top=top, nez=int(b.get("nz", 0)), tag=int(b.get("tag", pos + 1))))
    for L in layers:
        if L["nez"] <= 0:
            raise InputError("MESH_LAYERS_read: nz null, negative or missing")
        if L["tag"] < 1:
            raise InputError("MESH_LAYERS_read: tag null, negative or missing")
    layers = layers[::-1]  # bottom layer first
    nz = sum(L["nez"] for L in layers)
    if ezflt == -1:
        ezflt = nz // 2
    if ezflt >= nz:
        raise InputError("MESH_LAYERS_read: ezflt must be < nz")
    xmin, xmax, zmin = float(xlim[0]), float(xlim[1]), float(zmin)
    nxp = nx + 1
    xc = xmin + (xmax - xmin) / float(nxp - 1) * np.arange(nxp, dtype=np.float64)
    zrows = [np.full(nxp, zmin)]
    doms = []
    jlast = 0
    # a layer's rows interpolate between the previous layer's top profile (not its last
    # computed row, which may differ from it in the last bit) and its own top
    bot = zrows[0]
    for L in layers:
        top = _gen_cd(L["top"], np.stack([xc, np.zeros(nxp)], axis=1))
        if (top < zmin).any():
            raise InputError("MESH_LAYERS_build: zmin in MESH_LAYERED should be deeper than all "
                             "layer interfaces")
        for j in range(1, L["nez"] + 1):
            zrows.append(bot + float(j) / float(L["nez"]) * (top - bot))
            if len(zrows) - 1 == ezflt:
                zrows.append(zrows[-1])
        doms.append((L["tag"], [1, nx], [jlast + 1, jlast + L["nez"]]))
        jlast += L["nez"]
        bot = top
    zc = np.array(zrows)
    return dict(kind="layered", xmin=xmin, xmax=xmax, zmin=zmin, nx=nx, nz=nz, ezflt=max(ezflt, 0),
                fztag=fztag, fznz=fznz, domains=doms, ctrl_x=np.broadcast_to(xc, zc.shape).copy(),
                ctrl_z=zc)


def _rcm_perm(nx, nz):
    """Element permutation of MESH_STRUCTURED_renumber (SRC/mesh_structured.f90:204-269)."""
    jj, ii = np.meshgrid(np.arange(1, nz + 1), np.arange(1, nx + 1), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    offs = [(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]
    cols = []
    valid = []
    for di, dj in offs:
        i2, j2 = ii + di, jj + dj
        ok = (i2 >= 1) & (i2 <= nx) & (j2 >= 1) & (j2 <= nz)
        cols.append((j2 - 1) * nx + i2)
        valid.append(ok)
    cols = np.stack(cols, axis=1)
    valid = np.stack(valid, axis=1)
    counts = valid.sum(axis=1)
    adj = cols[valid].astype(np.int32)  # row-major -> per element in offs order
    adj_row = np.empty(nx * nz + 1, dtype=np.int32)
    adj_row[0] = 1
    np.cumsum(counts, out=adj_row[1:])
    adj_row[1:] += 1
    perm = np.zeros(nx * nz, dtype=np.int32)
    lib = load_host_lib()
    rc = lib.sem2d_genrcm(int(nx * nz), int(adj.size), adj_row.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          adj.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if rc != 0:
        raise RuntimeError(f"sem2d_genrcm failed ({rc})")
    return perm  # perm[k] (1-based values): old element index of new element k+1


# local numbering order inside the first element that touches a node
# (SRC/spec_grid.f90:252-285): interior, edges D,R,U,L counterclockwise, vertices
def _local_code(ngll=5):
    C = np.full((ngll, ngll), -1, dtype=np.int64)  # C[j, i]
    c = 0
    for j in range(1, ngll - 1):
        for i in range(1, ngll - 1):
            C[j, i] = c
            c += 1
    for i in range(1, ngll - 1):          # edge D, i increasing
        C[0, i] = c; c += 1
    for j in range(1, ngll - 1):          # edge R, j increasing
        C[j, ngll - 1] = c; c += 1
    for i in range(ngll - 2, 0, -1):      # edge U, i decreasing
        C[ngll - 1, i] = c; c += 1
    for j in range(ngll - 2, 0, -1):      # edge L, j decreasing
        C[j, 0] = c; c += 1
    for (i, j) in ((0, 0), (ngll - 1, 0), (ngll - 1, ngll - 1), (0, ngll - 1)):
        C[j, i] = c; c += 1
    return C


def _numbering(mesh, perm, xgll):
    ngll = len(xgll)
    deg = ngll - 1
    nx, nz, ezflt = mesh["nx"], mesh["nz"], mesh["ezflt"]
    nelem = nx * nz
    NX1 = deg * nx + 1
    NZ1 = deg * nz + 1 + (1 if ezflt > 0 else 0)
    old = perm.astype(np.int64) - 1
    ix = old % nx
    iz = old // nx
    shift = (iz >= ezflt).astype(np.int64) if ezflt > 0 else np.zeros_like(iz)
    loc = np.arange(ngll)
    # node-grid index of each (e, j, i)
    X = deg * ix[:, None, None] + loc[None, None, :]
    Z = deg * iz[:, None, None] + loc[None, :, None] + shift[:, None, None]
    G = Z * NX1 + X  # (nelem, ngll, ngll)
    npoin = NX1 * NZ1
    k = np.broadcast_to(np.arange(nelem)[:, None, None], G.shape)
    first = np.full(npoin, nelem, dtype=np.int64)
    np.minimum.at(first, G.ravel(), k.ravel())
    last = np.full(npoin, -1, dtype=np.int64)
    np.maximum.at(last, G.ravel(), k.ravel())
    if (first == nelem).any():
        raise RuntimeError("numbering: node not covered by any element")
    C = _local_code(ngll)
    code = np.empty(npoin, dtype=np.int64)
    sel = first[G] == k
    code[G[sel]] = np.broadcast_to(C[None], G.shape)[sel]
    order = np.lexsort((code, first))
    iglob = np.empty(npoin, dtype=np.int64)
    iglob[order] = np.arange(1, npoin + 1)
    ibool = iglob[G].astype(np.int32)  # [e, j, i], 1-based

    czr = iz + shift  # control row of the element's lower edge
    if "ctrl_z" in mesh:
        # layered mesh: control node coordinates per (row, column)
        X_, Z_ = mesh["ctrl_x"], mesh["ctrl_z"]
        cx = np.stack([X_[czr, ix], X_[czr, ix + 1], X_[czr + 1, ix + 1], X_[czr + 1, ix]], axis=1)
        cz = np.stack([Z_[czr, ix], Z_[czr, ix + 1], Z_[czr + 1, ix + 1], Z_[czr + 1, ix]], axis=1)
        return _numbering_coords(ibool, npoin, iglob, G, last, cx, cz, xgll, ix, iz, shift, NX1, NZ1)
    # control nodes (SRC/mesh_cartesian.f90:253-267)
    xc = mesh["xmin"] + (mesh["xmax"] - mesh["xmin"]) / float(nx) * np.arange(nx + 1, dtype=np.float64)
    if ezflt > 0:
        zi = np.concatenate([np.arange(0, ezflt + 1), np.arange(ezflt, nz + 1)]).astype(np.float64)
    else:
        zi = np.arange(0, nz + 1, dtype=np.float64)
    zc = mesh["zmin"] + (mesh["zmax"] - mesh["zmin"]) / float(nz) * zi
    # element corners (knods order 1..4 = (i,j),(i+1,j),(i+1,j+1),(i,j+1))
    cx = np.stack([xc[ix], xc[ix + 1], xc[ix + 1], xc[ix]], axis=1)
    cz = np.stack([zc[czr], zc[czr], zc[czr + 1], zc[czr + 1]], axis=1)
    return _numbering_coords(ibool, npoin, iglob, G, last, cx, cz, xgll, ix, iz, shift, NX1, NZ1)


def _numbering_coords(ibool, npoin, iglob, G, last, cx, cz, xgll, ix, iz, shift, NX1, NZ1):
    """GLL node coordinates of a structured grid from its element corners cx, cz."""
    nelem, ngll = G.shape[0], G.shape[1]
    loc = np.arange(ngll)
    k = np.broadcast_to(np.arange(nelem)[:, None, None], G.shape)
    # coordinates: written by the last element touching the node (SRC/spec_grid.f90:342-349)
    shp = _shape_q4(xgll)  # (ngll(j), ngll(i), 4)
    coord = np.empty((npoin, 2))
    selL = last[G] == k
    e_of = np.broadcast_to(np.arange(nelem)[:, None, None], G.shape)[selL]
    j_of = np.broadcast_to(loc[None, :, None], G.shape)[selL]
    i_of = np.broadcast_to(loc[None, None, :], G.shape)[selL]
    gidx = iglob[G[selL]] - 1
    s = shp[j_of, i_of]  # (m, 4)
    cxe = cx[e_of]
    cze = cz[e_of]
    coord[gidx, 0] = ((cxe[:, 0] * s[:, 0] + cxe[:, 1] * s[:, 1]) + cxe[:, 2] * s[:, 2]) + cxe[:, 3] * s[:, 3]
    coord[gidx, 1] = ((cze[:, 0] * s[:, 0] + cze[:, 1] * s[:, 1]) + cze[:, 2] * s[:, 2]) + cze[:, 3] * s[:, 3]
    return dict(ibool=ibool, coord=coord, npoin=npoin, cx=cx, cz=cz, ix=ix, iz=iz, shift=shift,
                G=G, iglob_of_grid=iglob, NX1=NX1, NZ1=NZ1)


def _read_mesh2d(g, fname_dir):
    """MESH2D_read (SRC/mesh_mesh2d.f90:45-111): a header line before each group, then
    "nel npel nnod nbc", nnod lines "id x z", nel lines "id n1 .. n_npel tag", and per
    boundary "tag nbel" and nbel lines "id element edge"."""
    if g is None:
        raise InputError("MESH2D_read: MESH_MESH2D input block not found")
    fname = str(g.get("file", "")).strip()
    if not fname:
        raise InputError("MESH2D_read: you must provide a mesh2d file name")
    path = fname if os.path.isabs(fname) else os.path.join(fname_dir or ".", fname)
    with open(path) as fh:
        lines = [ln for ln in fh.read().splitlines() if ln.strip()]
    vals = lambda ln: ln.replace("d", "e").replace("D", "e").split()  # noqa: E731
    pos = 1
    nel, npel, nnod, nb = (int(v) for v in vals(lines[pos])[:4])
    if npel != 4:
        raise InputError("MESH2D: only 4-node quadrilaterals are supported by this build")
    pos += 2
    coorg = np.array([[float(v) for v in vals(lines[pos + i])[1:3]] for i in range(nnod)])
    pos += nnod + 1
    knods = np.empty((nel, 4), dtype=np.int64)
    tag = np.empty(nel, dtype=np.int64)
    for e in range(nel):
        v = [int(float(x)) for x in vals(lines[pos + e])]
        knods[e] = v[1:5]
        tag[e] = v[5]
    pos += nel
    bounds = {}
    for _ in range(nb):
        pos += 1
        btag, nbel = (int(v) for v in vals(lines[pos])[:2])
        pos += 2
        rows = [[int(x) for x in vals(lines[pos + k])[:3]] for k in range(nbel)]
        pos += nbel
        bounds[btag] = ([r[1] for r in rows], [r[2] for r in rows])
    return dict(kind="mesh2d", nelem=nel, coorg=coorg, knods=knods - 1, tag=tag, bounds=bounds, nx=0, nz=0,
                ezflt=-1, fztag=0, fznz=0)


def _numbering_fem(mesh, xgll):
    """SE_init_numbering and SE_init_coord (SRC/spec_grid.f90:198-349) for an unstructured
    4-node mesh in element order: interior points first, then each edge not yet numbered
    (shared with the element across it, traversed in reverse there), then each vertex not
    yet numbered (shared with every element holding that mesh node).  Coordinates are
    written by every element touching a point, the last one in element order standing."""
    knods = mesh["knods"]
    nelem = len(knods)
    ngll = len(xgll)
    ivtx, jvtx = (0, ngll - 1, ngll - 1, 0), (0, 0, ngll - 1, ngll - 1)
    iedg = [_edge_ij(n, ngll) for n in (EDGE_D, EDGE_R, EDGE_U, EDGE_L)]  # (i[k], j[k]) counterclockwise
    # FE_SetConnectivity (SRC/fem_grid.f90:103-204): edge n joins vertices n and n+1
    edge_of = {}
    for e in range(nelem):
        for n in range(4):
            a, b = int(knods[e, n]), int(knods[e, (n + 1) % 4])
            edge_of.setdefault((min(a, b), max(a, b)), []).append((e, n))
    vconn = {}
    for e in range(nelem):
        for n in range(4):
            vconn.setdefault(int(knods[e, n]), []).append((e, n))
    ibool = np.zeros((nelem, ngll, ngll), dtype=np.int64)  # [e, j, i]
    npoin = 0
    for e in range(nelem):
        for j in range(1, ngll - 1):
            for i in range(1, ngll - 1):
                npoin += 1
                ibool[e, j, i] = npoin
        for n in range(4):
            ii, jj = iedg[n]
            if ibool[e, jj[1], ii[1]] > 0:
                continue
            a, b = int(knods[e, n]), int(knods[e, (n + 1) % 4])
            other = [x for x in edge_of[(min(a, b), max(a, b))] if x[0] != e]
            ee, nn = other[0] if other else (-1, -1)
            for k in range(1, ngll - 1):
                npoin += 1
                ibool[e, jj[k], ii[k]] = npoin
                if ee >= 0:
                    iR, jR = iedg[nn]
                    ibool[ee, jR[ngll - 1 - k], iR[ngll - 1 - k]] = npoin
        for n in range(4):
            if ibool[e, jvtx[n], ivtx[n]] > 0:
                continue
            npoin += 1
            for ee, nn in vconn[int(knods[e, n])]:
                ibool[ee, jvtx[nn], ivtx[nn]] = npoin
    cx = mesh["coorg"][knods, 0]
    cz = mesh["coorg"][knods, 1]
    shp = _shape_q4(xgll)
    coord = np.empty((npoin, 2))
    for e in range(nelem):
        for j in range(ngll):
            for i in range(ngll):
                sv = shp[j, i]
                g = ibool[e, j, i] - 1
                coord[g, 0] = ((cx[e, 0] * sv[0] + cx[e, 1] * sv[1]) + cx[e, 2] * sv[2]) + cx[e, 3] * sv[3]
                coord[g, 1] = ((cz[e, 0] * sv[0] + cz[e, 1] * sv[1]) + cz[e, 2] * sv[2]) + cz[e, 3] * sv[3]
    return dict(ibool=ibool.astype(np.int32), coord=coord, npoin=npoin, cx=cx, cz=cz,
                ix=np.zeros(nelem, dtype=np.int64), iz=np.zeros(nelem, dtype=np.int64))


def _elastic_coefs_general(DxiDx, DxiDz, DetaDx, DetaDz, w2J, Kx, Kz, lam, mu, mux, ndof):
    """General-geometry coefficients a(:,:,k) of SRC/mat_elastic.f90:327-358 (nelast 3 and
    10), arrays (n, ngll, ngll) -> (n, nelast, ngll, ngll).  Moduli Kx, Kz, lam, mu (= muz),
    mux as in SRC/mat_elastic.f90:307-321."""
    if ndof == 1:
        a = np.stack([mux * DxiDx * DxiDx + mu * DxiDz * DxiDz,
                      mux * DetaDx * DetaDx + mu * DetaDz * DetaDz,
                      mux * DxiDx * DetaDx + mu * DxiDz * DetaDz], axis=1)
    else:
        a = np.stack([Kx * DxiDx * DxiDx + mu * DxiDz * DxiDz,
                      lam * DxiDx * DetaDz + mu * DxiDz * DetaDx,
                      Kz * DetaDz * DetaDz + mu * DetaDx * DetaDx,
                      Kx * DetaDx * DetaDx + mu * DetaDz * DetaDz,
                      lam * DxiDz * DetaDx + mu * DxiDx * DetaDz,
                      Kz * DxiDz * DxiDz + mu * DxiDx * DxiDx,
                      Kx * DxiDx * DetaDx + mu * DxiDz * DetaDz,
                      (lam + mu) * DxiDx * DxiDz,
                      (lam + mu) * DetaDx * DetaDz,
                      Kz * DxiDz * DetaDz + mu * DxiDx * DetaDx], axis=1)
    return (-w2J)[:, None] * a


def _shape_q4(xgll):
    """Q4 shape functions at the GLL points (SRC/elem_q4.f90:37-53), [j, i, k]."""
    q = 0.25
    n = len(xgll)
    out = np.empty((n, n, 4))
    for j in range(n):
        t = xgll[j]
        for i in range(n):
            s = xgll[i]
            sp, sm, tp, tm = s + 1.0, s - 1.0, t + 1.0, t - 1.0
            out[j, i] = [q * sm * tm, -(q * sp * tm), q * sp * tp, -(q * sm * tp)]
    return out


def _dshape_q4(xgll):
    """Q4 shape derivatives (SRC/elem_q4.f90:57-78), [j, i, k, (dxi, deta)]."""
    q = 0.25
    n = len(xgll)
    out = np.empty((n, n, 4, 2))
    for j in range(n):
        t = xgll[j]
        for i in range(n):
            s = xgll[i]
            sp, sm, tp, tm = s + 1.0, s - 1.0, t + 1.0, t - 1.0
            out[j, i, :, 0] = [q * tm, -(q * tm), q * tp, -(q * tp)]
            out[j, i, :, 1] = [q * sm, -(q * sp), q * sp, -(q * sm)]
    return out


def _jacobians(cx, cz, xgll):
    """Jacobian entries at every (e, j, i): J[a][b] = sum_k coorg(a,k) dshape(k,b)."""
    ds = _dshape_q4(xgll)  # j,i,k,b

    def mm(c, b):
        d = ds[None, :, :, :, b]
        cc = c[:, None, None, :]
        return ((cc[..., 0] * d[..., 0] + cc[..., 1] * d[..., 1]) + cc[..., 2] * d[..., 2]) + cc[..., 3] * d[..., 3]

    J11 = mm(cx, 0)  # dx/dxi
    J12 = mm(cx, 1)  # dx/deta
    J21 = mm(cz, 0)  # dz/dxi
    J22 = mm(cz, 1)  # dz/deta
    return J11, J12, J21, J22


def _inv_jac(J11, J12, J21, J22):
    det = J11 * J22 - J12 * J21
    if (det <= 0).any():
        raise InputError("SE_InverseJacobian: undefined Jacobian")
    # SRC/utils.f90:92-111: B = [[J22, -J12], [-J21, J11]] / det
    return J22 / det, (-J12) / det, (-J21) / det, J11 / det  # DxiDx, DxiDz, DetaDx, DetaDz


# ---------------------------------------------------------------------------
# materials
# ---------------------------------------------------------------------------
def _read_hete1(par, group, fname_dir):
    """read_hete1_dist (SRC/distribution_hete1.f90): header 'ncol nx nz x0 z0 dx dz', then nx*nz rows."""
    fname = group.get("file", "")
    col = int(group.get("col", 1))
    path = fname if os.path.isabs(fname) else os.path.join(fname_dir or ".", fname)
    with open(path) as fh:
        head = fh.readline().replace("d", "e").replace("D", "e").split()
        ncol, nx, nz = int(float(head[0])), int(float(head[1])), int(float(head[2]))
        x0, z0, dx, dz = (float(v) for v in head[3:7])
        vals = np.loadtxt(fh, dtype=np.float64, ndmin=2, converters=None)
    if vals.shape[0] < nx * nz:
        raise InputError("read_hete1_dist: file too short")
    val = vals[:nx * nz, col - 1].reshape(nz, nx)  # [j][i]
    return ("hete1", dict(nx=nx, nz=nz, x0=x0, z0=z0, dx=dx, dz=dz, val=val))


def _gen_hete1(p, x, z):
    """generate_hete1_dist (SRC/distribution_hete1.f90): bilinear interpolation on a regular grid."""
    nx, nz, x0, z0, dx, dz, val = p["nx"], p["nz"], p["x0"], p["z0"], p["dx"], p["dz"], p["val"]
    if (x.min() < x0 - TINY_XABS or x.max() > x0 + (nx - 1) * dx + TINY_XABS or
            z.min() < z0 - TINY_XABS or z.max() > z0 + (nz - 1) * dz + TINY_XABS):
        raise InputError("generate_hete1_dist: input grid is too small")
    xi = (x - x0) / dx
    i = np.floor(xi).astype(np.int64)
    i = np.where(i >= nx - 1, nx - 2, i)
    i = np.where(i < 0, 0, i)
    xi = xi - i.astype(np.float64)
    eta = (z - z0) / dz
    j = np.floor(eta).astype(np.int64)
    j = np.where(j >= nz - 1, nz - 2, j)
    j = np.where(j < 0, 0, j)
    eta = eta - j.astype(np.float64)
    return ((((1.0 - xi) * (1.0 - eta)) * val[j, i] + (xi * (1.0 - eta)) * val[j, i + 1])
            + (xi * eta) * val[j + 1, i + 1]) + ((1.0 - xi) * eta) * val[j + 1, i]


def _materials(par, fname_dir=None):
    """MAT_read / MAT_ELAST_read (SRC/mat_gen.f90:101-200, SRC/mat_elastic.f90:69-188)."""
    mats = {}
    for g in par.all("MATERIAL"):
        tag = int(g.get("tag", 0))
        kind = g.get("kind", "ELAST")
        kinds = kind if isinstance(kind, list) else [kind]
        if tag <= 0:
            raise InputError("MAT_read: tag must be positive")
        for k in kinds:
            if k not in ("ELAST", "KV"):
                raise InputError(f"material kind {k!r} is not supported by this build")
        if "ELAST" not in kinds:
            raise InputError("this build needs an ELAST material (optionally with KV)")
        me = par.after(g, "MAT_ELASTIC")
        if me is None:
            raise InputError("MAT_ELAST_read: MAT_ELASTIC block not found")
        rho, cp, cs = float(me.get("rho", 0)), float(me.get("cp", 0)), float(me.get("cs", 0))
        rhoH, cpH, csH = me.get("rhoh", ""), me.get("cph", ""), me.get("csh", "")
        if rho <= 0 and not rhoH:
            raise InputError("MAT_ELAST_read: undefined density (rho)")
        cij = {k: float(me.get(k, 0)) for k in ("c11", "c13", "c33", "c55", "c66")}
        cijH = {k: str(me.get(k + "h", "")).strip() for k in cij}
        has = {k: cij[k] > 0 or bool(cijH[k]) for k in cij}
        if not ((cp > 0 or cpH) and (cs > 0 or csH)):
            # transverse anisotropy, vertical symmetry axis (SRC/mat_elastic.f90:131-152):
            # c11, c13, c33, c55 all given, or c55 and c66 (SH)
            if not ((has["c11"] and has["c13"] and has["c33"] and has["c55"])
                    or (has["c55"] and has["c66"])):
                raise InputError("MAT_ELAST_read: incomplete input")
            if rhoH or any(cijH.values()):
                raise InputError("MAT_ELASTIC distributions of anisotropic moduli are not supported "
                                 "by this build")
            # homogeneous (one shared coefficient array) only when all five moduli are set
            homog = all(v > 0 for v in cij.values())
            # wave speeds for the time step, absorbing boundaries and faults:
            # cp = sqrt(c11/rho), cs = sqrt(c55/rho) (SRC/mat_elastic.f90:224-230)
            mats[tag] = dict(homog=homog, aniso=True, rho=rho, cp=math.sqrt(cij["c11"] / rho),
                             cs=math.sqrt(cij["c55"] / rho), **cij)
        elif not (rhoH or cpH or csH):
            mu = rho * cs * cs
            kmod = rho * cp * cp
            lam = rho * (cp * cp - 2.0 * cs * cs)
            poiss = 0.5 * lam / (kmod - mu)
            if poiss < 0.0 or poiss > 0.5:
                raise InputError("Poisson's ratio out of range !")
            mats[tag] = dict(homog=True, rho=rho, cp=cp, cs=cs, lam=lam, mu=mu)
        else:
            # distributions follow MAT_ELASTIC in the order rho, cp, cs (SRC/mat_elastic.f90:105-113)
            it = iter(par.next_groups(me))
            spec = {}
            for nme, val, hname in (("rho", rho, rhoH), ("cp", cp, cpH), ("cs", cs, csH)):
                if not hname:
                    spec[nme] = ("const", val)
                    continue
                want = "DIST_" + str(hname).upper()
                gg = next(it, None)
                while gg is not None and gg.name != want:
                    gg = next(it, None)
                if gg is None:
                    raise InputError(f"{want} block not found")
                if want != "DIST_HETE1":
                    raise InputError(f"distribution {want} for materials is not supported by this build")
                spec[nme] = _read_hete1(par, gg, fname_dir)
            mats[tag] = dict(homog=False, spec=spec)
        if "KV" in kinds:  # MAT_KV_read (SRC/mat_kelvin_voigt.f90:72-100)
            kv = par.after(g, "MAT_KV")
            if kv is None:
                raise InputError("MAT_KV_read: MAT_KV input block not found")
            # eta, or etaH and the distribution block that follows MAT_KV
            # (MAT_setProp, SRC/mat_kelvin_voigt.f90:85)
            mats[tag]["kv_eta"] = _read_cd(par, iter(par.next_groups(kv)), kv.get("eta", 0.0),
                                           str(kv.get("etah", "")).strip())
            # NormalizeByDT is module-level: the last MAT_KV block read decides for all
            mats["_kv_by_dt"] = bool(kv.get("etaxdt", True))
    if not mats:
        raise InputError("MAT_read: MATERIAL block not found")
    return mats


def _eval_dist(d, x, z):
    if d[0] == "const":
        return np.full(x.shape, d[1])
    return _gen_hete1(d[1], x, z)


def _elastic_coefs(DxiDx, DetaDz, w2J, Kx, Kz, lam, mu, mux, ndof):
    """Flat-grid coefficients a(:,:,k) of SRC/mat_elastic.f90:324-339, one element; moduli
    as in _elastic_coefs_general."""
    return _elastic_coefs_batch(DxiDx[None], DetaDz[None], w2J[None], Kx[None], Kz[None],
                                lam[None], mu[None], mux[None], ndof)[0]


def _elastic_coefs_batch(DxiDx, DetaDz, w2J, Kx, Kz, lam, mu, mux, ndof):
    """_elastic_coefs for a batch of elements: arrays (n, ngll, ngll) -> (n, nelast, ngll, ngll)."""
    if ndof == 1:
        a = np.stack([mux * DxiDx * DxiDx, mu * DetaDz * DetaDz], axis=1)
    else:
        a = np.stack([Kx * DxiDx * DxiDx, lam * DxiDx * DetaDz, Kz * DetaDz * DetaDz,
                      mu * DetaDz * DetaDz, mu * DxiDx * DetaDz, mu * DxiDx * DxiDx], axis=1)
    return (-w2J)[:, None] * a


# ---------------------------------------------------------------------------
# boundaries
# ---------------------------------------------------------------------------
def _bound_elements(mesh):
    """Boundary element lists (old element indices, 1-based) of SRC/mesh_structured.f90:101-196,
    or, for a MESH2D mesh, the file's boundaries (element and edge of each boundary element)."""
    if "bounds" in mesh:
        return mesh["bounds"]
    nx, nz, ezflt = mesh["nx"], mesh["nz"], mesh["ezflt"]
    b = {}
    b[SIDE_D] = ([i + 1 for i in range(nx)], EDGE_D)
    b[SIDE_R] = ([j * nx + nx for j in range(nz)], EDGE_R)
    b[SIDE_U] = ([(nz - 1) * nx + i + 1 for i in range(nx)], EDGE_U)
    b[SIDE_L] = ([j * nx + 1 for j in range(nz)], EDGE_L)
    if ezflt > 0:
        b[FAULT_U] = ([ezflt * nx + i + 1 for i in range(nx)], EDGE_D)
        b[FAULT_D] = ([(ezflt - 1) * nx + i + 1 for i in range(nx)], EDGE_U)
    return b


def _edge_ij(edge, ngll):
    k = np.arange(ngll)
    if edge == EDGE_D:
        return k, np.zeros(ngll, int)
    if edge == EDGE_R:
        return np.full(ngll, ngll - 1), k
    if edge == EDGE_U:
        return k[::-1], np.full(ngll, ngll - 1)
    return np.zeros(ngll, int), k[::-1]


class _Bound:
    pass


def _bound_topo(tag, mesh, num, perm_inv):
    """BC_set_bulk_node (SRC/spec_grid.f90:676-822): node list sorted by coordinate."""
    elems_old, edge = _bound_elements(mesh)[tag]
    elem = perm_inv[np.array(elems_old) - 1]  # 0-based new element index
    edges = np.broadcast_to(np.asarray(edge, dtype=np.int64), (len(elem),)).copy()  # one per element
    ibool = num["ibool"]
    coord = num["coord"]
    ngll = ibool.shape[1]
    nodes = []
    bibool = np.empty((len(elem), ngll), dtype=np.int64)
    index = {}
    for n, e in enumerate(elem):
        ii, jj = _edge_ij(edges[n], ngll)
        for kloc in range(ngll):
            g = int(ibool[e, jj[kloc], ii[kloc]])
            if g not in index:
                index[g] = len(nodes)
                nodes.append(g)
            bibool[n, kloc] = index[g]
    nodes = np.array(nodes, dtype=np.int64)
    bc = coord[nodes - 1]
    Lx = bc[:, 0].max() - bc[:, 0].min()
    Lz = bc[:, 1].max() - bc[:, 1].min()
    key = bc[:, 0] if Lx > Lz else bc[:, 1]
    isort = np.argsort(key, kind="stable")
    back = np.empty_like(isort)
    back[isort] = np.arange(len(isort))
    t = _Bound()
    t.tag = tag
    t.elem = elem
    t.edges = edges
    t.node = nodes[isort]              # 1-based global nodes
    t.ibool = back[bibool]             # (nelem_b, ngll) -> index into t.node
    return t


def _normal_and_weights(t, jac, wgll, periodic=False):
    """BC_get_normal_and_weights (SRC/spec_grid.f90:961-1011).  ``periodic``: the boundary
    ends on a periodic boundary at both ends, whose two end nodes are then one node
    (normal and weight summed over both, SRC/spec_grid.f90:998-1003)."""
    J11, J12, J21, J22 = jac
    ngll = len(wgll)
    npb = len(t.node)
    NORM = np.zeros((npb, 2))
    W = np.zeros(npb)
    for n, e in enumerate(t.elem):
        edge = t.edges[n]
        ii, jj = _edge_ij(edge, ngll)
        dim_t = 0 if edge in (EDGE_D, EDGE_U) else 1
        sign = -1.0 if edge in (EDGE_U, EDGE_L) else 1.0
        for k in range(ngll):
            i, j = ii[k], jj[k]
            if dim_t == 0:
                dx, dz = J11[e, j, i], J21[e, j, i]
            else:
                dx, dz = J12[e, j, i], J22[e, j, i]
            jac1d = math.sqrt(dx * dx + dz * dz)
            tx, tz = sign * dx / jac1d, sign * dz / jac1d
            b = t.ibool[n, k]
            NORM[b, 0] += tz
            NORM[b, 1] += -tx
            W[b] += wgll[k] * jac1d
    if periodic:
        NORM[0] = NORM[0] + NORM[npb - 1]
        NORM[npb - 1] = NORM[0]
        W[0] = W[0] + W[npb - 1]
        W[npb - 1] = W[0]
    for n in range(len(t.elem)):
        for b in (t.ibool[n, 0], t.ibool[n, ngll - 1]):
            nrm = math.sqrt(NORM[b, 0] ** 2 + NORM[b, 1] ** 2)
            NORM[b] = NORM[b] / nrm
    return NORM, W


def _perio_init(bd, mesh, num, perm_inv, M):
    """BC_PERIO_init (SRC/bc_periodic.f90:45-73): master and slave boundaries (tags(1),
    tags(2)) with matching node lists (equal counts, one coordinate shift); the mass of
    each master node becomes the sum of both and the slave node takes it
    (BC_PERIO_set_field, SRC/bc_periodic.f90:78-90)."""
    tv = bd.get("tags", None)
    if not isinstance(tv, list) or len(tv) < 2:
        raise InputError("BC_DEF PERIOD needs tags = master, slave")
    tm = _bound_topo(int(tv[0]), mesh, num, perm_inv)
    ts = _bound_topo(int(tv[1]), mesh, num, perm_inv)
    if len(tm.elem) != len(ts.elem):
        raise InputError("bc_perio_init: number of boundary elements do not match")
    if len(tm.node) != len(ts.node):
        raise InputError("bc_perio_init: number of nodes on boundaries do not match")
    coord = num["coord"]
    cm, cs = coord[tm.node - 1], coord[ts.node - 1]
    shift = cm[0] - cs[0]
    if (np.abs(cm[:, 0] - cs[:, 0] - shift[0]) > TINY_XABS).any() or \
            (np.abs(cm[:, 1] - cs[:, 1] - shift[1]) > TINY_XABS).any():
        raise InputError("bc_perio_init: coordinates on boundaries do not match properly")
    im, is_ = tm.node - 1, ts.node - 1
    M[:, im] = M[:, im] + M[:, is_]
    M[:, is_] = M[:, im]
    return dict(master=tm.node.astype(np.int32), slave=ts.node.astype(np.int32))


def _perio_intersects(t, perio):
    """BC_PERIO_intersects (SRC/bc_periodic.f90:106-120): both end nodes of boundary t are
    on the periodic boundaries."""
    if perio is None:
        return False
    both = set(perio["master"].tolist()) | set(perio["slave"].tolist())
    return int(t.node[0]) in both and int(t.node[-1]) in both


def _abso_init(t, jac, wgll, props, M, coefA2Vrhs, ndof, stacey, periodic=False, let_wave=True):
    """BC_ABSO_init (SRC/bc_abso.f90:115-266), Clayton-Engquist; on a boundary that is not
    flat (normals not all along x or z) the P-SV coefficients are rho cp w|J| along the
    normal and rho cs w|J| along the boundary, applied with the node normals, and the mass
    is not modified.  Stacey term for flat P-SV boundaries when ``stacey``
    (SRC/bc_abso.f90:176-180,209-224):
    K[e, k, geo_tan] = -K[e, k, geo_nor] = w_k |J1D| * DLocDGlob(tan, tan) * rho cs (2 cs - cp)."""
    J11, J12, J21, J22 = jac
    NORM, WB = _normal_and_weights(t, jac, wgll, periodic)
    is_flat = True
    if np.all(np.abs(NORM[:, 0]) < TINY_XABS):
        geo_tan, geo_nor = 0, 1
    elif np.all(np.abs(NORM[:, 1]) < TINY_XABS):
        geo_tan, geo_nor = 1, 0
    else:
        # not flat (SRC/bc_abso.f90:163-166): cp along the local normal, cs along the
        # boundary, no implicit treatment in the mass, normals kept for BC_ABSO_apply
        is_flat = False
        geo_tan, geo_nor = 1, 0
    C = np.zeros((len(t.node), ndof))
    rho_p, cp_p, cs_p = props
    stacey = stacey and ndof == 2 and is_flat  # SRC/bc_abso.f90:176
    ngll = len(wgll)
    K = np.zeros((len(t.elem), ngll, 2)) if stacey else None
    cb, ce, cf = [], [], []  # contributions: boundary node, element, weight w_k |J1D|
    for n, e in enumerate(t.elem):
        dim_t = 0 if t.edges[n] in (EDGE_D, EDGE_U) else 1
        ii, jj = _edge_ij(t.edges[n], ngll)
        for k in range(ngll):
            i, j = ii[k], jj[k]
            rho = rho_p[e, j, i]
            c = [0.0, 0.0]
            c[geo_nor] = cp_p[e, j, i]
            c[geo_tan] = cs_p[e, j, i]
            if dim_t == 0:
                a, b = J11[e, j, i], J21[e, j, i]
            else:
                a, b = J12[e, j, i], J22[e, j, i]
            coef = wgll[k] * math.sqrt(a * a + b * b)
            bck = t.ibool[n, k]
            cb.append(bck)
            ce.append(e)
            cf.append(coef)
            if ndof == 1:
                C[bck, 0] = C[bck, 0] + rho * c[geo_tan] * coef
            else:
                C[bck, 0] = C[bck, 0] + rho * c[0] * coef
                C[bck, 1] = C[bck, 1] + rho * c[1] * coef
            if stacey:
                inv = _inv_jac(J11[e, j, i], J12[e, j, i], J21[e, j, i], J22[e, j, i])
                xji = inv[2 * dim_t + geo_tan]  # DLocDGlob(LocDimTan, GeoDimTan)
                K[n, k, 0] = coef * xji * rho * c[geo_tan] * (2.0 * c[geo_tan] - c[geo_nor])
    if stacey:
        K[:, :, 1] = K[:, :, 0]
        K[:, :, geo_tan] = -K[:, :, geo_tan]
    if periodic and is_flat:  # both end nodes are one periodic node (SRC/bc_abso.f90:231-235)
        C[0] = C[0] + C[-1]
        C[-1] = C[0]
    nodes0 = t.node - 1
    if is_flat:  # SRC/bc_abso.f90:243
        M[:, nodes0] = M[:, nodes0] + coefA2Vrhs * C.T
    # C[node, c] = sum over contributions of rho * c_c * weight, c_c = cp along the
    # normal (component geo_nor), cs along the boundary
    return dict(tag=t.tag, node=t.node.astype(np.int32), C=C, geo_nor=geo_nor,
                contrib=(np.array(cb, dtype=np.int64), np.array(ce, dtype=np.int64), np.array(cf)),
                stacey=(np.ascontiguousarray(t.ibool[:, :ngll], dtype=np.int32), K) if stacey else None,
                normals=None if (is_flat or ndof == 1) else NORM,
                # incident waves (SRC/bc_abso.f90:247-264): flat boundaries with let_wave
                let_wave=bool(let_wave and is_flat), wave_n=NORM, wave_B=WB,
                periodic_ends=bool(periodic and stacey))


# ---------------------------------------------------------------------------
# distributions used by the fault inputs (constant, PWCONR, LINEAR, ORDER0)
# ---------------------------------------------------------------------------
def _read_cd(par, group_iter, value, name_h):
    """DIST_CD_Read (SRC/distribution_cd.f90:26-63): a constant or the next &DIST_<name> block."""
    if not name_h:
        return ("const", float(value))
    want = "DIST_" + name_h.upper()
    g = next(group_iter, None)
    while g is not None and g.name != want:
        g = next(group_iter, None)
    if g is None:
        raise InputError(f"{want} block not found")
    if want == "DIST_PWCONR":
        num = int(g.get("num"))
        ref = g.get("ref", [0.0, 0.0])
        vals = " ".join(g.data_lines).replace("d", "e").replace("D", "e").split()
        vals = [float(v) for v in vals]
        if num < 2:
            raise InputError("read_pwconr_dist: needs more than 2 zones (num)")
        return ("pwconr", dict(num=num, ref=[float(ref[0]), float(ref[1])], rad=vals[:num - 1],
                               val=vals[num - 1:2 * num - 1]))
    if want == "DIST_LINEAR":  # SRC/distribution_linear.f90 read_linear_dist (no smoothing)
        if float(g.get("length", 0.0)) > 0:
            raise InputError("DIST_LINEAR with smoothing length is not supported by this build")
        dim = int(g.get("dim", 1))
        fname = g.get("file", "")
        if fname:
            tab = np.loadtxt(fname, ndmin=2)
        else:
            n = int(g.get("n", 0))
            vals = " ".join(g.data_lines).replace("d", "e").replace("D", "e").split()
            tab = np.array([float(v) for v in vals[:2 * n]]).reshape(n, 2)
        return ("linear", dict(dim=dim, x=tab[:, 0].copy(), val=tab[:, 1].copy()))
    if want == "DIST_GAUSSIAN":  # read_gaussian_dist (SRC/distribution_gaussian.f90:34-58)
        c = g.get("centered_at", [0.0, 0.0])
        ln = g.get("length", [1.0, 1.0])
        c = c if isinstance(c, list) else [c, 0.0]
        ln = ln if isinstance(ln, list) else [ln, 1.0]
        return ("gaussian", dict(x0=float(c[0]), z0=float(c[1]), lx=float(ln[0]), lz=float(ln[1]),
                                 level=float(g.get("offset", 0.0)), ampli=float(g.get("ampli", 1.0)),
                                 order=int(g.get("order", 1))))
    if want == "DIST_ORDER0":  # read_order0_dist (SRC/distribution_order0.f90:47-68)
        xn, zn = int(g.get("xn")), int(g.get("zn"))
        recs = list(g.data_lines)
        pos = 0
        xb, pos = _list_read(recs, pos, xn - 1) if xn > 1 else ([], pos)
        zb, pos = _list_read(recs, pos, zn - 1) if zn > 1 else ([], pos)
        val = []
        for _ in range(zn):  # val(:, iz), one read per z-zone
            row, pos = _list_read(recs, pos, xn)
            val.append(row)
        return ("order0", dict(xn=xn, zn=zn, xb=xb, zb=zb, val=val))
    raise InputError(f"distribution {want} is not supported by this build")


def _list_read(recs, pos, n):
    """A Fortran list-directed READ of n reals from the records recs[pos:]: it takes values
    from as many records as it needs and leaves the rest of its last record unread.
    Returns (values, index of the next record)."""
    out = []
    while len(out) < n:
        if pos >= len(recs):
            raise InputError("list-directed read: end of data")
        toks = recs[pos].replace("d", "e").replace("D", "e").replace(",", " ").split()
        pos += 1
        out.extend(float(t) for t in toks[:n - len(out)])
    return out, pos


def _order0_zone(c, nzones, bound):
    """zone (SRC/distribution_order0.f90:85-98): the first k with c < bound(k), else nzones"""
    for k in range(nzones - 1):
        if c < bound[k]:
            return k
    return nzones - 1


def _gen_cd(cd, coord):
    kind, p = cd
    if kind == "const":
        return np.full(coord.shape[0], p)
    if kind == "pwconr":  # SRC/distribution_pwconr.f90 generate_pwconr_dist
        out = np.empty(coord.shape[0])
        for n in range(coord.shape[0]):
            rad = math.sqrt((coord[n, 0] - p["ref"][0]) ** 2 + (coord[n, 1] - p["ref"][1]) ** 2)
            z = p["num"] - 1
            for iz in range(p["num"] - 1):
                if rad <= p["rad"][iz]:
                    z = iz
                    break
            out[n] = p["val"][z]
        return out
    if kind == "gaussian":  # generate_gaussian_dist (SRC/distribution_gaussian.f90:64-75)
        # the C library's exp, as the reference's Fortran runtime calls it
        out = np.empty(coord.shape[0])
        for n, (x, z) in enumerate(coord[:, :2]):
            a = (x - p["x0"]) / p["lx"]
            b = (z - p["z0"]) / p["lz"]
            out[n] = p["level"] + p["ampli"] * math.exp(-((a * a + b * b) ** p["order"]))
        return out
    if kind == "order0":  # generate_order0_dist (SRC/distribution_order0.f90:72-82)
        return np.array([p["val"][_order0_zone(z, p["zn"], p["zb"])][_order0_zone(x, p["xn"], p["xb"])]
                         for x, z in coord[:, :2]])
    if kind == "linear":  # interpol of SRC/distribution_linear.f90
        xi, yi = p["x"], p["val"]
        xo = coord[:, p["dim"] - 1]
        ni = len(xi)
        ipos = np.searchsorted(xi, xo, side="right")  # number of knots <= x (1-based position)
        ipos = np.clip(ipos, 1, ni - 1) - 1
        eta = (xo - xi[ipos]) / (xi[ipos + 1] - xi[ipos])
        return yi[ipos] + eta * (yi[ipos + 1] - yi[ipos])
    raise ValueError(kind)


def _dynflt_init(par, bcdef, mesh, num, perm_inv, jac, wgll, M, dt, beta, gamma, nt, ndof, scheme="newmark",
                 perio=None):
    """BC_DYNFLT_read + BC_DYNFLT_init (SRC/bc_dynflt.f90:102-520)."""
    g = par.after(bcdef, "BC_DYNFLT")
    if g is None:
        raise InputError("BC_DYNFLT_read: BC_DYNFLT input block not found")
    # tags(1), tags(2); one tag (or tags(2) = 0): a fault on a domain boundary with one
    # side (symmetry, SRC/bc_dynflt.f90:256-259)
    tag = int(bcdef.get("tag", 0))
    tv = bcdef.get("tags", None)
    if tag > 0:
        tags = [tag, 0]
    elif isinstance(tv, list) and len(tv) >= 2:
        tags = [int(tv[0]), int(tv[1])]
    else:
        tags = [int(tv if not isinstance(tv, list) else tv[0]), 0]
    two_sides = tags[1] > 0
    it = iter(par.next_groups(g))
    cd = {}
    # order of DIST_CD_Read calls: SRC/bc_dynflt.f90:166-174
    for key, dflt in (("cohesion", 0.0), ("tn", 0.0), ("tt", 0.0), ("sxx", 0.0), ("sxy", 0.0),
                      ("sxz", 0.0), ("syz", 0.0), ("szz", 0.0), ("v", 1e-12)):
        cd[key] = _read_cd(par, it, g.get(key, dflt), g.get(key + "h", ""))
    friction = g.get("friction", "SWF")
    fr = friction if isinstance(friction, list) else [friction]
    fr = [str(f).strip().upper() for f in fr if f]
    # friction(1:2) (SRC/bc_dynflt.f90:188-202): SWF or RSF, optionally with TWF (the
    # minimum of both coefficients), or TWF alone
    base = [f for f in fr if f != "TWF"]
    if len(fr) > 2 or len(base) > 1 or any(f not in ("SWF", "RSF", "TWF") for f in fr) or not fr:
        raise InputError("BC_DYNFLT friction must be 'SWF' or 'RSF', optionally with 'TWF', or 'TWF'")
    law = base[0] if base else "TWF"
    # friction blocks, read in the order of friction(:) from the input position after
    # the distributions: &BC_DYNFLT_SWF / _RSF (and their distributions), &BC_DYNFLT_TWF
    rest = list(it)
    fb = next((x for x in rest if x.name == "BC_DYNFLT_" + law), None) if law != "TWF" else None
    fr_it = iter(rest[rest.index(fb) + 1:]) if fb is not None else iter([])
    twf = None
    if "TWF" in fr:
        after = rest[rest.index(fb) + 1:] if (fb is not None and fr.index("TWF") > fr.index(law)) else rest
        tb = next((x for x in after if x.name == "BC_DYNFLT_TWF"), None)
        twf = _twf_read(tb)
    fp = {}
    if law == "TWF":
        pass
    elif law == "SWF":
        vals = {"kind": 1, "dc": 0.5, "mus": 0.6, "mud": 0.5, "alpha": 0.0, "p": 3.0, "healing": False}
        if fb is not None:
            for k in vals:
                if k in fb:
                    vals[k] = fb.get(k)
        if int(vals["kind"]) not in (1, 2, 3):
            raise InputError("BC_DYNFLT_SWF: invalid kind")
        fp["kind"] = int(vals["kind"])
        fp["healing"] = bool(vals["healing"])
        for k in ("dc", "mus", "mud", "alpha", "p"):
            fp[k + "_cd"] = _read_cd(par, fr_it, vals[k], fb.get(k + "h", "") if fb else "")
    else:
        vals = {"kind": 1, "dc": 0.5, "mus": 0.6, "a": 0.01, "b": 0.02, "vstar": 1.0, "theta": 0.0,
                "vc": 1e-6}
        if fb is not None:
            for k in vals:
                if k in fb:
                    vals[k] = fb.get(k)
        if int(vals["kind"]) not in (1, 2, 3, 4):
            raise InputError("BC_DYNFLT_RSF: invalid kind")
        fp["kind"] = int(vals["kind"])
        for k in ("dc", "mus", "a", "b", "vstar", "theta", "vc"):
            fp[k + "_cd"] = _read_cd(par, fr_it, vals[k], fb.get(k + "h", "") if fb else "")
    # normal stress response (normal_read / normal_init, SRC/bc_dynflt_normal.f90:40-118):
    # 0 cohesion (sigma stays the initial normal stress), 1 Coulomb, 2 Prakash-Clifton
    # with time scale T (coef = exp(-dt/T)), 3 with length scale L and velocity V
    # (coef = dt/L)
    nor = par.after(g, "BC_DYNFLT_NOR")
    nkind = int(nor.get("kind", 1)) if nor is not None else 1
    nT = float(nor.get("t", 1.0)) if nor is not None else 1.0
    nL = float(nor.get("l", 1.0)) if nor is not None else 1.0
    nV = float(nor.get("v", 1.0)) if nor is not None else 1.0
    if nkind > 3 or nkind < 0:
        raise InputError("BC_SWFF_init: invalid kind in BC_DYNFLT_NOR input block")
    if nL <= 0.0 or nV <= 0.0 or nT <= 0.0:
        raise InputError("BC_SWFF_init: L, V and T must be positive in BC_DYNFLT_NOR input block")
    nor_coef = math.exp(-dt / nT) if nkind == 2 else (dt / nL if nkind == 3 else 0.0)

    t1 = _bound_topo(tags[0], mesh, num, perm_inv)
    coord = num["coord"]
    if two_sides:
        t2 = _bound_topo(tags[1], mesh, num, perm_inv)
        if len(t1.elem) != len(t2.elem) or len(t1.node) != len(t2.node):
            raise InputError("bc_dynflt_init: boundaries do not match")
        keep = t1.node != t2.node
        node2 = t2.node[keep]
    else:
        keep = np.ones(len(t1.node), dtype=bool)
        node2 = None
    node1 = t1.node[keep]
    c1 = coord[node1 - 1]
    if two_sides and (np.abs(c1 - coord[node2 - 1]) > TINY_XABS).any():
        raise InputError("bc_dynflt_init: coordinates on boundaries do not match properly")
    NORM, W = _normal_and_weights(t1, jac, wgll, _perio_intersects(t1, perio))
    n1 = NORM[keep]
    B = np.repeat(W[keep][:, None], ndof, axis=1)
    # TIME_getCoefA2V / TIME_getCoefA2D (SRC/time.f90:425-458): gamma dt and beta dt^2 for
    # Newmark, dt and 0 for leapfrog
    coefA2V = dt if scheme == "leapfrog" else gamma * dt
    coefA2D = 0.0 if scheme == "leapfrog" else beta * dt * dt
    invM1 = 1.0 / M[:, node1 - 1].T
    if two_sides:
        invM2 = 1.0 / M[:, node2 - 1].T
        Z = 1.0 / (coefA2V * B * (invM1 + invM2))
    else:  # SRC/bc_dynflt.f90:353
        invM2 = None
        Z = 0.5 / (coefA2V * B * invM1)
    npf = len(node1)
    f = dict(tags=tags, node1=node1.astype(np.int32), node2=None if node2 is None else node2.astype(np.int32),
             coord=c1, n1=n1, B=B,
             invM1=invM1, invM2=invM2, Z=Z, coefA2V=coefA2V, coefA2D=coefA2D, law=law,
             opening=bool(g.get("opening", True)), nor_kind=nkind, nor_coef=nor_coef, nor_v=nV)
    f["twf"] = twf
    if law == "TWF":
        f["theta"] = np.zeros(npf)
    elif law == "SWF":
        f["swf_kind"] = fp["kind"]
        f["healing"] = fp["healing"]
        for k in ("dc", "mus", "mud", "alpha", "p"):
            f[k] = _gen_cd(fp[k + "_cd"], c1)
        f["theta"] = np.zeros(npf)
    else:
        f["rsf_kind"] = fp["kind"]
        for k in ("dc", "mus", "a", "b", "vstar", "theta", "vc"):
            f[k] = _gen_cd(fp[k + "_cd"], c1)
        f["Tc"] = f["dc"] / f["vstar"]
        f["coeft"] = np.exp(-dt / f["Tc"])
    V0 = np.zeros((npf, ndof))
    if law == "RSF":
        V0[:, 0] = _gen_cd(cd["v"], c1)
    f["V0"] = V0
    Tt0 = _gen_cd(cd["tt"], c1)
    Tn0 = _gen_cd(cd["tn"], c1)
    Sxx, Sxy, Sxz = _gen_cd(cd["sxx"], c1), _gen_cd(cd["sxy"], c1), _gen_cd(cd["sxz"], c1)
    Syz, Szz = _gen_cd(cd["syz"], c1), _gen_cd(cd["szz"], c1)
    nxv, nzv = n1[:, 0], n1[:, 1]
    Tx = Sxx * nxv + Sxz * nzv
    Ty = Sxy * nxv + Syz * nzv
    Tz = Sxz * nxv + Szz * nzv
    T0 = np.empty((npf, 2))
    T0[:, 0] = Tt0 + Ty if ndof == 1 else Tt0 + Tx * nzv - Tz * nxv
    T0[:, 1] = Tn0 + Tx * nxv + Tz * nzv
    f["T0"] = T0
    f["cohesion"] = _gen_cd(cd["cohesion"], c1)
    if (f["cohesion"] < 0).any():
        raise InputError("bc_dynflt_init: cohesion must be positive")
    # initial friction coefficient (SRC/bc_dynflt.f90:404-413)
    if law == "SWF":
        f["MU0"] = swf_mu(f["swf_kind"], f["theta"], f["dc"], f["mus"], f["mud"], f["p"], f["alpha"])
    elif law == "RSF":
        f["MU0"] = rsf_mu(f["rsf_kind"], V0[:, 0], f)
    if twf is not None:
        mu_t = twf_mu(twf, c1, 0.0, np.zeros(npf))
        f["MU0"] = mu_t if law == "TWF" else np.minimum(f["MU0"], mu_t)
    # outputs (SRC/bc_dynflt.f90:419-520)
    oxi = g.get("oxi", [1, 2 ** 31 - 1, 1])
    oix1 = max(int(oxi[0]), 1)
    oixn = min(int(oxi[1]), npf)
    oixd = int(oxi[2])
    oitd = max(1, _nint(float(g.get("otd", 0.0)) / dt))
    f.update(oix1=oix1, oixn=oixn, oixd=oixd, oitd=oitd, odt=dt * oitd,
             oit=_nint(float(g.get("ot1", 0.0)) / dt), osides=bool(g.get("osides", False)))
    return f


def _nint(x):
    """Fortran NINT: nearest integer, halves rounded away from zero."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


# exp, log, asinh and pow of the C library, as the reference's runtime calls them (the
# vectorised numpy versions may differ from them in the last bit)
_exp, _log, _asinh = np.vectorize(math.exp), np.vectorize(math.log), np.vectorize(math.asinh)
_pow = np.vectorize(math.pow)


def _twf_read(tb):
    """twf_read (SRC/bc_dynflt_twf.f90:53-107): kind, MuS, MuD, Mu0, X, Z, V, L, T, Dc."""
    v = dict(kind=1, mus=0.6, mud=0.5, mu0=0.6, x=0.0, z=0.0, v=1e3, l=1.0, t=sys.float_info.max,
             dc=sys.float_info.max)
    if tb is not None:
        for k in v:
            if k in tb:
                v[k] = tb.get(k)
    v = {k: (int(x) if k == "kind" else float(x)) for k, x in v.items()}
    if not 1 <= v["kind"] <= 3:
        raise InputError("BC_DYNFLT_TWF: kind must be 1, 2 or 3")
    if min(v["mus"], v["mud"], v["mu0"]) < 0 or min(v["l"], v["v"], v["t"], v["dc"]) <= 0:
        raise InputError("BC_DYNFLT_TWF: MuS, MuD, Mu0 must be positive, L, V, T, Dc strictly positive")
    return v


def twf_mu(tw, coord, time, d):
    """twf_mu (SRC/bc_dynflt_twf.f90:129-183) at the fault nodes (coord [n][2]) at time
    ``time`` with slip d; kind 3 leaves nodes beyond V*T unset in the reference, they
    get the huge value here (no weakening)."""
    big = sys.float_info.max
    mus, mud, L, V, T = tw["mus"], tw["mud"], tw["l"], tw["v"], tw["t"]
    dx, dz = coord[:, 0] - tw["x"], coord[:, 1] - tw["z"]
    dist = np.sqrt(dx * dx + dz * dz)
    if tw["kind"] in (1, 2):
        if tw["kind"] == 1:
            t = time + (mus - tw["mu0"]) * L / ((mus - mud) * V)
            if t > T:
                t = 0.0
            r = V * t
        else:
            t = time + 0.5 * T * (1.0 - math.sqrt(1.0 - 4.0 * (mus - tw["mu0"]) * L / ((mus - mud) * T * V)))
            t = min(t, T)
            r = V * t * (1.0 - t / T)
        r = dist - r
        return np.where(r < -L, mud, np.where(r <= L, mus + (mus - mud) / L * r, big))
    vt = V * time
    inside = dist <= V * T
    mu = np.where(dist < vt - L, mud, np.where((dist >= vt - L) & (dist <= vt), mus + (mus - mud) / L * (dist - vt),
                                                 big))
    return np.where(inside & (d <= tw["dc"]), mu, big)


def swf_mu(kind, theta, dc, mus, mud, p, alpha):
    """swf_mu (SRC/bc_dynflt_swf.f90:142-162)."""
    if kind == 1:
        mu = mus - (mus - mud) * np.minimum(theta / dc, 1.0)
    elif kind == 2:
        mu = mud - (mud - mus) * _exp(-theta / dc)
    else:
        mu = mud + (mus - mud) / _pow(1.0 + theta / dc, np.broadcast_to(p, np.shape(theta)))
    return mu + alpha * theta


def rsf_mu(kind, v, f):
    """rsf_mu (SRC/bc_dynflt_rsf.f90:165-184)."""
    a, b, mus, vstar, dc, th = f["a"], f["b"], f["mus"], f["vstar"], f["dc"], f["theta"]
    if kind == 1:
        return mus + a * np.abs(v) / (np.abs(v) + vstar) - b * th / (th + dc)
    if kind in (2, 3):
        return a * _asinh(np.abs(v) / (2.0 * vstar) * _exp((mus + b * _log(vstar * th / dc)) / a))
    return a * _asinh(np.abs(v) / (2.0 * vstar) * _exp((mus + b * _log(f["vc"] * th / dc + 1)) / a))


# ---------------------------------------------------------------------------
# sources and receivers
# ---------------------------------------------------------------------------
def _nearest_node(coord, p):
    """SE_find_nearest_node (SRC/spec_grid.f90:411-434): the last node at minimum distance."""
    d2 = (p[0] - coord[:, 0]) ** 2 + (p[1] - coord[:, 1]) ** 2
    m = d2.min()
    idx = np.nonzero(d2 <= m)[0][-1]
    return int(idx) + 1, math.sqrt(m)


def ricker(t, f0, t0, ampli):
    """ricker (SRC/stf_ricker.f90:89-101)."""
    arg = PI * f0 * (t - t0)
    arg = arg * arg
    return -ampli * (1.0 - 2.0 * arg) * math.exp(-arg)


def _f32(x):
    return float(np.float32(x))


def _sources(par, num, jac, hprime, ndof, dt, nt, fname_dir=None, t_shift=None):
    """SO_read / SO_init (SRC/src_gen.f90:96-215): one source at ``coord``, or one per
    line of ``file`` (x z [tdelay [ampli]]), all with the same STF and mechanism.
    ``t_shift``: the amplitudes are taken at it*dt + t_shift (HHT-alpha: t_alpha =
    time + (alpha-1) dt, SRC/solver.f90:114-115)."""
    g = par.first("SRC_DEF")
    if g is None:
        return []
    stf = g.get("stf", " ")
    mech = g.get("mechanism", " ")
    if str(stf).strip() == "" or str(mech).strip() == "":
        raise InputError("SO_read: stf/mechanism not set")
    fname = str(g.get("file", "none")).strip()
    if fname != "none":
        path = fname if os.path.isabs(fname) else os.path.join(fname_dir or ".", fname)
        tab = np.loadtxt(path, ndmin=2)
        if tab.shape[1] not in (2, 3, 4):
            raise InputError("SO_read: source file must have 2, 3 or 4 columns")
        rows = [(tab[i, :2], tab[i, 2] if tab.shape[1] > 2 else 0.0, tab[i, 3] if tab.shape[1] > 3 else 1.0)
                for i in range(tab.shape[0])]
    else:
        coord_in = g.get("coord")
        if coord_in is None:
            raise InputError("SO_read: coord not set")
        rows = [(coord_in, 0.0, 1.0)]
    return [_one_source(par, g, stf, mech, c, td, am, num, jac, hprime, ndof, dt, nt, t_shift)
            for c, td, am in rows]


def _one_source(par, g, stf, mech, coord_in, tdelay, ampli, num, jac, hprime, ndof, dt, nt, t_shift=None):
    if stf != "RICKER":
        raise InputError(f"stf {stf!r} is not supported by this build")
    # STF_RICKER reals are single precision (SRC/stf_ricker.f90:53-73)
    sr = par.after(g, "STF_RICKER")
    if sr is None:
        raise InputError("RICKER_read: STF_RICKER input block not found")
    f0 = _f32(sr.get("f0", 0.0))
    onset = _f32(sr.get("onset", 0.0))
    amp = _f32(sr.get("ampli", 1.0))
    if f0 <= 0:
        raise InputError("RICKER_read: f0 must be positive")
    if onset < 1.0 / f0:
        raise InputError("RICKER_read: Onset time too small (IO_warning aborts)")
    coord = num["coord"]
    ibool = num["ibool"]
    if mech == "WAVE":
        return dict(kind="wave", f0=f0, t0=onset, stf_ampli=amp, tdelay=float(tdelay), mechanism=mech,
                    wave=_wave_read(par, g))
    iglob, _ = _nearest_node(coord, coord_in)
    src = dict(iglob=iglob, f0=f0, t0=onset, stf_ampli=amp, tdelay=float(tdelay), ampli=float(ampli),
               mechanism=mech)
    if mech == "FORCE":
        fg = par.after(g, "SRC_FORCE")
        angle = float(fg.get("angle", 0.0)) if fg is not None else 0.0
        angle = angle * PI / 180.0
        src["kind"] = "force"
        src["dir"] = np.array([-math.sin(angle), math.cos(angle)])
    elif mech in ("EXPLOSION", "MOMENT", "DOUBLE_COUPLE"):
        Mt = np.zeros((2, ndof))
        if mech == "EXPLOSION":
            if ndof != 2:
                raise InputError("SRC_MOMENT_read: explosion only allowed in PSV (ndof=2)")
            Mt[:] = [[1.0, 0.0], [0.0, 1.0]]
        elif mech == "DOUBLE_COUPLE":
            dg = par.after(g, "SRC_DOUBLE_COUPLE")
            if dg is None:
                raise InputError("SRC_MOMENT_read: SRC_DOUBLE_COUPLE input block not found")
            dip = float(dg.get("dip", 90.0)) * PI / 180.0
            n = [math.sin(dip), math.cos(dip)]
            if ndof == 2:
                r = [-math.cos(dip), math.sin(dip)]
                Mt[0, 0] = 2.0 * r[0] * n[0]
                Mt[0, 1] = r[0] * n[1] + r[1] * n[0]
                Mt[1, 0] = Mt[0, 1]
                Mt[1, 1] = 2.0 * r[1] * n[1]
            else:
                Mt[0, 0] = n[0]
                Mt[1, 0] = n[1]
        else:
            mg = par.after(g, "SRC_MOMENT")
            if mg is None:
                raise InputError("SRC_MOMENT_read: SRC_MOMENT input block not found")
            if ndof == 2:
                Mt[:] = [[float(mg.get("mxx", 0)), float(mg.get("mxz", 0))],
                         [float(mg.get("mzx", 0)), float(mg.get("mzz", 0))]]
            else:
                Mt[:, 0] = [float(mg.get("myx", 0)), float(mg.get("myz", 0))]
        src.update(_moment_stencil(iglob, Mt, ibool, jac, hprime, ndof))
        src["kind"] = "moment"
    else:
        raise InputError(f"mechanism {mech!r} is not supported by this build")
    # amplitude table at t = it*dt, it = 1..nt (SRC/src_gen.f90:290-317), or at
    # t_alpha = it*dt + (alpha-1) dt for HHT-alpha
    if t_shift is None:
        src["stf_table"] = np.array([ricker(it * dt - src["tdelay"], f0, onset, amp) * src["ampli"]
                                     for it in range(1, nt + 1)])
    else:
        src["stf_table"] = np.array([ricker((it * dt + t_shift) - src["tdelay"], f0, onset, amp) * src["ampli"]
                                     for it in range(1, nt + 1)])
    return src


def _wave_read(par, g):
    """WAVE_read (SRC/src_wave.f90:33-70): incidence angle (folded into [-180, 180]) and phase."""
    wg = par.after(g, "SRC_WAVE")
    if wg is None:
        raise InputError("WAVE_read: SRC_WAVE input block not found")
    angle = float(wg.get("angle", 0.0))
    angle = ((angle + 180.0) % 360.0) - 180.0  # modulo(angle+180, 360) - 180
    phase = str(wg.get("phase", "S")).strip().upper()[:1]
    if phase not in ("P", "S"):
        raise InputError("WAVE_read: phase must be S or P")
    return dict(angle=angle, phase=phase)


def _wave_init(w, num, props):
    """WAVE_init (SRC/src_wave.f90:75-120).  The reference point wave%coord is never set
    there: the reference's phase origin and the node whose material gives the wave speed
    are those of its allocated (zeroed) value, (0, 0); the material of the first element
    holding that node (SE_node_belongs_to).  k = (-sin a, cos a), polarisation p = k (P)
    or (cos a, sin a) (S), slowness s = k / c."""
    rho_p, cp_p, cs_p, lam_p, mu_p = props
    ip, _ = _nearest_node(num["coord"], (0.0, 0.0))
    e, j, i = (int(v) for v in np.argwhere(num["ibool"] == ip)[0])
    cp, cs, mu, lam = cp_p[e, j, i], cs_p[e, j, i], mu_p[e, j, i], lam_p[e, j, i]
    a = PI * w["angle"] / 180.0
    k = np.array([-math.sin(a), math.cos(a)])
    if w["phase"] == "P":
        p, c = np.array([-math.sin(a), math.cos(a)]), cp
    else:
        p, c = np.array([math.cos(a), math.sin(a)]), cs
    w.update(k=k, p=p, c=c, s=k / c, cp=cp, cs=cs, mu=mu, lam=lam, x0=0.0, z0=0.0)
    return w


def _wave_dirs(w, N, ndof):
    """SRC_WAVE_get_VT (SRC/src_wave.f90:134-171) per boundary node, before the source time
    function: Vin and Tin for unit amplitude, N the node normals [n][2]."""
    K, P = w["k"], w["p"]
    KN = K[0] * N[:, 0] + K[1] * N[:, 1]
    n = N.shape[0]
    V = np.empty((n, ndof))
    T = np.empty((n, ndof))
    if ndof == 1:
        V[:, 0] = 1.0
        T[:, 0] = -w["mu"] / w["cs"] * KN
    else:
        V[:, 0], V[:, 1] = P[0], P[1]
        if w["phase"] == "P":
            T[:, 0] = -1.0 / w["cp"] * (w["lam"] * N[:, 0] + 2.0 * w["mu"] * KN * K[0])
            T[:, 1] = -1.0 / w["cp"] * (w["lam"] * N[:, 1] + 2.0 * w["mu"] * KN * K[1])
        else:
            PN = P[0] * N[:, 0] + P[1] * N[:, 1]
            T[:, 0] = -w["mu"] / w["cs"] * (PN * K[0] + KN * P[0])
            T[:, 1] = -w["mu"] / w["cs"] * (PN * K[1] + KN * P[1])
    return V, T


def _lsf_init(par, bd, mesh, num, perm_inv, jac, wgll, M, ndof, perio):
    """BC_LSF_read / BC_LSF_init (SRC/bc_lsf.f90:48-156): linear slip fault between the
    node lists of tags(1) and tags(2); stiffness from the compliance when given, huge(1d0)
    (welded) when neither is set (the reference leaves kt/kn unset then).  B the assembled
    weights of side 1, W1 = 1/(B (1 + M1/M2)), W2 = 1/(B (M2/M1 + 1))."""
    g = par.after(bd, "BC_LSF")
    if g is None:
        raise InputError("BC_LSF_read: BC_LSF input block not found")
    inf = sys.float_info.max
    ctang, cnorm = float(g.get("ctang", 0.0)), float(g.get("cnorm", 0.0))
    ktang, knorm = float(g.get("ktang", inf)), float(g.get("knorm", inf))
    if ctang != 0.0:
        ktang = 1.0 / ctang
    if cnorm != 0.0:
        knorm = 1.0 / cnorm
    tv = bd.get("tags", None)
    t1 = _bound_topo(int(tv[0]), mesh, num, perm_inv)
    t2 = _bound_topo(int(tv[1]), mesh, num, perm_inv)
    if len(t1.elem) != len(t2.elem):
        raise InputError("bc_lsf_init: number of boundary elements do not match")
    if len(t1.node) != len(t2.node):
        raise InputError("bc_lsf_init: number of nodes on boundaries do not match")
    coord = num["coord"]
    if (np.abs(coord[t1.node - 1] - coord[t2.node - 1]) > TINY_XABS).any():
        raise InputError("bc_lsf_init: coordinates on boundaries do not match properly")
    NORM, W = _normal_and_weights(t1, jac, wgll, _perio_intersects(t1, perio))
    B = np.repeat(W[:, None], ndof, axis=1)
    M1, M2 = M[:, t1.node - 1].T, M[:, t2.node - 1].T
    W1 = 1.0 / (B * (1.0 + M1 / M2))
    W2 = 1.0 / (B * (M2 / M1 + 1.0))
    return dict(node1=t1.node.astype(np.int32), node2=t2.node.astype(np.int32), n=NORM, B=B, W1=W1, W2=W2,
                kt=ktang, kn=knorm)


def _moment_stencil(iglob, Mt, ibool, jac, hprime, ndof):
    """SRC_MOMENT_init (SRC/src_moment.f90:170-213)."""
    J11, J12, J21, J22 = jac
    hits = np.argwhere(ibool == iglob)  # (e, j, i) sorted by e, then j, then i
    iglob_xi, iglob_eta, coef_xi, coef_eta = [], [], [], []
    for e, j, i in hits:
        iglob_xi.append(ibool[e, j, :].copy())
        iglob_eta.append(ibool[e, :, i].copy())
        a11, a12, a21, a22 = J11[e, j, i], J12[e, j, i], J21[e, j, i], J22[e, j, i]
        det = a11 * a22 - a12 * a21
        ji = np.array([[a22 / det, (-a12) / det], [(-a21) / det, a11 / det]])
        cx = np.zeros((ndof, ibool.shape[1]))
        ce = np.zeros((ndof, ibool.shape[1]))
        if ndof == 2:
            G = np.zeros((2, 2))
            jt = ji.T
            for r in range(2):
                for c in range(2):
                    G[r, c] = Mt[r, 0] * jt[0, c] + Mt[r, 1] * jt[1, c]
            cx[0] = G[0, 0] * hprime[:, i]
            ce[0] = G[0, 1] * hprime[:, j]
            cx[1] = G[1, 0] * hprime[:, i]
            ce[1] = G[1, 1] * hprime[:, j]
        else:
            G0 = ji[0, 0] * Mt[0, 0] + ji[0, 1] * Mt[1, 0]
            G1 = ji[1, 0] * Mt[0, 0] + ji[1, 1] * Mt[1, 0]
            cx[0] = G0 * hprime[:, i]
            ce[0] = G1 * hprime[:, j]
        coef_xi.append(cx)
        coef_eta.append(ce)
    return dict(iglob_xi=np.array(iglob_xi, dtype=np.int32), iglob_eta=np.array(iglob_eta, dtype=np.int32),
                coef_xi=np.array(coef_xi), coef_eta=np.array(coef_eta))


def _hgll(i, z, xgll):
    """HGLL (SRC/gll.f90:270-296): Lagrange interpolant through the GLL points."""
    dz = z - xgll[i]
    if abs(dz) < 1.0e-5:
        return 1.0
    n = len(xgll) - 1
    alfan = float(n) * (float(n) + 1.0)
    return -(((1.0 - z * z) * _pndleg(z, n)) / ((alfan * _pnleg(xgll[i], n)) * (z - xgll[i])))


def _q4(s, t):
    """Q4_getshape / Q4_getdershape (SRC/elem_q4.f90:37-78)."""
    q = 0.25
    sp, sm, tp, tm = s + 1.0, s - 1.0, t + 1.0, t - 1.0
    shape = [(q * sm) * tm, -((q * sp) * tm), (q * sp) * tp, -((q * sm) * tp)]
    der = [[q * tm, q * sm], [-(q * tm), -(q * sp)], [q * tp, q * sp], [-(q * tp), -(q * sm)]]
    return shape, der


def _find_point(p, cxe, cze, xi, eta):
    """FE_find_point (SRC/fem_grid.f90:443-500) on a Q4 element: Newton iterations for
    the local coordinates of point p from the initial guess (xi, eta)."""
    coorg = (cxe, cze)
    x1, x2 = cxe[0] - cxe[2], cze[0] - cze[2]
    y1, y2 = cxe[1] - cxe[3], cze[1] - cze[3]
    esize = math.sqrt((0.5 * abs(x1 * y2 - x2 * y1)) / PI)
    tolx, tolf = TINY_XABS, TINY_XABS * esize
    x = [xi, eta]
    n = 1
    while n <= 100:
        shap, dshap = _q4(x[0], x[1])
        fvec = [(((c[0] * shap[0] + c[1] * shap[1]) + c[2] * shap[2]) + c[3] * shap[3]) - q
                for c, q in zip(coorg, p)]
        fjac = [[((c[0] * dshap[0][k] + c[1] * dshap[1][k]) + c[2] * dshap[2][k]) + c[3] * dshap[3][k]
                 for k in range(2)] for c in coorg]
        if abs(fvec[0]) + abs(fvec[1]) <= tolf:
            break
        pv = [-fvec[0], -fvec[1]]
        det = fjac[0][0] * fjac[1][1] - fjac[0][1] * fjac[1][0]
        if det <= 0.0:
            raise InputError("SE_InverseJacobian: undefined Jacobian")
        B = [[fjac[1][1] / det, (-fjac[0][1]) / det], [(-fjac[1][0]) / det, fjac[0][0] / det]]
        pv = [B[0][0] * pv[0] + B[0][1] * pv[1], B[1][0] * pv[0] + B[1][1] * pv[1]]
        x = [x[0] + pv[0], x[1] + pv[1]]
        if abs(pv[0]) + abs(pv[1]) <= tolx:
            break
        n += 1
    if n >= 100:
        raise InputError("FE_find_point: did not converge")
    shap, _ = _q4(x[0], x[1])
    new = [((c[0] * shap[0] + c[1] * shap[1]) + c[2] * shap[2]) + c[3] * shap[3] for c in coorg]
    return x[0], x[1], new


def _receivers(par, coord, nt, num=None, xgll=None):
    g = par.first("REC_LINE")
    if g is None:
        return None
    number = int(g.get("number", 0))
    isamp = int(g.get("isamp", 1))
    field = str(g.get("field", "V")).strip()
    if field not in ("D", "V", "A"):
        raise InputError('REC_read: parameter field has wrong value [D,V,A]')
    atnode = bool(g.get("atnode", True))
    if g.get("file", "none") != "none":
        raise InputError("REC_LINE file= is not supported by this build")
    first = [float(v) for v in g.get("first")]
    last = [float(v) for v in g.get("last")]
    pts = []
    for i in range(1, number + 1):
        if number > 1:
            f = (i - 1) / float(number - 1)
            pts.append((first[0] + f * (last[0] - first[0]), first[1] + f * (last[1] - first[1])))
        else:
            pts.append((first[0], first[1]))
    if not atnode:
        # REC_posit, AtNode = F (SRC/receivers.f90:274-285): SE_find_point takes the first
        # element (element order, then j, then i) holding the nearest GLL node, starts
        # Newton from that node's (xi, eta), and -- FE_find_point tests the initial guess,
        # which is inside [-1,1] -- always keeps that element, extrapolating if needed.
        ibool = num["ibool"]
        elems, interp, rc = [], [], []
        for p in pts:
            ip, _ = _nearest_node(coord, p)
            e, j, i = (int(v) for v in np.argwhere(ibool == ip)[0])
            xi, eta, new = _find_point(p, num["cx"][e], num["cz"][e], xgll[i], xgll[j])
            ng = len(xgll)
            w = np.empty(ng * ng)
            for jj in range(ng):
                fj = _hgll(jj, eta, xgll)
                for ii in range(ng):
                    w[jj * ng + ii] = _hgll(ii, xi, xgll) * fj
            elems.append(e + 1)
            interp.append(w)
            rc.append(new)
        return dict(atnode=False, elem=np.array(elems, dtype=np.int32), interp=np.array(interp), isamp=isamp,
                    field=field, nsamp=nt // isamp + 1, coord=np.array(rc),
                    iglob=np.zeros(len(pts), dtype=np.int32))
    nodes = []
    for p in pts:  # REC_posit dedup (SRC/receivers.f90:247-272), incl. its "irec > 1" test
        ip, _ = _nearest_node(coord, p)
        if len(nodes) > 1 and ip in nodes:
            continue
        nodes.append(ip)
    return dict(atnode=True, iglob=np.array(nodes, dtype=np.int32), isamp=isamp, field=field,
                nsamp=nt // isamp + 1, coord=coord[np.array(nodes) - 1])


# ---------------------------------------------------------------------------
def build(par, nsteps=None, material=None, fname_dir=None, per_element_coefs=True, elem_props=None):
    """Build a :class:`Problem` from a parsed Par.inp (:func:`parfile.read_par`).

    ``material``, if given, is a callable ``(x, z) -> (rho, cp, cs)`` evaluated at the
    GLL nodes of every element; it replaces the MAT_ELASTIC blocks (one tag), as a
    distribution would.  ``elem_props = (rho, lam, mu)``, arrays over the cells
    (index iz*nx + ix), gives every element constant properties instead (the model
    space of the adjoint).  ``fname_dir`` is where relative data file names are read
    (default: the directory of the input file).  ``per_element_coefs=False`` skips
    the reference's per-element coefficient arrays for heterogeneous media (only the
    compact lambda/mu form is then built).
    """
    if isinstance(par, (str, os.PathLike)):
        if fname_dir is None:
            fname_dir = os.path.dirname(os.path.abspath(par))
        par = _pf.read_par(par)
    gen = par.first("GENERAL")
    if gen is None:
        raise InputError("GENERAL input block not found")
    ngll = int(gen.get("ngll", 9))
    ndof = int(gen.get("ndof", 2))
    fmax = float(gen.get("fmax", 0.0))
    if not 3 <= ngll <= 10:
        raise InputError("this build supports ngll from 3 to 10")
    if ndof not in (1, 2):
        raise InputError("ndof must be 1 or 2")
    # seismogenic width of 2.5D runs (SRC/input.f90:82-83,119,130): W < huge adds the
    # crustal-plane term of MAT_ELAST_add_25D_f (SRC/mat_elastic.f90:447-459)
    W25 = float(gen.get("w", 0.0)) if "w" in gen.values else None
    if W25 is not None and W25 <= 0.0:
        raise InputError("GENERAL input block: W must be positive")
    if W25 is not None and W25 >= sys.float_info.max:
        W25 = None
    md = par.first("MESH_DEF")
    method = str(md.get("method", "")).upper() if md is not None else ""
    if method == "CARTESIAN":
        mesh = _cart_mesh(par.first("MESH_CART"), par.all("MESH_CART_DOMAIN"))
    elif method == "MESH2D":
        mesh = _read_mesh2d(par.first("MESH_MESH2D"), fname_dir)
    elif method == "LAYERED":
        g = par.first("MESH_LAYERED")
        if g is None:
            raise InputError("MESH_LAYERS_read: input block not found")
        mesh = _layered_mesh(par, g, fname_dir)
    else:
        raise InputError("this build supports MESH_DEF method='CARTESIAN', 'LAYERED' or 'MESH2D'")
    # a MESH_CART box is a flat grid (SRC/mesh_cartesian.f90:242); MESH2D and LAYERED
    # meshes are not (SRC/mesh_layers.f90:369), which selects the general element
    # coefficients.  CARTESIAN and LAYERED share the structured numbering.
    flat = mesh.get("kind") not in ("mesh2d", "layered")
    structured = mesh.get("kind") != "mesh2d"

    pb = Problem()
    pb.ngll, pb.ndof = ngll, ndof
    # progress line interval (SRC/input.f90:95,122,131: default 100, must be positive)
    pb.itinfo = int(gen.get("itinfo", 100))
    if pb.itinfo <= 0:
        raise InputError("GENERAL input block: itInfo must be positive")
    pb.mesh = mesh
    xgll, wgll, H = gll_info(ngll)
    pb.xgll, pb.wgll, pb.hprime = xgll, wgll, H
    nx, nz = mesh["nx"], mesh["nz"]
    if structured:
        pb.nelem = nelem = nx * nz
        perm = _rcm_perm(nx, nz)
        perm_inv = np.empty(nelem, dtype=np.int64)
        perm_inv[perm - 1] = np.arange(nelem)
        num = _numbering(mesh, perm, xgll)
    else:  # MESH2D: the file's element order (no renumbering)
        if elem_props is not None:
            raise InputError("per-element model arrays need a MESH_CART grid")
        pb.nelem = nelem = mesh["nelem"]
        perm = np.arange(1, nelem + 1, dtype=np.int32)
        perm_inv = np.arange(nelem, dtype=np.int64)
        num = _numbering_fem(mesh, xgll)
    pb.perm = perm
    pb.ibool = num["ibool"]
    pb.coord = num["coord"]
    pb.npoin = npoin = num["npoin"]
    pb.elem_ix, pb.elem_iz = num["ix"], num["iz"]

    # element tags (CART_build, SRC/mesh_cartesian.f90:277-297): the domains in input
    # order (a later one overrides an earlier one), then the fault zone
    if structured:
        tags = np.zeros(nelem, dtype=np.int64)
        for tag, ex, ez in mesh["domains"]:
            inside = ((num["ix"] + 1 >= ex[0]) & (num["ix"] + 1 <= ex[1]) &
                      (num["iz"] + 1 >= ez[0]) & (num["iz"] + 1 <= ez[1]))
            tags[inside] = tag
    else:
        tags = mesh["tag"].copy()
    if structured and mesh["fztag"] > 0:
        j1 = max(mesh["ezflt"] + 1 - mesh["fznz"], 1)
        j2 = min(mesh["ezflt"] + mesh["fznz"], nz)
        rows = num["iz"] + 1
        tags[(rows >= j1) & (rows <= j2)] = mesh["fztag"]
    if (tags == 0).any():
        raise InputError("CART_build: Domain tags not entirely set")
    pb.tags = tags
    ibool0 = pb.ibool - 1
    ex = pb.coord[ibool0, 0]
    ez = pb.coord[ibool0, 1]
    # material properties at the GLL nodes of each element (MAT_init_prop,
    # SRC/mat_gen.f90:204-303; MAT_ELAST_init_elem_prop, SRC/mat_elastic.f90:194-234)
    shp = (nelem, ngll, ngll)
    rho_p, cp_p, cs_p, lam_p, mu_p = (np.empty(shp) for _ in range(5))
    homog_tag = {}
    aniso_tags = {}
    if elem_props is not None:
        cell = num["iz"] * nx + num["ix"]
        r, la, m = (np.asarray(a, dtype=np.float64)[cell][:, None, None] for a in elem_props)
        rho_p[:], lam_p[:], mu_p[:] = r, la, m
        cp_p[:] = np.sqrt((lam_p + 2.0 * mu_p) / rho_p)
        cs_p[:] = np.sqrt(mu_p / rho_p)
        homog_tag = {int(t): False for t in np.unique(tags)}
    elif material is not None:
        r, c1, c2 = material(ex, ez)
        rho_p[:], cp_p[:], cs_p[:] = r, c1, c2
        lam_p[:] = rho_p * (cp_p * cp_p - 2.0 * cs_p * cs_p)
        mu_p[:] = rho_p * cs_p * cs_p
        homog_tag = {int(t): False for t in np.unique(tags)}
    else:
        mats = _materials(par, fname_dir)
        kv_tags = {t: m["kv_eta"] for t, m in mats.items() if isinstance(t, int) and "kv_eta" in m}
        for t in np.unique(tags):
            if int(t) not in mats:
                raise InputError("MAT_init_prop: material tag missing")
            m = mats[int(t)]
            sel = tags == t
            homog_tag[int(t)] = m["homog"]
            if m.get("aniso"):
                rho_p[sel], cp_p[sel], cs_p[sel] = m["rho"], m["cp"], m["cs"]
                lam_p[sel], mu_p[sel] = m["c13"], m["c55"]
                aniso_tags[int(t)] = m
            elif m["homog"]:
                rho_p[sel], cp_p[sel], cs_p[sel] = m["rho"], m["cp"], m["cs"]
                lam_p[sel], mu_p[sel] = m["lam"], m["mu"]
            else:
                xs, zs = ex[sel], ez[sel]
                r = _eval_dist(m["spec"]["rho"], xs, zs)
                c1 = _eval_dist(m["spec"]["cp"], xs, zs)
                c2 = _eval_dist(m["spec"]["cs"], xs, zs)
                rho_p[sel], cp_p[sel], cs_p[sel] = r, c1, c2
                lam_p[sel] = r * (c1 * c1 - 2.0 * c2 * c2)
                mu_p[sel] = r * c2 * c2
    pb.rho, pb.cp, pb.cs, pb.lam, pb.mu = rho_p, cp_p, cs_p, lam_p, mu_p
    # moduli of the coefficient arrays (SRC/mat_elastic.f90:307-321): isotropic
    # Kx = Kz = lambda + 2 mu, mux = muz = mu; transversely anisotropic Kx = c11, Kz = c33,
    # la = c13, muz = c55, mux = c66 (lam_p and mu_p then hold c13 and c55)
    kx_p = lam_p + 2.0 * mu_p
    kz_p, mux_p = kx_p, mu_p
    if aniso_tags:
        kz_p, mux_p = kx_p.copy(), mu_p.copy()
        for t, m in aniso_tags.items():
            sel = tags == t
            kx_p[sel], kz_p[sel], mux_p[sel] = m["c11"], m["c33"], m["c66"]
    pb.anisotropic = bool(aniso_tags)
    mod = (kx_p, kz_p, lam_p, mu_p, mux_p)
    kv_eta0 = None
    if material is None and elem_props is None and kv_tags:
        kv_eta0 = np.zeros((nelem, ngll, ngll))
        for t, spec in kv_tags.items():
            sel = tags == t
            # MAT_KV_init_elem_prop: eta at the element's GLL points
            pts = np.stack([ex[sel].ravel(), ez[sel].ravel()], axis=1)
            kv_eta0[sel] = _gen_cd(spec, pts).reshape(-1, ngll, ngll)
    pb.homogeneous = all(homog_tag.values())

    jac = _jacobians(num["cx"], num["cz"], xgll)
    J11, J12, J21, J22 = jac
    DxiDx, DxiDz, DetaDx, DetaDz = _inv_jac(*jac)
    wgll2 = wgll[None, :] * wgll[:, None]  # [j, i] = wgll(i)*wgll(j)
    dvol = (J11 * J22 - J12 * J21) * wgll2[None]

    # CHECK_grid + TIME_init
    d1 = np.sqrt((ex[:, :-1, 1:] - ex[:, :-1, :-1]) ** 2 + (ez[:, :-1, 1:] - ez[:, :-1, :-1]) ** 2)
    d2 = np.sqrt((ex[:, 1:, :-1] - ex[:, :-1, :-1]) ** 2 + (ez[:, 1:, :-1] - ez[:, :-1, :-1]) ** 2)
    cvel = cp_p if ndof == 2 else cs_p
    grid_cfl = float((cvel[:, :-1, :-1] / np.minimum(d1, d2)).max())
    tg = par.first("TIME")
    kind = str(tg.get("kind", "leapfrog")).strip().lower() if tg is not None else "leapfrog"
    if kind not in ("newmark", "leapfrog", "hht-alpha"):
        raise InputError("this build supports TIME kind='newmark', 'leapfrog' or 'HHT-alpha' only")
    nbsteps = int(tg.get("nbsteps", 0))
    dt_in = float(tg.get("dt", 0.0))
    courant = float(tg.get("courant", 0.5))
    total = float(tg.get("totaltime", 0.0))
    if courant < 0 or courant > 0.6:
        raise InputError("TIME: Courant out of range [0,0.6]")
    if nbsteps * total != 0:
        raise InputError("TIME: bad combination of settings, NbSteps or TotalTime")
    # TIME_read (SRC/time.f90:195-215): alpha = 1, beta = 0, gamma = 1/2 unless a Newmark
    # run gives TIME_NEWMARK
    tn = par.after(tg, "TIME_NEWMARK") if kind == "newmark" else None
    beta = float(tn.get("beta", 0.0)) if tn is not None else 0.0
    gamma = float(tn.get("gamma", 0.5)) if tn is not None else 0.5
    alpha = 1.0
    if kind == "hht-alpha":
        # TIME_HHTA (SRC/time.f90:219-235): alpha and rho (defaults 0.5), gamma = 3/2 -
        # alpha, beta from rho (0 when alpha = 1); second-order explicit HHT
        th = par.after(tg, "TIME_HHTA")
        alpha = float(th.get("alpha", 0.5)) if th is not None else 0.5
        rho = float(th.get("rho", 0.5)) if th is not None else 0.5
        if alpha < 0.0 or alpha > 1.0:
            raise InputError("TIME_HHTA: alpha is out of range [0,1]")
        if rho < 0.5 or rho > 1.0:
            raise InputError("TIME_HHTA: rho is out of range [0.5,1]")
        gamma = 1.5 - alpha
        if alpha != 1.0:
            beta = 1.0 - alpha - rho ** 2 * (rho - 1.0) / ((1.0 - alpha) * (1.0 + rho) ** 3)
        else:
            beta = 0.0
    if dt_in > 0:
        if total > 0:
            nbsteps = math.ceil(total / dt_in)
        dt = dt_in
        courant = grid_cfl * dt
    else:
        dt = courant / grid_cfl
        if total > 0:
            nbsteps = math.ceil(total / dt)
    # Omega_max: 2 for leapfrog, sqrt(2/gamma) for Newmark, the HHT-alpha bound of rho
    # (SRC/time.f90:205-235)
    if kind == "hht-alpha":
        omega_max = math.sqrt(-4.0 * (1.0 + rho) ** 5 / (rho ** 5 + 9.0 * rho ** 4 - 34.0 * rho ** 3 + 6.0 * rho ** 2
                                                          - 15.0 * rho + 1.0))
    else:
        omega_max = 2.0 if kind == "leapfrog" else math.sqrt(2.0 / gamma)
    critical = omega_max * 3.0 / 7.0 / math.sqrt(2.0)
    if courant > critical:
        raise InputError("TIME_init: Courant number too high (IO_warning aborts)")
    pb.dt, pb.beta, pb.gamma, pb.alpha = dt, beta, gamma, alpha
    pb.scheme = "hht" if kind == "hht-alpha" else kind
    # Kelvin-Voigt viscosity per element point (MAT_KV_init_elem_work,
    # SRC/mat_kelvin_voigt.f90:121-135): eta, times dt when ETAxDT
    pb.kv_eta = None
    if kv_eta0 is not None:
        eta = kv_eta0
        pb.kv_eta = dt * eta if mats.get("_kv_by_dt", True) else eta
    pb.nt = int(nbsteps if nsteps is None else nsteps)
    pb.grid_cfl = grid_cfl

    # MAT_init_work: flat grid, elastic.  Homogeneous materials share the
    # array of the first element of their tag (SRC/mat_gen.f90:357-365); other
    # elements get their own array (SRC/mat_elastic.f90:242-360).
    pb.nelast = (2 if ndof == 1 else 6) if flat else (3 if ndof == 1 else 10)
    utags = np.unique(tags)
    first_of = {int(t): int(np.nonzero(tags == t)[0][0]) for t in utags}
    if not flat:
        # every element its own general-geometry array (SRC/mat_gen.f90:360-364 shares
        # arrays only on flat grids)
        pb.coef_mode = "per_element"
        coefs = np.empty((nelem, pb.nelast, ngll, ngll))
        for c0 in range(0, nelem, 65536):
            b = slice(c0, min(c0 + 65536, nelem))
            coefs[b] = _elastic_coefs_general(DxiDx[b], DxiDz[b], DetaDx[b], DetaDz[b], dvol[b],
                                              *(q[b] for q in mod), ndof)
        pb.coefs = coefs
    elif len(utags) == 1 and pb.homogeneous:
        e1 = 0
        pb.coefs = _elastic_coefs(DxiDx[e1], DetaDz[e1], dvol[e1], *(q[e1] for q in mod), ndof)
        pb.coef_mode = "shared"
    else:
        pb.coef_mode = "per_element"
        if per_element_coefs or pb.homogeneous or ngll != 5 or pb.anisotropic:
            coefs = np.empty((nelem, pb.nelast, ngll, ngll))
            for t in utags:
                sel = tags == t
                if homog_tag[int(t)]:
                    e1 = first_of[int(t)]
                    coefs[sel] = _elastic_coefs(DxiDx[e1], DetaDz[e1], dvol[e1], *(q[e1] for q in mod), ndof)
                else:
                    idx = np.nonzero(sel)[0]
                    for c0 in range(0, len(idx), 65536):
                        b = idx[c0:c0 + 65536]
                        coefs[b] = _elastic_coefs_batch(DxiDx[b], DetaDz[b], dvol[b], *(q[b] for q in mod), ndof)
            pb.coefs = coefs
        else:
            pb.coefs = None
    # compact form for heterogeneous flat grids: lambda, mu per GLL node and the three
    # geometric factors of each element, -detJ*DxiDx^2, -detJ*DetaDz^2, -detJ*DxiDx*DetaDz
    # (taken at the element's first GLL node); a(:,:,k) = wgll2 * material * factor.
    pb.geom = pb.lam_mu = None
    if flat and ngll == 5 and not (len(utags) == 1 and pb.homogeneous) and not pb.anisotropic:
        detJ = J11[:, 0, 0] * J22[:, 0, 0] - J12[:, 0, 0] * J21[:, 0, 0]
        gx, gz = DxiDx[:, 0, 0], DetaDz[:, 0, 0]
        pb.geom = np.stack([-detJ * gx * gx, -detJ * gz * gz, -detJ * gx * gz], axis=1)
        pb.lam_mu = np.stack([lam_p, mu_p], axis=1) if ndof == 2 else mu_p[:, None]
    pb.wgll2 = wgll2

    # MAT_MASS_init: sequential element-order assembly (np.add.at is sequential)
    pb.dvol = dvol  # w_i w_j detJ at every element-local point (dM/drho)
    # 2.5D: beta = dvol * mu * (pi/W)^2 (SH) or dvol * mu * (pi (1-nu)/W)^2 (P-SV), with
    # nu = lambda/(lambda+mu)/2 and pi = 4 atan(1) (MAT_ELAST_init_25D, SRC/mat_elastic.f90:363-384)
    pb.beta25 = None
    if W25 is not None:
        if elem_props is not None or material is not None:
            raise InputError("2.5D (W) with per-element model arrays is not supported by this build")
        if pb.anisotropic:
            # MAT_ELAST_init_25D reads the isotropic lambda and mu
            raise InputError("2.5D (W) with anisotropic materials is not supported")
        pi = 4.0 * math.atan(1.0)
        if ndof == 1:
            q = pi / W25
            beta25 = dvol * pb.mu * (q * q)
        else:
            nu = pb.lam / (pb.lam + pb.mu) / 2.0
            q = pi * (1.0 - nu) / W25
            beta25 = dvol * pb.mu * (q * q)
        if flat:
            # homogeneous materials on a flat grid share the work arrays (beta included)
            # of the first element of their tag (SRC/mat_gen.f90:357-365)
            for t in utags:
                if homog_tag[int(t)]:
                    beta25[tags == t] = beta25[first_of[int(t)]]
        pb.beta25 = beta25
    massloc = rho_p * dvol
    M = np.zeros((ndof, npoin))
    np.add.at(M[0], ibool0.ravel(), massloc.ravel())
    if ndof == 2:
        M[1] = M[0]
    pb.mass_raw = M.copy()

    # BC_init (SRC/bc_gen.f90:190-248): periodic boundaries first, then the others in
    # list order
    # TIME_getCoefA2Vrhs (SRC/time.f90:467-486): alpha gamma dt (Newmark, HHT-alpha), dt/2
    # (leapfrog, gamma = 1/2)
    coefA2Vrhs = 0.5 * dt if pb.scheme == "leapfrog" else pb.alpha * (gamma * dt)
    bcdefs = par.all("BC_DEF")
    present = set(_bound_elements(mesh).keys())
    pb.dirichlet = []
    pb.periodic = None
    pb.lsf = None

    def _tags_of(bd):
        tag = bd.get("tag", 0)
        tagsv = bd.get("tags", None)
        t1 = int(tag) if int(tag) > 0 else (int(tagsv[0]) if isinstance(tagsv, list) else
                                              (int(tagsv) if tagsv is not None else 0))
        t2 = int(tagsv[1]) if isinstance(tagsv, list) and len(tagsv) > 1 else 0
        if t1 <= 0:
            raise InputError("bc_read: tag(s) are null or not set")
        return t1, t2

    for bd in bcdefs:
        if bd.get("kind", " ") != "PERIOD":
            continue
        t1, t2 = _tags_of(bd)
        if t1 not in present or (t2 and t2 not in present):
            continue  # IS_EMPTY
        if pb.periodic is not None:
            raise InputError("more than one PERIOD boundary is not supported by this build")
        pb.periodic = _perio_init(bd, mesh, num, perm_inv, M)
    for bd in bcdefs:
        kind = bd.get("kind", " ")
        t1, t2 = _tags_of(bd)
        if t1 not in present or (t2 and t2 not in present):
            continue  # IS_EMPTY
        if kind == "PERIOD":
            continue
        if kind == "ABSORB":
            ab = par.after(bd, "BC_ABSORB")
            stacey = bool(ab.get("stacey", False)) if ab is not None else False
            let_wave = bool(ab.get("let_wave", True)) if ab is not None else True
            t = _bound_topo(t1, mesh, num, perm_inv)
            ad = _abso_init(t, jac, wgll, (rho_p, cp_p, cs_p), M, coefA2Vrhs, ndof, stacey,
                            _perio_intersects(t, pb.periodic), let_wave)
            ad["coord"] = num["coord"][t.node - 1]
            pb.abso.append(ad)
        elif kind == "LISFLT":
            if pb.lsf is not None:
                raise InputError("more than one LISFLT boundary is not supported by this build")
            pb.lsf = _lsf_init(par, bd, mesh, num, perm_inv, jac, wgll, M, ndof, pb.periodic)
        elif kind == "DIRNEU":
            # BC_DIRNEU_read (SRC/bc_dirneu.f90:50-112): Neumann with no source time
            # function is a free boundary, BC_DIRNEU_apply adds nothing
            dn = par.after(bd, "BC_DIRNEU")
            if dn is None:
                raise InputError("bc_DIRNEU_read: no BC_DIRNEU block found")
            mask = 0
            for bit, comp in enumerate(("h", "v")):
                val = str(dn.get(comp, "N")).upper()
                if val not in ("N", "D"):
                    raise InputError(f"bc_DIRNEU_read: {comp} must be N or D")
                if val == "D" and bit < ndof:  # SH: only h (SRC/bc_dirneu.f90:161)
                    mask |= 1 << bit
                elif str(dn.get(comp + "stf", "none")).lower() != "none":
                    raise InputError("BC_DIRNEU time-dependent tractions are not supported by this build")
            if mask:
                # BC_DIRNEU_init (SRC/bc_dirneu.f90:117-143): a flat, vertical or horizontal boundary
                t = _bound_topo(t1, mesh, num, perm_inv)
                nrm, _ = _normal_and_weights(t, jac, wgll, _perio_intersects(t, pb.periodic))
                if not ((np.abs(nrm[:, 0]) < TINY_XABS).all() or (np.abs(nrm[:, 1]) < TINY_XABS).all()):
                    raise InputError("BC_DIRNEU_init: boundary is not vertical or horizontal")
                # applied in list order among the non-absorbing conditions (SRC/bc_gen.f90:273-281)
                pb.dirichlet.append(dict(tag=t1, node=t.node.astype(np.int32), comp=mask,
                                         after_fault=pb.fault is not None))
        elif kind == "DYNFLT":
            if pb.fault is not None:
                raise InputError("more than one DYNFLT boundary is not supported by this build")
            pb.fault = _dynflt_init(par, bd, mesh, num, perm_inv, jac, wgll, M, dt, beta, gamma, pb.nt, ndof,
                                    scheme=pb.scheme, perio=pb.periodic)
        else:
            raise InputError(f"BC kind {kind!r} is not supported by this build")

    pb.receivers = _receivers(par, pb.coord, pb.nt, num, xgll)
    srcs = _sources(par, num, jac, H, ndof, dt, pb.nt, fname_dir,
                    t_shift=(pb.alpha - 1.0) * dt if pb.scheme == "hht" else None)
    # incident plane waves (mechanism 'WAVE') enter through the absorbing boundaries
    # (SO_add skips them; BC_ABSO_init takes the first wave source, SRC/bc_abso.f90:249-258)
    waves = [x for x in srcs if x["kind"] == "wave"]
    pb.sources = [x for x in srcs if x["kind"] != "wave"]
    pb.wave = None
    if waves:
        if pb.anisotropic:
            # WAVE_init reads the isotropic lambda and mu of its reference node
            raise InputError("WAVE sources with anisotropic materials are not supported")
        w = _wave_init(dict(waves[0]["wave"]), num, (rho_p, cp_p, cs_p, lam_p, mu_p))
        w.update(f0=waves[0]["f0"], t0=waves[0]["t0"], stf_ampli=waves[0]["stf_ampli"],
                 tdelay=waves[0]["tdelay"])
        pb.wave = w
        for ad in pb.abso:
            if ad["let_wave"]:
                ad["wave_V"], ad["wave_T"] = _wave_dirs(w, ad["wave_n"], ndof)
    if pb.lsf is not None and (pb.fault is not None or pb.dirichlet):
        raise InputError("LISFLT with DYNFLT or Dirichlet boundaries is not supported by this build")
    pb.rmass = 1.0 / M
    pb.mass = M
    return pb
