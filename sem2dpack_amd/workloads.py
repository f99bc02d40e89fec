"""Synthetic inputs of the benchmark configurations.

The reference has no random-field generator, so the heterogeneous media of the
large in-plane throughput case are produced here (seeded, numpy PCG64) and fed to
the solver exactly as the reference would read them: as a ``DIST_HETE1`` grid
(``SRC/distribution_hete1.f90``) interpolated bilinearly at the GLL nodes.

* :func:`layered_random_grid` -- 8 layers of vp from 4000 to 7000 m/s (top to
  bottom) times (1 + 0.1 g), g a unit-variance Gaussian random field with a
  Gaussian correlation function of length 2 km; vs = vp/sqrt(3);
  rho = 770 + 0.32 vp (kg/m^3).
* :func:`par_throughput` -- the 200 km x 100 km, 2000x1000-element P-SV input with
  an explosive Ricker source (f0 = 1 Hz) and absorbing boundaries.
"""
import math
import os

import numpy as np

RHO_A, RHO_B = 770.0, 0.32  # rho = RHO_A + RHO_B * vp


def gaussian_random_field(nx, nz, dx, corr_len, seed):
    """Unit-variance field with correlation exp(-r^2/corr_len^2), by FFT filtering."""
    rng = np.random.Generator(np.random.PCG64(seed))
    px, pz = int(2 ** math.ceil(math.log2(nx + 8 * corr_len / dx))), int(2 ** math.ceil(math.log2(nz + 8 * corr_len / dx)))
    w = rng.standard_normal((pz, px))
    kx = np.fft.fftfreq(px, d=dx) * 2 * np.pi
    kz = np.fft.fftfreq(pz, d=dx) * 2 * np.pi
    K2 = kz[:, None] ** 2 + kx[None, :] ** 2
    # spectral amplitude of a Gaussian correlation exp(-r^2/L^2): ~ exp(-k^2 L^2 / 8)
    filt = np.exp(-K2 * corr_len ** 2 / 8.0)
    g = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))[:nz, :nx]
    g = (g - g.mean()) / g.std()
    return g


def layered_random_grid(xlim, zlim, dx=100.0, nlayers=8, vp_top=4000.0, vp_bot=7000.0, std=0.10,
                        corr_len=2000.0, seed=20260201):
    """(x0, z0, dx, rho, vp, vs) on a regular grid [nz][nx] covering the box."""
    nx = int(round((xlim[1] - xlim[0]) / dx)) + 1
    nz = int(round((zlim[1] - zlim[0]) / dx)) + 1
    z = zlim[0] + dx * np.arange(nz)
    depth = zlim[1] - z
    H = (zlim[1] - zlim[0]) / nlayers
    layer = np.minimum((depth / H).astype(np.int64), nlayers - 1)
    vp_layer = vp_top + (vp_bot - vp_top) * layer / (nlayers - 1)
    g = gaussian_random_field(nx, nz, dx, corr_len, seed)
    vp = vp_layer[:, None] * (1.0 + std * g)
    vs = vp / math.sqrt(3.0)
    rho = RHO_A + RHO_B * vp
    return xlim[0], zlim[0], dx, rho, vp, vs


def write_hete1(path, x0, z0, dx, rho, vp, vs):
    """DIST_HETE1 file: 'ncol nx nz x0 z0 dx dz' then one row per node, x fastest."""
    nz, nx = vp.shape
    with open(path, "w") as fh:
        fh.write(f"3 {nx} {nz} {x0!r} {z0!r} {dx!r} {dx!r}\n")
        data = np.stack([rho.ravel(), vp.ravel(), vs.ravel()], axis=1)
        np.savetxt(fh, data, fmt="%.17g")


def hete1_material(x0, z0, dx, rho, vp, vs):
    """Callable (x, z) -> (rho, cp, cs) that interpolates the grid as DIST_HETE1 does."""
    from .builder import _gen_hete1
    nz, nx = vp.shape
    grids = [dict(nx=nx, nz=nz, x0=x0, z0=z0, dx=dx, dz=dx, val=v) for v in (rho, vp, vs)]

    def mat(x, z):
        return tuple(_gen_hete1(g, x, z) for g in grids)
    return mat


THROUGHPUT_PAR = """&GENERAL iexec=1, ngll=5, fmax=1d0, ndof=2, title='{title}', verbose='1101', ItInfo=1000 /
&MESH_DEF method='CARTESIAN' /
&MESH_CART xlim={x0!r},{x1!r}, zlim={z0!r},{z1!r}, nelem={nx},{nz} /
&MATERIAL tag=1, kind='ELAST' /
&MAT_ELASTIC rhoH='HETE1', cpH='HETE1', csH='HETE1' /
&DIST_HETE1 file='{hete}', col=1 /
&DIST_HETE1 file='{hete}', col=2 /
&DIST_HETE1 file='{hete}', col=3 /
&BC_DEF tag=1, kind='ABSORB' /
&BC_DEF tag=2, kind='ABSORB' /
&BC_DEF tag=3, kind='ABSORB' /
&BC_DEF tag=4, kind='ABSORB' /
&TIME kind='newmark', NbSteps={nt}, Courant=0.3d0 /
&SRC_DEF stf='RICKER', coord={sx!r},{sz!r}, mechanism='EXPLOSION' /
&STF_RICKER f0={f0}, onset={onset}, ampli=1d0 /
&REC_LINE number={nrec}, first={r0x!r},{rz!r}, last={r1x!r},{rz!r}, field='V' /
&SNAP_DEF itd=1000000, ps=F, bin=F /
"""


def par_throughput(nx=2000, nz=1000, lx=200e3, lz=100e3, nt=5000, hete="c3_media.tab", f0="1d0",
                   onset="1d0", nrec=16, title="P-SV throughput: layered random medium, explosion"):
    return THROUGHPUT_PAR.format(title=title, x0=0.0, x1=lx, z0=0.0, z1=lz, nx=nx, nz=nz, hete=hete, nt=nt,
                                 sx=lx / 2, sz=lz / 2, f0=f0, onset=onset, nrec=nrec, r0x=lx * 0.1,
                                 r1x=lx * 0.9, rz=lz * 0.99)


def build_throughput(nx=2000, nz=1000, lx=200e3, lz=100e3, nt=5000, seed=20260201, per_element_coefs=False,
                     dx=100.0):
    """Build the throughput problem in memory (same media as the HETE1 file route)."""
    from . import builder, parfile
    grid = layered_random_grid((0.0, lx), (0.0, lz), dx=dx, seed=seed)
    par = parfile.parse_par(par_throughput(nx, nz, lx, lz, nt))
    return builder.build(par, nsteps=nt, material=hete1_material(*grid), per_element_coefs=per_element_coefs)


# The reference's own CPU-runnable case: a 10 km x 10 km homogeneous P-SV box of
# 100 x 100 elements, vertical point force (Ricker f0 = 2 Hz, onset 0.5 s) at the
# centre, four absorbing sides, Courant 0.3, 2000 Newmark steps, 9 velocity
# receivers on the line z = 7 km, 2 km above the source.
C1_PAR = """&GENERAL iexec=1, ngll=5, fmax=5d0, ndof=2, title='P-SV box, point force', verbose='1101', ItInfo=500 /
&MESH_DEF method='CARTESIAN' /
&MESH_CART xlim=0d0,10d3, zlim=0d0,10d3, nelem=100,100 /
&MATERIAL tag=1, kind='ELAST' /
&MAT_ELASTIC rho=2670d0, cp=6000d0, cs=3464d0 /
&BC_DEF tag=1, kind='ABSORB' /
&BC_DEF tag=2, kind='ABSORB' /
&BC_DEF tag=3, kind='ABSORB' /
&BC_DEF tag=4, kind='ABSORB' /
&TIME kind='newmark', NbSteps={nt}, Courant=0.3d0 /
&SRC_DEF stf='RICKER', coord=5d3,5d3, mechanism='FORCE' /
&STF_RICKER f0=2d0, onset=0.5d0, ampli=1d0 /
&SRC_FORCE angle=0d0 /
&REC_LINE number=9, first=3d3,7d3, last=7d3,7d3, field='V' /
&SNAP_DEF itd=100000, ps=F, bin=F /
"""


def par_c1(nt=2000):
    return C1_PAR.format(nt=nt)


def build_c1(nt=2000):
    """Build the 100x100-element P-SV point-force box in memory."""
    from . import builder, parfile
    return builder.build(parfile.parse_par(par_c1(nt)), nsteps=nt)


# ---------------------------------------------------------------------------
# Batched rupture ensemble
# ---------------------------------------------------------------------------
# Each scenario: P-SV box of nx x nz elements of 100 m (nz = nx/2), homogeneous
# medium (rho 2670, cp 6000, cs 3464), a horizontal split-node fault through the
# middle (ezflt = -1), rate-and-state aging law (kind 2) with a ~ U(0.008, 0.016),
# b = 0.012, Dc = 0.02 m, V* = 1e-6 m/s, f* = 0.6, normal stress -100 MPa, initial
# slip rate 1e-9 m/s and state Dc/V, four absorbing boundaries, and an initial shear
# stress 60 MPa * (1 + 0.05 g(x)), g a unit-variance 1D Gaussian random field of
# correlation length 1 km, given to the solver as a DIST_LINEAR table sampled every
# 100 m.  A fixed time step of 0.86 ms (Courant 0.299) is shared by all scenarios
# (a Courant number would give each mesh a dt differing in the last bits).  Fault
# outputs every 10 steps at every 4th fault node.
RUPTURE_DX = 100.0
RUPTURE_SEED = 20260401


def gaussian_random_field_1d(n, dx, corr_len, seed):
    """Unit-variance 1D field with correlation exp(-r^2/corr_len^2), by FFT filtering."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = int(2 ** math.ceil(math.log2(n + 8 * corr_len / dx)))
    w = rng.standard_normal(p)
    k = np.fft.fftfreq(p, d=dx) * 2 * np.pi
    g = np.real(np.fft.ifft(np.fft.fft(w) * np.exp(-k ** 2 * corr_len ** 2 / 8.0)))[:n]
    return (g - g.mean()) / g.std()


def rupture_ensemble(n=256, seed=RUPTURE_SEED, nx_min=100, nx_max=400):
    """Scenario parameters: list of dicts (nx, nz, a, seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nxs = rng.integers(nx_min, nx_max + 1, size=n)
    avals = rng.uniform(0.008, 0.016, size=n)
    return [dict(nx=int(nx), nz=int(nx) // 2, a=float(a), seed=seed + 1 + s)
            for s, (nx, a) in enumerate(zip(nxs, avals))]


def par_rupture(nx, nz, a, seed, nt=3000, dx=RUPTURE_DX, tt_mean=60e6, tt_std=0.05, corr_len=1000.0,
                dt="8.6d-4", otd="8.6d-3", oxd=4):
    """Par.inp text of one ensemble scenario (see the block comment above)."""
    lx, lz = nx * dx, nz * dx
    xs = -0.5 * lx + dx * np.arange(nx + 1)
    tt = tt_mean * (1.0 + tt_std * gaussian_random_field_1d(nx + 1, dx, corr_len, seed))
    table = "\n".join(f" {float(x)!r} {float(t)!r}" for x, t in zip(xs, tt))
    bcs = "".join(f"&BC_DEF tag={t}, kind='ABSORB' /\n" for t in (1, 2, 3, 4))
    return (f"&GENERAL iexec=1, ngll=5, fmax=1d0, ndof=2, title='RSF ensemble scenario {seed}', "
            f"verbose='0000', ItInfo=100000 /\n"
            f"&MESH_DEF method='CARTESIAN' /\n"
            f"&MESH_CART xlim={-0.5 * lx!r},{0.5 * lx!r}, zlim={-0.5 * lz!r},{0.5 * lz!r}, "
            f"nelem={nx},{nz}, ezflt=-1 /\n"
            f"&MATERIAL tag=1, kind='ELAST' /\n"
            f"&MAT_ELASTIC rho=2670d0, cp=6000d0, cs=3464d0 /\n"
            + bcs +
            f"&BC_DEF tags=5,6, kind='DYNFLT' /\n"
            f"&BC_DYNFLT friction='RSF', Tn=-100d6, TtH='LINEAR', V=1d-9, otd={otd}, oxi=1,{2 ** 31 - 1},{oxd} /\n"
            f"&DIST_LINEAR n={nx + 1}, dim=1 /\n{table}\n"
            f"&BC_DYNFLT_RSF kind=2, Dc=0.02d0, MuS=0.6d0, a={a!r}, b=0.012d0, Vstar=1d-6, theta=2d7 /\n"
            f"&TIME kind='newmark', NbSteps={nt}, Dt={dt} /\n")


def build_rupture(sc, nt=3000):
    """Problem of one ensemble scenario (parameters from :func:`rupture_ensemble`)."""
    from . import builder, parfile
    par = parfile.parse_par(par_rupture(sc["nx"], sc["nz"], sc["a"], sc["seed"], nt=nt))
    return builder.build(par)


# ---------------------------------------------------------------------------
# Antiplane (SH) rupture: 60 km x 30 km, 600 x 300 elements, planar split fault at mid
# depth, linear slip weakening (MuS 0.677, MuD 0.525, Dc 0.4 m), normal stress 120 MPa,
# initial shear stress 70 MPa raised to 81.6 MPa within 1.5 km of the centre, absorbing
# sides, Courant 0.3, fault records every step at every node.
# ---------------------------------------------------------------------------
def par_sh_rupture(nx=600, nz=300, lx=60e3, lz=30e3, nt=4000, courant="0.3d0"):
    bcs = "".join(f"&BC_DEF tag={t}, kind='ABSORB' /\n" for t in (1, 2, 3, 4))
    return (f"&GENERAL iexec=1, ngll=5, fmax=3d0, ndof=1, title='SH rupture {nx}x{nz}', verbose='0000', "
            f"ItInfo=100000 /\n"
            f"&MESH_DEF method='CARTESIAN' /\n"
            f"&MESH_CART xlim={-0.5 * lx!r},{0.5 * lx!r}, zlim={-0.5 * lz!r},{0.5 * lz!r}, "
            f"nelem={nx},{nz}, ezflt=-1 /\n"
            f"&MATERIAL tag=1, kind='ELAST' /\n"
            f"&MAT_ELASTIC rho=2670d0, cp=6000d0, cs=3464d0 /\n"
            + bcs +
            f"&BC_DEF tags=5,6, kind='DYNFLT' /\n"
            f"&BC_DYNFLT friction='SWF', Tn=-120d6, TtH='PWCONR' /\n"
            f"&DIST_PWCONR num=2, ref=0d0,0d0 /\n1.5d3\n81.6d6 70d6\n"
            f"&BC_DYNFLT_SWF Dc=0.4d0, MuS=0.677d0, MuD=0.525d0 /\n"
            f"&TIME kind='newmark', NbSteps={nt}, Courant={courant} /\n")


def build_sh_rupture(nx=600, nz=300, nt=4000):
    from . import builder, parfile
    return builder.build(parfile.parse_par(par_sh_rupture(nx, nz, nt=nt)))


# ---------------------------------------------------------------------------
# Adjoint (waveform gradient) problem
# ---------------------------------------------------------------------------
# nx x nz elements of 100 m, free surface on top, absorbing bottom, left and right
# boundaries; nsrc vertical point forces (Ricker f0, fired together, one source file)
# 1 km below the surface (a quarter of the depth on small boxes); nrec velocity
# receivers along the surface; fixed Dt.  Models are per-element (rho, lambda, mu):
# the smooth background rho 2670, cp 6000, cs 3464 and a true model with seeded
# Gaussian anomalies of up to +-5 % in rho, cp and cs.
ADJ_BACKGROUND = (2670.0, 6000.0, 3464.0)


def par_adjoint(nx, nz, nt, nrec, f0=3.0, onset=None, dx=100.0, dt="8.6d-4", src_file="sources.tab"):
    lx, lz = nx * dx, nz * dx
    onset = 1.2 / f0 if onset is None else onset
    rho, cp, cs = ADJ_BACKGROUND
    return (f"&GENERAL iexec=1, ngll=5, fmax={f0 * 2.5!r}, ndof=2, title='adjoint problem {nx}x{nz}', "
            f"verbose='0000', ItInfo=100000 /\n"
            f"&MESH_DEF method='CARTESIAN' /\n"
            f"&MESH_CART xlim=0d0,{lx!r}, zlim=0d0,{lz!r}, nelem={nx},{nz} /\n"
            f"&MATERIAL tag=1, kind='ELAST' /\n"
            f"&MAT_ELASTIC rho={rho!r}, cp={cp!r}, cs={cs!r} /\n"
            f"&BC_DEF tag=1, kind='ABSORB' /\n"
            f"&BC_DEF tag=2, kind='ABSORB' /\n"
            f"&BC_DEF tag=4, kind='ABSORB' /\n"
            f"&TIME kind='newmark', NbSteps={nt}, Dt={dt} /\n"
            f"&SRC_DEF stf='RICKER', mechanism='FORCE', file='{src_file}' /\n"
            f"&STF_RICKER f0={f0!r}, onset={onset!r}, ampli=1d0 /\n"
            f"&SRC_FORCE angle=0d0 /\n"
            f"&REC_LINE number={nrec}, first={2 * dx!r},{lz!r}, last={lx - 2 * dx!r},{lz!r}, field='V' /\n")


def write_adjoint_sources(path, nx, nz, nsrc, dx=100.0):
    lx, lz = nx * dx, nz * dx
    depth = min(1000.0, 0.25 * lz)
    with open(path, "w") as fh:
        for k in range(nsrc):
            fh.write(f"{(k + 0.5) * lx / nsrc!r} {lz - depth!r}\n")


def adjoint_models(nx, nz, seed=20260501, amp=0.05, nblob=8, dx=100.0):
    """(background, true) models, each (rho, lam, mu) over the cells (iz*nx + ix)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lx, lz = nx * dx, nz * dx
    xc = (np.arange(nx) + 0.5) * dx
    zc = (np.arange(nz) + 0.5) * dx
    X, Z = np.meshgrid(xc, zc)
    fields = []
    for _ in range(3):
        g = np.zeros_like(X)
        for _b in range(nblob):
            cx, cz = rng.uniform(0, lx), rng.uniform(0, lz)
            w = rng.uniform(0.04, 0.08) * max(lx, lz)
            g += rng.uniform(-1.0, 1.0) * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / w ** 2)
        fields.append(amp * g / np.abs(g).max())
    rho0, cp0, cs0 = ADJ_BACKGROUND

    def lame(rho, cp, cs):
        return rho.ravel(), (rho * (cp * cp - 2.0 * cs * cs)).ravel(), (rho * cs * cs).ravel()

    one = np.ones_like(X)
    back = lame(rho0 * one, cp0 * one, cs0 * one)
    true = lame(rho0 * (1 + fields[0]), cp0 * (1 + fields[1]), cs0 * (1 + fields[2]))
    return back, true


def adjoint_setup(workdir, nx=400, nz=400, nt=3000, nsrc=16, nrec=64, f0=3.0, onset=None, seed=20260501):
    """Write the input of the adjoint problem to workdir; returns (par_path, background, true)."""
    import os
    os.makedirs(workdir, exist_ok=True)
    write_adjoint_sources(os.path.join(workdir, "sources.tab"), nx, nz, nsrc)
    path = os.path.join(workdir, "Par.inp")
    with open(path, "w") as fh:
        fh.write(par_adjoint(nx, nz, nt, nrec, f0=f0, onset=onset))
    back, true = adjoint_models(nx, nz, seed=seed)
    return path, back, true
