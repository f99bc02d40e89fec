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
