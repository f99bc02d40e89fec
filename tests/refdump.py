"""Readers for the FP64 dumps written by oracle/ref_harness.f90 (test helper)."""
import os

import numpy as np


def read_grid(d):
    with open(os.path.join(d, "h_grid.bin"), "rb") as f:
        ngll, nelem, npoin, ndof, nt = np.fromfile(f, np.int32, 5)
        dt = np.fromfile(f, np.float64, 1)[0]
        ibool = np.fromfile(f, np.int32, ngll * ngll * nelem).reshape(nelem, ngll, ngll)
        coord = np.fromfile(f, np.float64, 2 * npoin).reshape(npoin, 2)
        rmass = np.fromfile(f, np.float64, npoin * ndof).reshape(ndof, npoin)
        H = np.fromfile(f, np.float64, ngll * ngll).reshape(ngll, ngll).T
        xg = np.fromfile(f, np.float64, ngll)
        wg = np.fromfile(f, np.float64, ngll)
        nb = np.fromfile(f, np.int32, 1)[0]
        bounds = {}
        for _ in range(nb):
            tag, npb, neb = np.fromfile(f, np.int32, 3)
            node = np.fromfile(f, np.int32, npb)
            elem = np.fromfile(f, np.int32, neb)
            edge = np.fromfile(f, np.int32, neb)
            bounds[int(tag)] = dict(node=node, elem=elem, edge=edge)
    return dict(ngll=int(ngll), nelem=int(nelem), npoin=int(npoin), ndof=int(ndof), nt=int(nt), dt=dt,
                ibool=ibool, coord=coord, rmass=rmass, H=H, xgll=xg, wgll=wg, bounds=bounds)


def read_fint(d, npoin, ndof):
    a = np.fromfile(os.path.join(d, "h_fint.bin"), np.float64)
    n = npoin * ndof
    return a[:n].reshape(ndof, npoin), a[n:2 * n].reshape(ndof, npoin)


def read_nodes(d, ndof):
    with open(os.path.join(d, "h_nodes.bin"), "rb") as f:
        nrec, n = np.fromfile(f, np.int32, 2)
        nodes = np.fromfile(f, np.int32, n)
        rest = np.fromfile(f, np.float64)
    per = 3 * n * ndof
    nstep = rest.size // per
    rest = rest[:nstep * per].reshape(nstep, 3, ndof, n)
    return dict(nrec=int(nrec), nodes=nodes, d=rest[:, 0], v=rest[:, 1], a=rest[:, 2])


def read_final(d, npoin, ndof):
    a = np.fromfile(os.path.join(d, "h_final.bin"), np.float64).reshape(3, ndof, npoin)
    return a[0], a[1], a[2]
