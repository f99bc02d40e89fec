"""Write the MESH2D meshes of the golden cases mesh2d_psv/box.mesh2d (also used by
mesh2d_sh) and mesh2d_curved/curved.mesh2d: a 6 km x 4 km domain of 30 x 20 four-node quadrilaterals in
the MESH2D format (SRC/mesh_mesh2d.f90:17-36), with the interior mesh nodes moved off the
lines of a regular grid (smooth sine displacements, so that no element is a rectangle),
elements and nodes listed in a shuffled order, two material tags (1 below z = 1.6 km
before the distortion, 2 above) and four boundaries: 1 bottom, 2 right, 3 top, 4 left
(edges 1..4 of their elements, counterclockwise from the bottom).  In the curved mesh the
right side bulges out by up to 400 m (x += 400 sin(pi z / 4 km) x / 6 km at every node),
so that boundary 2 is not flat."""
import os

import numpy as np

NX, NZ, LX, LZ = 30, 20, 6000.0, 4000.0


def main(path, bulge=0.0):
    rng = np.random.default_rng(20261202)
    xs, zs = np.linspace(0.0, LX, NX + 1), np.linspace(0.0, LZ, NZ + 1)
    X, Z = np.meshgrid(xs, zs)  # [iz, ix]
    inner = np.zeros_like(X, dtype=bool)
    inner[1:-1, 1:-1] = True
    dx, dz = LX / NX, LZ / NZ
    X2 = X + inner * 0.18 * dx * np.sin(2 * np.pi * Z / LZ * 1.5) * np.sin(np.pi * X / LX * 3)
    Z2 = Z + inner * 0.18 * dz * np.sin(2 * np.pi * X / LX * 2.0) * np.cos(np.pi * Z / LZ * 2)
    X2 = X2 + bulge * np.sin(np.pi * Z / LZ) * X / LX
    nnod = (NX + 1) * (NZ + 1)
    nid = rng.permutation(nnod) + 1  # shuffled node ids: nid[grid node] = file id
    order = np.argsort(nid)         # file id - 1 -> grid node
    gx, gz = X2.ravel(), Z2.ravel()
    elems = []
    for iz in range(NZ):
        for ix in range(NX):
            n = [iz * (NX + 1) + ix, iz * (NX + 1) + ix + 1, (iz + 1) * (NX + 1) + ix + 1, (iz + 1) * (NX + 1) + ix]
            tag = 1 if (iz + 0.5) * dz < 1600.0 else 2
            elems.append(([int(nid[k]) for k in n], tag, ix, iz))
    eorder = rng.permutation(len(elems))
    eid = np.empty(len(elems), dtype=int)
    eid[eorder] = np.arange(1, len(elems) + 1)  # file id of grid element k
    bnd = {1: [], 2: [], 3: [], 4: []}
    for k, (_, _, ix, iz) in enumerate(elems):
        if iz == 0:
            bnd[1].append((eid[k], 1))
        if ix == NX - 1:
            bnd[2].append((eid[k], 2))
        if iz == NZ - 1:
            bnd[3].append((eid[k], 3))
        if ix == 0:
            bnd[4].append((eid[k], 4))
    with open(path, "w") as fh:
        fh.write("NEL NPEL NNOD NBC\n")
        fh.write(f"{len(elems)} 4 {nnod} 4\n")
        fh.write("NID X Y\n")
        for i in range(nnod):
            g = order[i]
            fh.write(f"{i + 1} {gx[g]:.17g} {gz[g]:.17g}\n")
        fh.write("EID NODES TAG\n")
        for fid, k in enumerate(eorder):
            n, tag, _, _ = elems[k]
            fh.write(f"{fid + 1} {n[0]} {n[1]} {n[2]} {n[3]} {tag}\n")
        for t in (1, 2, 3, 4):
            rows = sorted(bnd[t])
            fh.write("BCTAG NBEL\n")
            fh.write(f"{t} {len(rows)}\n")
            fh.write("BID EID EDGE\n")
            for b, (e, edge) in enumerate(rows):
                fh.write(f"{b + 1} {e} {edge}\n")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(os.path.join(here, "mesh2d_psv", "box.mesh2d"))
    main(os.path.join(here, "mesh2d_sh", "box.mesh2d"))
    main(os.path.join(here, "mesh2d_curved", "curved.mesh2d"), bulge=400.0)
