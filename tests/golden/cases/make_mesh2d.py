"""Write the MESH2D meshes of the golden cases mesh2d_psv/box.mesh2d (also used by
mesh2d_sh) and mesh2d_curved/curved.mesh2d: a 6 km x 4 km domain of 30 x 20 four-node quadrilaterals in
the MESH2D format (SRC/mesh_mesh2d.f90:17-36), with the interior mesh nodes moved off the
lines of a regular grid (smooth sine displacements, so that no element is a rectangle),
elements and nodes listed in a shuffled order, two material tags (1 below z = 1.6 km
before the distortion, 2 above) and four boundaries: 1 bottom, 2 right, 3 top, 4 left
(edges 1..4 of their elements, counterclockwise from the bottom).  In the curved mesh the
right side bulges out by up to 400 m (x += 400 sin(pi z / 4 km) x / 6 km at every node),
so that boundary 2 is not flat.  mesh2d_flt/fault.mesh2d splits the mesh along node row
iz = 10, moved onto a kinked line (z = 2 km + 60 m x / 3 km up to x = 3 km, then falling
back at half that slope): the nodes of that row are duplicated for the elements above it,
with boundaries 5 (top edges of the elements below) and 6 (bottom edges of those above)."""
import os

import numpy as np

NX, NZ, LX, LZ = 30, 20, 6000.0, 4000.0


def main(path, bulge=0.0, fault=False):
    rng = np.random.default_rng(20261202)
    xs, zs = np.linspace(0.0, LX, NX + 1), np.linspace(0.0, LZ, NZ + 1)
    X, Z = np.meshgrid(xs, zs)  # [iz, ix]
    inner = np.zeros_like(X, dtype=bool)
    inner[1:-1, 1:-1] = True
    dx, dz = LX / NX, LZ / NZ
    X2 = X + inner * 0.18 * dx * np.sin(2 * np.pi * Z / LZ * 1.5) * np.sin(np.pi * X / LX * 3)
    Z2 = Z + inner * 0.18 * dz * np.sin(2 * np.pi * X / LX * 2.0) * np.cos(np.pi * Z / LZ * 2)
    X2 = X2 + bulge * np.sin(np.pi * Z / LZ) * X / LX
    IZF = NZ // 2
    if fault:  # kinked fault row, regular in x
        xf = xs
        zf = 2000.0 + np.where(xf <= 3000.0, 60.0 * xf / 3000.0, 60.0 - 30.0 * (xf - 3000.0) / 3000.0)
        X2[IZF, :] = xf
        Z2[IZF, :] = zf
    ngrid = (NX + 1) * (NZ + 1)
    gx, gz = X2.ravel(), Z2.ravel()
    # a split row gets a second copy of its nodes (grid ids ngrid + ix), used above it
    nnod = ngrid + (NX + 1 if fault else 0)
    if fault:
        gx = np.concatenate([gx, X2[IZF, :]])
        gz = np.concatenate([gz, Z2[IZF, :]])
    nid = rng.permutation(nnod) + 1  # shuffled node ids: nid[grid node] = file id
    order = np.argsort(nid)         # file id - 1 -> grid node

    def gnode(iz, ix, above):
        if fault and iz == IZF and above:
            return ngrid + ix
        return iz * (NX + 1) + ix
    elems = []
    for iz in range(NZ):
        for ix in range(NX):
            n = [gnode(iz, ix, True), gnode(iz, ix + 1, True), gnode(iz + 1, ix + 1, False),
                 gnode(iz + 1, ix, False)]
            tag = 1 if (iz + 0.5) * dz < 1600.0 else 2
            elems.append(([int(nid[k]) for k in n], tag, ix, iz))
    eorder = rng.permutation(len(elems))
    eid = np.empty(len(elems), dtype=int)
    eid[eorder] = np.arange(1, len(elems) + 1)  # file id of grid element k
    bnd = {1: [], 2: [], 3: [], 4: []}
    if fault:
        bnd[5], bnd[6] = [], []
    for k, (_, _, ix, iz) in enumerate(elems):
        if iz == 0:
            bnd[1].append((eid[k], 1))
        if ix == NX - 1:
            bnd[2].append((eid[k], 2))
        if iz == NZ - 1:
            bnd[3].append((eid[k], 3))
        if ix == 0:
            bnd[4].append((eid[k], 4))
        if fault and iz == IZF - 1:
            bnd[5].append((eid[k], 3))
        if fault and iz == IZF:
            bnd[6].append((eid[k], 1))
    with open(path, "w") as fh:
        fh.write("NEL NPEL NNOD NBC\n")
        fh.write(f"{len(elems)} 4 {nnod} {len(bnd)}\n")
        fh.write("NID X Y\n")
        for i in range(nnod):
            g = order[i]
            fh.write(f"{i + 1} {gx[g]:.17g} {gz[g]:.17g}\n")
        fh.write("EID NODES TAG\n")
        for fid, k in enumerate(eorder):
            n, tag, _, _ = elems[k]
            fh.write(f"{fid + 1} {n[0]} {n[1]} {n[2]} {n[3]} {tag}\n")
        for t in sorted(bnd):
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
    main(os.path.join(here, "mesh2d_flt", "fault.mesh2d"), fault=True)
