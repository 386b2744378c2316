"""Generate the golden fixtures under tests/golden/ (run from the repo root):

    python tests/golden/make_golden.py [--mesh-dir <HyTeG data/meshes/3D>]

* p1_bruteforce.npz -- independent numpy assembly (bruteforce.py) of the P1
  stiffness matrix on the unit cube split into 6 tetrahedra (meshCuboid node
  table) and on the 8-tetrahedra regular octahedron, levels 1..3; stores micro
  vertex coordinates, a seeded vector x and A @ x.  Comparisons go through the
  coordinates, so the fixture is independent of any DoF numbering.
* octahedron_8el.npz -- vertex coordinates and tetrahedra of HyTeG's test mesh
  data/meshes/3D/regular_octahedron_8el.msh (MSH 2.2), vertex ids mapped to
  0..n-1 in id order, cell vertex order as in the file.
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import bruteforce as bf  # noqa: E402

# the six tetrahedra of a hexahedron (vertex numbering 0..7 of a unit hex,
# node order of MeshInfo::meshCuboid) and the unit-cube vertices in z-y-x order
HEX_TETS = [[4, 5, 7, 3], [0, 3, 1, 4], [4, 1, 5, 3], [5, 6, 7, 2], [1, 3, 2, 5], [7, 2, 3, 5]]
HEX_NODES = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def cube_mesh():
    coords = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], dtype=float)
    gid = lambda n: n[2] * 4 + n[1] * 2 + n[0]  # noqa: E731
    cells = np.array([[gid(HEX_NODES[k]) for k in t] for t in HEX_TETS], dtype=np.int64)
    return coords, cells


def read_msh22(path):
    toks = open(path).read().split()
    i = toks.index("$Nodes") + 1
    nn = int(toks[i])
    i += 1
    ids, X = [], []
    for _ in range(nn):
        ids.append(int(toks[i]))
        X.append([float(t) for t in toks[i + 1:i + 4]])
        i += 4
    order = np.argsort(ids)
    idmap = {ids[j]: r for r, j in enumerate(order)}
    X = np.array(X)[order]
    i = toks.index("$Elements") + 1
    ne = int(toks[i])
    i += 1
    cells = []
    for _ in range(ne):
        et, nt = int(toks[i + 1]), int(toks[i + 2])
        j = i + 3 + nt
        nv = {1: 2, 2: 3, 4: 4, 15: 1}[et]
        if et == 4:
            cells.append([idmap[int(t)] for t in toks[j:j + 4]])
        i = j + nv
    return X, np.array(cells, dtype=np.int64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mesh-dir", default="/root/reference/data/meshes/3D")
    args = ap.parse_args()
    octa = os.path.join(HERE, "octahedron_8el.npz")
    msh = os.path.join(args.mesh_dir, "regular_octahedron_8el.msh")
    if os.path.exists(msh):
        X, cells = read_msh22(msh)
        np.savez(octa, coords=X, cells=cells)
    o = np.load(octa)
    meshes = {"cube": cube_mesh(), "octa": (o["coords"], o["cells"])}
    out = {}
    rng = np.random.default_rng(20261202)
    for name, (coords, cells) in meshes.items():
        for L in (1, 2, 3):
            A, P = bf.assemble(coords, cells, L)
            x = rng.uniform(-1, 1, A.shape[0])
            out[f"{name}_L{L}_coords"] = P
            out[f"{name}_L{L}_x"] = x
            out[f"{name}_L{L}_Ax"] = A @ x
            out[f"{name}_L{L}_diag"] = A.diagonal()
    np.savez_compressed(os.path.join(HERE, "p1_bruteforce.npz"), **out)
    print("wrote", sorted(out)[:4], "...")


if __name__ == "__main__":
    main()
