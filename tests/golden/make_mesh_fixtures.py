"""Writes tests/golden/{ecore,team7,d800kva}_mesh.npz: the reference's COMSOL
meshes (data/Ecore.mphtxt, data/Team7.mphtxt, data/D800kVA.mphtxt) read with
magpackage_amd.mesh_io.read_mphtxt.

Run in a checkout that has the reference at /root/reference:

    python tests/golden/make_mesh_fixtures.py

Before saving, each mesh is checked against an independent parse of the file
(the counts on its '# number of ...' lines and the numbers between the
markers, read with numpy), so the fixture does not just restate the reader.
Stored: P (Nn, 3) float64, T (Ne, 4) int32 1-based, tags (Ne,) int32,
faces (Nf, 3) int32 1-based, face_tags (Nf,) int32, and the sha256 of the file.
"""
import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from magpackage_amd.mesh_io import read_mphtxt  # noqa: E402

REF = Path("/root/reference/data")


def independent(path):
    """Block-by-block parse by marker line numbers (single tet/tri blocks)."""
    lines = path.read_text().splitlines()
    nv = next(int(l.split()[0]) for l in lines if "# number of mesh vertices" in l)
    i = next(k for k, l in enumerate(lines) if "# Mesh vertex coordinates" in l)
    P = np.array([float(x) for l in lines[i + 1: i + 1 + nv] for x in l.split()]).reshape(nv, 3)
    out = {"P": P}
    for kind, npe in (("tet", 4), ("tri", 3)):
        k = [j for j, l in enumerate(lines) if "# type name" in l and f" {kind} " in l]
        if not k:
            continue
        j = k[0]
        ne = next(int(l.split()[0]) for l in lines[j:] if "# number of elements" in l)
        e0 = next(m for m in range(j, len(lines)) if "# Elements" in lines[m])
        E = np.array([int(x) for l in lines[e0 + 1: e0 + 1 + ne] for x in l.split()])
        g0 = next(m for m in range(e0, len(lines)) if "# Geometric entity indices" in lines[m])
        G = np.array([int(l.split()[0]) for l in lines[g0 + 1: g0 + 1 + ne]])
        out[kind] = (E.reshape(ne, npe) + 1, G)
    return out


def main():
    for name in ("Ecore", "Team7", "D800kVA"):
        path = REF / f"{name}.mphtxt"
        m = read_mphtxt(str(path))
        ind = independent(path)
        assert np.array_equal(m.P, ind["P"])
        assert np.array_equal(m.T, ind["tet"][0]) and np.array_equal(m.RegionTags, ind["tet"][1])
        if "tri" in ind:
            assert np.array_equal(m.Faces, ind["tri"][0])
            assert np.array_equal(m.FaceTags, ind["tri"][1])
        sha = hashlib.sha256(path.read_bytes()).hexdigest()
        out = ROOT / "tests" / "golden" / f"{name.lower()}_mesh.npz"
        np.savez_compressed(out, P=m.P, T=m.T.astype(np.int32), tags=m.RegionTags.astype(np.int32),
                            faces=m.Faces.astype(np.int32),
                            face_tags=m.FaceTags.astype(np.int32), sha256=np.array(sha))
        print(f"{out.name}: {m.P.shape[0]} nodes, {m.T.shape[0]} tets, {m.Faces.shape[0]} faces, "
              f"{out.stat().st_size} bytes")


if __name__ == "__main__":
    main()
