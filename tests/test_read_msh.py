"""read_msh (src/mesh_io/read_msh.m) on Gmsh files written here: a Kuhn box's
nodes in two blocks, point, line, triangle and tetrahedron blocks in two
volume entities. The reader returns the file's coordinates, the tetrahedra
with their entity tags, and the triangles, and skips the point and line
blocks by their sizes. No .msh file ships with the reference, so the round
trip is the check."""
import warnings

import numpy as np
import pytest

from magpackage_amd.meshgen import kuhn_box
from magpackage_amd.mesh_io import read_msh


def _write(path, P, T, tags, tris, extra_block=None):
    nn = P.shape[0]
    half = nn // 2
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$Nodes",
             f"2 {nn} 1 {nn}"]
    for lo, hi in ((0, half), (half, nn)):
        lines.append(f"3 1 0 {hi - lo}")
        lines += [f"{i + 1} {P[i, 0]:.17g} {P[i, 1]:.17g} {P[i, 2]:.17g}" for i in range(lo, hi)]
    lines.append("$EndNodes")
    blocks = []
    blocks.append(["0 1 15 2", "1 1", "2 5"])
    blocks.append(["1 1 1 2", "3 1 2", "4 2 3"])
    blocks.append([f"2 7 2 {tris.shape[0]}"] + [f"{100 + k} " + " ".join(map(str, t))
                                                for k, t in enumerate(tris)])
    eid = 1000
    for ent in np.unique(tags):
        Ts = T[tags == ent]
        blk = [f"3 {ent} 4 {Ts.shape[0]}"]
        for t in Ts:
            blk.append(f"{eid} " + " ".join(map(str, t)))
            eid += 1
        blocks.append(blk)
    if extra_block is not None:
        blocks.append(extra_block)
    lines += ["$Elements", f"{len(blocks)} {eid} 1 {eid}"]
    for b in blocks:
        lines += b
    lines.append("$EndElements")
    path.write_text("\n".join(lines) + "\n")


def test_round_trip(tmp_path):
    km = kuhn_box(3, 2, 2, lengths=(0.3, 0.2, 0.2), jitter=0.15, seed=6)
    tags = np.where(km.centroids()[:, 0] < 0.15, 11, 12).astype(np.int32)
    tris = km.T[:5, :3]
    f = tmp_path / "box.msh"
    _write(f, km.P, km.T, tags, tris)
    raw = read_msh(str(f))
    order = np.argsort(tags, kind="stable")  # blocks are written per entity
    assert np.array_equal(raw.P, km.P)
    assert np.array_equal(raw.T, km.T[order])
    assert np.array_equal(raw.RegionTags, tags[order])
    assert np.array_equal(raw.Faces, tris) and np.all(raw.FaceTags == 7)


def test_unknown_type_warns_and_missing_file(tmp_path):
    km = kuhn_box(1, 1, 1, jitter=0.0)
    f = tmp_path / "odd.msh"
    _write(f, km.P, km.T, np.ones(km.T.shape[0], dtype=np.int32), km.T[:1, :3],
           extra_block=["2 3 3 0"])
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        raw = read_msh(str(f))
    assert any("unknown element type 3" in str(x.message) for x in w)
    assert raw.T.shape == (6, 4)
    with pytest.raises(FileNotFoundError):
        read_msh(str(tmp_path / "none.msh"))
