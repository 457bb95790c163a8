"""COMSOL mesh reader (src/mesh_io/read_mphtxt.m) and the reference-mesh
fixtures under tests/golden/ (written by tests/golden/make_mesh_fixtures.py
from the reference's data/Ecore.mphtxt and data/Team7.mphtxt)."""
from pathlib import Path

import numpy as np
import pytest

from magpackage_amd.mesh_io import read_mphtxt
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"

# a two-tet mesh with a vertex block (skipped), a triangle block and a tet block,
# in the layout COMSOL writes
SMALL = """# Created by COMSOL Multiphysics.

# Major & minor version
0 1
1 # number of tags
# Tags
5 mesh1

3 # sdim
5 # number of mesh vertices
0 # lowest mesh vertex index

# Mesh vertex coordinates
0 0 0
1000 0 0
0 1000 0
0 0 1000
1000 1000 1000

3 # number of element types

# Type #0

3 vtx # type name


1 # number of vertices per element
2 # number of elements
# Elements
0
1

2 # number of geometric entity indices
# Geometric entity indices
1
2

# Type #1

3 tri # type name


3 # number of vertices per element
1 # number of elements
# Elements
0 1 2

1 # number of geometric entity indices
# Geometric entity indices
7

# Type #2

3 tet # type name


4 # number of vertices per element
2 # number of elements
# Elements
0 1 2 3
1 2 3 4

2 # number of geometric entity indices
# Geometric entity indices
1
2
"""


def test_read_small(tmp_path):
    f = tmp_path / "small.mphtxt"
    f.write_text(SMALL)
    m = read_mphtxt(str(f), "mm")
    assert m.P.shape == (5, 3) and m.P[4, 2] == 1.0 and m.P[1, 0] == 1.0
    assert np.array_equal(m.T, [[1, 2, 3, 4], [2, 3, 4, 5]])
    assert np.array_equal(m.RegionTags, [1, 2])
    assert np.array_equal(m.Faces, [[1, 2, 3]]) and np.array_equal(m.FaceTags, [7])
    assert read_mphtxt(str(f)).P[4, 2] == 1000.0


def test_read_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        read_mphtxt(str(tmp_path / "missing.mphtxt"))
    f = tmp_path / "short.mphtxt"
    f.write_text(SMALL.replace("5 # number of mesh vertices", "6 # number of mesh vertices"))
    with pytest.raises(ValueError, match="coordinate"):
        read_mphtxt(str(f))
    f.write_text(SMALL)
    with pytest.warns(UserWarning, match="unknown unit"):
        m = read_mphtxt(str(f), "furlong")
    assert m.P[4, 2] == 1000.0


@pytest.mark.parametrize("name,nn,ne,nf", [("ecore", 4324, 23569, 0),
                                            ("team7", 10007, 57892, 5936),
                                            ("d800kva", 11225, 65394, 0)])
def test_reference_mesh_fixtures(name, nn, ne, nf):
    """The counts the files' headers state; every node used; a closed
    tetrahedral volume (Euler characteristic of a ball: V - E + F - T = 1)."""
    d = np.load(GOLDEN / f"{name}_mesh.npz")
    P, T, tags = d["P"], d["T"], d["tags"]
    assert P.shape == (nn, 3) and T.shape == (ne, 4) and tags.shape == (ne,)
    assert d["faces"].shape == (nf, 3)
    assert T.min() == 1 and T.max() == nn and np.unique(T).size == nn
    edges, _, _ = O.generate_edges(T)
    faces = np.unique(np.sort(T[:, [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]].reshape(-1, 3),
                              axis=1), axis=0)
    assert nn - edges.shape[0] + faces.shape[0] - ne == 1
