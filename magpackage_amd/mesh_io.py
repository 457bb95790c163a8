"""COMSOL text mesh reader (src/mesh_io/read_mphtxt.m:1-158) and the Gmsh
reader (``read_msh``, src/mesh_io/read_msh.m:1-128, described at the function).

``read_mphtxt(filename, unit_tag="m")`` returns a ``MeshRaw``:
- P: (Nn, 3) float64, the vertex coordinates scaled to metres;
- T: (Ne, 4) int32, 1-based node ids of the tetrahedra;
- RegionTags: (Ne,) int32, the file's geometric entity indices, unchanged;
- Faces: (Nf, 3) int32, 1-based, and FaceTags: (Nf,) int32, from triangle blocks.

The parse follows the reference reader's markers.
- A line containing ``# number of mesh vertices`` gives the vertex count.
- ``# Mesh vertex coordinates`` is followed by 3 x count numbers.
- A ``# type name`` line opens an element block; it is kept when it names ``tet``
  or ``tri``. Other blocks (vertex, edge) are skipped.
- Inside a block:
  - ``# number of ... per element`` and ``# number of elements`` give the sizes;
  - ``# Elements`` is followed by the 0-based connectivity, which is made 1-based;
  - ``# Geometric entity indices`` is followed by one tag per element.

Like the reference, the reader:
- raises on a missing file and on a short coordinate block;
- warns on an unknown unit and reads it as metres;
- warns, and drops the block, when a block's connectivity comes before its sizes.

Blocks of the same kind are concatenated in file order.
"""
from __future__ import annotations

import os
import warnings
from dataclasses import dataclass, field

import numpy as np

UNIT_SCALE = {"m": 1.0, "cm": 1e-2, "mm": 1e-3, "um": 1e-6, "nm": 1e-9, "in": 0.0254,
              "ft": 0.3048}


@dataclass
class MeshRaw:
    P: np.ndarray
    T: np.ndarray
    RegionTags: np.ndarray
    Faces: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), dtype=np.int32))
    FaceTags: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))


def _leading_int(line: str):
    """First integer of a line (textscan(line, '%d') keeps the leading number)."""
    tok = line.split()
    if not tok:
        return None
    try:
        return int(tok[0])
    except ValueError:
        return None


class _Lines:
    """Line cursor with fscanf-like numeric reads across lines."""

    def __init__(self, lines):
        self.lines = lines
        self.i = 0

    def eof(self) -> bool:
        return self.i >= len(self.lines)

    def next(self) -> str:
        line = self.lines[self.i].strip()
        self.i += 1
        return line

    def numbers(self, count: int, kind=float) -> list:
        """Up to `count` numbers from the following lines (fscanf): stops at
        the first token that is not a number or once `count` are read; a line
        not read to its end is left to be read again as text."""
        out = []
        while len(out) < count and not self.eof():
            tok = self.lines[self.i].split()
            for k, t in enumerate(tok):
                try:
                    out.append(kind(t))
                except ValueError:
                    return out
                if len(out) == count:
                    if k + 1 == len(tok):
                        self.i += 1
                    return out
            self.i += 1
        return out


def _element_block(cur: _Lines):
    """parse_element_block_v2 (read_mphtxt.m:105-157)."""
    num_elems = 0
    per_elem = 0
    elems, tags = None, None
    while not cur.eof():
        pos = cur.i
        line = cur.next()
        if not line:
            continue
        if "# type name" in line:
            cur.i = pos
            break
        if "# number of" in line and "per element" in line:
            v = _leading_int(line)
            if v is not None:
                per_elem = v
        if "# number of elements" in line:
            v = _leading_int(line)
            if v is not None:
                num_elems = v
        if "# Elements" in line and elems is None:
            if num_elems == 0 or per_elem == 0:
                warnings.warn("element block: connectivity before its sizes; block dropped")
                break
            raw = cur.numbers(per_elem * num_elems, int)
            if len(raw) == per_elem * num_elems:
                elems = np.asarray(raw, dtype=np.int64).reshape(num_elems, per_elem) + 1
        if "# Geometric entity indices" in line and tags is None:
            if num_elems == 0:
                break
            raw = cur.numbers(num_elems, int)
            if len(raw) == num_elems:
                tags = np.asarray(raw, dtype=np.int64)
                break
    return elems, tags


def read_mphtxt(filename: str, unit_tag: str = "m") -> MeshRaw:
    scale = UNIT_SCALE.get(unit_tag.lower())
    if scale is None:
        warnings.warn(f'unknown unit tag "{unit_tag}": read as metres (scale 1.0)')
        scale = 1.0
    if not os.path.isfile(filename):
        raise FileNotFoundError(f"mesh file not found: {filename}")
    with open(filename, "r") as fh:
        cur = _Lines(fh.read().splitlines())
    nv = 0
    P = np.zeros((0, 3))
    tets, tet_tags, tris, tri_tags = [], [], [], []
    while not cur.eof():
        line = cur.next()
        if not line:
            continue
        if "# number of mesh vertices" in line:
            v = _leading_int(line)
            if v is not None:
                nv = v
        if "# Mesh vertex coordinates" in line:
            if nv == 0:
                warnings.warn("vertex count not found before the coordinates")
            data = cur.numbers(3 * nv, float)
            if len(data) != 3 * nv:
                raise ValueError(f"read {len(data)} coordinate values, expected {3 * nv}")
            P = np.asarray(data, dtype=np.float64).reshape(nv, 3) * scale
        if "# type name" in line:
            if " tet " in line:
                e, t = _element_block(cur)
                if e is not None:
                    tets.append(e)
                    tet_tags.append(t if t is not None else np.zeros(0, dtype=np.int64))
            elif " tri " in line:
                e, t = _element_block(cur)
                if e is not None:
                    tris.append(e)
                    tri_tags.append(t if t is not None else np.zeros(0, dtype=np.int64))
    T = (np.concatenate(tets) if tets else np.zeros((0, 4), dtype=np.int64)).astype(np.int32)
    tags = (np.concatenate(tet_tags) if tet_tags else np.zeros(0)).astype(np.int32)
    F = (np.concatenate(tris) if tris else np.zeros((0, 3), dtype=np.int64)).astype(np.int32)
    ftags = (np.concatenate(tri_tags) if tri_tags else np.zeros(0)).astype(np.int32)
    return MeshRaw(P=P, T=T, RegionTags=tags, Faces=F, FaceTags=ftags)


class _Stream:
    """MATLAB's fgetl / fscanf over one text buffer: ``line()`` returns the rest
    of the current line (after what fscanf consumed), ``numbers(k)`` reads up to
    k whitespace-separated numbers across lines and stops at the first token
    that is not a number, leaving the position right after the last number."""

    def __init__(self, text: str):
        self.s = text
        self.i = 0

    def eof(self) -> bool:
        return self.i >= len(self.s)

    def line(self) -> str:
        j = self.s.find("\n", self.i)
        if j < 0:
            j = len(self.s)
        out = self.s[self.i:j]
        self.i = j + 1
        return out

    def numbers(self, k: int) -> list:
        out = []
        n = len(self.s)
        while len(out) < k:
            while self.i < n and self.s[self.i].isspace():
                self.i += 1
            if self.i >= n:
                break
            j = self.i
            while j < n and not self.s[j].isspace():
                j += 1
            try:
                out.append(float(self.s[self.i:j]))
            except ValueError:
                break
            self.i = j
        return out


def _sscanf_floats(line: str) -> list:
    """sscanf(line, '%f'): the leading run of numbers."""
    out = []
    for tok in line.split():
        try:
            out.append(float(tok))
        except ValueError:
            break
    return out


def read_msh(filename: str) -> MeshRaw:
    """Gmsh mesh reader (src/mesh_io/read_msh.m:1-128), the reference's reading
    of the format:
    - ``$Nodes``: a line (blocks, nodes, ...), then per block a header line
      whose 4th number is the block's node count, followed by that many
      (tag, x, y, z) quadruples. Coordinates are kept in file order;
    - ``$Elements``: a line (blocks, ...), then per block a header line
      (dim, entity tag, element type, count). Type 2 (3-node triangle, rows
      of id + 3 nodes) goes to Faces/FaceTags and type 4 (tetrahedron, id + 4
      nodes) to T/RegionTags, tagged by the block's entity tag. Types 1 and
      15 are skipped by their sizes; other types are warned about and not
      skipped, as in the reference.
    Node ids are used as given (1-based indices into the coordinates)."""
    if not os.path.isfile(filename):
        raise FileNotFoundError(f"File not found: {filename}")
    with open(filename, "r") as f:
        st = _Stream(f.read())

    def header():
        while not st.eof():
            line = st.line().strip()
            if not line:
                continue
            v = _sscanf_floats(line)
            if v:
                return v
        return []

    P = np.zeros((0, 3))
    T_list, tag_list, F_list, ftag_list = [], [], [], []
    while not st.eof():
        line = st.line().strip()
        if not line:
            continue
        if line == "$Nodes":
            info = header()
            if not info:
                continue
            nblocks, nnodes = int(info[0]), int(info[1])
            P = np.zeros((nnodes, 3))
            cnt = 0
            for _ in range(nblocks):
                b = header()
                k = int(b[3])
                data = st.numbers(4 * k)
                if len(data) != 4 * k:
                    raise ValueError(f"node block: expected {4 * k} numbers, read {len(data)}")
                d = np.asarray(data).reshape(k, 4)
                P[cnt:cnt + k] = d[:, 1:4]
                cnt += k
        elif line == "$Elements":
            info = header()
            if not info:
                continue
            for _ in range(int(info[0])):
                b = header()
                ent, etype, k = int(b[1]), int(b[2]), int(b[3])
                if etype == 2:
                    data = st.numbers(4 * k)
                    if len(data) != 4 * k:
                        raise ValueError("triangle block: short read")
                    F_list.append(np.asarray(data).reshape(k, 4)[:, 1:4])
                    ftag_list.append(np.full(k, ent))
                elif etype == 4:
                    data = st.numbers(5 * k)
                    if len(data) != 5 * k:
                        raise ValueError("tetrahedron block: short read")
                    T_list.append(np.asarray(data).reshape(k, 5)[:, 1:5])
                    tag_list.append(np.full(k, ent))
                elif etype in (1, 15):
                    st.numbers((3 if etype == 1 else 2) * k)
                else:
                    warnings.warn(f"unknown element type {etype}; not skipped")
    T = (np.concatenate(T_list).astype(np.int32) if T_list
         else np.zeros((0, 4), dtype=np.int32))
    tags = (np.concatenate(tag_list).astype(np.int32) if tag_list
            else np.zeros(0, dtype=np.int32))
    F = (np.concatenate(F_list).astype(np.int32) if F_list
         else np.zeros((0, 3), dtype=np.int32))
    ftags = (np.concatenate(ftag_list).astype(np.int32) if ftag_list
             else np.zeros(0, dtype=np.int32))
    return MeshRaw(P=P, T=T, RegionTags=tags, Faces=F, FaceTags=ftags)
