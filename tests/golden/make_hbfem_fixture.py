"""Writes tests/golden/ecore_hbfem_run.npz: tutorials/run_ecore_hbfem.m's
harmonic-balance run (E-core mesh, primary of 300 turns through 100 ohm,
-76j V at 50 Hz, harmonics 1, 3, 5, 7, the tutorial's 23-point curve,
MatrixK_Add = eps M with eps = 1e-7 |K|_1 / |M|_1, Tolerance 1e-3) solved by
tests/hbfem_ref.py, the CPU restatement of HBFEMCoupledSolver with its
monolithic Jacobian factored by sparse LU.  Stored: the current phasors, the
residual history and the Newton iteration count.  Takes ~10 minutes.

    python tests/golden/make_hbfem_fixture.py
"""
import sys
import time
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402
from magpackage_amd.coil import CoilGeometryUtils as CG  # noqa: E402
from magpackage_amd.hbfem import AFT  # noqa: E402
from magpackage_amd.material import create_linear, create_nonlinear  # noqa: E402
from tests.hbfem_ref import hbfem_coupled_cpu  # noqa: E402
from tests.helpers import mass_tables, stiffness_tables  # noqa: E402
from tests.test_gpu_hbfem import B_DATA, H_DATA  # noqa: E402


def main():
    d = np.load(ROOT / "tests" / "golden" / "ecore_mesh.npz")
    tags = d["tags"].copy()
    tags[np.isin(tags, [5, 6, 8, 9])] = 50
    tags[np.isin(tags, [3, 4, 7, 10])] = 60
    km = SimpleNamespace(P=d["P"], T=d["T"].astype(np.int64), RegionTags=tags,
                         NumElements=len(tags), num_elems=len(tags))
    c, _, area, ax = CG.autoDetectCircular(km, 50)
    dirm = -CG.computeCircularDirection(km, 50, c, ax)
    lib = {1: create_linear(1.0), 2: create_nonlinear(B_DATA, H_DATA), 50: create_linear(1.0),
           60: create_linear(1.0)}
    mi = O.mesh_inputs(km)
    n = mi["ndof"]
    qw, curl, _ = stiffness_tables()
    qm, val, _ = mass_tables(2)
    K = O.sparse_scipy(*O.curl_curl(mi["P"], mi["T"], mi["Dofs"], mi["Signs"],
                                    np.full(len(tags), 1 / (4e-7 * np.pi)), qw, curl), n, n)
    M = O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], None, qm, val), n, n)
    eps = spla.norm(K, 1) * 1e-7 / spla.norm(M, 1)
    fixed = O.boundary_edges(km.T, mi["edges"])
    emap = np.zeros(mi["edges"].shape[0], dtype=np.int64)
    emap[mi["t2e"].ravel() - 1] = mi["cell"].ravel()
    fx = np.zeros(n, dtype=bool)
    fd = emap[fixed]
    fx[fd[fd > 0] - 1] = True
    t0 = time.time()
    A, I, it, hist = hbfem_coupled_cpu(km, lib, AFT([1, 3, 5, 7], None, 50),
                                       dict(region=50, turns=300, area=area,
                                            dir=np.ascontiguousarray(dirm.T)),
                                       100.0, [-76j, 0, 0, 0], fx, (eps * M).tocsc(), tol=1e-3)
    out = ROOT / "tests" / "golden" / "ecore_hbfem_run.npz"
    np.savez_compressed(out, current=I, residuals=np.asarray(hist), iterations=np.array(it),
                        a_norms=np.linalg.norm(A, axis=0))
    print(f"wrote {out}: {it} iterations, |I| {np.abs(I)}, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
