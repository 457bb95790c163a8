"""LossCalculator and ParameterExtractor (src/post/LossCalculator.m,
src/post/ParameterExtractor.m) on the GPU:
- ohmic losses 0.5 w^2 Re(A^H M_sigma A) (single frequency and harmonic
  sums) against M_sigma from the oracle's mass kernel;
- element volumes summing to the box's volume; Steinmetz and Bertotti iron
  losses against the oracle's element B;
- the inductance matrix against the gauge-free penalised solve of
  tests/circuit_ref.py (a route independent of the engine's projected load and
  auxiliary-space CG)."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import core_box_mesh, mass_tables, stiffness_tables

pytestmark = pytest.mark.gpu


def _setup(km):
    from magpackage_amd.assembler import Assembler, DofHandler, FunctionSpace, Mesh
    from magpackage_amd.post import LossCalculator
    mesh = Mesh.from_kuhn(km)
    dh = DofHandler(mesh)
    sp = FunctionSpace("Nedelec", 1)
    dh.distributeDofs(sp)
    asm = Assembler(mesh, dh)
    return sp, asm, LossCalculator(asm), dh.NumGlobalDofs


def test_ohmic_and_iron_losses(gpu):
    km = core_box_mesh(8, seed=7)
    sp, asm, lc, n = _setup(km)
    mi = O.mesh_inputs(km)
    qm, val, _ = mass_tables(2)
    _, curl, _ = stiffness_tables()
    sigma = {1: 0.0, 2: 2.0e6}
    sig_e = np.where(km.RegionTags == 2, 2.0e6, 0.0)
    M = O.sparse_scipy(*O.mass(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], sig_e, qm, val), n, n)
    rng = np.random.default_rng(12)
    A = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    f = 50.0
    P1 = lc.computeOhmicLoss_Frequency(A[:, 0], sigma, f, sp)
    ref1 = 0.5 * (2 * np.pi * f) ** 2 * np.real(np.conj(A[:, 0]) @ (M @ A[:, 0]))
    assert abs(P1 - ref1) <= 1e-12 * abs(ref1)
    aft = SimpleNamespace(BaseFreq=f, Harmonics=np.array([0, 1, 3]))
    Ph = lc.computeOhmicLoss_HBFEM(A, sigma, aft, sp)
    refh = sum(0.5 * (h * 2 * np.pi * f) ** 2 * np.real(np.conj(A[:, k]) @ (M @ A[:, k]))
               for k, h in enumerate([0, 1, 3]) if h)
    assert abs(Ph - refh) <= 1e-12 * abs(refh)
    vols = lc.computeElementVolumes()
    assert abs(vols.sum() - 1.0) <= 1e-12  # unit box
    a = 1e-2 * rng.standard_normal(n)
    B = np.linalg.norm(O.element_b(mi["P"], mi["T"], mi["Dofs"], mi["Signs"], a, curl), axis=1)
    kh, al, ke = 1e-2, 1.8, 5e-5
    Pi = lc.computeIronLoss_Steinmetz(a, f, kh, al, ke, sp)
    refi = float(((kh * f * B ** al + ke * f ** 2 * B ** 2) * vols).sum())
    assert abs(Pi - refi) <= 1e-12 * refi
    Ar = np.stack([np.zeros(n), a, 0.5 * a], axis=1)
    Pih = lc.computeIronLoss_HBFEM(Ar, aft, kh, al, ke, sp)
    refih = sum(float(((kh * h * f * (c * B) ** al + ke * (h * f) ** 2 * (c * B) ** 2)
                       * vols).sum()) for h, c in ((1, 1.0), (3, 0.5)))
    assert abs(Pih - refih) <= 1e-12 * refih
    from magpackage_amd.material import create_linear
    iron = create_linear(1000.0)
    iron.extra.update(Lam_d=0.35e-3, Lam_sigma=2.0e6, Loss_Kexc=0.1)
    Pb = lc.computeBertottiLoss_TimeStep(a, {1: create_linear(1.0), 2: iron})
    m = km.RegionTags == 2
    cl = float((2.0e6 * 0.35e-3 ** 2 / 12.0 * B[m] ** 2 * vols[m]).sum())
    exc = float((0.1 * B[m] ** 1.5 * vols[m]).sum())
    print(f"ohmic {P1:.6e} W, harmonics {Ph:.6e} W, Steinmetz {Pi:.6e} W, Bertotti {Pb}",
          flush=True)
    assert abs(Pb["Classic"] - cl) <= 1e-12 * cl and abs(Pb["Excess"] - exc) <= 1e-12 * exc


def test_inductance_matrix(gpu):
    from magpackage_amd.circuit import Winding
    from magpackage_amd.meshgen import ecore_box
    from magpackage_amd.post import ParameterExtractor
    from tests.circuit_ref import inductance_matrix
    em = ecore_box(6, 6, 4)
    km = em.mesh
    sp, asm, _, n = _setup(km)
    mu0 = 4e-7 * np.pi
    nu_map = {1: 1.0 / mu0, 2: 1.0 / (500.0 * mu0), 3: 1.0 / mu0}
    nu = np.array([nu_map[int(t)] for t in km.RegionTags])
    ydir = np.where((km.RegionTags == 3)[:, None], np.array([0.0, 1.0, 0.0]), 0.0)
    w1 = Winding("A", 3, 200, 3.0, 4.8e-4)
    w1.setDirectionField(em.js_dir.T)
    w2 = Winding("B", 3, 120, 1.5, 4.8e-4)
    w2.setDirectionField(ydir.T)
    pe = ParameterExtractor(asm)
    Lm, Rm = pe.extractMatrices(sp, nu_map, [w1, w2])
    ref = inductance_matrix(km, nu, [dict(region=3, turns=200, area=4.8e-4, dir=em.js_dir),
                                     dict(region=3, turns=120, area=4.8e-4, dir=ydir)])
    ref = 0.5 * (ref + ref.T)
    print(f"L {Lm.tolist()}\nref {ref.tolist()}; CG {pe.Iterations}", flush=True)
    assert np.array_equal(Rm, np.diag([3.0, 1.5]))
    assert np.abs(Lm - ref).max() <= 1e-7 * np.abs(ref).max()
