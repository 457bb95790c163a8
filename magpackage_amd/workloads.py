"""Seeded synthetic workloads of the engine's benchmark configurations.

Each returns host arrays (a KuhnMesh and per-element data) so the benchmark,
the full-size tests and the CPU baseline build exactly the same input.

- ``plate_box``: the frequency-domain eddy-current operator's box: a
  jittered n^3-cube Kuhn box (6 tets per cube) with a conducting plate
  (element centroids with z in [0.45, 0.55], sigma = 3.5e7 S/m), mu_r = 1
  everywhere, 50 Hz.  n = 128: 12,582,912 tets, 14,827,904 edges.
- ``newton_core_box``: the nonlinear Newton-set box: 100 x 100 x 67 cubes of
  edge 0.01, an iron core (centroids in [0.3,0.7] x [0.3,0.7] x [0.2,0.47],
  Marrocco curve, sigma 2e6 S/m) in air: 4,020,000 tets, 4,760,467 edges.
- ``newton_solution``: the Newton set's edge solution: A ~ N(0, 1) (seeded),
  scaled so the largest element |B| is 2.2 T.
"""
from __future__ import annotations

import numpy as np

from .material import MU0, create_linear, marrocco
from .meshgen import kuhn_box, tag_box, tag_slab

PLATE_SIGMA = 3.5e7
PLATE_FREQ_HZ = 50.0
CORE_SIGMA = 2e6


def plate_box(n: int = 128, seed: int = 2):
    """(mesh, nu, sigma, omega) of the eddy-current operator K + j w M_sigma."""
    km = kuhn_box(n, n, n, seed=seed)
    tag_slab(km, 2, 0.45, 0.55, 2)
    nu = np.full(km.num_elems, 1.0 / MU0)
    sigma = np.where(km.RegionTags == 2, PLATE_SIGMA, 0.0)
    return km, nu, sigma, 2 * np.pi * PLATE_FREQ_HZ


def newton_core_box(nx: int = 100, ny: int = 100, nz: int = 67, seed: int = 3):
    """(mesh, material library, sigma) of the fused Newton set."""
    km = kuhn_box(nx, ny, nz, lengths=(nx / 100.0, ny / 100.0, nz / 100.0), seed=seed)
    tag_box(km, (0.3, 0.3, 0.2), (0.7, 0.7, 0.47), 2)
    lib = {1: create_linear(1.0), 2: marrocco()}
    sigma = np.where(km.RegionTags == 2, CORE_SIGMA, 0.0)
    return km, lib, sigma


def newton_solution(asm, space, n: int, seed: int = 3, bmax: float = 2.2):
    """Device edge solution A ~ N(0, 1) scaled so that max_e |B_e| = bmax."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    a = torch.from_numpy(np.random.default_rng(seed).normal(size=n)).to(dev)
    b = float(asm.computeElementB(a, space).norm(dim=1).max())
    return a * (bmax / b)
