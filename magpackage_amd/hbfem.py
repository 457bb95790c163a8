"""Harmonic balance (HBFEM) host pieces: the alternating frequency-time
transform (src/utils/AFT.m:19-75) and the assembly entry point that the
reference's HBFEM solvers call (HBFEMSolver.m:78,96: Assembler.assembleHBFEM,
absent from the reference's Assembler) on the engine's HBFEM kernel
(mpk_hbfem_coo, the restatement of src/mex/assemble_hbfem_kernel_mex.cpp).
"""
from __future__ import annotations

import math
import warnings

import numpy as np


class AFT:
    """D_matrix [Nt x K] = exp(1j h theta) (frequency -> time), P_matrix
    [K x Nt] = f_h exp(-1j h theta)^T with f_h = 2/Nt (1/Nt for h = 0), theta =
    2 pi (0..Nt-1)/Nt.  Default Nt = max(32, next power of two >= 2 max(h) + 1);
    a given Nt below 2 max(h) + 1 is raised to it with a warning."""

    def __init__(self, harmonics, n_time_steps=None, base_freq: float = 50.0):
        self.Harmonics = np.asarray(harmonics, dtype=np.float64).ravel()
        self.NumHarmonics = int(self.Harmonics.size)
        min_steps = int(2 * self.Harmonics.max() + 1)
        if n_time_steps is None:
            self.NumTimeSteps = max(32, 1 << max(0, math.ceil(math.log2(min_steps))))
        elif np.size(n_time_steps) != 1:
            warnings.warn("TimeSteps argument is a vector: using the default steps")
            self.NumTimeSteps = max(32, 1 << max(0, math.ceil(math.log2(min_steps))))
        elif int(n_time_steps) < min_steps:
            warnings.warn(f"TimeSteps ({int(n_time_steps)}) is insufficient for the harmonics: "
                          f"adjusted to {min_steps}")
            self.NumTimeSteps = min_steps
        else:
            self.NumTimeSteps = int(n_time_steps)
        self.BaseFreq = float(base_freq)
        N = self.NumTimeSteps
        theta = np.arange(N) * (2 * np.pi / N)
        self.D_matrix = np.exp(1j * np.outer(theta, self.Harmonics))
        f = np.where(self.Harmonics == 0, 1.0 / N, 2.0 / N)
        self.P_matrix = f[:, None] * np.exp(-1j * np.outer(self.Harmonics, theta))

    def freq2time(self, coeffs):
        return np.real(self.D_matrix @ coeffs)

    def time2freq(self, vals):
        return self.P_matrix @ vals

    def jacobian_scalings(self) -> np.ndarray:
        """Block scalings of the DC-reluctivity Jacobian: 1 for h = 0, 2 for AC
        (assemble_hbfem_kernel.m:28-31)."""
        return np.where(self.Harmonics > 0, 2.0, 1.0)
