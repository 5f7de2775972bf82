"""Input-field generators (host side, NumPy): nodal material/control fields fed to the hot path.

FourierControl follows fol/controls/fourier_control.py:49-59; the two-phase Gaussian random field
is a build-defined generator (seeded white noise, Gaussian filter in Fourier space, threshold).
"""
from __future__ import annotations

import numpy as np


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


class FourierControl:
    """k(x) = (max-min) sigmoid(beta (c0/2 + sum_i c_i cos(fx pi x) cos(fy pi y) cos(fz pi z) - 0.5)) + min.

    Frequencies ordered as meshgrid(x_freqs, y_freqs, z_freqs, indexing='ij').ravel()
    (fourier_control.py:45-46); variable vector = [c0, c_1, ..., c_n]."""

    def __init__(self, control_name, control_settings, fe_mesh):
        self.name = control_name
        self.settings = control_settings
        self.fe_mesh = fe_mesh
        self.initialized = False

    def Initialize(self, reinitialize=False):
        if self.initialized and not reinitialize:
            return
        s = self.settings
        self.min = s.get("min", 1e-6)
        self.max = s.get("max", 1.0)
        self.beta = s["beta"]
        fx, fy, fz = np.meshgrid(np.asarray(s["x_freqs"]), np.asarray(s["y_freqs"]), np.asarray(s["z_freqs"]),
                                 indexing="ij")
        self.frequencies = np.stack([fx.ravel(), fy.ravel(), fz.ravel()], axis=1).astype(np.float64)
        self.num_control_vars = self.frequencies.shape[0] + 1
        self.num_controlled_vars = self.fe_mesh.GetNumberOfNodes()
        self.initialized = True

    def GetName(self):
        return self.name

    def GetNumberOfVariables(self):
        return self.num_control_vars

    def GetNumberOfControlledVariables(self):
        return self.num_controlled_vars

    def ComputeControlledVariables(self, variable_vector):
        v = np.asarray(variable_vector, dtype=np.float64)
        X = self.fe_mesh.GetNodesCoordinates()
        K = v[0] / 2.0 + np.zeros(X.shape[0])
        for f, cf in zip(self.frequencies, v[1:]):
            K = K + cf * np.cos(f[0] * np.pi * X[:, 0]) * np.cos(f[1] * np.pi * X[:, 1]) * np.cos(f[2] * np.pi * X[:, 2])
        return (self.max - self.min) * _sigmoid(self.beta * (K - 0.5)) + self.min

    def ComputeBatchControlledVariables(self, batch_variable_vector):
        return np.stack([self.ComputeControlledVariables(v) for v in np.atleast_2d(batch_variable_vector)])


def fourier_field(coords, coeffs, freqs=(1, 2, 3, 4), vmin=0.1, vmax=1.0):
    """Nodal field 0.1 + 0.9 sigmoid(sum c cos(fx pi x) cos(fy pi y)) style generator (c0 = 1, beta = 1)."""
    fx, fy = np.meshgrid(np.asarray(freqs), np.asarray(freqs), indexing="ij")
    x, y = coords[:, 0], coords[:, 1]
    K = np.zeros(coords.shape[0])
    for a, b, cf in zip(fx.ravel(), fy.ravel(), np.asarray(coeffs, dtype=np.float64)):
        K += cf * np.cos(a * np.pi * x) * np.cos(b * np.pi * y)
    return (vmax - vmin) * _sigmoid(K) + vmin


def grf_two_phase(shape, corr_len, seed, lo=1.0, hi=10.0):
    """Two-phase field on a node lattice `shape` (nz, ny, nx): N(0,1) white noise (seed), Gaussian
    filter exp(-corr_len^2 |k|^2 / 2) in Fourier space, value `hi` where positive else `lo`."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    ks = [np.fft.fftfreq(s) * 2.0 * np.pi for s in shape]
    K2 = sum(np.meshgrid(*[k ** 2 for k in ks], indexing="ij"))
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.exp(-0.5 * corr_len ** 2 * K2)))
    return np.where(f > 0.0, hi, lo).ravel()
