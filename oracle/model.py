"""FP64 forward model and gradient built from the restated operators (TEST INFRASTRUCTURE ONLY).

L(delta, beta) = sum(|D(exp(i k R delta - k R beta))|^2 - d)^2 with R =
``oracle.usfft.tomo_R`` (processing/tomo.py:121-151) and D = ``oracle.fresnel.D``
(processing/propagation.py:47-64).  The gradient is assembled from the oracle's
own adjoints (``tomo_RT``, ``fresnel.DT``) by the chain rule:
    r = 4 (|u|^2 - d) u, G = DT r,
    dL/dp = -k Im(conj(G) psi), dL/db = -k Re(conj(G) psi),
    dL/ddelta = RT dL/dp, dL/dbeta = RT dL/db.
Parity pin: ``fd_grad`` (central finite differences of ``loss`` in FP64) agrees
with ``grad`` to ~1e-8 (tests/test_oracle_model.py).
"""
from __future__ import annotations

import numpy as np

from oracle import fresnel, usfft


class Model:
    def __init__(self, theta, n, nz, wavelength, voxelsize, z, k, pad="symmetric"):
        self.theta = np.asarray(theta, dtype=np.float32)
        self.k = float(k)
        self.pad = pad
        self.fk = fresnel.fresnel_kernel_2d(n, nz, wavelength, voxelsize, [z], np.float64)[0]

    def psi_u(self, delta, beta):
        p = usfft.tomo_R(delta, self.theta, np.float64)
        b = usfft.tomo_R(beta, self.theta, np.float64)
        psi = np.exp(1j * self.k * p - self.k * b)
        return p, b, psi, fresnel.D(psi, self.fk, np.float64, pad=self.pad)

    def intensity(self, delta, beta):
        return np.abs(self.psi_u(delta, beta)[3]) ** 2

    def loss(self, delta, beta, data):
        u = self.psi_u(delta, beta)[3]
        return float(np.sum((np.abs(u) ** 2 - data) ** 2))

    def grad(self, delta, beta, data):
        _, _, psi, u = self.psi_u(delta, beta)
        r = 4.0 * (np.abs(u) ** 2 - data) * u
        G = fresnel.DT(r, self.fk, np.float64, pad=self.pad)
        q = np.conj(G) * psi
        gp = -self.k * q.imag
        gb = -self.k * q.real
        return (usfft.tomo_RT(gp, self.theta, dtype=np.float64),
                usfft.tomo_RT(gb, self.theta, dtype=np.float64))

    def fd_grad(self, delta, beta, data, which, idx, h):
        """Central differences of ``loss`` w.r.t. delta (which=0) or beta (which=1) at flat indices idx."""
        out = []
        for i in idx:
            vols = [delta.copy(), beta.copy()]
            vols[which].flat[i] += h
            lp = self.loss(vols[0], vols[1], data)
            vols[which].flat[i] -= 2 * h
            lm = self.loss(vols[0], vols[1], data)
            out.append((lp - lm) / (2 * h))
        return np.array(out)
