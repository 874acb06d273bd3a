"""Pins the FP64 model gradient of oracle/model.py by central finite differences."""
import numpy as np

from oracle.model import Model


def test_oracle_gradient_matches_finite_differences():
    n, nz, nt = 32, 8, 12
    theta = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(0)
    d_true = np.abs(rng.standard_normal((nz, n, n)))
    m = Model(theta, n, nz, 4.13e-11, 1.4e-6, 0.05, k=0.5)
    data = m.intensity(d_true, 0.02 * d_true)
    delta = d_true + 0.1 * rng.standard_normal(d_true.shape)
    beta = 0.02 * d_true + 0.01 * rng.standard_normal(d_true.shape)
    gd, gb = m.grad(delta, beta, data)
    idx = rng.choice(delta.size, 6, replace=False)
    fd_d = m.fd_grad(delta, beta, data, 0, idx, 1e-5)
    fd_b = m.fd_grad(delta, beta, data, 1, idx, 1e-5)
    assert np.allclose(fd_d, gd.flat[idx], rtol=1e-6, atol=1e-8 * np.abs(gd).max())
    assert np.allclose(fd_b, gb.flat[idx], rtol=1e-6, atol=1e-8 * np.abs(gb).max())
    # a dropped factor (e.g. 2 instead of 4 in the residual) would fail
    assert not np.allclose(0.5 * gd.flat[idx], fd_d, rtol=1e-3)
