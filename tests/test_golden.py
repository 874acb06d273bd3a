"""The oracle still reproduces the committed golden fixtures (guards against oracle drift)."""
import os

import numpy as np

from oracle import fresnel, usfft

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_radon_small_fixture():
    g = load("radon_small")
    assert rel(usfft.tomo_R(g["obj"], g["theta"], np.float64), g["R"]) < 1e-6
    assert rel(usfft.tomo_RT(g["y"], g["theta"], dtype=np.float64), g["RT"]) < 1e-6
    assert rel(usfft.tomo_RT(g["yc"], g["theta"], dtype=np.float64), g["RTc"]) < 1e-6


def test_radon_c1_fixture_nudft_error_stated():
    g = load("radon_c1")
    # FP64 USFFT vs exact polar NUDFT: the window's design error, O(eps)
    err = rel(g["R"][:, :1], g["R_nudft"])
    assert err < 2e-3, err


def test_fresnel_small_fixture():
    g = load("fresnel_small")
    n, nz = g["psi"].shape[2], g["psi"].shape[1]
    fk = fresnel.fresnel_kernel_2d(n, nz, float(g["wavelength"]), float(g["voxelsize"]), g["dist"], np.float64)
    for j in range(len(g["dist"])):
        assert rel(fresnel.D(g["psi"], fk[j], np.float64), g["D"][j]) < 1e-6
        assert rel(fresnel.DT(g["psi"], fk[j], np.float64), g["DT"][j]) < 1e-6
