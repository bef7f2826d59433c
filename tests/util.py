"""Helpers shared by the tests: synthetic inputs in the reference layout and
norm-wise error measures."""
import numpy as np


def rel_inf(a, b):
    """||a - b||_inf / ||b||_inf (norm-wise; element-wise relative error is
    ill-conditioned where a field crosses zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if b.size else 0.0
    if den == 0.0:
        return num
    return num / den


def random_hydro(rng, n, drho=1e-3, du=1e-3):
    """rho = 1 + drho N(0,1), u = du N(0,1) per component (flat, reference layout)."""
    rho = 1.0 + drho * rng.standard_normal(n)
    u = du * rng.standard_normal(3 * n)
    return rho, u


def isotropic(rng, num):
    v = rng.standard_normal((3, num))
    return v / np.linalg.norm(v, axis=0)


def readme_mobility(tau, a=0.1, hydro=1.5):
    """(1/a - 1/hydro)/(6 pi eta) with eta = tau/3, evaluated as README.md:88-106 writes it."""
    eta = tau / 3
    return (1.0 / a - 1.0 / hydro) / (6.0 * np.pi * eta)
