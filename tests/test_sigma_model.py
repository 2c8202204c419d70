"""The Sigma SGS model of the oracle (computeSgsSigma, IdealGasSolverSgs.cpp:
298-360, on calcSvd33, SvdNew.hpp:24-124), pinned by what the mathematics
fixes: the singular values of the one-sided Jacobi SVD equal numpy's, and the
model's operator D_sigma = s3 (s1 - s2) (s2 - s3) / s1^2 vanishes exactly
where the model is built to vanish (pure shear, solid-body rotation, any
two-dimensional flow, isotropic and axisymmetric strain) and equals the
closed form on numpy's singular values elsewhere."""
import numpy as np
import pytest

from oracle import oracle as orc


def dp(a):
    return orc.dptr(a)


def dsigma(G):
    d = np.ascontiguousarray(np.asarray(G, dtype=np.float64).reshape(9))
    return orc.lib().og_sigma_dsigma(dp(d))


def test_svd33_singular_values_match_numpy():
    rng = np.random.default_rng(3)
    L = orc.lib()
    for trial in range(400):
        A = rng.normal(size=(3, 3)) * 10.0 ** rng.uniform(-3, 3)
        if trial % 4 == 1:
            A[:, 2] = A[:, 0] * 0.5 + A[:, 1]  # rank 2
        if trial % 4 == 2:
            A = np.outer(rng.normal(size=3), rng.normal(size=3))  # rank 1
        Acm = np.ascontiguousarray(A.T.reshape(9))
        U, V, s = np.zeros(9), np.zeros(9), np.zeros(3)
        it = L.og_svd33(dp(Acm), dp(U), dp(V), dp(s))
        assert 1 <= it <= 101
        ref = np.linalg.svd(A, compute_uv=False)
        scale = ref[0]
        assert np.allclose(np.sort(s)[::-1], ref, rtol=0, atol=1e-13 * scale), (A, s, ref)
        # U S V^T reassembles A
        Um, Vm = U.reshape(3, 3).T, V.reshape(3, 3).T
        assert np.allclose(Um @ np.diag(s) @ Vm.T, A, rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("name,G", [
    ("pure shear", [[0, 2.5, 0], [0, 0, 0], [0, 0, 0]]),
    ("solid rotation", [[0, -1.3, 0], [1.3, 0, 0], [0, 0, 0]]),
    ("isotropic expansion", np.eye(3) * 0.7),
    ("axisymmetric strain", np.diag([2.0, -1.0, -1.0]) * 0.4),
    ("2-D flow", [[0.3, -1.1, 0], [0.8, -0.3, 0], [0, 0, 0]]),
    ("zero gradient", np.zeros((3, 3))),
])
def test_sigma_operator_vanishes(name, G):
    G = np.asarray(G, dtype=float)
    scale = max(np.abs(G).max(), 1e-300)
    assert abs(dsigma(G)) <= 1e-14 * scale, name


def test_sigma_operator_closed_form():
    rng = np.random.default_rng(8)
    for _ in range(300):
        G = rng.normal(size=(3, 3))
        s = np.linalg.svd(G, compute_uv=False)
        ref = s[2] * (s[0] - s[1]) * (s[1] - s[2]) / s[0] ** 2
        assert dsigma(G) == pytest.approx(ref, rel=1e-10, abs=1e-15 * s[0])
    # homogeneous of degree one in the gradient
    G = rng.normal(size=(3, 3))
    assert dsigma(3.0 * G) == pytest.approx(3.0 * dsigma(G), rel=1e-12)
