"""Pins the oracle's Smith Neo-Hookean response (batched_smith_neo_hookean.cpp:29-269).

- the reference's own golden data (5 FEniCS cases, E=1, nu=0.4, rel tol 1e-4,
  mochi_core/test/batched_materials_test.cpp:154,413);
- finite differences of energy -> PK1 -> tangent;
- the closed form at F = I: the tangent is linear elasticity with the
  unmodified Lame constants (lambda, mu);
- the PSD projection against numpy's 9x9 eigen-decomposition, clamped at
  eps(float32).
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "smith_nh_golden.json")


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


@pytest.mark.parametrize("fp64", [True, False])
def test_golden_cases(oracle, fp64):
    d = json.load(open(GOLD))
    lam, mu = lame(d["youngs_modulus"], d["poisson_ratio"])
    tol = d["rel_tol"]
    for c in d["cases"]:
        e, P, C = oracle.smith_nh(np.array(c["F"]), mu, lam, fp64=fp64, project=False)
        assert abs(e - c["energy"]) <= tol * abs(c["energy"]) + 1e-6
        Pg = np.array(c["P"])
        Cg = np.array(c["tangent"])
        np.testing.assert_allclose(P.ravel(), Pg, rtol=tol, atol=tol * np.abs(Pg).max())
        np.testing.assert_allclose(C.ravel(), Cg, rtol=tol, atol=tol * np.abs(Cg).max())


def test_finite_differences(oracle):
    rng = np.random.default_rng(3)
    lam, mu = lame(1e5, 0.4)
    for _ in range(5):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        e, P, C = oracle.smith_nh(F, mu, lam, fp64=True)
        h = 1e-6
        Pfd = np.zeros((3, 3))
        Cfd = np.zeros((3, 3, 3, 3))
        for k in range(3):
            for l in range(3):
                Fp, Fm = F.copy(), F.copy()
                Fp[k, l] += h
                Fm[k, l] -= h
                ep, Pp, _ = oracle.smith_nh(Fp, mu, lam, fp64=True)
                em, Pm, _ = oracle.smith_nh(Fm, mu, lam, fp64=True)
                Pfd[k, l] = (ep - em) / (2 * h)
                Cfd[:, :, k, l] = (Pp - Pm) / (2 * h)
        np.testing.assert_allclose(P, Pfd, rtol=1e-5, atol=1e-5 * np.abs(P).max())
        np.testing.assert_allclose(C, Cfd, rtol=1e-5, atol=1e-5 * np.abs(C).max())


def test_rest_state_is_linear_elasticity(oracle):
    lam, mu = lame(1e5, 0.4)
    e, P, C = oracle.smith_nh(np.eye(3), mu, lam, fp64=True)
    np.testing.assert_allclose(P, 0.0, atol=1e-9)
    I = np.eye(3)
    Clin = lam * np.einsum("ij,kl->ijkl", I, I) + mu * (np.einsum("ik,jl->ijkl", I, I) +
                                                         np.einsum("il,jk->ijkl", I, I))
    np.testing.assert_allclose(C, Clin, rtol=1e-12, atol=1e-9 * mu)


def test_projection_matches_numpy_eigen(oracle):
    rng = np.random.default_rng(7)
    lam, mu = lame(1e5, 0.4)
    projected = 0
    for _ in range(40):
        F = np.eye(3) + 0.6 * rng.standard_normal((3, 3))
        _, _, C = oracle.smith_nh(F, mu, lam, fp64=True, project=False)
        _, _, Cp = oracle.smith_nh(F, mu, lam, fp64=True, project=True)
        A = C.reshape(9, 9)
        A = 0.5 * (A + A.T)
        w, Q = np.linalg.eigh(A)
        if w.min() >= 0:
            # definite tangents are returned unchanged
            np.testing.assert_allclose(Cp.reshape(9, 9), C.reshape(9, 9), rtol=1e-12, atol=1e-9 * mu)
            continue
        projected += 1
        ref = (Q * np.maximum(w, np.finfo(np.float32).eps)) @ Q.T
        np.testing.assert_allclose(Cp.reshape(9, 9), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
        assert np.linalg.eigvalsh(Cp.reshape(9, 9)).min() > -1e-8 * np.abs(ref).max()
    assert projected > 5


def test_inverted_element_is_projected(oracle):
    lam, mu = lame(1e5, 0.4)
    F = np.diag([1.0, 1.0, -0.5])
    _, _, C = oracle.smith_nh(F, mu, lam, fp64=True, project=False)
    _, _, Cp = oracle.smith_nh(F, mu, lam, fp64=True, project=True)
    assert np.linalg.eigvalsh(C.reshape(9, 9)).min() < 0.0
    w = np.linalg.eigvalsh(Cp.reshape(9, 9))
    assert w.min() >= -1e-12 * w.max()


def test_polyrelu_c2_ramp(oracle):
    t, shift = 0.005, 0.004
    f, df, ddf = oracle.polyrelu(shift - t, t, shift)
    assert abs(f) < 1e-15 and abs(df) < 1e-12
    f, df, ddf = oracle.polyrelu(shift + t, t, shift)
    assert abs(f - t) < 1e-12 and abs(df - 1) < 1e-12 and abs(ddf) < 1e-9
    xs = np.linspace(shift - 2 * t, shift + 2 * t, 401)
    vals = np.array([oracle.polyrelu(x, t, shift) for x in xs])
    # f' and f'' are the derivatives of f
    h = xs[1] - xs[0]
    np.testing.assert_allclose(np.gradient(vals[:, 0], h)[2:-2], vals[2:-2, 1], atol=2e-2)
    np.testing.assert_allclose(vals[xs > shift + t, 0], xs[xs > shift + t] - shift, rtol=1e-12)
    assert np.all(vals[xs < shift - t, 0] == 0)
