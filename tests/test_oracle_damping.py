"""Pins the oracle's stiffness (viscous) damping, StressDampingWork
(fem_stress_damping.h:102-501), by what its definition fixes (FP64, no GPU).

- the rest stiffness C0 = dS/dE at F = I of the Smith material is the Lame
  tensor lambda d_ij d_kl + mu (d_ik d_jl + d_il d_jk) with zero rest stress
  (the oracle's Smith tangent at F = I), which is what the damping uses;
- St. Venant-Kirchhoff equivalence (fem_stress_damping_test.cpp:706-743): with
  kappa = 1 and a zero stage start, the damping energy is the StVK energy
  vol (lambda/2 tr(E)^2 + mu E:E) evaluated here from its own formula, its
  residual is that energy's gradient (central differences of the formula),
  and its tangent with the geometric term is the exact derivative of its
  residual;
- the static state (:440-478): stage start equal to u gives zero energy and
  residual;
- geometric-term gating: tangent(geo) - tangent(no geo) = vol (dN_f . S dN_g) I
  with S the viscous PK2 computed here; with projection, S is replaced by its
  eigen-clamped (at eps) form where it is indefinite;
- the scene: the objective's derivative equals the residual with damping on,
  and damping slows a released stretched bar (the speed after one step is
  lower than without it) while a rigid translation is not damped.
"""
import numpy as np
import pytest

from project_superdex_amd import _abi, mesh

X_TET = np.array([[0.0, 0.0, 0.0], [0.02, 0.001, 0.0], [0.002, 0.018, 0.001], [0.001, 0.002, 0.021]])
LAM, MU = 4.0e5, 1.5e5


def basis_grads(X):
    D = (X[1:] - X[0]).T  # columns X_f - X_0
    Dinv = np.linalg.inv(D)
    dN = np.zeros((4, 3))
    dN[1:] = Dinv
    dN[0] = -Dinv.sum(0)
    return dN, np.linalg.det(D) / 6.0


def deformation(X, u):
    dN, vol = basis_grads(X)
    return np.eye(3) + u.reshape(4, 3).T @ dN, dN, vol


def stvk_energy(X, u, lam=LAM, mu=MU):
    F, _, vol = deformation(X, u)
    E = 0.5 * (F.T @ F - np.eye(3))
    return vol * (0.5 * lam * np.trace(E) ** 2 + mu * np.sum(E * E))


def viscous_pk2(X, u, us, kappa, lam=LAM, mu=MU):
    F, dN, vol = deformation(X, u)
    Fs, _, _ = deformation(X, us)
    dE = 0.5 * (F.T @ F - Fs.T @ Fs)
    return kappa * (lam * np.trace(dE) * np.eye(3) + 2 * mu * dE), dN, vol


def test_smith_rest_stiffness_is_the_lame_tensor(oracle):
    e, P, Cm = oracle.smith_nh(np.eye(3), MU, LAM)
    assert np.abs(P).max() < 1e-9 * MU  # no rest stress
    d = np.eye(3)
    lame = LAM * np.einsum("ij,kl->ijkl", d, d) + MU * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d))
    np.testing.assert_allclose(Cm, lame, rtol=1e-12, atol=1e-9 * MU)


def test_st_venant_kirchhoff_equivalence(oracle):
    rng = np.random.default_rng(71)
    for trial in range(3):
        u = rng.uniform(-0.05, 0.05, 12) * 0.02
        e, r, H = oracle.stress_damping_element(X_TET, u, np.zeros(12), 1.0, MU, LAM, geo=True)
        ref = stvk_energy(X_TET, u)
        assert abs(e - ref) <= 1e-10 * abs(ref)
        h = 1e-7
        g = np.array([(stvk_energy(X_TET, u + h * np.eye(12)[k]) - stvk_energy(X_TET, u - h * np.eye(12)[k])) / (2 * h)
                      for k in range(12)])
        np.testing.assert_allclose(r, g, rtol=1e-6, atol=1e-6 * np.abs(g).max())
        # exact tangent with the geometric term: central differences of the residual
        J = np.zeros((12, 12))
        for k in range(12):
            rp = oracle.stress_damping_element(X_TET, u + h * np.eye(12)[k], np.zeros(12), 1.0, MU, LAM, geo=True)[1]
            rm = oracle.stress_damping_element(X_TET, u - h * np.eye(12)[k], np.zeros(12), 1.0, MU, LAM, geo=True)[1]
            J[:, k] = (rp - rm) / (2 * h)
        np.testing.assert_allclose(H, J, rtol=1e-5, atol=1e-6 * np.abs(J).max())
        np.testing.assert_allclose(H, H.T, rtol=0, atol=1e-9 * np.abs(H).max())


def test_static_state_is_undamped(oracle):
    u = np.random.default_rng(11).uniform(-0.05, 0.05, 12) * 0.02
    e, r, _ = oracle.stress_damping_element(X_TET, u, u, 37.0, MU, LAM)
    assert e == 0.0 and np.all(r == 0.0)


@pytest.mark.parametrize("project", [False, True])
def test_geometric_term_gating(oracle, project):
    rng = np.random.default_rng(5)
    us = rng.uniform(-1, 1, 12) * 4e-4
    for u in (us + rng.uniform(-1, 1, 12) * 4e-4,  # mixed signs: S indefinite
              us + 0.02 * X_TET.reshape(-1),        # a dilation: S positive definite
              us):                                  # no strain increment: S = 0
        kappa = 3.0
        _, _, Hn = oracle.stress_damping_element(X_TET, u, us, kappa, MU, LAM, geo=False, project=project)
        _, _, Hg = oracle.stress_damping_element(X_TET, u, us, kappa, MU, LAM, geo=True, project=project)
        S, dN, vol = viscous_pk2(X_TET, u, us, kappa)
        if project:
            w, V = np.linalg.eigh(S)
            if w.min() <= 0:
                S = V @ np.diag(np.maximum(w, np.finfo(float).eps)) @ V.T
        expect = np.kron(vol * dN @ S @ dN.T, np.eye(3))
        # (against the roundoff of the material block, which the difference inherits)
        np.testing.assert_allclose(Hg - Hn, expect, rtol=1e-7, atol=1e-13 * np.abs(Hn).max())
        # the material block alone is positive semi-definite
        assert np.linalg.eigvalsh(0.5 * (Hn + Hn.T)).min() >= -1e-9 * np.abs(Hn).max()


def bar(oracle, beta, geo=0):
    X, tets = mesh.hex_grid_tets(2, 0.01)
    X = X * np.array([2.0, 1.0, 1.0])
    sc = oracle.OracleScene(fp64=True)
    sc.add_soft_tet_mesh(X, tets, _abi.default_soft_params(E=1e5, nu=0.3, penalty=1e6, stiffness_damping=beta,
                                                            stiffness_damping_geometric=geo))
    return sc, X


def test_objective_gradient_with_stiffness_damping(oracle):
    sc, X = bar(oracle, 0.02, geo=1)
    nv = X.shape[0]
    rng = np.random.default_rng(9)
    u0 = 1e-3 * rng.standard_normal((nv, 3))
    u = u0 + 1e-3 * rng.standard_normal((nv, 3))
    p = _abi.pinned_solver_params()
    p.gravity[1] = 0.0
    dt = 1e-3
    sc.set_displacements(0, u)
    res = sc.assemble(0, dt, p, u0=u0.reshape(-1), v0=np.zeros(3 * nv))[0]
    off, _ = bar(oracle, 0.0)
    off.set_displacements(0, u)
    res_off = off.assemble(0, dt, p, u0=u0.reshape(-1), v0=np.zeros(3 * nv))[0]
    assert np.linalg.norm(res - res_off) > 1e-3 * np.linalg.norm(res)
    for trial in range(3):
        d = rng.standard_normal(3 * nv)
        d /= np.linalg.norm(d)
        h = 1e-7
        e = []
        for s in (h, -h):
            sc.set_displacements(0, u.reshape(-1) + s * d)
            e.append(sc.objective(0, dt, p, u0=u0.reshape(-1), v0=np.zeros(3 * nv)))
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(fd - res @ d) <= 1e-6 * np.linalg.norm(res), (trial, fd, res @ d)


def test_damping_slows_a_released_bar_but_not_a_translation(oracle):
    p = _abi.pinned_solver_params(4, lin_rel_tol=1e-12)
    p.gravity[1] = 0.0
    speeds = {}
    for beta in (0.0, 0.01):
        sc, X = bar(oracle, beta)
        u = np.zeros_like(X)
        u[:, 0] = 0.05 * X[:, 0]  # 5% stretch along the bar
        sc.set_displacements(0, u)
        sc.step(1e-3, p)
        speeds[beta] = np.abs(sc.velocities(0)).max()
        # a rigid translation: equal with and without damping
        sc2, _ = bar(oracle, beta)
        sc2.set_velocities(0, np.tile([0.3, -0.1, 0.2], (X.shape[0], 1)))
        sc2.step(1e-3, p)
        if beta == 0.0:
            v_free = sc2.velocities(0)
        else:
            np.testing.assert_allclose(sc2.velocities(0), v_free, rtol=1e-9, atol=1e-12)
    assert speeds[0.01] < 0.5 * speeds[0.0]
