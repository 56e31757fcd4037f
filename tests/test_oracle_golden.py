"""The oracle against the reference's own element code (golden vectors made by
tests/golden/make_golden.py from the reference compiled from source), plus
invariants the mathematics fixes."""
import os

import numpy as np
import pytest

import pyoracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "element_golden.npz"))
NREC = GOLD["Xpts"].shape[0]


def _rec(i):
    m = O.material(E=GOLD["E"][i], nu=GOLD["nu"][i], rho=GOLD["rho"][i], cte=GOLD["cte"][i])
    return (m, GOLD["thickness"][i], GOLD["temperature"][i], GOLD["Xpts"][i], GOLD["vars"][i], GOLD["psi"][i])


def rel(a, b):
    return np.abs(np.asarray(a) - b).max() / np.abs(b).max()


@pytest.mark.parametrize("i", range(NREC))
def test_residual_and_stiffness_match_reference(i):
    m, t, T, X, q, psi = _rec(i)
    assert rel(O.elem_residual(0, m, t, T, X, q), GOLD["res_lin"][i]) < 1e-13
    assert rel(O.elem_residual(1, m, t, T, X, q), GOLD["res_nl"][i]) < 1e-13
    assert rel(O.elem_stiffness(0, m, t, T, X, q), GOLD["K_lin"][i]) < 1e-13
    _, Kn = O.elem_jacobian(1, m, t, T, X, q)
    assert rel(Kn, GOLD["K_nl"][i]) < 1e-13


@pytest.mark.parametrize("i", range(NREC))
def test_geometric_fd_matches_reference(i):
    m, t, T, X, q, psi = _rec(i)
    G = O.elem_geometric_fd(m, t, T, X, q)
    # both sides are a central difference with step 1e-4: their round-off
    # differs; with a temperature the step along the state is ~1e-4/|T| and the
    # difference carries more cancellation.
    tol = 1e-11 if T == 0.0 else 1e-7
    assert rel(G, GOLD["G"][i]) < tol


@pytest.mark.parametrize("i", range(NREC))
def test_thickness_adjoint_matches_reference_derivative(i):
    m, t, T, X, q, psi = _rec(i)
    a_lin = O.elem_adj_res_dv(0, m, t, T, psi, X, q)
    a_nl = O.elem_adj_res_dv(1, m, t, T, psi, X, q)
    assert abs(a_lin - GOLD["dRdt_lin"][i]) <= 1e-12 * abs(GOLD["dRdt_lin"][i])
    assert abs(a_nl - GOLD["dRdt_nl"][i]) <= 1e-12 * abs(GOLD["dRdt_nl"][i])
    if T == 0.0:
        # the reference's own addAdjResProduct agrees when no thermal strain
        assert abs(a_lin - GOLD["adj_lin"][i]) <= 1e-12 * abs(GOLD["adj_lin"][i])


@pytest.mark.parametrize("i", range(NREC))
def test_energy_matches_reference(i):
    m, t, T, X, q, psi = _rec(i)
    assert abs(O.elem_energy(0, m, t, T, X, q) - GOLD["energy_lin"][i]) <= 1e-13 * abs(GOLD["energy_lin"][i])


def test_linear_model_invariants():
    """K symmetric, R = K q (T = 0), K annihilates rigid translations, and
    R = dU/dq by central differences of the energy."""
    rng = np.random.default_rng(7)
    m = O.material()
    for k in range(3):
        X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float) * 0.4 + 0.05 * rng.standard_normal((4, 3))
        q = 1e-3 * rng.standard_normal(24)
        K = O.elem_stiffness(0, m, 0.003, 0.0, X, q)
        assert np.abs(K - K.T).max() <= 1e-14 * np.abs(K).max()
        R = O.elem_residual(0, m, 0.003, 0.0, X, q)
        assert rel(K @ q, R) < 1e-13
        for d in range(3):
            r = np.zeros(24)
            r[d::6][:4] = 1.0
            assert np.abs(K @ r).max() <= 1e-9 * np.abs(K).max()
        # energy gradient: U is quadratic in q, central differences are exact
        h = 1e-6
        for j in rng.choice(24, 4, replace=False):
            e = np.zeros(24)
            e[j] = h
            dU = (O.elem_energy(0, m, 0.003, 0.0, X, q + e) - O.elem_energy(0, m, 0.003, 0.0, X, q - e)) / (2 * h)
            assert abs(dU - R[j]) <= 1e-6 * np.abs(R).max()


def test_nonlinear_tangent_is_residual_derivative():
    """K_nl = dR_nl/dq (central differences, O(h^2) exact for the quadratic
    strain up to round-off)."""
    rng = np.random.default_rng(8)
    m = O.material()
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float) * 0.4 + 0.05 * rng.standard_normal((4, 3))
    q = 1e-2 * rng.standard_normal(24)
    _, K = O.elem_jacobian(1, m, 0.003, 0.0, X, q)
    h = 1e-7
    for j in range(24):
        e = np.zeros(24)
        e[j] = h
        col = (O.elem_residual(1, m, 0.003, 0.0, X, q + e) - O.elem_residual(1, m, 0.003, 0.0, X, q - e)) / (2 * h)
        assert np.abs(col - K[:, j]).max() <= 1e-6 * np.abs(K).max()


@pytest.mark.parametrize("i", range(NREC))
def test_mass_matches_reference(i):
    m, t, T, X, q, psi = _rec(i)
    M = O.elem_mass(m, t, X)
    assert rel(M, GOLD["M"][i]) < 1e-13


def _plate(a=0.7, b=0.4, z=0.3):
    # flat rectangle in the plane z = const, tensor node order
    return np.array([[0, 0, z], [a, 0, z], [0, b, z], [a, b, z]], float)


def test_mass_rigid_motion_energies():
    """Closed forms on a flat a x b plate: a rigid translation carries
    rho t a b, an in-plane rotation vector rho t^3/12 a b per unit square
    component, and a rotation about the normal (the drilling direction,
    where the director d = q x n does not move) carries none."""
    rho, t, a, b = 2700.0, 0.01, 0.7, 0.4
    m = O.material(rho=rho)
    M = O.elem_mass(m, t, _plate(a, b))
    assert np.abs(M - M.T).max() <= 1e-14 * np.abs(M).max()
    area = a * b
    for d in range(3):
        u = np.zeros(24)
        u[d::6] = 1.0
        assert abs(u @ M @ u - rho * t * area) <= 1e-12 * rho * t * area
    mr = rho * t ** 3 / 12.0
    for d, want in [(3, mr * area), (4, mr * area), (5, 0.0)]:
        u = np.zeros(24)
        u[d::6] = 1.0
        assert abs(u @ M @ u - want) <= 1e-12 * mr * area
    # translations and rotations do not couple (m1 = 0 for tOffset = 0)
    tr = [6 * n + d for n in range(4) for d in range(3)]
    rot = [6 * n + 3 + d for n in range(4) for d in range(3)]
    assert np.abs(M[np.ix_(tr, rot)]).max() == 0.0
    # consistent bilinear mass: the translational block of node pairs is
    # rho t a b / 36 [4 2 2 1] pattern
    base = np.array([[4, 2, 2, 1], [2, 4, 1, 2], [2, 1, 4, 2], [1, 2, 2, 4]]) * rho * t * area / 36.0
    Mx = M[0::6, 0::6]
    assert np.abs(Mx - base).max() <= 1e-12 * base.max()


@pytest.mark.parametrize("i", range(NREC))
def test_output_data_matches_reference(i):
    """getOutputData: strains, stresses, failure and design values at the
    4 nodes (the element zone of an .f5 file); the leading node location and
    displacement groups are the element's own inputs."""
    m, t, T, X, q, psi = _rec(i)
    out = O.elem_output(0, m, t, float(GOLD["ys"]), X, q)
    ref = GOLD["output_lin"][i]
    assert np.array_equal(ref[:, 0:3], X.reshape(4, 3))
    assert np.array_equal(ref[:, 3:9], q.reshape(4, 6))
    assert rel(out, ref[:, 9:40]) < 1e-13
    assert rel(out[:, 18], ref[:, 27]) < 1e-13  # failure0 on its own scale


@pytest.mark.parametrize("i", range(NREC))
def test_geometric_long_double_matches_reference(i):
    """The long-double evaluation of the oracle's geometric stiffness against
    the reference's own central difference (its step 1e-4, so its round-off
    sets the tolerance as above), against the double evaluation at the
    accurate step 1, and free of the step: the tangent is exactly quadratic
    in the state, so every step gives the same matrix once round-off is
    gone."""
    m, t, T, X, q, psi = _rec(i)
    Gl = O.elem_geometric_step_ld(m, t, T, X, q, 1.0)
    tol = 1e-11 if T == 0.0 else 1e-7
    assert rel(Gl, GOLD["G"][i]) < tol
    # with a temperature the difference direction includes it (a larger
    # perturbation of the stresses): more round-off on both evaluations
    assert rel(O.elem_geometric_step(m, t, T, X, q, 1.0), Gl) < (1e-12 if T == 0.0 else 1e-11)
    for dh in (0.1, 10.0):
        assert rel(O.elem_geometric_step_ld(m, t, T, X, q, dh), Gl) < (1e-15 if T == 0.0 else 1e-13)
