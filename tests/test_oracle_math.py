"""Closed-form and invariant pins of the oracle (CPU only).

Each check is chosen so that a plausible slip in the restatement -- a dropped
term, a wrong sign or index, a transposed operand -- fails it:
  * B-spline stencil: partition of unity, zero first moment, second moment
    1/4 cell^2, the fraction-0 and fraction-1/2 special cases;
  * SVD: reconstruction, orthonormal V, singular values = numpy's;
  * Hencky stress of a diagonal F: tau_kk = 2 mu ln F_kk + lambda sum ln F_jj;
  * Drucker-Prager: tension -> sigma = 1, cone -> |eps_hat| = limit,
    projection idempotent, elastic states untouched;
  * P2G: mass, momentum and stress-force sums (sum w d = 0);
  * G2P: an affine grid velocity field v = a + B x is reproduced exactly
    (v_p = a + B x_p, C = B);
  * grid update: free fall adds g dt; floor Coulomb rule by hand;
  * fixed-point mass: every particle deposits exactly its quanta;
  * M1 single-node exchange by hand; rigid integrator closed form.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import MASS_FIXED, Body, Collider, OracleSim, Params, make_particles

DX = 0.1


def sim(prec="f64", **kw):
    return OracleSim(DX, prec, params=Params(**kw) if kw else Params())


@pytest.mark.parametrize("x", [0.0123, 0.2, 0.25, 0.3456, 0.05, 0.15, -0.137])
def test_bspline_moments(x):
    s = sim()
    fr, base, w = s.weights((x, x + 0.011, x + 0.027))
    for a in range(3):
        xa = (x, x + 0.011, x + 0.027)[a]
        d = (base[a] + np.arange(3)) - xa / DX  # node i at i*dx
        assert abs(w[a].sum() - 1.0) < 1e-14
        assert abs((w[a] * d).sum()) < 1e-13
        assert abs((w[a] * d * d).sum() - 0.25) < 1e-13


def test_bspline_special_fractions():
    s = sim()
    # x/dx - 1/2 integral -> fraction 0 -> weights (1/2, 1/2, 0)
    _, b, w = s.weights((0.25, 0.25, 0.25))
    assert np.allclose(w[0], [0.5, 0.5, 0.0], atol=1e-12) and b[0] == 2
    # fraction 1/2 -> (1/8, 3/4, 1/8)
    _, b, w = s.weights((0.3, 0.3, 0.3))
    assert np.allclose(w[0], [0.125, 0.75, 0.125], atol=1e-12) and b[0] == 2


def test_svd_reconstruction_and_singular_values():
    rng = np.random.default_rng(0)
    s = sim()
    for _ in range(300):
        F = rng.normal(size=(3, 3))
        U, S, V = s.svd(F)
        assert np.allclose(U @ np.diag(S) @ V.T, F, atol=1e-9 * max(1, np.abs(F).max()))
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.allclose(np.sort(S), np.sort(np.linalg.svd(F, compute_uv=False)), rtol=1e-6, atol=1e-9)


def test_svd_fp32_matches_numpy_near_identity():
    rng = np.random.default_rng(1)
    s = OracleSim(DX, "f32")
    for _ in range(200):
        F = np.eye(3) + 0.05 * rng.normal(size=(3, 3))
        U, S, V = s.svd(F)
        assert np.allclose(U @ np.diag(S) @ V.T, F, atol=5e-6)
        assert np.allclose(np.sort(S), np.sort(np.linalg.svd(F, compute_uv=False)), atol=1e-5)


def test_hencky_diagonal_closed_form():
    E, nu = 3.5e5, 0.3
    s = sim(E=E, nu=nu)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    d = np.array([1.02, 0.97, 1.005])
    tau = s.hencky_tau(np.diag(d))
    want = 2 * mu * np.log(d) + lam * np.log(d).sum()
    assert np.allclose(np.diag(tau), want, rtol=1e-10)
    assert np.allclose(tau - np.diag(np.diag(tau)), 0, atol=1e-6)
    # rotation covariance: tau(R F) = R tau(F) R^T for a rotation R
    th = 0.3
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Fr = R @ np.diag(d)
    assert np.allclose(s.hencky_tau(Fr), R @ tau @ R.T, atol=1e-6 * np.abs(tau).max())


def _koh(E, nu):
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return (3 * lam + 2 * mu) / (2 * mu)


def test_dp_tension_goes_to_rotation():
    s = sim(alpha=0.4)
    F = np.diag([1.1, 1.05, 1.02])
    Fn = s.plasticity(F, np.zeros((3, 3)), 1e-3)
    assert np.allclose(np.linalg.svd(Fn, compute_uv=False), 1.0, atol=1e-12)


def test_dp_cone_projection_and_idempotence():
    E, nu, alpha = 1e5, 0.3, 0.3
    s = sim(E=E, nu=nu, alpha=alpha)
    F = np.diag([1.2, 0.7, 1.0])  # tr ln < 0, strongly deviatoric
    Fn = s.plasticity(F, np.zeros((3, 3)), 1e-3)
    sv = np.linalg.svd(Fn, compute_uv=False)
    e = np.log(sv)
    tr = e.sum()
    assert tr < 0
    assert abs(np.linalg.norm(e - tr / 3) - (-alpha * _koh(E, nu) * tr)) < 1e-10
    assert abs(tr - np.log([1.2, 0.7, 1.0]).sum()) < 1e-12  # volume preserved by the return map
    Fnn = s.plasticity(Fn, np.zeros((3, 3)), 1e-3)
    assert np.allclose(Fnn, Fn, atol=1e-12)


def test_dp_elastic_state_untouched():
    s = sim(alpha=0.4)
    F = np.diag([0.99, 0.985, 0.99])
    Fn = s.plasticity(F, np.zeros((3, 3)), 1e-3)
    assert np.allclose(Fn, F, atol=1e-13)


def _cloud(n=50, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    v = rng.normal(size=(n, 3))
    P = make_particles(x, v, mass=rng.uniform(0.5, 2.0, size=n), rho=1500.0)
    P["F"] = np.eye(3) + 0.05 * rng.normal(size=(n, 3, 3))
    P["C"] = rng.normal(size=(n, 3, 3))
    return P


def test_p2g_conservation():
    s = sim()
    P = _cloud()
    s.set_particles(P)
    s.p2g()
    G = s.grid_dense((-2, -2, -2), (16, 16, 16))
    assert abs(G["mass"].sum() - P["m"].sum()) < 1e-12 * P["m"].sum()
    mom = G["vel"].reshape(-1, 3).sum(0)
    assert np.allclose(mom, (P["m"][:, None] * P["v"]).sum(0), atol=1e-11)
    assert np.allclose(G["force"].reshape(-1, 3).sum(0), 0.0, atol=1e-6 * np.abs(G["force"]).max())


def test_fixed_point_mass_is_exact():
    P = _cloud(200, seed=4)
    P["m"] = np.full(200, np.float32(7.6293945e-4), np.float64)
    e = int(math.floor(math.log2(P["m"].max()))) - 39
    s = OracleSim(DX, "f32", MASS_FIXED, e)
    s.set_particles(P)
    s.p2g()
    q = np.rint(np.ldexp(P["m"], -e)).astype(np.int64)
    assert s.grid_mass_q_total() == int(q.sum())
    G = s.grid_dense((-2, -2, -2), (16, 16, 16))
    assert np.array_equal(G["active"], G["mass_q"] != 0)
    assert np.allclose(G["mass"], np.ldexp(G["mass_q"].astype(np.float64), e), rtol=1e-7)


def test_fixed_point_node_quanta_round_the_weights_to_nearest():
    """Per node, the fixed-point deposit of one particle is w q_p rounded to
    the nearest integer, at least one quantum on every node with w > 0 and
    none where w = 0 (a particle on a node plane), and the centre takes the
    remainder.  Rounding up (or down) every node would bias the rim of each
    stencil against its centre, and fails the half-quantum bound; dropping
    the minimum would leave a touched node empty."""
    rng = np.random.default_rng(11)
    e = -30
    # few quanta, so that the FP32 weights' own rounding (~1e-7 of w) stays far
    # below half a quantum and the rounding rule is what the bound sees
    qp = 5_003
    m = float(np.ldexp(float(qp), e))
    xs = list(rng.uniform(0.3, 0.6, size=(6, 3)))
    xs.append(np.array([10.5, 12.25, 11.0]) * DX)  # f = x/dx - 0.5 integral on x: w = 0 on one plane
    xs.append(np.array([10.5, 12.5, 11.5]) * DX)   # ... on all three axes
    xs.append(np.array([10.5 + 1e-2, 12.5 + 2e-2, 11.5 + 3e-3]) * DX)  # tiny weights: the one-quantum minimum
    tiny = 0
    for x in xs:
        s = OracleSim(DX, "f32", MASS_FIXED, e)
        P = make_particles([x], [[0.0, 0.0, 0.0]], mass=m)
        s.set_particles(P)
        s.p2g()
        f = P["x"][0].astype(np.float64) / DX - 0.5  # the stored FP32 position
        b = np.floor(f).astype(int)
        fr = f - b
        w1 = np.stack([0.5 * (1 - fr) ** 2, 0.75 - (fr - 0.5) ** 2, 0.5 * fr ** 2])  # (3 offsets, 3 axes)
        G = s.grid_dense(tuple(b), (3, 3, 3))
        q = G["mass_q"]
        w = w1[:, 0][:, None, None] * w1[:, 1][None, :, None] * w1[:, 2][None, None, :]
        assert int(q.sum()) == qp
        others = np.ones((3, 3, 3), bool)
        others[1, 1, 1] = False
        assert np.array_equal(q[others] > 0, w[others] > 0), x
        wq = w[others] * qp
        big = wq >= 0.5
        assert (np.abs(q[others][big] - wq[big]) <= 0.5 + 1e-3).all(), (x, q[others] - wq)
        assert (q[others][~big & (w[others] > 0)] == 1).all()
        tiny += int((~big & (w[others] > 0)).sum())
        assert q[1, 1, 1] == qp - q[others].sum()
    assert tiny > 0  # the minimum was exercised


def test_g2p_reproduces_affine_field():
    s = sim(gravity=(0.0, 0.0, 0.0), alpha=10.0)
    a = np.array([0.3, -0.2, 0.1])
    B = np.array([[0.1, 0.2, -0.3], [0.05, -0.1, 0.2], [0.3, 0.1, 0.0]])
    org, dims = (0, 0, 0), (12, 12, 12)
    idx = np.indices(dims).reshape(3, -1).T
    xn = (idx + np.array(org)) * DX  # P2G/G2P node i at i*dx
    vel = (a + xn @ B.T).reshape(dims + (3,))
    s.grid_set_dense(org, np.ones(dims), vel, np.ones(dims, np.uint8))
    rng = np.random.default_rng(2)
    x = rng.uniform(0.3, 0.8, size=(20, 3))
    s.set_particles(make_particles(x, np.zeros((20, 3)), mass=1.0))
    s.g2p(1e-9)
    Pn = s.get_particles()
    assert np.allclose(Pn["v"], a + x @ B.T, atol=1e-10)
    assert np.allclose(Pn["C"], np.broadcast_to(B, (20, 3, 3)), atol=1e-9)


def test_free_fall_one_substep():
    g = -9.81
    s = sim(gravity=(0.0, g, 0.0))
    s.set_particles(make_particles([[0.5, 0.9, 0.5]], [[0.0, 0.0, 0.0]], mass=1.0))
    dt = 1e-3
    s.step(dt)
    Pn = s.get_particles()
    assert abs(Pn["v"][0, 1] - g * dt) < 1e-12
    assert abs(Pn["x"][0, 1] - (0.9 + g * dt * dt)) < 1e-12


def test_floor_coulomb_rule():
    s = sim(gravity=(0.0, 0.0, 0.0), mu_floor=0.4)
    # one particle near the floor moving down and sideways; nodes j <= 0 see the floor
    s.set_particles(make_particles([[0.5, 0.02, 0.5]], [[1.0, -2.0, 0.0]], mass=1.0))
    s.p2g()
    s.update(1e-3)
    G = s.grid_dense((0, -2, 0), (10, 6, 10))
    for j in range(6):
        jj = j - 2
        v = G["vel"][:, j, :][G["active"][:, j, :]]
        if jj <= 0:
            # v_y clamped to 0, tangential reduced by mu * 2 (from |v_t| = 1)
            assert np.allclose(v[:, 1], 0.0)
            assert np.allclose(v[:, 0], 1.0 - 0.4 * 2.0, atol=1e-12)
        else:
            assert np.allclose(v[:, 1], -2.0, atol=1e-12)


def test_m1_single_node_exchange():
    s = sim()
    m, v = 0.5, np.array([0.0, -1.0, 0.0])
    s.grid_set_dense((0, 0, 0), np.full((1, 1, 1), m), v.reshape(1, 1, 1, 3), np.ones((1, 1, 1), np.uint8))
    he = (0.1, 0.1, 0.1)
    M = 4.0
    col = Collider(center=(0.05, 0.05, 0.05), half_extents=he, lin_vel=(0.0, 0.5, 0.0), mass=M)
    f, t, n = s.couple(1, col, 1e-3)
    approx = max(1, round(8 * 0.1 ** 3 / DX ** 3))
    mb = M / approx
    vb = np.array([0.0, 0.5, 0.0])
    vnew = (m * v + mb * vb) / (m + mb)
    # only node (0,0,0) exists in the grid's single block but the block has 64 nodes; the
    # inactive ones inside the box are claimed at v_body and contribute nothing
    assert n >= 1
    assert np.allclose(f, (vnew - vb) * mb / 1e-3, rtol=1e-12)
    G = s.grid_dense((0, 0, 0), (1, 1, 1))
    assert np.allclose(G["vel"][0, 0, 0], vnew) and abs(G["mass"][0, 0, 0] - (m + mb)) < 1e-15


def test_rigid_integrator_closed_form():
    s = sim()
    b = Body(x=(0, 0, 0), v=(0, 0, 0), inv_mass=0.5, lin_damp=0.0, ang_damp=0.0, force=(2.0, 0.0, -4.0))
    b2 = s.integrate_body(b, 0.1)
    assert np.allclose(b2.v, (0.1, 0.0, -0.2)) and np.allclose(b2.x, (0.01, 0.0, -0.02))
    assert np.allclose(b2.force, 0.0)
