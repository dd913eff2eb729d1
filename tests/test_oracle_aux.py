"""Closed-form checks of the oracle's updatesolution, cfl and L2energysquared
(no GPU).  The meshes are bricks with a different element count per
direction, so the metric terms differ per direction and a mixed-up field
index changes the result."""
import numpy as np
import pytest

from oracle import oracle as O

from states import small_setup


def uniform_state(s, rho=1.2, theta=300.0, vel=None):
    Q = s.zeros_state()
    d = s.dim
    Q[:, d, :] = rho          # rho
    Q[:, d + 1, :] = rho * theta  # rho theta
    if vel is not None:
        for k, u in enumerate(vel):
            Q[:, k, :] = rho * u
    return np.asfortranarray(Q)


def widths(s, L=1000.0, Ne=3):
    # small_setup: Ne + k elements along direction k over [0, L]
    return [L / (Ne + k) for k in range(s.dim)]


@pytest.mark.parametrize("dim", [2, 3])
def test_cfl_closed_form(dim):
    s = small_setup(0, dim, 3)
    ph = s.phys
    rho, theta = 1.2, 300.0
    P = ph.p0 * (ph.R_gas * rho * theta / ph.p0) ** (ph.c_p / ph.c_v)
    cs = np.sqrt(ph.gamma * P / rho)
    # at rest: 2 rho / (rho cs) on every node
    dt0 = O.euler_cfl(s, uniform_state(s, rho, theta))
    assert abs(dt0 - 2 / cs) <= 1e-13 * (2 / cs)
    # wind along direction k: |U . grad xi_k| = rho |u| 2 / h_k in that direction only
    h = widths(s)
    for k in range(dim):
        vel = [0.0] * dim
        vel[k] = 40.0
        dt = O.euler_cfl(s, uniform_state(s, rho, theta, vel))
        want = 2 / (2 * 40.0 / h[k] + cs)
        assert abs(dt - want) <= 1e-12 * want, (k, dt, want)


@pytest.mark.parametrize("eqn,dim", [(0, 2), (0, 3), (1, 2)])
def test_energy_closed_form(eqn, dim):
    """Sum of MJ q^2 of a uniform state = domain volume x sum of q_s^2."""
    s = small_setup(eqn, dim, 4)
    vals = [0.3, -1.7, 2.5, 4.0, 0.9][:s.nstate]
    Q = s.zeros_state()
    for k, v in enumerate(vals):
        Q[:, k, :] = v
    Q = np.asfortranarray(Q)
    vol = (1.0 if eqn == 1 else 1000.0) ** dim
    want = vol * sum(v * v for v in vals)
    got = O.l2energysquared(s, Q)
    assert abs(got - want) <= 1e-12 * want


@pytest.mark.parametrize("eqn,dim", [(0, 2), (0, 3), (1, 2)])
def test_updatesolution_closed_form(eqn, dim):
    """Q gains rkb dt rhs / MJ (MJ from its own vgeo field, MJI = 1/MJ);
    rhs becomes rka rhs."""
    s = small_setup(eqn, dim, 3)
    rng = np.random.default_rng(4)
    Q = np.asfortranarray(rng.uniform(1.0, 2.0, s.zeros_state().shape))
    rhs = np.asfortranarray(rng.standard_normal(Q.shape))
    imj = 9 if dim == 3 else 4
    MJ = s.vgeo[:, imj, :]
    Qo, ro = Q.copy(order="F"), rhs.copy(order="F")
    O.updatesolution(s, ro, Qo, 0.25, 0.75, 0.5)
    assert np.array_equal(ro, 0.25 * rhs)
    for k in range(s.nstate):
        want = 0.75 * 0.5 * rhs[:, k, :] / MJ
        # rounding Q + dQ costs up to an ulp of Q (Q is in [1, 2]); the
        # increments are 1e-7..1e-2, so a wrong field or factor still shows
        np.testing.assert_allclose(Qo[:, k, :] - Q[:, k, :], want, rtol=1e-9, atol=5e-16)
        assert np.max(np.abs(want)) > 1e-8


def test_swe_advection_freezes_velocity():
    s = small_setup(1, 2, 3)
    rng = np.random.default_rng(5)
    Q = np.asfortranarray(rng.uniform(1.0, 2.0, s.zeros_state().shape))
    rhs = np.asfortranarray(rng.standard_normal(Q.shape))
    rhs[:, 3, :] = 0.0
    Qo, ro = Q.copy(order="F"), rhs.copy(order="F")
    O.updatesolution(s, ro, Qo, 0.5, 0.5, 0.1, advection=True)
    H0, H1 = Q[:, 2, :] + Q[:, 3, :], Qo[:, 2, :] + Qo[:, 3, :]
    assert not np.allclose(H0, H1)  # the depth moved
    for k in (0, 1):  # velocity U / H kept
        np.testing.assert_allclose(Qo[:, k, :] / H1, Q[:, k, :] / H0, rtol=1e-14)
    assert np.array_equal(Qo[:, 3, :], Q[:, 3, :])  # b has no tendency
