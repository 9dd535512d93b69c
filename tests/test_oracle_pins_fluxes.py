"""Closed-form pins of the oracle's convective, pressure and energy fluxes (src/Solver.cpp:170-418).

Each state below makes every term of one equation a polynomial whose central differences are exact, so the
slope at every interior node is known in closed form. The coefficients are distinct per axis, so a flipped
sign, a dropped term, a transposed axis or a wrong constant in any flux fails at least one check.
"""
import numpy as np
import pytest

import nsfsi
import oracle


def periodic_box(N=10, Re=200.0):
    return nsfsi.make_config(NI=N, NJ=N + 1, NK=N + 2, L_x=1.0, L_y=1.1, L_z=1.2, bc=["periodic"] * 6, body_type="none",
                             ic_type="stagnation", M_0=0.3, Re=Re, conv_stability_limit=0.5,
                             stop_criterion="timesteps", stop_time_level=10, filter_intervals=[0],
                             filter_change_times=[])


def _setup(cfg):
    o = oracle.Oracle(cfg)
    o.initialize()
    sp, of = o.geometry()
    k, j, i = np.meshgrid(np.arange(cfg.NK), np.arange(cfg.NJ), np.arange(cfg.NI), indexing="ij")
    X = [(i - of[0]) * sp[0], (j - of[1]) * sp[1], (k - of[2]) * sp[2]]
    return o, X, sp, o.node_types() == 0


def _set(o, cfg, rho, m, rE):
    shape = (cfg.NK, cfg.NJ, cfg.NI)
    for name, f in zip(["rho", "rho_u", "rho_v", "rho_w", "rho_E"], [rho, m[0], m[1], m[2], rE]):
        o.set_field(name, np.broadcast_to(f, shape))
    o.derive_all()


def test_pressure_gradient_at_rest():
    """Fluid at rest with T' = 0 and a linear pressure p' = a x + b y + c z (so rho' = gamma p'):
    k_rho_d = -dp'/dx_d exactly (the pressure term of each momentum flux, src/Solver.cpp:225-227, 272-274,
    320-323); continuity and energy slopes vanish."""
    cfg = periodic_box()
    o, X, _, m = _setup(cfg)
    a, b, c = 0.031, -0.047, 0.013
    p = 0.002 + a * X[0] + b * X[1] + c * X[2]
    rho = cfg.gamma * p
    zero = np.zeros_like(p)
    _set(o, cfg, rho, [zero, zero, zero], p / (cfg.gamma - 1))
    assert np.abs(o.field("T")).max() < 1e-15
    k = o.slopes()
    assert np.all(k[0][m] == 0.0)
    for d, g in enumerate((a, b, c)):
        assert np.allclose(k[1 + d][m], -g, rtol=1e-10, atol=0), f"pressure term of momentum {d}"
    assert np.abs(k[4][m]).max() < 1e-12


def test_advection_by_uniform_velocity():
    """Uniform velocity U0 = (U, V, W), T' = 0 (p' = rho'/gamma) and a linear rho': with
    s = U rho_x + V rho_y + W rho_z, k_rho = -s, k_rho_u_d = -(U_d s + rho_d/gamma) (the rho u u_d and p
    terms of each momentum flux) and k_rho_E = -s (1/(gamma-1) + |U0|^2/2), the inviscid energy flux
    (rhoH0 + rho E' + p') u (src/Solver.cpp:378-380)."""
    cfg = periodic_box()
    o, X, _, m = _setup(cfg)
    U0 = np.array([0.21, -0.13, 0.08])
    g = np.array([0.017, 0.029, -0.011])
    rho = 0.001 + g[0] * X[0] + g[1] * X[1] + g[2] * X[2]
    p = rho / cfg.gamma
    q2 = float(U0 @ U0)
    _set(o, cfg, rho, [(1 + rho) * U0[d] for d in range(3)], p / (cfg.gamma - 1) + (1 + rho) / 2 * q2)
    k = o.slopes()
    s = float(U0 @ g)
    assert np.allclose(k[0][m], -s, rtol=1e-11, atol=0)
    for d in range(3):
        assert np.allclose(k[1 + d][m], -(U0[d] * s + g[d] / cfg.gamma), rtol=1e-9, atol=1e-15), f"momentum {d}"
    assert np.allclose(k[4][m], -s * (1 / (cfg.gamma - 1) + q2 / 2), rtol=1e-9, atol=0)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_energy_flux_of_linear_velocity(d):
    """u_d = a x_d (one component), rho' = p' = T' = 0, rho E' = u_d^2/2: the inviscid energy flux gives
    -(H0 a + a^3 (3 x_d^2 + h_d^2)/2) with H0 = 1/(gamma-1) (central differences of a cubic) and the viscous
    energy term d/dx_d(4/3 mu u_d du_d/dx_d) gives 4/3 mu a^2 exactly; the momentum flux rho u_d u_d gives
    k_rho_u_d = -2 a^2 x_d and its viscous term 4/3 mu u'' = 0."""
    cfg = periodic_box()
    o, X, h, m = _setup(cfg)
    a = 0.5
    zero = np.zeros_like(X[0])
    mom = [zero, zero, zero]
    mom[d] = a * X[d]
    _set(o, cfg, zero, mom, 0.5 * (mom[d] * mom[d]))
    assert np.all(o.field("p") == 0.0) and np.all(o.field("T") == 0.0)
    k = o.slopes()
    mu = 1.0 / cfg.Re
    H0 = 1.0 / (cfg.gamma - 1)
    x = X[d]
    expect_E = -(H0 * a + a**3 * (3 * x * x + h[d] ** 2) / 2) + 4.0 / 3.0 * mu * a * a
    assert np.allclose(k[4][m], expect_E[m], rtol=1e-11, atol=1e-13)
    assert np.allclose(k[1 + d][m], -2 * a * a * x[m], rtol=1e-10, atol=1e-14)
    assert np.allclose(k[0][m], -a, rtol=1e-12, atol=0)


@pytest.mark.parametrize("d,e", [(0, 1), (0, 2), (1, 2)])
def test_energy_viscous_cross_terms(d, e):
    """u_d = a x_d and u_e = b x_e (rho' = p' = 0): the energy equation's viscous part is
    4/3 mu (a^2 + b^2) - 2/3 mu a b - 2/3 mu b a (the -2/3 mixed terms eA, eB, eE of src/Solver.cpp:345-359
    in both directions), and the inviscid part the exact central difference of (H0 + |u|^2/2) u."""
    cfg = periodic_box()
    o, X, h, m = _setup(cfg)
    a, b = 0.4, -0.3
    zero = np.zeros_like(X[0])
    mom = [zero, zero, zero]
    mom[d] = a * X[d]
    mom[e] = b * X[e]
    _set(o, cfg, zero, mom, 0.5 * (mom[d] * mom[d] + mom[e] * mom[e]))
    assert np.all(o.field("p") == 0.0)
    k = o.slopes()
    mu = 1.0 / cfg.Re
    H0 = 1.0 / (cfg.gamma - 1)
    xd, xe = X[d], X[e]
    inv = -(H0 * (a + b) + a**3 * (3 * xd**2 + h[d] ** 2) / 2 + a * b * b * xe**2 / 2
            + b**3 * (3 * xe**2 + h[e] ** 2) / 2 + b * a * a * xd**2 / 2)
    visc = 4.0 / 3.0 * mu * (a * a + b * b) - 4.0 / 3.0 * mu * a * b
    assert np.allclose(k[4][m], (inv + visc)[m], rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("d,e", [(0, 1), (0, 2), (1, 2)])
def test_energy_shear_terms(d, e):
    """Shear flow u_d = a x_e, u_e = b x_d (divergence free): the viscous energy part is
    mu (a + b)^2 (the mu u_d du_d/dx_e diagonal terms plus the +1 mixed terms eC, eD, eF of
    src/Solver.cpp:345-359), the inviscid part -a b x_d x_e (a + b)."""
    cfg = periodic_box()
    o, X, h, m = _setup(cfg)
    a, b = 0.35, 0.2
    zero = np.zeros_like(X[0])
    mom = [zero, zero, zero]
    mom[d] = a * X[e]
    mom[e] = b * X[d]
    _set(o, cfg, zero, mom, 0.5 * (mom[d] * mom[d] + mom[e] * mom[e]))
    k = o.slopes()
    mu = 1.0 / cfg.Re
    expect = -a * b * X[d] * X[e] * (a + b) + mu * (a + b) ** 2
    assert np.allclose(k[4][m], expect[m], rtol=1e-10, atol=1e-13)
    # momentum d: -d/dx_e(rho u_d u_e) = -a b x_d (the d/dx_d(rho u_d u_d) part is 0: u_d is constant along x_d)
    assert np.allclose(k[1 + d][m], -a * b * X[d][m], rtol=1e-10, atol=1e-14)
