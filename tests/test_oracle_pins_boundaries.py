"""Closed-form pins of the oracle's boundary treatment, filters, norm of change and time-step size.

Covers the immersed filter (src/Boundary.cpp:618-652), the ghost-node interpolation and mirror
(:655-883), the inlet's in-plane density filter (:142-245), the periodic density filter and its schedule
(src/Mesh.cpp:85-140), the norm of change (src/Mesh.cpp:171-178) and the step size of a moving fluid
(src/Solver.cpp:80-127). Every expected value follows from the mathematics of the written rule (exactness of
central averages on polynomials, Fourier symbols, closed-form maxima), not from a second transcription.
"""
import numpy as np
import pytest

import cases
import nsfsi
import oracle


def _coords(o, cfg):
    sp, of = o.geometry()
    k, j, i = np.meshgrid(np.arange(cfg.NK), np.arange(cfg.NJ), np.arange(cfg.NI), indexing="ij")
    return [(i - of[0]) * sp[0], (j - of[1]) * sp[1], (k - of[2]) * sp[2]], sp, of


def _set(o, cfg, fields):
    shape = (cfg.NK, cfg.NJ, cfg.NI)
    for q, f in enumerate(fields):
        o.set_field(q, np.broadcast_to(f, shape))
    o.derive_all()


def _sphere(N=24):
    return cases.sphere48(N=N, steps=1)


# ------------------------------------------------------------------ immersed filter

def test_immersed_filter_on_quadratic_field():
    """(6 P + sum of the 6 neighbours)/12 = P + (h_x^2 f_xx + h_y^2 f_yy + h_z^2 f_zz)/12 exactly for a
    quadratic f: applied to rho' and rho E' at every filter node (both read before either is written)."""
    cfg = _sphere()
    o = oracle.Oracle(cfg)
    o.initialize()
    X, h, _ = _coords(o, cfg)
    x, y, z = X
    f = 0.01 + 0.03 * x * x - 0.02 * y * y + 0.05 * z * z + 0.04 * x * y - 0.01 * y * z + 0.02 * x
    g = 2.0 + 0.07 * x * x + 0.01 * y * y - 0.03 * z * z + 0.02 * x * z
    zero = np.zeros_like(x)
    _set(o, cfg, [f, zero, zero, zero, g])
    o.apply_bcs(0.0)
    F = o.filter_nodes()
    assert len(F) > 100
    i, j, k = F[:, 0], F[:, 1], F[:, 2]
    cf = (h[0] ** 2 * 0.06 - h[1] ** 2 * 0.04 + h[2] ** 2 * 0.10) / 12
    cg = (h[0] ** 2 * 0.14 + h[1] ** 2 * 0.02 - h[2] ** 2 * 0.06) / 12
    rho, rE = o.field("rho"), o.field("rho_E")
    assert np.allclose(rho[k, j, i], f[k, j, i] + cf, rtol=0, atol=1e-15)
    assert np.allclose(rE[k, j, i], g[k, j, i] + cg, rtol=1e-15, atol=1e-15)
    # momenta are not filtered
    assert np.all(o.field("rho_u")[k, j, i] == 0.0)


# ------------------------------------------------------------------ ghost interpolation and mirror

def _ghost_geometry(o, cfg):
    ijk, bi, ip, box = o.ghosts()
    nt = o.node_types()
    X, h, of = _coords(o, cfg)
    # corners in the reference's list order: k fastest, i slowest (src/SmallVectors.h:91-103)
    offs = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    corners = np.stack([box + np.array(d) for d in offs], axis=1)  # (n, 8, 3) as (i, j, k)
    ctype = nt[corners[..., 2], corners[..., 1], corners[..., 0]]
    return ijk, bi, ip, corners, ctype, h, of


def _trilinear(c, p):
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    return c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * x * z + c[6] * y * z + c[7] * x * y * z


def test_ghost_values_reproduce_trilinear_fields_and_mirror():
    """Fluid data from span{1,x,y,z,xy,xz,yz,xyz}: where all 8 nodes around the image point are fluid, the
    ghost gets rho'_g = f(IP), m_g = -m(IP), rho E'_g = e(IP) (the z-then-y-then-x lerp is exact on trilinear
    data, src/Boundary.cpp:732-757, mirror :777-785), and its primitives follow from the EOS."""
    cfg = _sphere(32)
    o = oracle.Oracle(cfg)
    o.initialize()
    X, _, _ = _coords(o, cfg)
    P = np.stack(X, axis=-1)
    cr = np.array([0.02, 0.01, -0.015, 0.005, 0.003, -0.002, 0.004, -0.001])
    cm = [np.array([0.1, 0.02, -0.01, 0.03, 0.005, 0.002, -0.004, 0.001]),
          np.array([-0.05, 0.01, 0.02, -0.02, 0.001, 0.003, 0.002, -0.002]),
          np.array([0.03, -0.01, 0.005, 0.01, -0.003, 0.001, 0.001, 0.0005])]
    ce = np.array([2.5, 0.02, 0.01, -0.03, 0.002, 0.001, -0.001, 0.0003])
    fields = [_trilinear(cr, P)] + [_trilinear(c, P) for c in cm] + [_trilinear(ce, P)]
    _set(o, cfg, fields)
    o.apply_bcs(1.0)
    ijk, bi, ip, corners, ctype, h, of = _ghost_geometry(o, cfg)
    allfluid = np.all(ctype <= 2, axis=1)
    assert allfluid.sum() > 50 and (~allfluid).sum() > 10
    gi, gj, gk = ijk[allfluid, 0], ijk[allfluid, 1], ijk[allfluid, 2]
    IP = ip[allfluid]
    got = [o.field(q)[gk, gj, gi] for q in range(5)]
    assert np.allclose(got[0], _trilinear(cr, IP), rtol=0, atol=1e-15)
    for d in range(3):
        assert np.allclose(got[1 + d], -_trilinear(cm[d], IP), rtol=0, atol=1e-15), f"momentum {d}"
    assert np.allclose(got[4], _trilinear(ce, IP), rtol=1e-15, atol=1e-14)
    # primitives at the ghost follow from its conserved values (derivePrimitive, src/Boundary.cpp:447-455)
    u = o.field("u")[gk, gj, gi]
    assert np.allclose(u, got[1] / (1 + got[0]), rtol=1e-15, atol=1e-18)


def test_ghost_values_vandermonde_boxes():
    """Image points with ghost nodes among their 8 neighbours (src/Boundary.cpp:788-849): a ghost corner
    enters at its body intercept, with value 0 in a Dirichlet row (momenta) and a zero normal derivative
    in a Neumann row (rho', rho E'). Constants are reproduced through Neumann rows; the momentum value
    equals the interpolant that numpy's solver gives on the same 8 conditions (to cond * eps)."""
    cfg = _sphere(32)
    o = oracle.Oracle(cfg)
    o.initialize()
    X, _, _ = _coords(o, cfg)
    P = np.stack(X, axis=-1)
    cm = np.array([0.1, 0.02, -0.01, 0.03, 0.005, 0.002, -0.004, 0.001])
    zero = np.zeros_like(X[0])
    _set(o, cfg, [zero + 0.013, _trilinear(cm, P), zero, zero, zero + 2.6])
    o.apply_bcs(1.0)
    ijk, bi, ip, corners, ctype, h, of = _ghost_geometry(o, cfg)
    gmap = {tuple(g): n for n, g in enumerate(ijk)}
    vd = ~np.all(ctype <= 2, axis=1)
    assert vd.sum() > 10
    rho, ru = o.field("rho"), o.field("rho_u")
    worst = 0.0
    for n in np.nonzero(vd)[0]:
        g = ijk[n]
        # Neumann rows reproduce constants
        assert rho[g[2], g[1], g[0]] == pytest.approx(0.013, rel=1e-9)
        rows, data = [], []
        for c in range(8):
            ci = corners[n, c]
            if ctype[n, c] == 4:
                q = bi[gmap[tuple(ci)]]
                data.append(0.0)
            else:
                assert ctype[n, c] <= 2
                q = P[ci[2], ci[1], ci[0]]
                data.append(_trilinear(cm, q))
            rows.append([1, q[0], q[1], q[2], q[0] * q[1], q[0] * q[2], q[1] * q[2], q[0] * q[1] * q[2]])
        V = np.array(rows)
        coef = np.linalg.solve(V, np.array(data))
        p = ip[n]
        val = np.array([1, p[0], p[1], p[2], p[0] * p[1], p[0] * p[2], p[1] * p[2], p[0] * p[1] * p[2]]) @ coef
        tol = np.linalg.cond(V) * 4e-16 * max(1.0, np.abs(coef).max() * 10)
        err = abs(ru[g[2], g[1], g[0]] - (-val))
        worst = max(worst, err / tol)
        assert err <= tol, (n, ru[g[2], g[1], g[0]], -val, tol)
    assert np.allclose(o.field("rho_E")[ijk[vd, 2], ijk[vd, 1], ijk[vd, 0]], 2.6, rtol=1e-9)


# ------------------------------------------------------------------ inlet in-plane filter

@pytest.mark.parametrize("axis", [0, 1, 2])
def test_inlet_density_filter_on_linear_field(axis):
    """rho' = a + b x + c y + d z everywhere. The inlet extrapolates rho'_b = 2 rho'_1 - rho'_2 (exact on a
    linear field), then replaces it by (sum of in-plane neighbours inside the owned box + n rho'_P)/(2n)
    (src/Boundary.cpp:142-245): unchanged at nodes with 4 neighbours, shifted by -/+ g h/6 at an edge of the
    owned box (3 neighbours) and by (-/+ g1 h1 -/+ g2 h2)/4 at its corners (2 neighbours)."""
    # every other face extrapolates linearly (outlet), so the planes the inlet extrapolates from stay linear
    # after the faces applied before it
    bc = ["outlet"] * 6
    bc[2 * axis] = "inlet"
    cfg = nsfsi.make_config(NI=14, NJ=13, NK=12, L_x=1.3, L_y=1.1, L_z=1.0, bc=bc, body_type="none",
                            ic_type="stagnation", M_0=0.2, Re=100.0, filter_intervals=[0], filter_change_times=[],
                            stop_criterion="timesteps")
    o = oracle.Oracle(cfg)
    o.initialize()
    X, h, _ = _coords(o, cfg)
    grad = np.array([0.011, -0.023, 0.017])
    rho = 0.004 + grad[0] * X[0] + grad[1] * X[1] + grad[2] * X[2]
    zero = np.zeros_like(rho)
    _set(o, cfg, [rho, zero, zero, zero, rho / cfg.gamma / (cfg.gamma - 1)])
    o.apply_bcs(3.0)
    got = o.field("rho")
    t1, t2 = [a for a in range(3) if a != axis]
    Ns = [cfg.NI, cfg.NJ, cfg.NK]
    # the inlet owns its plane minus the rims claimed by the faces pushed before it (z, then y, then x)
    lo = [1, 1, 1]
    hi = [Ns[0] - 2, Ns[1] - 2, Ns[2] - 2]
    if axis == 0:
        lo[0] = hi[0] = 0
    elif axis == 1:
        lo[0], hi[0] = 0, Ns[0] - 1
        lo[1] = hi[1] = 0
    else:
        lo[0], hi[0] = 0, Ns[0] - 1
        lo[1], hi[1] = 0, Ns[1] - 1
        lo[2] = hi[2] = 0
    checked = 0
    for a in range(lo[t1], hi[t1] + 1):
        for b in range(lo[t2], hi[t2] + 1):
            idx = [0, 0, 0]
            idx[axis] = 0
            idx[t1], idx[t2] = a, b
            shift, nb = 0.0, 4
            for ax, v in ((t1, a), (t2, b)):
                if v == lo[ax]:
                    shift += grad[ax] * h[ax]
                    nb -= 1
                elif v == hi[ax]:
                    shift -= grad[ax] * h[ax]
                    nb -= 1
            exact = rho[idx[2], idx[1], idx[0]]
            expect = exact + shift / (2 * nb)
            assert got[idx[2], idx[1], idx[0]] == pytest.approx(expect, rel=0, abs=2e-16), (idx, nb)
            checked += 1
    assert checked == (hi[t1] - lo[t1] + 1) * (hi[t2] - lo[t2] + 1)


# ------------------------------------------------------------------ periodic density filter

def _periodic(N=16, intervals=(1,), change_times=()):
    return nsfsi.make_config(NI=N, NJ=N + 2, NK=N + 4, L_x=1.0, L_y=1.3, L_z=0.9, bc=["periodic"] * 6,
                             body_type="none", ic_type="stagnation", M_0=0.2, Re=100.0,
                             filter_intervals=list(intervals), filter_change_times=list(change_times),
                             stop_criterion="timesteps")


def test_periodic_density_filter_fourier_symbol():
    """rho' = A cos(k.x + phi), periodic on the mesh: the filter (6 P + sum of 6 neighbours)/12 multiplies it
    by (6 + 2 sum_d cos(k_d h_d))/12 at interior nodes (src/Mesh.cpp:107-140) and copies every other node."""
    cfg = _periodic()
    o = oracle.Oracle(cfg)
    o.initialize()
    X, h, _ = _coords(o, cfg)
    L = [cfg.L_x, cfg.L_y, cfg.L_z]
    kv = [2 * np.pi * m / L[d] for d, m in enumerate((2, 3, 1))]
    A, phi = 1e-3, 0.4
    f = A * np.cos(kv[0] * X[0] + kv[1] * X[1] + kv[2] * X[2] + phi)
    zero = np.zeros_like(f)
    _set(o, cfg, [f, zero, zero, zero, zero])
    o.set_time(0.0, 0)
    assert o.rho_filter()
    got = o.field("rho")
    interior = o.node_types() == 0
    sym = (6 + 2 * sum(np.cos(kv[d] * h[d]) for d in range(3))) / 12
    assert 0.5 < sym < 0.99
    assert np.allclose(got[interior], sym * f[interior], rtol=0, atol=1e-17)
    assert np.array_equal(got[~interior], f[~interior])


@pytest.mark.parametrize("t,level,fires", [
    (0.5, 3, False),    # before the first change time: interval 0 = off
    (1.0, 4, False),    # t > change time is strict
    (1.5, 4, True),     # interval 5: (4 + 1) % 5 == 0
    (1.5, 5, False),
    (2.5, 8, True),     # interval 3: (8 + 1) % 3 == 0
    (2.5, 9, False),
    (3.5, 0, False),    # past every change time: more thresholds than intervals -> off
])
def test_filter_schedule(t, level, fires):
    """Mesh::checkFilterCondition (src/Mesh.cpp:85-104): count the change times strictly below t, take that
    entry of filterIntervals if there is one, fire when it is positive and divides timeLevel + 1."""
    cfg = _periodic(N=8, intervals=(0, 5, 3), change_times=(1.0, 2.0, 3.0))
    o = oracle.Oracle(cfg)
    o.initialize()
    o.set_time(t, level)
    assert o.filter_now() == fires
    before = o.field("rho").copy()
    o.set_field("rho", before + 0.01 * np.cos(np.arange(before.size)).reshape(before.shape))
    ref = o.field("rho").copy()
    assert o.rho_filter() == fires
    assert np.array_equal(o.field("rho"), ref) != fires


# ------------------------------------------------------------------ norm of change

def test_norm_of_change_closed_forms():
    """sqrt(dx dy dz * sum over FluidInterior nodes of (old - new)^2) (src/Mesh.cpp:171-178): a change c at
    one interior node gives |c| sqrt(dx dy dz); a change at non-interior nodes gives 0; a uniform change c
    gives |c| sqrt(dx dy dz n_interior). The norm stored by a step is this norm of that step's change."""
    cfg = _sphere(20)
    o = oracle.Oracle(cfg)
    o.initialize()
    (dx, dy, dz), _ = o.geometry()
    nt = o.node_types()
    interior = nt == 0
    nI = int(interior.sum())
    assert nI == o.counts()[2]
    for q in range(5):
        cur = o.field(q)
        d = np.zeros_like(cur)
        d[7, 9, 11] = 0.3
        assert interior[7, 9, 11]
        o.set_old_field(q, cur + d)
        assert o.norm_of_change(q) == pytest.approx(0.3 * np.sqrt(dx * dy * dz), rel=1e-12)
        d = np.where(interior, 0.0, 1.0)
        o.set_old_field(q, cur + d)
        assert o.norm_of_change(q) == 0.0
        o.set_old_field(q, cur - 0.02)
        assert o.norm_of_change(q) == pytest.approx(0.02 * np.sqrt(dx * dy * dz * nI), rel=1e-12)
    o.march(1)
    for q in range(5):
        assert o.norm_history(q)[-1] == o.norm_of_change(q)
        old, new = o.old_field(q), o.field(q)
        ref = np.sqrt(dx * dy * dz * np.sum((old[interior] - new[interior]) ** 2))
        assert o.norm_history(q)[-1] == pytest.approx(ref, rel=1e-12)


# ------------------------------------------------------------------ step size of a moving fluid

def test_dt_of_moving_fluid():
    """dt_conv = CFL / (sx/dx + sy/dy + sz/dz) with s_d the maximum over ALL nodes of c + |u_d| taken per axis
    (src/Solver.cpp:91-110). The state is at rest except u = 0.5 at an interior node, v = -0.7 at a node of
    the x-min face and p' = 0.3 at a third node (c = sqrt(1 + 0.3 gamma)): sx = 1.5, sy = 1.7,
    sz = sqrt(1.42). At Re = 1e4 the viscous limit does not bind; at Re = 1 it does, with nu set by the
    warm node: mu = (1 + T')^1.5 (1 + S)/(Re (T' + 1 + S)), T' = 0.3 gamma, nu = mu max(4/3, gamma/Pr)."""
    for Re in (1e4, 1.0):
        cfg = nsfsi.make_config(NI=12, NJ=11, NK=10, L_x=1.0, L_y=0.8, L_z=0.9, bc=cases.SIDE_WALLS,
                                body_type="none", ic_type="stagnation", M_0=0.2, Re=Re, conv_stability_limit=0.7,
                                filter_intervals=[0], filter_change_times=[], stop_criterion="timesteps")
        o = oracle.Oracle(cfg)
        o.initialize()
        (dx, dy, dz), _ = o.geometry()
        shape = (cfg.NK, cfg.NJ, cfg.NI)
        rho, ru, rv, rw, rE = (np.zeros(shape) for _ in range(5))
        ru[4, 5, 6] = 0.5
        rE[4, 5, 6] = 0.5 * 0.25
        rv[3, 4, 0] = -0.7
        rE[3, 4, 0] = 0.5 * 0.49
        assert o.node_types()[3, 4, 0] == 1
        rE[6, 2, 8] = 0.3 / (cfg.gamma - 1)
        _set(o, cfg, [rho, ru, rv, rw, rE])
        c3 = np.sqrt(1 + cfg.gamma * 0.3)
        dtc = cfg.conv_stability_limit / (1.5 / dx + 1.7 / dy + c3 / dz)
        T = cfg.gamma * 0.3
        S1 = 1 + cfg.sutherlands_C2 / cfg.T_0
        mu = (1 + T) ** 1.5 * S1 / (Re * (T + S1))
        dtv = abs(cfg.visc_stability_limit) / (mu * max(4 / 3, cfg.gamma / cfg.Pr) * (1 / dx**2 + 1 / dy**2 + 1 / dz**2))
        assert (dtc < dtv) == (Re > 1)
        assert o.state()[2] == pytest.approx(min(dtc, dtv), rel=1e-13)
        assert o.update_dt(0.0) == pytest.approx(min(dtc, dtv), rel=1e-13)


def test_dt_end_time_clamp():
    """With the end_time criterion, a step that would pass t_end is shortened to end there
    (src/Solver.cpp:86-87)."""
    cfg = _periodic(N=8, intervals=(0,))
    cfg.stop_criterion = 1
    cfg.t_end = 1.0
    o = oracle.Oracle(cfg)
    o.initialize()
    full = o.update_dt(0.0)
    assert o.update_dt(1.0 - 0.25 * full) == pytest.approx(0.25 * full, rel=1e-12)
    assert o.update_dt(1.0 - 2 * full) == full
