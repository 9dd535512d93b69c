"""Pins of the serial FP64 oracle (oracle/ns_oracle.cpp) against what mathematics fixes.

The reference ships no tests, fixtures or golden vectors, and cannot be built here (SFML, libconfig++ and
Eigen are absent), so every check below comes from a closed form, an invariant, a library routine or a
brute-force restatement of the written rules. Each is chosen so that a plausible transcription mistake
(dropped term, wrong sign, transposed index, wrong constant) fails it.
"""
import numpy as np
import pytest

import cases
import nsfsi
import oracle


def periodic_box(N=12, ic="uniform_oblique", amp=0.0, seed=0, **kw):
    return nsfsi.make_config(NI=N, NJ=N + 1, NK=N + 2, L_x=1.0, L_y=1.1, L_z=1.2, bc=["periodic"] * 6, body_type="none",
                             ic_type=ic, ic_amplitude=amp, ic_seed=seed, M_0=0.3, Re=200.0,
                             conv_stability_limit=0.5, stop_criterion="timesteps", stop_time_level=10,
                             filter_intervals=[0], filter_change_times=[], **kw)


def interior_mask(o):
    return o.node_types() == 0


# ------------------------------------------------------------------ closed forms

def test_uniform_flow_is_a_fixed_point():
    """Uniform flow on a periodic box without a body: every slope is exactly zero, the state is kept bitwise."""
    cfg = periodic_box()
    o = oracle.Oracle(cfg)
    o.initialize()
    U0 = o.conserved()
    assert np.all(o.slopes() == 0.0)
    o.march(5)
    assert np.array_equal(o.conserved(), U0)


def test_viscosity_at_reference_temperature():
    """mu(T'=0) = 1/Re (Sutherland law normalised at T_0), kappa = mu / ((gamma-1) Pr)."""
    cfg = periodic_box()
    o = oracle.Oracle(cfg)
    o.initialize()
    mu, ka = o.field("mu"), o.field("kappa")
    assert np.allclose(mu, 1.0 / cfg.Re, rtol=2e-16, atol=0)
    assert np.allclose(ka, mu / ((cfg.gamma - 1) * cfg.Pr), rtol=2e-16, atol=0)


def test_eos_round_trip():
    """derivePrimitive(deriveConserved(q)) = q to a few ulp on a perturbed state."""
    cfg = periodic_box(ic="freestream_gaussian", amp=0.05, seed=3)
    o = oracle.Oracle(cfg)
    o.initialize()
    u, v, w = o.field("u"), o.field("v"), o.field("w")
    rho, ru, rv, rw, rE = o.conserved()
    assert np.allclose(ru / (1 + rho), u, rtol=1e-15, atol=1e-17)
    assert np.allclose(rv / (1 + rho), v, rtol=1e-15, atol=1e-17)
    p = (cfg.gamma - 1) * (rE - (1 + rho) / 2 * (u * u + v * v + w * w))
    assert np.allclose(p, o.field("p"), rtol=0, atol=1e-16)


def _set_state(o, cfg, rho, ru, rv, rw, rE):
    for name, f in zip(["rho", "rho_u", "rho_v", "rho_w", "rho_E"], [rho, ru, rv, rw, rE]):
        o.set_field(name, np.broadcast_to(f, (cfg.NK, cfg.NJ, cfg.NI)))
    o.derive_all()


def _coords(o, cfg):
    sp, of = o.geometry()
    k, j, i = np.meshgrid(np.arange(cfg.NK), np.arange(cfg.NJ), np.arange(cfg.NI), indexing="ij")
    return (i - of[0]) * sp[0], (j - of[1]) * sp[1], (k - of[2]) * sp[2], sp


def test_continuity_of_linear_momentum():
    """rho*u = a x, rho*v = b y, rho*w = c z gives k_rho = -(a + b + c) at every interior node."""
    cfg = periodic_box(N=10)
    o = oracle.Oracle(cfg)
    o.initialize()
    x, y, z, _ = _coords(o, cfg)
    a, b, c = 0.3, -0.7, 1.9
    zero = np.zeros_like(x)
    _set_state(o, cfg, zero, a * x, b * y, c * z, zero + 0.2)
    k = o.slopes()
    m = interior_mask(o)
    assert np.allclose(k[0][m], -(a + b + c), rtol=1e-12, atol=0)


def test_viscous_term_of_quadratic_velocity():
    """u = c y^2 with v = w = 0, p' = T' = 0 (constant mu): k_rho_u = mu u_yy = 2 c mu, k_rho = 0, k_rho_v = 0."""
    cfg = periodic_box(N=10)
    o = oracle.Oracle(cfg)
    o.initialize()
    x, y, z, _ = _coords(o, cfg)
    cc = 0.4
    u = cc * y * y
    zero = np.zeros_like(x)
    _set_state(o, cfg, zero, u, zero, zero, 0.5 * (u * u))
    assert np.all(o.field("p") == 0.0)
    k = o.slopes()
    m = interior_mask(o)
    mu = 1.0 / cfg.Re
    assert np.allclose(k[1][m], 2 * cc * mu, rtol=1e-9, atol=0)
    assert np.all(k[0][m] == 0.0)
    assert np.allclose(k[2][m], 0.0, atol=1e-15)


def test_cross_viscous_terms_of_bilinear_field():
    """v = a x (others zero, constant mu): the x-momentum cross term d/dy(mu dv/dx) = 0 and the
    y-momentum diagonal term d/dx(mu dv/dx) = 0 (dv/dx constant), so k_rho_u = k_rho_v = 0 up to roundoff,
    while w = b x y gives k_rho_w = 0 and k_rho_u = mu d/dz(dw/dx) - 2/3 mu d/dx(dw/dz) = 0 and
    k_rho_v = mu d/dz(dw/dy) - 2/3 ... = 0; a field with d2/dxdy != 0 (u = x y) gives
    k_rho_v = mu d/dx(du/dy) - 2/3 mu d/dy(du/dx) = mu/3 y-independent of the sign convention."""
    cfg = periodic_box(N=10)
    o = oracle.Oracle(cfg)
    o.initialize()
    x, y, z, _ = _coords(o, cfg)
    zero = np.zeros_like(x)
    c0 = 0.2
    u = c0 * x * y
    _set_state(o, cfg, zero, u, zero, zero, 0.5 * (u * u))
    k = o.slopes()
    m = interior_mask(o)
    mu = 1.0 / cfg.Re
    # y-momentum: d/dx(mu du/dy) [+1] - 2/3 d/dy(mu du/dx) = mu c0 - 2/3 mu c0 = mu c0 / 3, plus the
    # convective -d/dx(rho_v u) - d/dy(rho_v v + p) - ... = 0 (v = 0, p = 0)
    assert np.allclose(k[2][m], mu * c0 / 3, rtol=1e-9)


@pytest.mark.parametrize("vel,axes,target", [
    (1, (0, 1), 0),   # v = c x y  -> k_rho_u = d/dy(mu dv/dx) - 2/3 d/dx(mu dv/dy) = mu c / 3
    (0, (0, 2), 2),   # u = c x z  -> k_rho_w = d/dx(mu du/dz) - 2/3 d/dz(mu du/dx) = mu c / 3
    (2, (1, 2), 1),   # w = c y z  -> k_rho_v = d/dz(mu dw/dy) - 2/3 d/dy(mu dw/dz) = mu c / 3
    (2, (0, 2), 0),   # w = c x z  -> k_rho_u = d/dz(mu dw/dx) - 2/3 d/dx(mu dw/dz) = mu c / 3
    (1, (1, 2), 2),   # v = c y z  -> k_rho_w = d/dy(mu dv/dz) - 2/3 d/dz(mu dv/dy) = mu c / 3
])
def test_cross_viscous_terms(vel, axes, target):
    """Bilinear velocity component (constant mu, rest elsewhere): the mixed-derivative terms of the
    target momentum equation sum to mu c / 3; the convective terms vanish only up to O(c^2) products,
    so c is small and the check is to O(c)."""
    cfg = periodic_box(N=10)
    o = oracle.Oracle(cfg)
    o.initialize()
    x, y, z, _ = _coords(o, cfg)
    X = [x, y, z]
    zero = np.zeros_like(x)
    c0 = 1e-3
    vels = [zero, zero, zero]
    vels[vel] = c0 * X[axes[0]] * X[axes[1]]
    q2 = vels[0] ** 2 + vels[1] ** 2 + vels[2] ** 2
    _set_state(o, cfg, zero, vels[0], vels[1], vels[2], 0.5 * q2)
    k = o.slopes()
    m = interior_mask(o)
    mu = 1.0 / cfg.Re
    assert np.allclose(k[1 + target][m], mu * c0 / 3, rtol=2e-2, atol=0)


def test_heat_flux_of_quadratic_temperature():
    """Fluid at rest, p' = 0, T' = c x^2 (rho' = -T'/(1+T')): k_rho_E = d/dx(kappa dT/dx) ~ 2 c kappa."""
    cfg = periodic_box(N=10)
    o = oracle.Oracle(cfg)
    o.initialize()
    x, y, z, _ = _coords(o, cfg)
    c0 = 1e-4
    T = c0 * x * x
    rho = -T / (1 + T)
    zero = np.zeros_like(x)
    _set_state(o, cfg, rho, zero, zero, zero, zero)
    assert np.allclose(o.field("T"), T, rtol=1e-12, atol=1e-20)
    k = o.slopes()
    m = interior_mask(o)
    kappa = 1.0 / cfg.Re / ((cfg.gamma - 1) * cfg.Pr)
    assert np.allclose(k[4][m], 2 * c0 * kappa, rtol=1e-3)
    assert np.all(k[0][m] == 0.0)


def test_initial_dt_at_rest():
    """Stagnation IC: c = 1, u = 0, mu = 1/Re, so dt_conv = CFL / (1/dx + 1/dy + 1/dz) and
    dt_visc = |sigma| / (max(4/3, gamma/Pr) / Re * (1/dx^2 + 1/dy^2 + 1/dz^2)); dt = min of both."""
    cfg = periodic_box(ic="stagnation")
    o = oracle.Oracle(cfg)
    o.initialize()
    (dx, dy, dz), _ = o.geometry()
    dtc = cfg.conv_stability_limit / (1 / dx + 1 / dy + 1 / dz)
    dtv = abs(cfg.visc_stability_limit) / (max(4 / 3, cfg.gamma / cfg.Pr) / cfg.Re * (1 / dx**2 + 1 / dy**2 + 1 / dz**2))
    assert o.state()[2] == pytest.approx(min(dtc, dtv), rel=1e-14)


@pytest.mark.parametrize("t,expected", [(0.0, 0.0), (5.0, 0.5), (10.0, 1.0), (20.0, 1.0)])
def test_inlet_ramp(t, expected):
    """Inlet velocity min(1, t/10) * M_0, T' = 0, p' = rho'/gamma (src/Boundary.cpp:267-295)."""
    cfg = cases.sphere48(N=16, steps=1)
    o = oracle.Oracle(cfg)
    o.initialize()
    o.apply_bcs(t)
    u = o.field("u")
    assert np.allclose(u[1:-1, 1:-1, 0], expected * cfg.M_0, rtol=1e-15, atol=0)
    assert np.all(o.field("T")[1:-1, 1:-1, 0] == 0.0)
    rho = o.field("rho")[1:-1, 1:-1, 0]
    assert np.allclose(o.field("p")[1:-1, 1:-1, 0], rho / cfg.gamma, rtol=1e-15, atol=1e-18)


def test_symmetry_and_outlet_boundaries():
    """Symmetry: primitives mirrored from two nodes in with the normal velocity negated; outlet: linear
    extrapolation of rho and momenta, p' = 0."""
    cfg = cases.sphere48(N=16, steps=1)
    o = oracle.Oracle(cfg)
    o.initialize()
    o.march(1)
    u, v, w, p = o.field("u"), o.field("v"), o.field("w"), o.field("p")
    # y-min plane (j = 0) for interior i, k: mirror of j = 2
    assert np.array_equal(v[1:-1, 0, 1:-1], -v[1:-1, 2, 1:-1])
    assert np.array_equal(u[1:-1, 0, 1:-1], u[1:-1, 2, 1:-1])
    assert np.array_equal(p[1:-1, 0, 1:-1], p[1:-1, 2, 1:-1])
    # z-max plane mirrors k = NK-3 with w negated
    assert np.array_equal(w[-1, :, :], -w[-3, :, :])
    rho = o.field("rho")
    assert np.array_equal(rho[1:-1, 1:-1, -1], 2 * rho[1:-1, 1:-1, -2] - rho[1:-1, 1:-1, -3])
    assert np.all(p[1:-1, 1:-1, -1] == 0.0)


# ------------------------------------------------------------------ invariants

def test_conservation_on_periodic_box():
    """Fully periodic box without a body: all fluxes are in telescoping form, so the interior sum of every
    conserved variable is constant to roundoff."""
    cfg = periodic_box(N=12, ic="freestream_gaussian", amp=0.05, seed=11)
    o = oracle.Oracle(cfg)
    o.initialize()
    # the reference applies no boundary condition before step 0 (src/Solver.cpp:22-28): make the seeded
    # ghost planes periodic copies first, else stage 1 of step 0 sees a non-periodic field
    o.apply_bcs(0.0)
    m = interior_mask(o)
    U0 = o.conserved()
    s0 = np.array([U0[q][m].sum() for q in range(5)])
    scale = np.array([np.abs(U0[q][m]).sum() for q in range(5)]) + 1.0
    o.march(10)
    U1 = o.conserved()
    s1 = np.array([U1[q][m].sum() for q in range(5)])
    assert np.all(np.abs(s1 - s0) <= 1e-12 * scale)
    assert np.abs(U1 - U0).max() > 1e-6  # the state did evolve


def test_mirror_symmetry_about_body_midplanes():
    """Stagnation IC with a sphere centred in y and z: rho, u, E even in y and z; v odd in y; w odd in z."""
    cfg = nsfsi.make_config(NI=20, NJ=21, NK=21, L_x=2.0, L_y=2.0, L_z=2.0, bc=cases.SIDE_WALLS, body_type="sphere",
                            body_centre=(0.83, 1.0, 1.0), body_radius=0.37, ic_type="stagnation", M_0=0.3, Re=300.0,
                            conv_stability_limit=0.5, stop_criterion="timesteps", stop_time_level=8,
                            filter_intervals=[0], filter_change_times=[])
    o = oracle.Oracle(cfg)
    o.initialize()
    nt = o.node_types()
    assert np.array_equal(nt, nt[:, ::-1, :]) and np.array_equal(nt, nt[::-1, :, :])
    o.march(8)
    rho, ru, rv, rw, rE = o.conserved()
    sc = np.abs(ru).max()
    assert sc > 1e-3
    for f in (rho, ru, rE):
        assert np.allclose(f, f[:, ::-1, :], atol=1e-13 * sc)
        assert np.allclose(f, f[::-1, :, :], atol=1e-13 * sc)
    assert np.allclose(rv, -rv[:, ::-1, :], atol=1e-13 * sc)
    assert np.allclose(rw, -rw[::-1, :, :], atol=1e-13 * sc)
    assert np.allclose(rv, rv[::-1, :, :], atol=1e-13 * sc)


def test_spatial_order_two():
    """One slope evaluation of a smooth periodic field: the error of k_rho against the exact divergence
    falls by ~4 when the grid is refined by 2."""
    errs = []
    for N in (16, 32):
        cfg = nsfsi.make_config(NI=N + 2, NJ=N + 2, NK=N + 2, L_x=1.0, L_y=1.0, L_z=1.0, bc=["periodic"] * 6,
                                body_type="none", ic_type="stagnation", M_0=0.1, Re=100.0,
                                filter_intervals=[0], filter_change_times=[], stop_criterion="timesteps")
        o = oracle.Oracle(cfg)
        o.initialize()
        x, y, z, _ = _coords(o, cfg)
        tp = 2 * np.pi
        ru = 0.1 * np.sin(tp * x)
        rv = 0.1 * np.cos(tp * y)
        rw = 0.1 * np.sin(tp * z)
        zero = np.zeros_like(x)
        _set_state(o, cfg, zero, ru, rv, rw, zero + 0.01)
        k = o.slopes()
        exact = -0.1 * tp * (np.cos(tp * x) - np.sin(tp * y) + np.cos(tp * z))
        m = interior_mask(o)
        errs.append(np.abs(k[0][m] - exact[m]).max())
    assert 3.6 < errs[0] / errs[1] < 4.4


# ------------------------------------------------------------------ textbook routines

def _vandermonde(X, Y, Z):
    return np.stack([np.ones_like(X), X, Y, Z, X * Y, X * Z, Y * Z, X * Y * Z], axis=1)


def test_lu8_against_numpy_random():
    rng = np.random.default_rng(5)
    for _ in range(50):
        A = rng.normal(size=(8, 8))
        b = rng.normal(size=8)
        x = oracle.lu8_solve(A, b)
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_lu8_on_global_coordinate_vandermonde():
    """The ill-conditioned matrices of the ghost interpolation (global coordinates of a far-from-origin
    box): LU agrees with numpy to cond * eps, and the interpolant reproduces every f in
    span{1, x, y, z, xy, xz, yz, xyz} at an interior point."""
    rng = np.random.default_rng(6)
    h = 0.0157
    for _ in range(30):
        base = rng.uniform(2.0, 6.0, size=3)
        X = base[0] + h * np.array([0, 0, 0, 0, 1, 1, 1, 1]) + rng.uniform(-0.3, 0.3, 8) * h
        Y = base[1] + h * np.array([0, 0, 1, 1, 0, 0, 1, 1]) + rng.uniform(-0.3, 0.3, 8) * h
        Z = base[2] + h * np.array([0, 1, 0, 1, 0, 1, 0, 1]) + rng.uniform(-0.3, 0.3, 8) * h
        V = _vandermonde(X, Y, Z)
        coef = rng.normal(size=8)
        f = V @ coef
        x = oracle.lu8_solve(V, f)
        cond = np.linalg.cond(V)
        xn = np.linalg.solve(V, f)
        assert np.allclose(x, xn, rtol=cond * 1e-14, atol=cond * 1e-14 * np.abs(xn).max())
        p = base + h * rng.uniform(0.2, 0.8, 3)
        b = np.array([1, p[0], p[1], p[2], p[0] * p[1], p[0] * p[2], p[1] * p[2], p[0] * p[1] * p[2]])
        assert abs(b @ x - b @ coef) <= 1e-9 * max(1.0, np.abs(f).max())


def test_lu8_neumann_rows_reproduce_constants():
    """Rows replaced by normal-derivative rows with zero data: a constant is still reproduced exactly."""
    rng = np.random.default_rng(7)
    h = 0.05
    for _ in range(20):
        base = rng.uniform(1.0, 3.0, size=3)
        X = base[0] + h * np.array([0, 0, 0, 0, 1, 1, 1, 1])
        Y = base[1] + h * np.array([0, 0, 1, 1, 0, 0, 1, 1])
        Z = base[2] + h * np.array([0, 1, 0, 1, 0, 1, 0, 1])
        V = _vandermonde(X, Y, Z)
        f = np.full(8, 0.37)
        for r in rng.choice(8, size=rng.integers(1, 4), replace=False):
            nv = rng.normal(size=3)
            nv /= np.linalg.norm(nv)
            V[r] = [0, nv[0], nv[1], nv[2], X[r] * nv[1] + Y[r] * nv[0], X[r] * nv[2] + Z[r] * nv[0],
                    Y[r] * nv[2] + Z[r] * nv[1], X[r] * Y[r] * nv[2] + X[r] * Z[r] * nv[1] + Y[r] * Z[r] * nv[0]]
            f[r] = 0.0
        c = oracle.lu8_solve(V, f)
        p = base + 0.5 * h
        b = np.array([1, p[0], p[1], p[2], p[0] * p[1], p[0] * p[2], p[1] * p[2], p[0] * p[1] * p[2]])
        assert b @ c == pytest.approx(0.37, rel=1e-8)


# ------------------------------------------------------------------ brute-force classification

def brute_force_classes(cfg):
    """Node classes, ghost set and filter set by direct enumeration of the written rules
    (src/Mesh.cpp:25-82, src/Boundary.cpp:18-71, 461-615, 887-1103), in numpy, arrays [i, j, k]."""
    NI, NJ, NK = cfg.NI, cfg.NJ, cfg.NK
    L = [cfg.L_x, cfg.L_y, cfg.L_z]
    Ns = [NI, NJ, NK]
    h, off = [], []
    for a in range(3):
        tmin, tmax = cfg.bc[2 * a], cfg.bc[2 * a + 1]
        gmin = 1 if tmin in (2, 3) else 0
        gmax = 1 if tmax == 3 else 0
        h.append(L[a] / (Ns[a] - 1 - gmin - gmax))
        off.append(gmin)
    face_type = [1 if cfg.bc[q] in (0, 1) else 2 for q in range(6)]
    t = np.zeros((NI, NJ, NK), dtype=np.int64)
    # claims in push order; a node keeps the first claim
    claimed = np.zeros_like(t, dtype=bool)
    for a, side in [(2, 0), (2, 1), (1, 0), (1, 1), (0, 0), (0, 1)]:
        sl = [slice(None)] * 3
        sl[a] = 0 if side == 0 else Ns[a] - 1
        sl = tuple(sl)
        sel = ~claimed[sl]
        t[sl][sel] = face_type[2 * a + side]
        claimed[sl] = True
    ii, jj, kk = np.meshgrid(np.arange(NI), np.arange(NJ), np.arange(NK), indexing="ij")
    X = (ii - off[0]) * h[0]
    Y = (jj - off[1]) * h[1]
    Z = (kk - off[2]) * h[2]
    mpb = np.finfo(float).eps * (cfg.L_x + cfg.L_y + cfg.L_z)
    bodies = [(cfg.body_type, cfg.body_axis, list(cfg.body_centre), cfg.body_radius)]
    for q in range(cfg.n_extra_bodies):
        e = cfg.extra_bodies[q]
        bodies.append((e.type, e.axis, list(e.centre), e.radius))
    filt = np.zeros_like(t, dtype=bool)
    for btype, bax, c, r in bodies:  # one body after the other (src/Mesh.cpp:73-74)
        if btype == 2:
            d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
            fw = 2 * max(h)
            axes = [0, 1, 2]
        else:
            comps = [(X - c[0]) ** 2, (Y - c[1]) ** 2, (Z - c[2]) ** 2]
            d = np.sqrt(sum(comps[q] for q in range(3) if q != bax))
            fw = 2 * max(h[q] for q in range(3) if q != bax)
            axes = [q for q in range(3) if q != bax]
        inbox = np.ones_like(t, dtype=bool)
        for a in axes:
            rad = int(np.ceil(r / h[a])) + 3
            ci = int(np.round(c[a] / h[a]))
            idx = [ii, jj, kk][a]
            inbox &= (idx >= ci - rad) & (idx <= ci + rad)
        solid = inbox & (d < r - mpb)
        filt |= inbox & ~solid & (d < r + fw) & (t == 0)
        t[solid] = 3
        # ghost rule: 18 neighbours
        fluid = np.pad(t == 0, 1)
        has = np.zeros_like(solid)
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    nz = (di != 0) + (dj != 0) + (dk != 0)
                    if nz in (0, 3):
                        continue
                    has |= fluid[1 + di:1 + di + NI, 1 + dj:1 + dj + NJ, 1 + dk:1 + dk + NK]
        mine = solid & has  # this body's ghosts
        t[mine] = 4
        # iterative augmentation: every SolidInactive node of an image-point box of this body's ghosts becomes
        # one of its ghosts
        cvec = np.array(c)
        while True:
            g = np.argwhere(mine)
            pos = (g - np.array(off)) * np.array(h)
            c2g = pos - cvec
            if btype == 1:
                c2g[:, bax] = 0
            ln = np.sqrt((c2g ** 2).sum(axis=1))
            probe = c2g * ((r - ln) / ln)[:, None]
            ip = pos + 2 * probe
            nxt = np.ceil(ip / np.array(h)) + np.array(off)
            nxt = np.clip(nxt, 1, np.array(Ns)).astype(int)
            new = 0
            for q in range(len(g)):
                for a in (0, 1):
                    for b in (0, 1):
                        for e in (0, 1):
                            n = (nxt[q, 0] - 1 + a, nxt[q, 1] - 1 + b, nxt[q, 2] - 1 + e)
                            if t[n] == 3:
                                t[n] = 4
                                mine[n] = True
                                new += 1
            if new == 0:
                break
    return t, filt


@pytest.mark.parametrize("name", ["sphere", "cylinder_z", "cylinder_x", "offcentre", "two_bodies", "three_bodies"])
def test_classification_brute_force(name):
    if name == "sphere":
        cfg = cases.sphere48(N=24, steps=1)
    elif name == "cylinder_z":
        cfg = cases.cylinder256(N=64, steps=1)
    elif name == "cylinder_x":
        cfg = nsfsi.make_config(NI=30, NJ=26, NK=28, L_x=3.0, L_y=2.0, L_z=2.2, bc=cases.SIDE_WALLS,
                                body_type="cylinder", body_axis="x", body_centre=(0.0, 1.03, 1.11), body_radius=0.41,
                                filter_intervals=[0], filter_change_times=[])
    elif name == "two_bodies":
        cfg = cases.two_bodies()
    elif name == "three_bodies":
        cfg = cases.three_bodies()
    else:
        cfg = nsfsi.make_config(NI=33, NJ=29, NK=31, L_x=3.0, L_y=2.5, L_z=2.7, bc=cases.SIDE_WALLS, body_type="sphere",
                                body_centre=(1.31, 1.27, 1.33), body_radius=0.61, filter_intervals=[0],
                                filter_change_times=[])
    o = oracle.Oracle(cfg)
    o.initialize()
    t, filt = brute_force_classes(cfg)
    assert np.array_equal(o.node_types(), t.transpose(2, 1, 0))
    f_ijk = o.filter_nodes()
    got = np.zeros_like(filt)
    got[f_ijk[:, 0], f_ijk[:, 1], f_ijk[:, 2]] = True
    assert np.array_equal(got, filt)
    g = o.counts()[0]
    assert g == int((t == 4).sum())


def test_two_bodies_overlay_single_body_runs():
    """With disjoint boxes, the two-body classes are the overlay of each body classified alone, the ghost and
    filter lists are the single-body lists one after the other, in body order."""
    cfg = cases.two_bodies()
    o = oracle.Oracle(cfg)
    o.initialize()
    first = nsfsi.make_config(**{k: getattr(cfg, k) for k, _ in nsfsi.NSConfig._fields_
                                 if k not in ("extra_bodies", "n_extra_bodies", "pad_bodies")})
    e = cfg.extra_bodies[0]
    second = nsfsi.make_config(**{k: getattr(cfg, k) for k, _ in nsfsi.NSConfig._fields_
                                  if k not in ("extra_bodies", "n_extra_bodies", "pad_bodies", "body_type",
                                               "body_axis", "body_centre", "body_radius")})
    second.body_type, second.body_axis, second.body_radius = e.type, e.axis, e.radius
    for a in range(3):
        second.body_centre[a] = e.centre[a]
    runs = []
    for c in (first, second):
        r = oracle.Oracle(c)
        r.initialize()
        runs.append(r)
    t0, t1 = runs[0].node_types(), runs[1].node_types()
    body0, body1 = t0 >= 3, t1 >= 3
    assert not np.any(body0 & body1)
    expect = np.where(body0, t0, t1)
    assert np.array_equal(o.node_types(), expect)
    ijk, bi, ip, box = o.ghosts()
    n0 = runs[0].counts()[0]
    for mine, part in zip((ijk[:n0], bi[:n0], ip[:n0], box[:n0]), runs[0].ghosts()):
        assert np.array_equal(mine, part)
    for mine, part in zip((ijk[n0:], bi[n0:], ip[n0:], box[n0:]), runs[1].ghosts()):
        assert np.array_equal(mine, part)
    assert np.array_equal(o.filter_nodes(), np.concatenate([runs[0].filter_nodes(), runs[1].filter_nodes()]))


def test_mirror_symmetry_of_two_mirrored_spheres():
    """Two equal spheres mirrored about the y midplane (all coordinates multiples of 1/16, so the mirrored
    distances are exact): classes and the state after 8 steps are mirror images to rounding, v odd. Each body's ghosts
    and filter nodes must be applied with its own geometry for this to hold."""
    cfg = nsfsi.make_config(NI=33, NJ=35, NK=19, L_x=2.0, L_y=2.0, L_z=1.0, bc=cases.SIDE_WALLS, body_type="sphere",
                            body_centre=(0.75, 0.5, 0.5), body_radius=0.22,
                            extra_bodies=[dict(type="sphere", centre=(0.75, 1.5, 0.5), radius=0.22)],
                            ic_type="freestream_uniform", M_0=0.3, Re=300.0, conv_stability_limit=0.5,
                            stop_criterion="timesteps", stop_time_level=8, filter_intervals=[0],
                            filter_change_times=[])
    o = oracle.Oracle(cfg)
    o.initialize()
    nt = o.node_types()
    assert (nt == 4).sum() > 0 and np.array_equal(nt, nt[:, ::-1, :])
    o.march(8)
    rho, ru, rv, rw, rE = o.conserved()
    sc = np.abs(ru).max()
    assert sc > 1e-3
    # rounding of the mirrored stencil sums (the j+1 and j-1 terms swap places) stays near 1e-12 of each field
    for f in (rho, ru, rw, rE):
        assert np.allclose(f, f[:, ::-1, :], atol=1e-11 * np.abs(f).max())
    assert np.allclose(rv, -rv[:, ::-1, :], atol=1e-11 * sc)
    # the second sphere matters: the one-sphere run differs near it
    one = nsfsi.make_config(**{k: getattr(cfg, k) for k, _ in nsfsi.NSConfig._fields_
                               if k not in ("extra_bodies", "n_extra_bodies", "pad_bodies")})
    o1 = oracle.Oracle(one)
    o1.initialize()
    o1.march(8)
    assert np.abs(o1.field(0) - rho).max() > 1e-4


def test_sphere_solid_volume():
    """Solid node count (ghost + inactive) approximates (4/3) pi r^3 / (dx dy dz)."""
    cfg = cases.sphere48(N=48, steps=1)
    o = oracle.Oracle(cfg)
    o.initialize()
    nt = o.node_types()
    (dx, dy, dz), _ = o.geometry()
    n = int(((nt == 3) | (nt == 4)).sum())
    expect = 4 / 3 * np.pi * cfg.body_radius ** 3 / (dx * dy * dz)
    assert abs(n - expect) / expect < 0.05


def test_perturbation_generator_statistics():
    """The counter-based uniform generator is uniform on [0,1) (mean 1/2, variance 1/12)."""
    u = np.array([oracle.uniform01(42, c) for c in range(20000)])
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.005


def test_pow15_correctly_rounded():
    """The oracle's correctly rounded x^1.5 (its crPow mode, the one the device's FP64 path uses) equals x*sqrt(x)
    evaluated to 50 digits and rounded once, on 20000 inputs around 1 + T' and some far from 1; the reference's
    std::pow (glibc's pow is within 0.52 ulp) is at most one ulp away."""
    from decimal import Decimal, getcontext

    getcontext().prec = 50
    rng = np.random.default_rng(15)
    xs = np.concatenate([1 + rng.uniform(-0.02, 0.02, 20000), rng.uniform(0.5, 4.0, 2000), [1.0, 4.0, 0.25, 2.0]])
    diff_glibc = 0
    for x in xs:
        d = Decimal(float(x))
        exact = float(d * d.sqrt())  # one rounding of the 50-digit value
        assert oracle.pow15(x, True) == exact, x
        g = oracle.pow15(x, False)
        assert abs(g - exact) <= np.spacing(exact)
        diff_glibc += g != exact
    print(f"std::pow(x, 1.5) differs from the correctly rounded value at {diff_glibc} of {xs.size} inputs")
