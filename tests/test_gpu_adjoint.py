"""Discrete adjoint of the drag objective (ns_adjoint_drag) against central finite differences.

J = mean C_D over time levels n0..n1; the adjoint differentiates the run at its own step sizes, so every
finite-difference evaluation runs at the same prescribed step sequence (ns_set_dt_schedule). Directions
are random over all conserved fields of all nodes of the initial state, over the inlet profile, and Re.
Central differences with eps = 1e-4 (FP64) have a truncation error far below the 1e-6 relative bar
used here at these sizes (the bar the gradients must meet is 1e-4).
"""
import numpy as np
import pytest
import torch

import cases
import nsfsi
import nsfsi_torch as nt

pytestmark = pytest.mark.gpu

EPS = 1e-4
TOL = 1e-6


def small_case(N=24):
    return cases.sphere48(N=N, steps=10)  # FP64, sphere r=0.5 in [-2,2]^3, inlet ramp, symmetry walls


def base(cfg):
    s = nsfsi.Solver(cfg)
    s.initialize()
    U = s.mesh.conserved()
    q = np.ones(s.inletPlaneShape())
    s.close()
    return U, q


def J_at(cfg, U, q, re, nsteps, n0, n1, dts):
    return nt.drag_and_gradients(cfg, U, q, re, nsteps, n0, n1, checkpoint=nsteps, dts=dts)["J"]


def check_u0(nsteps, n0, n1, ckpt, ndir=3, N=24, cfg=None):
    cfg = cfg if cfg is not None else small_case(N)
    U, q = base(cfg)
    print(f"adjoint grid={cfg.NI}x{cfg.NJ}x{cfg.NK} nsteps={nsteps} window={n0}..{n1} checkpoint={ckpt}", flush=True)
    r = nt.drag_and_gradients(cfg, U, q, cfg.Re, nsteps, n0, n1, ckpt)
    assert np.all(np.isfinite(r["grad_u0"]))
    rng = np.random.default_rng(nsteps * 100 + n0)
    worst = 0.0
    for _ in range(ndir):
        d = rng.standard_normal(U.shape)
        d /= np.abs(d).max()
        fd = (J_at(cfg, U + EPS * d, q, cfg.Re, nsteps, n0, n1, r["dts"]) -
              J_at(cfg, U - EPS * d, q, cfg.Re, nsteps, n0, n1, r["dts"])) / (2 * EPS)
        ad = float(np.sum(r["grad_u0"] * d))
        rel = abs(fd - ad) / max(abs(fd), 1e-30)
        print(f"  J={r['J']:.10e} FD={fd:.10e} adjoint={ad:.10e} rel={rel:.2e}", flush=True)
        worst = max(worst, rel)
    assert worst <= TOL, worst
    return r


def test_objective_at_initial_state():
    """n0 = n1 = 0: only the force of the initial state (the transposed face fluxes)."""
    check_u0(1, 0, 0, 1)


def test_one_step():
    check_u0(1, 1, 1, 1)


def test_several_steps_two_segments():
    check_u0(7, 3, 7, 4)


def test_other_face_axes():
    """Inflow along y (inlet y-min, outlet y-max, symmetry on x and z) and inflow along z with periodic x:
    the transposed edge faces of the other axes and the periodic copy."""
    for bc, dims in ((["symmetry", "symmetry", "inlet", "outlet", "symmetry", "symmetry"], (20, 24, 18)),
                     (["periodic", "periodic", "symmetry", "symmetry", "inlet", "outlet"], (18, 20, 24))):
        ni, nj, nk = dims
        cfg = nsfsi.make_config(
            NI=ni, NJ=nj, NK=nk, L_x=2.0, L_y=2.2, L_z=2.1, bc=bc, body_type="sphere", body_centre=(1.0, 1.1, 1.05),
            body_radius=0.45, stop_criterion="timesteps", stop_time_level=6, filter_intervals=[0],
            filter_change_times=[], ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=19, M_0=0.3, Re=200.0,
            conv_stability_limit=0.5, inlet_ramp_time=0.02)
        check_u0(6, 2, 6, 3, cfg=cfg)


def test_inflow_gradient():
    """A short inlet ramp and enough steps for the inlet to reach the body; a coherent profile direction
    (the profile scales the inlet velocity, so a random-sign direction mostly cancels)."""
    cfg = small_case()
    cfg.inlet_ramp_time = 0.05
    U, q = base(cfg)
    nsteps, n0, n1 = 24, 16, 24
    r = nt.drag_and_gradients(cfg, U, q, cfg.Re, nsteps, n0, n1, 8)
    rng = np.random.default_rng(7)
    dq = 1.0 + 0.5 * rng.standard_normal(q.shape)
    dq /= np.abs(dq).max()
    fd = (J_at(cfg, U, q + EPS * dq, cfg.Re, nsteps, n0, n1, r["dts"]) -
          J_at(cfg, U, q - EPS * dq, cfg.Re, nsteps, n0, n1, r["dts"])) / (2 * EPS)
    ad = float(np.sum(r["grad_inflow"] * dq))
    print(f"inflow J={r['J']:.10e} FD={fd:.10e} adjoint={ad:.10e}", flush=True)
    assert abs(fd - ad) <= TOL * abs(fd)


def test_reynolds_gradient():
    cfg = small_case()
    U, q = base(cfg)
    nsteps, n0, n1 = 8, 4, 8
    r = nt.drag_and_gradients(cfg, U, q, cfg.Re, nsteps, n0, n1, 3)
    h = EPS * cfg.Re
    fd = (J_at(cfg, U, q, cfg.Re + h, nsteps, n0, n1, r["dts"]) -
          J_at(cfg, U, q, cfg.Re - h, nsteps, n0, n1, r["dts"])) / (2 * h)
    print(f"Re FD={fd:.10e} adjoint={r['grad_re']:.10e}", flush=True)
    assert abs(fd - r["grad_re"]) <= TOL * abs(fd)


def test_torch_function_backward():
    cfg = small_case()
    U, q = base(cfg)
    u0 = torch.tensor(U, dtype=torch.float64, requires_grad=True)
    qq = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    re = torch.tensor(cfg.Re, dtype=torch.float64, requires_grad=True)
    J = nt.DragObjective.apply(u0, qq, re, cfg, 4, 2, 4, 2)
    (2.0 * J).backward()
    r = nt.drag_and_gradients(cfg, U, q, cfg.Re, 4, 2, 4, 2)
    assert J.item() == r["J"]
    assert np.allclose(u0.grad.numpy(), 2.0 * r["grad_u0"], rtol=1e-12, atol=0)
    assert np.allclose(qq.grad.numpy(), 2.0 * r["grad_inflow"], rtol=1e-12, atol=0)
    assert re.grad.item() == pytest.approx(2.0 * r["grad_re"], rel=1e-12)


def test_adjoint_leaves_final_state():
    """After the run the handle holds the forward run's final state (as ns_step would)."""
    cfg = small_case()
    s = nsfsi.Solver(cfg)
    s.initialize()
    r = s.adjointDrag(5, 3, 5, 2)
    U = s.mesh.conserved()
    assert s.time()[1] == 5
    s.close()
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setDtSchedule(r["dts"])
    s.marchTimeStep(5)
    assert np.array_equal(s.mesh.conserved(), U)
    s.close()


def test_fourth_order_differences():
    """Fourth-order central differences (eps = 1e-3: truncation O(eps^4), roundoff ~1e-13 / eps) pin the
    adjoint to ~1e-9 relative, three orders below the second-order bar: the reverse sweep is the exact
    transpose of the forward step, not an approximation of it (measured 4e-11 and 6e-10)."""
    cfg = small_case()
    U, q = base(cfg)
    nsteps, n0, n1 = 7, 3, 7
    r = nt.drag_and_gradients(cfg, U, q, cfg.Re, nsteps, n0, n1, 4)
    rng = np.random.default_rng(3)
    eps = 1e-3
    for _ in range(3):
        d = rng.standard_normal(U.shape)
        d /= np.abs(d).max()
        J = lambda a: J_at(cfg, U + a * eps * d, q, cfg.Re, nsteps, n0, n1, r["dts"])
        fd = (8 * (J(1) - J(-1)) - (J(2) - J(-2))) / (12 * eps)
        ad = float(np.sum(r["grad_u0"] * d))
        rel = abs(fd - ad) / abs(ad)
        print(f"4th-order FD={fd:.14e} adjoint={ad:.14e} rel={rel:.2e}", flush=True)
        assert rel <= 5e-9


def test_fp32_forward_adjoint():
    """FP32 forward (compensated combine) with the FP64 reverse sweep, at the FP64 run's step sizes: the
    gradients are those of the FP64 map linearised along the FP32 trajectory, so they agree with central
    differences of the FP64 run and with the FP64 adjoint to the FP32 rounding of the states they are
    evaluated at (bars 1e-4; measured <= 2.5e-8 against the differences, 1.7e-8 against the FP64 adjoint). Recomputation from checkpoints restores the FP32 state,
    tables and carry exactly: every checkpoint interval gives the same J and gradients bit for bit."""
    cfg = small_case()
    c32 = cases.sphere48(N=24, steps=10, precision="fp32")
    c32.kahan = 1
    U, q = base(cfg)
    nsteps, n0, n1 = 12, 6, 12
    r64 = nt.drag_and_gradients(cfg, U, q, cfg.Re, nsteps, n0, n1, 5)
    r32 = nt.drag_and_gradients(c32, U, q, cfg.Re, nsteps, n0, n1, 5, dts=r64["dts"])
    assert np.array_equal(r32["dts"], r64["dts"])
    for ck in (1, nsteps):
        o = nt.drag_and_gradients(c32, U, q, cfg.Re, nsteps, n0, n1, ck, dts=r64["dts"])
        print(f"  checkpoint {ck}: J {o['J']:.15e}", flush=True)
        assert o["J"] == r32["J"] and o["grad_re"] == r32["grad_re"], ck
        assert np.array_equal(o["grad_u0"], r32["grad_u0"]), ck
        assert np.array_equal(o["grad_inflow"], r32["grad_inflow"]), ck
    relJ = abs(r32["J"] - r64["J"]) / abs(r64["J"])
    g = np.linalg.norm(r32["grad_u0"] - r64["grad_u0"]) / np.linalg.norm(r64["grad_u0"])
    gre = abs(r32["grad_re"] - r64["grad_re"]) / abs(r64["grad_re"])
    print(f"fp32 forward: J rel {relJ:.2e}, grad_u0 rel L2 {g:.2e}, grad_re rel {gre:.2e}", flush=True)
    assert relJ <= 1e-5 and g <= 1e-4 and gre <= 1e-4
    rng = np.random.default_rng(32)
    for _ in range(3):
        d = rng.standard_normal(U.shape)
        d /= np.abs(d).max()
        fd = (J_at(cfg, U + EPS * d, q, cfg.Re, nsteps, n0, n1, r64["dts"]) -
              J_at(cfg, U - EPS * d, q, cfg.Re, nsteps, n0, n1, r64["dts"])) / (2 * EPS)
        ad = float(np.sum(r32["grad_u0"] * d))
        rel = abs(fd - ad) / abs(fd)
        print(f"  FP64 FD={fd:.10e} FP32-forward adjoint={ad:.10e} rel={rel:.2e}", flush=True)
        assert rel <= 1e-4


def test_gradients_are_reproducible():
    """Every transposed boundary piece sums in a fixed order (no atomic additions), so the gradients of a
    run repeat bit for bit, and do not depend on the checkpoint interval (recomputation is exact): inlet
    ramp reaching the body, inlet profile and Re gradients included."""
    cfg = small_case()
    cfg.inlet_ramp_time = 0.05
    U, q = base(cfg)
    nsteps, n0, n1 = 10, 5, 10
    runs = [nt.drag_and_gradients(cfg, U, q, cfg.Re, nsteps, n0, n1, ck) for ck in (4, 4, 1, nsteps)]
    for r in runs[1:]:
        assert r["J"] == runs[0]["J"] and r["grad_re"] == runs[0]["grad_re"]
        assert np.array_equal(r["grad_u0"], runs[0]["grad_u0"])
        assert np.array_equal(r["grad_inflow"], runs[0]["grad_inflow"])


def periodic_sphere_adjoint():
    """A sphere in a box periodic in every direction (24x22x20, 6 steps, J over levels 3..6): the case, its
    initial state and the device adjoint's results."""
    cfg = nsfsi.make_config(
        NI=24, NJ=22, NK=20, L_x=2.0, L_y=2.0, L_z=2.0, bc=["periodic"] * 6, body_type="sphere",
        body_centre=(1.03, 0.98, 1.01), body_radius=0.42, ic_type="freestream_gaussian", ic_amplitude=1e-3,
        ic_seed=5, M_0=0.3, Re=300.0, stop_criterion="timesteps", stop_time_level=6, filter_intervals=[0],
        filter_change_times=[], conv_stability_limit=0.5)
    nsteps, n0, n1 = 6, 3, 6
    s = nsfsi.Solver(cfg)
    s.initialize()
    U = s.mesh.conserved()
    s.close()
    r = nt.drag_and_gradients(cfg, U, None, cfg.Re, nsteps, n0, n1, 3)
    return cfg, U, r, (nsteps, n0, n1)


def test_gradient_against_differences_of_the_oracle():
    """The adjoint gradient against central differences of the serial FP64 oracle, not of the device: in a box
    periodic in every direction the drag is minus the node volume times the sum of the momentum right-hand
    sides over the FluidInterior nodes (the reference's own per-node expression, src/Solver.cpp:170-418), so
    J(U0) = mean C_D over levels n0..n1 is computed from the oracle's state after each of its steps, at the
    device run's step sizes (prescribed in both). J agrees with the device's to 1e-10, and the adjoint's
    directional derivatives with the oracle's differences to 1e-6 on three random directions."""
    import oracle

    cfg, U, r, (nsteps, n0, n1) = periodic_sphere_adjoint()
    area = np.pi * cfg.body_radius ** 2

    def J_oracle(U0):
        o = oracle.Oracle(cfg)
        o.initialize()
        for q in range(5):
            o.set_field(q, U0[q])
        o.derive_all()
        o.set_dt_schedule(r["dts"])
        (sp, _) = o.geometry()
        cd = []
        for level in range(1, nsteps + 1):
            o.march(1)
            if level >= n0:
                fx = -sp.prod() * o.slopes()[1].sum()
                cd.append(fx / (0.5 * cfg.M_0 ** 2 * area))
        return float(np.mean(cd))

    j0 = J_oracle(U)
    print(f"J device {r['J']:.15e} oracle {j0:.15e}", flush=True)
    assert abs(j0 - r["J"]) <= 1e-10 * abs(j0)
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(3):
        d = rng.standard_normal(U.shape)
        d /= np.abs(d).max()
        fd = (J_oracle(U + EPS * d) - J_oracle(U - EPS * d)) / (2 * EPS)
        ad = float(np.sum(r["grad_u0"] * d))
        rel = abs(fd - ad) / abs(fd)
        print(f"  oracle FD={fd:.10e} adjoint={ad:.10e} rel={rel:.2e}", flush=True)
        worst = max(worst, rel)
    assert worst <= TOL, worst


def test_gradient_against_the_tangent_linear_oracle():
    """The adjoint gradient against the tangent-linear form of the serial FP64 oracle (the oracle's own step
    on value/tangent pairs, oracle.TangentOracle): on the periodic box above, the derivative of J along d is
    the mean over levels n0..n1 of minus the node volume times the sum of the tangents of the oracle's
    x-momentum slopes, from one tangent run started at (U0, d) at the device run's step sizes. The values of
    that run give J itself (the plain oracle's, bit for bit); the derivatives equal the adjoint's <grad, d>
    to 1e-10 on three random directions (measured 1.4e-13 to 8.2e-13), with no difference step involved;
    a tangent on the Reynolds number alone gives dJ/dRe against the adjoint's."""
    import oracle

    cfg, U, r, (nsteps, n0, n1) = periodic_sphere_adjoint()
    area = np.pi * cfg.body_radius ** 2
    o = oracle.Oracle(cfg)
    o.initialize()
    vol = o.geometry()[0].prod()
    o.close()
    scale = 0.5 * cfg.M_0 ** 2 * area
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(3):
        d = rng.standard_normal(U.shape)
        d /= np.abs(d).max()
        tl = oracle.TangentOracle(cfg)
        tl.initialize()
        tl.set_state(U, d)
        tl.set_dt_schedule(r["dts"])
        cd, dcd = [], []
        for level in range(1, nsteps + 1):
            tl.march(1)
            if level >= n0:
                K, dK = tl.slopes()
                cd.append(-vol * K[1].sum() / scale)
                dcd.append(-vol * dK[1].sum() / scale)
        tl.close()
        J, dJ = float(np.mean(cd)), float(np.mean(dcd))
        assert abs(J - r["J"]) <= 1e-10 * abs(J)
        ad = float(np.sum(r["grad_u0"] * d))
        rel = abs(dJ - ad) / abs(dJ)
        print(f"  tangent-linear oracle dJ={dJ:.12e} adjoint={ad:.12e} rel={rel:.2e}", flush=True)
        worst = max(worst, rel)
    # the Reynolds number: a tangent on Re alone
    tl = oracle.TangentOracle(cfg)
    tl.initialize()
    tl.set_re_tangent(1.0)
    tl.set_state(U)
    tl.set_dt_schedule(r["dts"])
    dcd = []
    for level in range(1, nsteps + 1):
        tl.march(1)
        if level >= n0:
            dcd.append(-vol * tl.slopes()[1][1].sum() / scale)
    tl.close()
    dJ = float(np.mean(dcd))
    rel = abs(dJ - r["grad_re"]) / abs(dJ)
    print(f"  tangent-linear oracle dJ/dRe={dJ:.12e} adjoint={r['grad_re']:.12e} rel={rel:.2e}", flush=True)
    worst = max(worst, rel)
    assert worst <= 1e-10, worst
