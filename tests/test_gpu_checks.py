"""Error paths and monitors of the device library, and the two workloads the other GPU suites do not run at
their own geometry: the 512^3 sphere (r = 0.6 at (-1, 0, 0) of [-4, 4]^3, FP32) and the 96^3 adjoint.

* a prescribed step-size schedule that does not cover a run is refused before any step runs, by every
  path that steps (ns_step, ns_profile, ns_adjoint_drag);
* ns_profile over more steps than the device history ring holds keeps every norm-of-change entry;
* a non-finite value in the state is reported as NS_ERR_NONFINITE (the reference prints it,
  src/Mesh.h:48-61), with the end-of-step scans inside the stage kernel and as a separate pass, in FP64
  and FP32, fused and unfused;
* ns_totals equals the oracle's sum over FluidInterior nodes times dx dy dz.
"""
import ctypes as C

import numpy as np
import pytest

import cases
import nsfsi
import nsfsi_torch as nt
import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def solver(cfg, **opts):
    print(f"case NI={cfg.NI} NJ={cfg.NJ} NK={cfg.NK} body={cfg.body_type} prec={cfg.precision} opts={opts}", flush=True)
    s = nsfsi.Solver(cfg)
    s.initialize()
    for k, v in opts.items():
        s.setOption(k, v)
    return s


# ------------------------------------------------------------------ step-size schedule guard

def test_short_schedule_refused_by_every_stepping_path():
    cfg = cases.sphere48(N=24, steps=10)
    s = solver(cfg)
    dts = np.full(6, 1e-3)  # covers levels 0..5: at most 5 steps from level 0
    s.setDtSchedule(dts)
    for call in (lambda: s.marchTimeStep(6), lambda: s.profile(6), lambda: s.adjointDrag(6, 1, 6, 3)):
        with pytest.raises(nsfsi.NSError) as e:
            call()
        assert e.value.status == 5  # NS_ERR_ARGUMENT
        assert s.time()[1] == 0  # nothing ran
    s.marchTimeStep(5)
    assert s.time()[1] == 5
    s.close()


def test_adjoint_schedule_of_length_nsteps_is_refused():
    """ns_adjoint_drag needs the step sizes of levels 0..nsteps: a schedule of nsteps entries is one short."""
    cfg = cases.sphere48(N=24, steps=10)
    s = solver(cfg)
    s.setDtSchedule(np.full(4, 2e-3))
    N = s.mesh.nNodesTotal
    gu = np.zeros(5 * N)
    J = C.c_double()
    st = nsfsi.lib().ns_adjoint_drag(s._h, 4, 1, 4, 2, C.byref(J), gu.ctypes.data_as(C.POINTER(C.c_double)), None,
                                     None, None)
    assert st == 5
    assert s.time()[1] == 0
    s.close()


# ------------------------------------------------------------------ profile drains the history ring

def test_profile_longer_than_history_ring_keeps_every_norm():
    # a periodic box (stable over the 1100 steps; the coarse sphere cases are not)
    cfg = nsfsi.make_config(NI=16, NJ=17, NK=18, L_x=1.0, L_y=1.1, L_z=1.2, bc=["periodic"] * 6, body_type="none",
                            ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=7, M_0=0.3, Re=200.0,
                            conv_stability_limit=0.5, stop_criterion="timesteps", stop_time_level=1200,
                            filter_intervals=[0], filter_change_times=[], conv_history="all")
    a = solver(cfg)
    a.marchTimeStep(1100)
    b = solver(cfg)
    ms, ln = b.profile(1100)
    assert ms[0] > 0 and ln[0] == 4
    ha, hb = a.getConvergenceHistory(), b.getConvergenceHistory()
    for k in ha:
        assert len(ha[k]) == len(hb[k]) == 1100, k
        assert np.array_equal(ha[k], hb[k]), k
    assert np.array_equal(a.mesh.conserved(), b.mesh.conserved())
    a.close()
    b.close()


# ------------------------------------------------------------------ non-finite detection

@pytest.mark.parametrize("precision,kernel,fused_scans", [
    ("fp64", 1, 1), ("fp64", 1, 0), ("fp64", 0, 0), ("fp32", 1, 1), ("fp32", 1, 0), ("fp32", 0, 0)])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_nonfinite_state_is_reported(precision, kernel, fused_scans, bad):
    cfg = cases.sphere48(N=24, steps=10, precision=precision)
    s = solver(cfg, stage_kernel=kernel, fused_scans=fused_scans)
    U = s.mesh.conserved()
    nt_ = s.mesh.nodeType
    k, j, i = 12, 11, 4
    assert nt_[k, j, i] == 0
    s.marchTimeStep(2)  # finite steps are fine
    U = s.mesh.conserved()
    U[1, k, j, i] = bad
    s.setInitialField(U.reshape(-1))
    with pytest.raises(nsfsi.NSError) as e:
        s.marchTimeStep(1)
    assert e.value.status == 3  # NS_ERR_NONFINITE
    s.close()
    # the oracle reports the same state as non-finite at the next step
    o = oracle.Oracle(cases.sphere48(N=24, steps=10))
    o.initialize()
    o.set_field(1, U[1])
    with pytest.raises(oracle.OracleError) as eo:
        o.march(1)
    assert eo.value.status == 3


# ------------------------------------------------------------------ conservation totals

@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_totals_equal_oracle_interior_sums(precision):
    cfg = cases.sphere48(N=24, steps=12, precision=precision)
    s = solver(cfg)
    s.marchTimeStep(12)
    tot = s.mesh.totals()
    U = s.mesh.conserved()
    interior = s.mesh.nodeType == 0
    vol = s.mesh.dx * s.mesh.dy * s.mesh.dz
    own = np.array([U[q][interior].sum() * vol for q in range(5)])
    s.close()
    o = oracle.Oracle(cases.sphere48(N=24, steps=12))
    o.initialize()
    o.march(12)
    V = o.conserved()
    (dx, dy, dz), _ = o.geometry()
    ref = np.array([V[q][o.node_types() == 0].sum() * dx * dy * dz for q in range(5)])
    scale = np.array([np.abs(V[q][o.node_types() == 0]).sum() * dx * dy * dz for q in range(5)])
    print(f"{precision} totals {tot} oracle {ref}")
    # the device sum of its own state, in any order
    assert np.all(np.abs(tot - own) <= 1e-13 * scale)
    tol = 1e-13 if precision == "fp64" else 1e-6
    assert np.all(np.abs(tot - ref) <= tol * scale), (tot, ref)


# ------------------------------------------------------------------ the 512^3 sphere workload

def test_sphere512_geometry_fp64_against_oracle_at_96():
    """The 512^3 workload's geometry and physics (sphere r = 0.6 at (-1,0,0) of [-4,4]^3, Re_D = 300,
    M = 0.3, seed 3) at 96^3 in FP64: classes bit-exact, 10 steps within max(1e-12, 4x the oracle's
    one-ulp sensitivity) per field."""
    from test_gpu_parity import ulp_sensitivity

    cfg = cases.sphere512(N=96, steps=10, precision="fp64")
    s = solver(cfg)
    o = oracle.Oracle(cfg)
    o.initialize()
    assert np.array_equal(s.mesh.nodeType, o.node_types())
    s.marchTimeStep(10)
    o.march(10)
    U, V = s.mesh.conserved(), o.conserved()
    l2 = np.array([rel_l2(U[q], V[q]) for q in range(5)])
    floor = np.array(ulp_sensitivity(cfg, 10))
    print(f"sphere512@96 FP64 rel L2 {l2} floor {floor}")
    assert np.all(l2 <= np.maximum(1e-12, 4 * floor)), (l2, floor)
    s.close()


def test_sphere512_full_size_fp32_against_fp64():
    """Full 512^3 for the specified 1000 steps: the FP32 compensated form against the FP64 form on the device
    (which matches the oracle per the tests above): relative L2 <= 1e-5 per conserved field (measured
    <= 1.4e-6), and the same step count."""
    steps = 1000
    out = {}
    for prec in ("fp64", "fp32"):
        cfg = cases.sphere512(N=512, steps=steps, precision=prec)
        cfg.kahan = 1
        s = solver(cfg, graph=1)
        for done in range(0, steps, 250):
            s.marchTimeStep(250)
            print(f"  {prec}: {done + 250} steps", flush=True)
        out[prec] = (s.mesh.conserved(), s.time())
        s.close()
    U64, (t64, l64, _) = out["fp64"]
    U32, (t32, l32, _) = out["fp32"]
    assert l64 == l32 == steps
    assert abs(t32 - t64) <= 1e-5 * t64
    l2 = np.array([rel_l2(U32[q], U64[q]) for q in range(5)])
    linf = np.array([np.abs(U32[q] - U64[q]).max() / np.abs(U64[q]).max() for q in range(5)])
    print(f"sphere512 FP32 compensated vs FP64 after {steps} steps: rel L2 {l2} rel Linf {linf}")
    assert np.all(l2 <= 1e-5), l2


def test_cylinder256_1000_steps_fp32_against_fp64():
    """The precision study as specified: the 256^3 z-cylinder (Re 200, Mach 0.2), 1000 steps, FP32 with the
    compensated combine against FP64 on the device. Relative L2 <= 1e-5 on rho, rho_u, rho_v, rho_E; rho_w is
    zero up to the seeded perturbation in this spanwise-uniform flow (its norm is ~5e-4 of the momentum
    field's), so its error is measured against the norm of the whole momentum field (<= 1e-5); the drag
    coefficient of the last step to 1e-5 (measured: <= 4.7e-7, 3.1e-8 and 9.9e-8)."""
    steps = 1000
    out = {}
    for prec in ("fp64", "fp32"):
        cfg = cases.cylinder256(N=256, steps=steps, precision=prec)
        cfg.kahan = 1
        print(f"cylinder256 {prec}, {steps} steps", flush=True)
        s = solver(cfg, graph=1)
        s.setOption("body_force", 1)
        s.marchTimeStep(steps)
        out[prec] = (s.mesh.conserved(), s.time(), s.dragCoefficient()[-1])
        s.close()
    U64, (t64, l64, _), cd64 = out["fp64"]
    U32, (t32, l32, _), cd32 = out["fp32"]
    assert l64 == l32 == steps
    assert abs(t32 - t64) <= 1e-5 * t64
    l2 = {q: rel_l2(U32[q], U64[q]) for q in (0, 1, 2, 4)}
    m = np.sqrt(sum(np.linalg.norm(U64[q]) ** 2 for q in (1, 2, 3)))
    w = np.linalg.norm(U32[3] - U64[3]) / m
    cd = abs(cd32 - cd64) / abs(cd64)
    print(f"cylinder256 FP32 compensated vs FP64: rel L2 {l2}, rho_w vs |m| {w:.2e}, C_D {cd:.2e}", flush=True)
    assert all(v <= 1e-5 for v in l2.values()), l2
    assert w <= 1e-5 and cd <= 1e-5


# ------------------------------------------------------------------ the 96^3 adjoint workload

def test_adjoint_sphere96_against_central_differences():
    """96^3, 400 FP64 steps, J = mean C_D over levels 300..400, checkpoint every 20 steps: the adjoint
    directional derivative agrees with central differences (eps = 1e-4, same frozen step sizes) to 1e-4
    relative on the specified 16 random initial-state directions (measured <= 7.1e-7; the directions whose
    derivative is small lose most to the differences' rounding) and on Re (7e-8)."""
    cfg = cases.sphere96(N=96, steps=400)
    s = solver(cfg)
    U = s.mesh.conserved()
    q = np.ones(s.inletPlaneShape())
    s.close()
    r = nt.drag_and_gradients(cfg, U, q, cfg.Re, 400, 300, 400, 20)
    eps = 1e-4

    def J(Ux, re):
        return nt.drag_and_gradients(nt._config_with_re(cfg, re), Ux, q, re, 400, 300, 400, 400, dts=r["dts"])["J"]

    rng = np.random.default_rng(96)
    worst = 0.0
    for d in range(16):
        dU = rng.standard_normal(U.shape)
        dU /= np.abs(dU).max()
        fd = (J(U + eps * dU, cfg.Re) - J(U - eps * dU, cfg.Re)) / (2 * eps)
        ad = float(np.sum(r["grad_u0"] * dU))
        rel = abs(fd - ad) / abs(fd)
        print(f"direction {d}: FD {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    hre = 1e-4 * cfg.Re
    fd = (J(U, cfg.Re + hre) - J(U, cfg.Re - hre)) / (2 * hre)
    rel = abs(fd - r["grad_re"]) / abs(fd)
    print(f"Re: FD {fd:.10e} adjoint {r['grad_re']:.10e} rel {rel:.2e}")
    worst = max(worst, rel)
    assert worst <= 1e-4, worst


def test_adjoint_fp32_forward_against_fp64():
    """The 96^3 adjoint with the FP32 forward (compensated combine; forces, J and step sizes reduced in FP64)
    and the FP64 reverse sweep, at the FP64 run's step sizes: J agrees with the FP64 J to 1e-5 relative (also
    as the FP32 force history of a plain run gives it), and the directional derivatives on the 16 random
    initial-state directions and dJ/dRe agree with the FP64 adjoint's (pinned to central differences in the
    test above) to 1e-4 relative (measured <= 7.4e-7; J: 3.8e-8). Central differences of the FP32 run itself are no check: its rounding
    (~1e-7 relative per step) is far above what eps = 1e-4 differences resolve."""
    cfg = cases.sphere96(N=96, steps=400)
    s = solver(cfg)
    U = s.mesh.conserved()
    q = np.ones(s.inletPlaneShape())
    s.close()
    r = nt.drag_and_gradients(cfg, U, q, cfg.Re, 400, 300, 400, 20)
    c32 = cases.sphere96(N=96, steps=400, precision="fp32")
    c32.kahan = 1
    r32 = nt.drag_and_gradients(c32, U, q, cfg.Re, 400, 300, 400, 20, dts=r["dts"])
    s = solver(c32, body_force=1, graph=1)
    s.setInitialField(U.reshape(-1))
    s.setDtSchedule(r["dts"])
    s.marchTimeStep(400)
    f = s.getForceHistory()
    s.close()
    sel = (f[:, 0] >= 300) & (f[:, 0] <= 400)
    assert sel.sum() == 101
    J32 = float(f[sel, 1].mean() / (0.5 * c32.M_0 ** 2 * np.pi * c32.body_radius ** 2))
    rel = abs(J32 - r["J"]) / abs(r["J"])
    print(f"J FP64 {r['J']:.12e} FP32 {J32:.12e} (adjoint run {r32['J']:.12e}) rel {rel:.2e}")
    assert rel <= 1e-5, rel
    assert abs(r32["J"] - J32) <= 1e-12 * abs(J32)
    rng = np.random.default_rng(96)
    worst = 0.0
    for d in range(16):
        dU = rng.standard_normal(U.shape)
        dU /= np.abs(dU).max()
        a64, a32 = float(np.sum(r["grad_u0"] * dU)), float(np.sum(r32["grad_u0"] * dU))
        rel = abs(a32 - a64) / abs(a64)
        print(f"direction {d}: FP64 adjoint {a64:.10e} FP32-forward adjoint {a32:.10e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    rel = abs(r32["grad_re"] - r["grad_re"]) / abs(r["grad_re"])
    print(f"Re: FP64 adjoint {r['grad_re']:.10e} FP32-forward adjoint {r32['grad_re']:.10e} rel {rel:.2e}")
    worst = max(worst, rel)
    assert worst <= 1e-4, worst


# ------------------------------------------------------------------ fields beyond 4 GiB

@pytest.mark.parametrize("precision,nk", [("fp32", 44500), ("fp64", 29700)])
def test_field_beyond_4gib_matches_a_thin_slab(precision, nk):
    """A fully periodic box whose initial field varies in x and y only stays z-uniform exactly (every z-derivative
    is a difference of equal values), so every plane of a run with more than 4 GiB per field (the fused stage
    kernel's global-access form, offsets from each tile's first plane) must equal, bit for bit, the plane of the
    same run on 5 planes (z spacing 1/16 in both, so the step sizes agree too). 3 steps of the default form
    (FP32: compensated combine)."""
    dz = 0.0625

    def make(NK):
        return nsfsi.make_config(
            NI=126, NJ=126, NK=NK, L_x=4.0, L_y=4.0, L_z=dz * (NK - 2), bc=["periodic"] * 6, body_type="none",
            stop_criterion="timesteps", stop_time_level=3, filter_intervals=[0], filter_change_times=[],
            M_0=0.3, Re=300.0, conv_stability_limit=0.5, precision=precision)

    def plane_state(cfg):
        # x_i = (i - 1) hx on the periodic axes; a smooth field of one period in x and in y
        hx, hy = cfg.L_x / (cfg.NI - 2), cfg.L_y / (cfg.NJ - 2)
        x = (np.arange(cfg.NI) - 1) * hx
        y = (np.arange(cfg.NJ) - 1) * hy
        X, Y = np.meshgrid(x, y)  # [j, i]
        sx, sy = np.sin(2 * np.pi * X / cfg.L_x), np.sin(2 * np.pi * Y / cfg.L_y)
        rho = 0.01 * sx * sy
        u = 0.3 + 0.05 * np.cos(2 * np.pi * Y / cfg.L_y)
        v = 0.05 * sx
        p = 0.002 * np.cos(2 * np.pi * X / cfg.L_x)
        m = 1 + rho
        return np.stack([rho, m * u, m * v, np.zeros_like(rho), p / (cfg.gamma - 1) + m / 2 * (u * u + v * v)])

    big, thin = make(nk), make(5)
    eb = 4 if precision == "fp32" else 8
    # device row: x-segments of 62 nodes in 64 slots, padded to whole segments (FP32) or 128 bytes (FP64)
    slots = (big.NI - 1) + 2 * ((big.NI - 1) // 62) + 2
    unit = 64 if precision == "fp32" else 128 // eb
    field_bytes = nk * big.NJ * (-(-slots // unit) * unit) * eb
    print(f"{precision} 126x126x{nk}: {field_bytes / 2**30:.2f} GiB per field", flush=True)
    assert field_bytes > 2**32
    p0 = plane_state(thin)
    t = solver(thin)
    t.setInitialField(np.broadcast_to(p0[:, None], (5, 5, 126, 126)))
    t.marchTimeStep(3)
    ref = [t.mesh.field(nsfsi.FIELDS[q])[2] for q in range(5)]
    tt = t.time()
    t.close()
    s = solver(big, graph=1)
    U = np.empty((5, nk, 126, 126))
    U[:] = p0[:, None]
    s.setInitialField(U)
    del U
    s.marchTimeStep(3)
    assert s.time() == tt
    for q in range(5):
        f = s.mesh.field(nsfsi.FIELDS[q])
        for k in (1, 2, nk // 3, nk // 2, nk - 2):
            assert np.array_equal(f[k], ref[q]), (nsfsi.FIELDS[q], k)
        del f
    s.close()
