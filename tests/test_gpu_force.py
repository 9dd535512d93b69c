"""Force on the body (body_force option) against the oracle's momentum balance.

In a box that is periodic in every direction the outer faces contribute nothing to the sum of the
momentum right-hand sides over the FluidInterior nodes, so the force on the body equals minus that
sum times the node volume. The oracle's slopes are the reference's per-node expression
(computeRK4slopes, src/Solver.cpp:170-418), not the face-flux form the device kernel sums: a dropped
term, a wrong sign or a wrong face orientation in the device kernel fails this comparison.
"""
import numpy as np
import pytest

import nsfsi
import oracle

pytestmark = pytest.mark.gpu


def periodic_sphere(precision="fp64"):
    return nsfsi.make_config(
        NI=32, NJ=30, NK=28, L_x=2.0, L_y=2.0, L_z=2.0, bc=["periodic"] * 6, body_type="sphere",
        body_centre=(1.03, 0.98, 1.01), body_radius=0.42, ic_type="freestream_gaussian", ic_amplitude=1e-3,
        ic_seed=3, M_0=0.3, Re=300.0, stop_criterion="timesteps", stop_time_level=5, filter_intervals=[0],
        filter_change_times=[], conv_stability_limit=0.5, precision=precision)


def device_forces(cfg, steps, kernel=1):
    print(f"force run prec={cfg.precision} kernel={kernel} steps={steps}", flush=True)
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", kernel)
    s.setOption("body_force", 1)
    s.marchTimeStep(steps)
    f = s.getForceHistory()
    s.close()
    return f


def oracle_force(cfg, steps):
    o = oracle.Oracle(cfg)
    o.initialize()
    o.march(steps)
    sp, _ = o.geometry()
    S = o.slopes()
    return np.array([-sp.prod() * S[q].sum() for q in (1, 2, 3)])


@pytest.mark.parametrize("kernel", [1, 0], ids=["fused", "unfused"])
def test_force_matches_momentum_balance(kernel):
    cfg = periodic_sphere()
    f = device_forces(cfg, 5, kernel)
    assert f.shape == (5, 4)
    assert np.array_equal(f[:, 0], np.arange(1, 6))
    ref = oracle_force(cfg, 5)
    print("device", f[-1, 1:], "oracle", ref)
    assert ref[0] > 0.1  # drag along the +x stream
    assert np.all(np.abs(f[-1, 1:] - ref) <= 1e-9 * np.abs(ref[0]))


def test_force_at_intermediate_step():
    cfg = periodic_sphere()
    f = device_forces(cfg, 3)
    ref = oracle_force(cfg, 3)
    assert np.all(np.abs(f[-1, 1:] - ref) <= 1e-9 * np.abs(ref[0]))


def test_force_fp32_close_to_fp64():
    f64 = device_forces(periodic_sphere(), 5)
    f32 = device_forces(periodic_sphere("fp32"), 5)
    rel = np.abs(f32[:, 1] - f64[:, 1]) / np.abs(f64[:, 1])
    print("FP32 drag relative error per step", rel)
    assert np.all(rel < 1e-4)


def test_force_option_leaves_state_unchanged():
    cfg = periodic_sphere()
    out = []
    for on in (0, 1):
        s = nsfsi.Solver(cfg)
        s.initialize()
        s.setOption("body_force", on)
        s.setOption("graph", 1)
        s.marchTimeStep(4)
        out.append(s.mesh.conserved())
        if on:
            assert s.getForceHistory().shape == (4, 4)
        else:
            assert s.getForceHistory().shape == (0, 4)
        s.close()
    assert np.array_equal(out[0], out[1])
