"""The tangent-linear form of the oracle (oracle/ns_oracle.cpp instantiated on value/tangent pairs,
ns_oracle_tl_*): its values equal the plain oracle's run bit for bit, and its tangents equal central
differences of the plain oracle. Two cases between them run every part of the step: a sphere in inflow
(inlet with its in-plane density filter, outlet, symmetry faces, the immersed filter and the Vandermonde
ghost solves, the periodic density filter every other step) and a sphere in a box periodic in every
direction. Step sizes are prescribed (the tangent model is that of the step at given step sizes)."""
import numpy as np
import pytest

import cases
import nsfsi
import oracle


def inflow_sphere():
    c = cases.sphere48(N=20, steps=6)
    c.n_filter_intervals = 1
    c.filter_intervals[0] = 2
    c.n_filter_change_times = 0
    c.inlet_ramp_time = 0.05  # the inlet velocity ramps within the run
    return c


def periodic_sphere():
    return nsfsi.make_config(
        NI=16, NJ=15, NK=14, L_x=2.0, L_y=2.0, L_z=2.0, bc=["periodic"] * 6, body_type="sphere",
        body_centre=(1.03, 0.98, 1.01), body_radius=0.42, ic_type="freestream_gaussian", ic_amplitude=1e-3,
        ic_seed=5, M_0=0.3, Re=300.0, stop_criterion="timesteps", stop_time_level=6, filter_intervals=[0],
        filter_change_times=[], conv_stability_limit=0.5)


CASES = {"inflow_sphere": inflow_sphere, "periodic_sphere": periodic_sphere}
STEPS = 5


def plain_run(cfg, U0, dts, steps):
    o = oracle.Oracle(cfg)
    o.initialize()
    for q in range(5):
        o.set_field(q, U0[q])
    o.derive_all()
    o.set_dt_schedule(dts)
    o.march(steps)
    out = o.conserved(), o.slopes(), o.state()
    o.close()
    return out


def setup(name):
    cfg = CASES[name]()
    o = oracle.Oracle(cfg)
    o.initialize()
    U0 = o.conserved()
    dts = [o.state()[2]]
    for _ in range(STEPS):
        o.march(1)
        dts.append(o.state()[2])
    o.close()
    assert o is not None and len(dts) == STEPS + 1
    return cfg, U0, np.array(dts)


@pytest.mark.parametrize("name", sorted(CASES))
def test_tangent_values_are_the_plain_oracle(name):
    cfg, U0, dts = setup(name)
    d = np.random.default_rng(3).standard_normal(U0.shape)
    tl = oracle.TangentOracle(cfg)
    tl.initialize()
    tl.set_state(U0, d)
    tl.set_dt_schedule(dts)
    tl.march(STEPS)
    v, _ = tl.state_pair()
    sv, _ = tl.slopes()
    U, K, st = plain_run(cfg, U0, dts, STEPS)
    assert tl.state() == st
    assert np.array_equal(v, U)
    assert np.array_equal(sv, K)
    tl.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_tangents_match_central_differences_of_the_oracle(name):
    cfg, U0, dts = setup(name)
    rng = np.random.default_rng(11)
    for _ in range(2):
        d = rng.standard_normal(U0.shape)
        d /= np.abs(d).max()
        tl = oracle.TangentOracle(cfg)
        tl.initialize()
        tl.set_state(U0, d)
        tl.set_dt_schedule(dts)
        tl.march(STEPS)
        _, dU = tl.state_pair()
        _, dK = tl.slopes()
        tl.close()
        eps = 1e-5  # measured: <= 2e-9 per field here, the differences' own rounding and truncation
        Up, Kp, _ = plain_run(cfg, U0 + eps * d, dts, STEPS)
        Um, Km, _ = plain_run(cfg, U0 - eps * d, dts, STEPS)
        for tan, fd in ((dU, (Up - Um) / (2 * eps)), (dK, (Kp - Km) / (2 * eps))):
            for q in range(5):
                rel = np.linalg.norm(tan[q] - fd[q]) / np.linalg.norm(tan[q])
                assert rel <= 5e-8, (name, q, rel)


@pytest.mark.parametrize("name", sorted(CASES))
def test_reynolds_tangent_matches_central_differences(name):
    """A tangent on Re alone (zero state tangent): the run's tangents are the derivatives with respect to
    the Reynolds number of the viscosity law, against central differences of the plain oracle in Re."""
    cfg, U0, dts = setup(name)
    tl = oracle.TangentOracle(cfg)
    tl.initialize()
    tl.set_re_tangent(1.0)
    tl.set_state(U0)
    tl.set_dt_schedule(dts)
    tl.march(STEPS)
    _, dU = tl.state_pair()
    _, dK = tl.slopes()
    tl.close()
    h = 1e-4 * cfg.Re  # measured: <= 6e-8 per field here, the differences' own rounding (it grows as h shrinks)
    runs = []
    for re in (cfg.Re + h, cfg.Re - h):
        c = CASES[name]()
        c.Re = re
        runs.append(plain_run(c, U0, dts, STEPS))
    for tan, fd in ((dU, (runs[0][0] - runs[1][0]) / (2 * h)), (dK, (runs[0][1] - runs[1][1]) / (2 * h))):
        for q in range(5):
            rel = np.linalg.norm(tan[q] - fd[q]) / np.linalg.norm(tan[q])
            assert rel <= 2e-7, (name, q, rel)
