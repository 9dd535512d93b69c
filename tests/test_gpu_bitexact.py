"""The FP64 unfused stage form (stored primitive fields, the reference's own combine and expression order,
no FMA contraction, correctly rounded x^1.5) against the oracle with a correctly rounded x^1.5: the conserved,
primitive and transport fields, t, timeLevel and dt after every run are bit-identical, on every geometry of the
parity suite, the reference's CPU-runnable configuration (48^3 sphere, 200 steps) and the reference's shipped
case with its periodic density filter. The norm-of-change history agrees to summation order (rtol 1e-13).

Against the oracle with the reference's std::pow (glibc: within 0.52 ulp, not correctly rounded) the same runs
agree to the oracle's own one-ulp sensitivity, as the fused form does (test_gpu_parity.py)."""
import numpy as np
import pytest

import cases
import nsfsi
import oracle
from test_gpu_parity import CLASSIFY_CASES, rel_l2, ulp_sensitivity

pytestmark = pytest.mark.gpu


def run_pair(cfg, steps, cr_pow=True):
    print(f"stored-primitive FP64 NI={cfg.NI} NJ={cfg.NJ} NK={cfg.NK} body={cfg.body_type} steps={steps}", flush=True)
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", 0)
    o = oracle.Oracle(cfg)
    o.set_correctly_rounded_pow(cr_pow)
    o.initialize()
    s.marchTimeStep(steps)
    o.march(steps)
    return s, o


def assert_bit_identical(s, o, steps):
    t, lv, dt = s.time()
    ot, olv, odt = o.state()
    assert lv == olv == steps
    assert t == ot and dt == odt, (t, ot, dt, odt)
    for q in range(12):
        a, b = s.mesh.field(nsfsi.FIELDS[q]), o.field(q)
        nd = int(np.sum(a != b))
        assert nd == 0, f"{nsfsi.FIELDS[q]}: {nd} nodes differ, max |diff| {np.abs(a - b).max():.3e}"
    hist = s.getConvergenceHistory()
    for q, name in enumerate(nsfsi.FIELDS[:5]):
        assert np.allclose(hist[name], o.norm_history(q), rtol=1e-13, atol=0), name


@pytest.mark.parametrize("name,steps", [("sphere24", 20), ("sphere48", 20), ("cylinder_z64", 10),
                                        ("sphere_offcentre33", 15), ("sphere_seam100", 15), ("sphere_inflow_y", 15),
                                        ("cylinder_x_inflow_z", 15), ("two_bodies", 15), ("three_bodies", 10)])
def test_bit_identical_trajectory(name, steps):
    cfg = CLASSIFY_CASES[name]()
    s, o = run_pair(cfg, steps)
    assert_bit_identical(s, o, steps)


def test_bit_identical_reference_cpu_configuration_200_steps():
    """The reference's CPU-runnable configuration: 48^3 sphere, 200 RK4 steps."""
    cfg = cases.sphere48(N=48, steps=200)
    s, o = run_pair(cfg, 200)
    assert_bit_identical(s, o, 200)


def test_bit_identical_reference_default_case_with_filter():
    """The shipped 161x83x3 case for 60 steps: the periodic density filter fires at step 50."""
    cfg = cases.reference_default()
    cfg.stop_criterion = 0
    cfg.stop_time_level = 60
    s, o = run_pair(cfg, 60)
    assert_bit_identical(s, o, 60)


def test_bit_identical_with_graph_replay_and_filter_every_three_steps():
    cfg = cases.sphere48(N=24, steps=12)
    cfg.n_filter_intervals = 1
    cfg.filter_intervals[0] = 3
    cfg.n_filter_change_times = 0
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", 0)
    s.setOption("graph", 1)
    s.marchTimeStep(12)
    o = oracle.Oracle(cfg)
    o.set_correctly_rounded_pow(True)
    o.initialize()
    o.march(12)
    assert_bit_identical(s, o, 12)


def test_end_time_run_bit_identical():
    cfg = cases.sphere48(N=24, steps=1)
    cfg.stop_criterion = 1
    cfg.t_end = 0.3
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", 0)
    while s.time()[0] < cfg.t_end:  # Simulation::run's loop; the last step is clamped to end at t_end
        s.marchTimeStep(1)
    assert s.time()[0] == cfg.t_end
    o = oracle.Oracle(cfg)
    o.set_correctly_rounded_pow(True)
    o.initialize()
    o.run()
    assert_bit_identical(s, o, o.state()[1])


def test_against_reference_pow_within_floor():
    """With the reference's std::pow in the oracle, the stored-primitive form is within the oracle's own
    one-ulp sensitivity, per field (the rounding differences of std::pow are what remains)."""
    cfg = cases.sphere48(N=48, steps=200)
    s, o = run_pair(cfg, 200, cr_pow=False)
    floor = np.array(ulp_sensitivity(cfg, 200))
    U, V = s.mesh.conserved(), o.conserved()
    errs = np.array([rel_l2(U[q], V[q]) for q in range(5)])
    print(f"vs std::pow oracle: rel L2 {errs} floor {floor}")
    assert np.all(errs <= np.maximum(1e-12, 4 * floor))
