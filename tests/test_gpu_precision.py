"""FP32 forms against the serial FP64 oracle.

Every reduced-precision result states its relative L2 and Linf error per conserved field against the
oracle run on the same configuration and seeded initial field. The compensated combine (increment
form + TwoSum, rounding error carried between steps in bfloat16) must not be less accurate than the
plain FP32 combine, and both FP32 stage kernels must implement it alike.
"""
import numpy as np
import pytest

import cases
import nsfsi
import oracle

pytestmark = pytest.mark.gpu


def errors(U, V):
    l2, linf = [], []
    for q in range(5):
        n2 = np.linalg.norm(V[q])
        ni = np.abs(V[q]).max()
        l2.append(np.linalg.norm(U[q] - V[q]) / (n2 if n2 > 0 else 1.0))
        linf.append(np.abs(U[q] - V[q]).max() / (ni if ni > 0 else 1.0))
    return np.array(l2), np.array(linf)


def run_device(cfg, steps, kernel=1, **opts):
    print(f"device run NI={cfg.NI} prec={cfg.precision} kahan={cfg.kahan} kernel={kernel} steps={steps}", flush=True)
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", kernel)
    for k, v in opts.items():
        s.setOption(k, v)
    s.marchTimeStep(steps)
    U = s.mesh.conserved()
    t = s.time()
    s.close()
    return U, t


@pytest.fixture(scope="module")
def sphere48_200():
    """The 48^3 sphere case for 200 steps: oracle state and the device FP32 states."""
    steps = 200
    cfg = cases.sphere48(N=48, steps=steps, precision="fp32")
    o = oracle.Oracle(cases.sphere48(N=48, steps=steps))
    o.initialize()
    o.march(steps)
    ref = o.conserved()
    out = {"ref": ref, "t": o.state()[0]}
    for kahan in (1, 0):
        for kernel in (1, 0):
            cfg.kahan = kahan
            out[(kahan, kernel)] = run_device(cfg, steps, kernel)
    return out


# Per-field bars on the relative L2 error against the oracle after 200 steps of the 48^3 sphere: a small
# multiple (2.5x) of the largest error measured per form (compensated combine 8.1e-7, plain 4.5e-6), so a
# regression of either form shows, and both under the 1e-5 the FP32 form is specified to.
FP32_BAR = {1: 2e-6, 0: 1e-5}


@pytest.mark.parametrize("kernel", [1, 0], ids=["fused", "unfused"])
def test_fp32_error_vs_oracle(sphere48_200, kernel):
    ref = sphere48_200["ref"]
    for kahan in (1, 0):
        U, (t, lv, dt) = sphere48_200[(kahan, kernel)]
        assert lv == 200
        assert abs(t - sphere48_200["t"]) <= 1e-5 * sphere48_200["t"]
        l2, linf = errors(U, ref)
        print(f"kernel={kernel} kahan={kahan} rel L2 {l2} rel Linf {linf} bar {FP32_BAR[kahan]}")
        assert np.all(l2 < FP32_BAR[kahan]), (kahan, l2)


@pytest.fixture(scope="module")
def sphere48_200_floor():
    """The oracle's own 1-ulp sensitivity on the 200-step 48^3 sphere (see test_gpu_parity.ulp_sensitivity)."""
    from test_gpu_parity import ulp_sensitivity
    return np.array(ulp_sensitivity(cases.sphere48(N=48, steps=200), 200))


@pytest.mark.parametrize("kernel", [1, 0], ids=["fused", "unfused"])
def test_fp64_matches_oracle_200_steps(sphere48_200, sphere48_200_floor, kernel):
    """The 48^3 sphere for 200 steps (the reference's CPU-runnable configuration) in FP64 on the device:
    same time level and time as the oracle, and per conserved field within 1e-12 relative L2 of it or
    within 4x the oracle's own change under a 1-ulp perturbation, whichever is larger (the transverse
    momenta, whose norms are three orders below the axial one, sit at that floor: about 2e-12)."""
    cfg = cases.sphere48(N=48, steps=200, precision="fp64")
    U, (t, lv, dt) = run_device(cfg, 200, kernel)
    assert lv == 200
    assert abs(t - sphere48_200["t"]) <= 1e-12 * sphere48_200["t"]
    l2, linf = errors(U, sphere48_200["ref"])
    floor = sphere48_200_floor
    print(f"FP64 kernel={kernel} rel L2 {l2} rel Linf {linf} oracle 1-ulp sensitivity {floor}")
    assert np.all(l2 <= np.maximum(1e-12, 4 * floor)), (l2, floor)


def test_kahan_not_less_accurate(sphere48_200):
    ref = sphere48_200["ref"]
    k_l2, _ = errors(sphere48_200[(1, 1)][0], ref)
    p_l2, _ = errors(sphere48_200[(0, 1)][0], ref)
    print("compensated", k_l2, "plain", p_l2)
    # summed over the fields: the compensated form removes the rounding of U + increment
    assert k_l2.sum() <= p_l2.sum()


def test_fused_and_unfused_compensated_agree(sphere48_200):
    a = sphere48_200[(1, 1)][0]
    b = sphere48_200[(1, 0)][0]
    l2, _ = errors(a, b)
    assert np.all(l2 < 1e-5), l2


def test_compensated_graph_replay_bitwise():
    """The compensation buffer persists across steps: graph replays must equal stream launches."""
    cfg = cases.sphere48(N=32, steps=6, precision="fp32")
    cfg.kahan = 1
    a, _ = run_device(cfg, 6, 1, graph=0)
    b, _ = run_device(cfg, 6, 1, graph=1)
    assert np.array_equal(a, b)


def test_uniform_flow_fixed_point_compensated():
    """Uniform flow with no body is a fixed point of the step; the compensation must stay zero."""
    cfg = nsfsi.make_config(NI=24, NJ=20, NK=18, L_x=2.0, L_y=2.0, L_z=2.0, bc=["periodic"] * 6, body_type="none",
                            ic_type="freestream_uniform", ic_amplitude=0.0, M_0=0.3, stop_criterion="timesteps",
                            stop_time_level=5, filter_intervals=[0], filter_change_times=[], precision="fp32")
    cfg.kahan = 1
    s = nsfsi.Solver(cfg)
    s.initialize()
    U0 = s.mesh.conserved()
    s.marchTimeStep(5)
    assert np.array_equal(s.mesh.conserved(), U0)
    s.close()
