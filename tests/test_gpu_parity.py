"""Device path vs the serial FP64 oracle on the same configuration and seeded initial field.

Integer outputs (node classes, ghost and filter sets, image-point boxes) must match bit for bit; the
FP64 geometry (body intercept and image points) must match bit for bit; the FP64 trajectory must match
to <= 1e-12 relative L2 per conserved field after N steps.
"""
import numpy as np
import pytest

import cases
import nsfsi
import oracle

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _pair(cfg):
    print(f"case NI={cfg.NI} NJ={cfg.NJ} NK={cfg.NK} body={cfg.body_type} prec={cfg.precision}", flush=True)
    s = nsfsi.Solver(cfg)
    s.initialize()
    o = oracle.Oracle(cfg)
    o.initialize()
    return s, o


CLASSIFY_CASES = {
    "sphere24": lambda: cases.sphere48(N=24, steps=10),
    "sphere48": lambda: cases.sphere48(N=48, steps=10),
    "cylinder_z64": lambda: cases.cylinder256(N=64, steps=10),
    "reference_default": cases.reference_default,
    "sphere_offcentre33": lambda: nsfsi.make_config(
        NI=33, NJ=29, NK=31, L_x=3.0, L_y=2.5, L_z=2.7, bc=cases.SIDE_WALLS, body_type="sphere",
        body_centre=(1.31, 1.27, 1.33), body_radius=0.61, stop_criterion="timesteps", stop_time_level=5,
        filter_intervals=[0], filter_change_times=[], ic_type="freestream_gaussian", ic_amplitude=1e-3, ic_seed=7,
        M_0=0.3, Re=300.0, conv_stability_limit=0.5),
}


@pytest.mark.parametrize("name", sorted(CLASSIFY_CASES))
def test_classification_bit_exact(name):
    cfg = CLASSIFY_CASES[name]()
    s, o = _pair(cfg)
    assert np.array_equal(s.mesh.nodeType, o.node_types())
    sg, og = s.mesh.counts(), o.counts()
    assert sg == og
    ijk, bi, ip, box = s.mesh.ghosts()
    oijk, obi, oip, obox = o.ghosts()
    key = lambda a: (a[:, 0].astype(np.int64) * cfg.NJ + a[:, 1]) * cfg.NK + a[:, 2]
    order = np.argsort(key(oijk), kind="stable")
    assert np.array_equal(ijk, oijk[order])
    assert np.array_equal(bi, obi[order])  # bit-exact FP64 geometry
    assert np.array_equal(ip, oip[order])
    assert np.array_equal(box, obox[order])
    assert np.array_equal(s.mesh.filterNodes(), o.filter_nodes())
    sp, of = o.geometry()
    assert s.mesh.dx == sp[0] and s.mesh.dy == sp[1] and s.mesh.dz == sp[2]
    assert np.array_equal(s.mesh.positionOffset, of)


@pytest.mark.parametrize("name", ["sphere24", "sphere48", "cylinder_z64", "sphere_offcentre33"])
def test_initial_state_bit_exact(name):
    cfg = CLASSIFY_CASES[name]()
    s, o = _pair(cfg)
    assert np.array_equal(s.mesh.conserved(), o.conserved())
    assert s.dt == o.state()[2]


@pytest.mark.parametrize("name,steps", [("sphere24", 20), ("sphere48", 20), ("cylinder_z64", 10),
                                        ("sphere_offcentre33", 15)])
def test_trajectory_fp64(name, steps):
    cfg = CLASSIFY_CASES[name]()
    s, o = _pair(cfg)
    s.marchTimeStep(steps)
    o.march(steps)
    t, lv, dt = s.time()
    ot, olv, odt = o.state()
    assert lv == olv == steps
    assert abs(t - ot) <= 1e-12 * abs(ot)
    assert abs(dt - odt) <= 1e-12 * abs(odt)
    U, V = s.mesh.conserved(), o.conserved()
    errs = [rel_l2(U[q], V[q]) for q in range(5)]
    print("rel L2 per field", errs)
    assert max(errs) <= TOL_FP64
    hist = s.getConvergenceHistory()
    for q, name_ in enumerate(nsfsi.FIELDS[:5]):
        oh = o.norm_history(q)
        assert len(hist[name_]) == len(oh) == steps
        assert np.allclose(hist[name_], oh, rtol=1e-10, atol=0)


def test_geometry_error_matches_oracle():
    """At 40^3 the z-cylinder's ghosts on the top symmetry plane get an image-point box that leaves the
    mesh (getSurroundingNodesBox clamps to N, src/SmallVectors.h:120-122): both paths must refuse."""
    cfg = cases.cylinder256(N=40, steps=2)
    o = oracle.Oracle(cfg)
    with pytest.raises(oracle.OracleError):
        o.initialize()
    s = nsfsi.Solver(cfg)
    with pytest.raises(nsfsi.NSError) as e:
        s.initialize()
    assert e.value.status == 2
