"""Batched cases (ns_batch_*) against the same cases run one at a time.

The batched kernels run the per-node arithmetic of the single-case kernels (same device functions, same
tile shape), each case with its own dt, so every case's state, time and dt after N steps must equal its
single-case run bit for bit; the norm histories differ only in the order of their FP64 sums.
"""
import numpy as np
import pytest

import cases
import nsfsi

pytestmark = pytest.mark.gpu


def small_sweep(n=6, steps=8):
    cfgs = cases.sweep64(n=n, steps=steps)
    # a second body shape and boundary set in the mix: z-cylinder case with periodic z
    cyl = cases.cylinder256(N=40, steps=steps, precision="fp32")
    cyl.NK = 20
    cyl.bc[4] = cyl.bc[5] = nsfsi.BC["periodic"]
    # inflow along y, and along z with periodic x: faces of other axes in the batched face groups, and
    # outlets that the stage kernel leaves to the edge-face pass beside x-outlets that it applies itself
    yin = nsfsi.make_config(
        NI=30, NJ=34, NK=26, L_x=2.2, L_y=2.6, L_z=2.0, bc=["symmetry", "symmetry", "inlet", "outlet", "symmetry",
                                                         "symmetry"],
        body_type="sphere", body_centre=(1.08, 1.21, 0.97), body_radius=0.45, stop_criterion="timesteps",
        stop_time_level=steps, filter_intervals=[0], filter_change_times=[], ic_type="freestream_gaussian",
        ic_amplitude=1e-3, ic_seed=13, M_0=0.3, Re=250.0, conv_stability_limit=0.5, precision="fp32")
    zin = nsfsi.make_config(
        NI=28, NJ=26, NK=36, L_x=2.0, L_y=2.0, L_z=2.8, bc=["periodic", "periodic", "symmetry", "symmetry", "inlet",
                                                         "outlet"],
        body_type="cylinder", body_axis="x", body_centre=(1.0, 0.98, 1.31), body_radius=0.4, stop_criterion="timesteps",
        stop_time_level=steps, filter_intervals=[0], filter_change_times=[], ic_type="freestream_gaussian",
        ic_amplitude=1e-3, ic_seed=17, M_0=0.25, Re=200.0, conv_stability_limit=0.5, precision="fp32")
    for c in (yin, zin):
        c.kahan = cfgs[0].kahan
    return cfgs + [cyl, yin, zin]


def single(cfg, steps, tile):
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("tile_ty", tile[0])
    s.setOption("tile_kc", tile[1])
    s.marchTimeStep(steps)
    out = (s.mesh.conserved(), s.time(), s.getConvergenceHistory())
    s.close()
    return out


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_equals_single_cases(graph):
    steps = 8
    cfgs = small_sweep(steps=steps)
    print("batch of", [(c.NI, c.NJ, c.NK) for c in cfgs], "graph", graph, flush=True)
    b = nsfsi.Batch(cfgs)
    b.initialize()
    b.setOption("graph", graph)
    tile = b.tile()
    b.step(3)
    b.step(steps - 3)  # two calls: the buffer parity carries over
    for c, cfg in enumerate(cfgs):
        U, (t, lv, dt), hist = single(cfg, steps, tile)
        bt, blv, bdt = b.time(c)
        assert blv == lv == steps
        assert bt == t and bdt == dt, (c, bt, t, bdt, dt)
        assert np.array_equal(b.conserved(c), U), c
        bh = b.getConvergenceHistory(c)
        for name in nsfsi.FIELDS[:5]:
            assert np.allclose(bh[name], hist[name], rtol=1e-12, atol=0)
    b.close()


@pytest.mark.parametrize("graph", [1, 0])
def test_persistent_stage_launches_equal_grid_launches(graph):
    """The persistent form (resident blocks pulling tiles through counters that each launch leaves at zero)
    computes every tile as the grid launch does: equal states and times over several steps and graph
    replays, so the counters come back to zero after every launch."""
    cfgs = small_sweep(n=5, steps=7)
    out = []
    for persistent in (0, 1):
        b = nsfsi.Batch(cfgs)
        b.initialize()
        b.setOption("graph", graph)
        b.setOption("persistent", persistent)
        b.step(2)
        b.step(5)
        out.append([(b.conserved(c), b.time(c)) for c in range(len(cfgs))])
        b.close()
    for c, ((U0, t0), (U1, t1)) in enumerate(zip(out[0], out[1])):
        assert t0 == t1, c
        assert np.array_equal(U0, U1), c


def test_persistent_launch_with_fewer_tiles_than_ranges():
    """One 24^3 case at 14-row x 128-plane tiles is a single tile (its two y-bands share one half-width
    tile): seven of the eight XCD ranges are empty and all but one resident block find no tile; the state
    still equals the grid launch's and the counters come back to zero (later steps would skip tiles
    otherwise)."""
    cfg = cases.sphere48(N=24, steps=6, precision="fp32")
    out = []
    for persistent in (0, 1):
        b = nsfsi.Batch([cfg])
        b.initialize()
        b.setOption("graph", 1)
        b.setOption("tile_ty", 14)
        b.setOption("tile_kc", 128)
        b.setOption("persistent", persistent)
        b.step(6)
        out.append((b.conserved(0), b.time(0)))
        b.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_batch_cases_advance_at_their_own_dt():
    cfgs = small_sweep(n=4, steps=5)
    b = nsfsi.Batch(cfgs)
    b.initialize()
    b.step(5)
    dts = [b.time(c)[2] for c in range(len(cfgs))]
    assert len(set(dts)) == len(dts)  # different grids, Mach numbers, Reynolds numbers
    b.close()


def test_batch_autotune_keeps_results():
    cfgs = small_sweep(n=4, steps=4)
    out = []
    for tune in (0, 1):
        b = nsfsi.Batch(cfgs)
        b.initialize()
        if tune:
            b.autotune(1)
        b.step(4)
        out.append([b.conserved(c) for c in range(len(cfgs))])
        b.close()
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)


def test_batch_plain_combine():
    cfgs = small_sweep(n=3, steps=4)
    for c in cfgs:
        c.kahan = 0
    b = nsfsi.Batch(cfgs)
    b.initialize()
    tile = b.tile()
    b.step(4)
    for c, cfg in enumerate(cfgs):
        U, _, _ = single(cfg, 4, tile)
        assert np.array_equal(b.conserved(c), U)
    b.close()


@pytest.mark.parametrize("graph", [1, 0])
def test_batch_with_periodic_density_filter(graph):
    """Cases with the periodic density filter (src/Mesh.cpp:85-140) in a batch beside cases without: the
    shipped 161x83x3 case (intervals 50/200/1000 switching at t = 30/60/100; fires at step 50) and a 24^3
    sphere filtered every 3 steps, 60 steps; every case equals its single-case run bit for bit."""
    steps = 60
    ref = nsfsi.make_config(precision="fp32", stop_criterion="timesteps", stop_time_level=steps)
    sph = cases.sphere48(N=24, steps=steps, precision="fp32")
    sph.n_filter_intervals = 1
    sph.filter_intervals[0] = 3
    sph.n_filter_change_times = 0
    plain = cases.sweep64(n=1, steps=steps)[0]
    cfgs = [ref, plain, sph]
    for c in cfgs:
        c.kahan = 1
    print("batch with filters, graph", graph, flush=True)
    b = nsfsi.Batch(cfgs)
    b.initialize()
    b.setOption("graph", graph)
    tile = b.tile()
    b.step(steps)
    for c, cfg in enumerate(cfgs):
        U, (t, lv, dt), _ = single(cfg, steps, tile)
        bt, blv, bdt = b.time(c)
        assert blv == lv == steps and bt == t and bdt == dt, c
        assert np.array_equal(b.conserved(c), U), c
    filtered = b.conserved(0)
    b.close()
    # the filter changed the run: without it the shipped case ends elsewhere
    nofilter = nsfsi.make_config(precision="fp32", stop_criterion="timesteps", stop_time_level=steps,
                                 filter_intervals=[0], filter_change_times=[])
    nofilter.kahan = 1
    U0, _, _ = single(nofilter, steps, tile)
    assert not np.array_equal(U0, filtered)


def test_full_sweep_of_64_cases_500_steps():
    """The sweep as specified: 64 seeded cases (grid edges 32-128), 500 steps each at its own dt, in one
    batch. Four cases of this draw (31: 32^3, Mach 0.38, sphere radius 0.53; 52: 48^3, r 0.31; 39 and 42:
    64^3, r 0.32-0.34, Re 1700-3800) go non-finite within the 500 steps: the serial FP64 oracle does too,
    at steps 70, 205, 243 and 322 (the reference's non-dissipative central scheme on spheres of 4-5 cells
    radius; its periodic density filter does not change this). As the reference's checkFinity
    (src/Mesh.h:48-61) reports and goes on, the batch finishes the steps of every case and names the first
    such case; every other case stays finite, and three of them (64^3, 96^3, 128^3) equal their single-case
    runs bit for bit. Case 31 in the device's FP64 form first holds a non-finite value at the oracle's
    step."""
    import oracle

    steps = 500
    cfgs = cases.sweep64(n=64, steps=steps)
    b = nsfsi.Batch(cfgs)
    b.initialize()
    tile = b.tile()
    print("sweep64 batch, 500 steps, tile", tile, flush=True)
    with pytest.raises(nsfsi.NSError) as e:
        b.step(steps)
    assert e.value.status == 3 and "of case 31" in str(e.value), str(e.value)
    bad = [c for c in range(64) if not np.all(np.isfinite(b.conserved(c)))]
    assert bad == [31, 39, 42, 52], bad
    for c in range(64):
        assert b.time(c)[1] == steps
    picked = {}
    for c, cfg in enumerate(cfgs):
        if c not in bad and cfg.NI in (64, 96, 128):
            picked.setdefault(cfg.NI, c)
    for N, c in sorted(picked.items()):
        print(f"case {c} ({N}^3) against its single run", flush=True)
        U, (t, lv, dt), _ = single(cfgs[c], steps, tile)
        assert b.time(c) == (t, lv, dt)
        assert np.array_equal(b.conserved(c), U), c
    b.close()

    def first_nonfinite(step, finite, limit):
        for n in range(1, limit + 1):
            step()
            if not finite():
                return n
        return None

    c31 = cases.sweep64(n=64, steps=steps, precision="fp64")[31]
    o = oracle.Oracle(c31)
    o.initialize()
    def oracle_step():
        try:
            o.march(1)
        except oracle.OracleError:  # reports the non-finite state of the step before and goes on
            pass

    n_oracle = first_nonfinite(oracle_step, lambda: np.all(np.isfinite(o.field(0))), 80)
    o.close()
    sv = nsfsi.Solver(c31)
    sv.initialize()

    def dev_step():
        try:
            sv.marchTimeStep(1)
        except nsfsi.NSError:
            pass

    n_dev = first_nonfinite(dev_step, lambda: np.all(np.isfinite(sv.mesh.field("rho"))), 80)
    sv.close()
    print(f"case 31 FP64: first non-finite step oracle {n_oracle}, device {n_dev}", flush=True)
    assert n_oracle == n_dev == 70
