"""Checkpoint and restart (ns_save / ns_restore): a run saved after k steps and restored into a fresh handle
continues bit for bit as the uninterrupted run does — state, stored primitive fields, t, dt, timeLevel, norm
and force histories — in every stage form (FP64 fused, FP64 stored-primitive, FP32 compensated fused, FP32
unfused), with the periodic density filter firing across the restart and with graph replay; and a checkpoint
is refused by a case of another grid or physics, and when truncated."""
import os

import numpy as np
import pytest

import cases
import nsfsi

pytestmark = pytest.mark.gpu

FORMS = {
    "fp64_fused": ("fp64", 1, 0),
    "fp64_stored_primitives": ("fp64", 0, 0),
    "fp32_compensated_fused": ("fp32", 1, 1),
    "fp32_unfused": ("fp32", 0, 0),
}


def make(form, steps=12):
    prec, kernel, kahan = FORMS[form]
    cfg = cases.sphere48(N=24, steps=steps, precision=prec)
    cfg.kahan = kahan
    cfg.n_filter_intervals = 1
    cfg.filter_intervals[0] = 4 if prec == "fp64" else 0  # the batched/FP32 forms run it too; FP64 checks it
    cfg.n_filter_change_times = 0
    return cfg, kernel


def solver(cfg, kernel, graph):
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", kernel)
    s.setOption("body_force", 1)
    s.setOption("graph", graph)
    return s


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("form", sorted(FORMS))
def test_restart_is_bitwise_continuation(tmp_path, form, graph):
    cfg, kernel = make(form)
    print(f"checkpoint {form} graph={graph}", flush=True)
    a = solver(cfg, kernel, graph)
    a.marchTimeStep(12)
    b = solver(cfg, kernel, graph)
    b.marchTimeStep(5)
    path = tmp_path / "run.ckpt"
    b.save(path)
    b.close()
    cfg2, _ = make(form, steps=40)  # another stopping setting is allowed
    c = solver(cfg2, kernel, graph)
    c.restore(path)
    assert c.time()[1] == 5
    c.marchTimeStep(7)
    assert a.time() == c.time()
    for name in nsfsi.FIELDS:
        assert np.array_equal(a.mesh.field(name), c.mesh.field(name)), name
    ha, hc = a.getConvergenceHistory(), c.getConvergenceHistory()
    for k in ha:
        assert len(hc[k]) == 12 and np.array_equal(ha[k], hc[k]), k
    assert np.array_equal(a.getForceHistory(), c.getForceHistory())
    a.close()
    c.close()


@pytest.mark.parametrize("form", ["fp64_fused", "fp32_compensated_fused"])
def test_restart_with_two_bodies_and_a_body_through_the_symmetry_faces(tmp_path, form):
    """The two-body case (its z-cylinder crosses both symmetry z-faces, whose nodes inside the body hold the
    constant mirrored values of untouched solid nodes): restored from a checkpoint written by a run with
    another initial field seed, the run still continues bit for bit (the checkpoint carries those values)."""
    prec, kernel, kahan = FORMS[form]

    def cfg_of(seed, steps=12):
        c = cases.two_bodies(steps=steps, precision=prec)
        c.kahan = kahan
        c.ic_seed = seed
        return c

    print(f"checkpoint two bodies {form}", flush=True)
    a = solver(cfg_of(23), kernel, 1)
    a.marchTimeStep(9)
    b = solver(cfg_of(23), kernel, 1)
    b.marchTimeStep(4)
    path = tmp_path / "two.ckpt"
    b.save(path)
    b.close()
    c = solver(cfg_of(99, steps=30), kernel, 1)
    c.restore(path)
    c.marchTimeStep(5)
    assert a.time() == c.time()
    for name in nsfsi.FIELDS[:5]:
        assert np.array_equal(a.mesh.field(name), c.mesh.field(name)), name
    a.close()
    c.close()


def test_checkpoint_refused_by_another_case(tmp_path):
    cfg, kernel = make("fp64_fused")
    s = solver(cfg, kernel, 0)
    s.marchTimeStep(2)
    path = tmp_path / "run.ckpt"
    s.save(path)
    s.close()
    for change in ({"NI": 26}, {"Re": 123.0}, {"precision": 1}, {"body_radius": 0.45}):
        other = cases.sphere48(N=24, steps=12)
        for k, v in change.items():
            setattr(other, k, v)
        o = nsfsi.Solver(other)
        o.initialize()
        with pytest.raises(nsfsi.NSError) as e:
            o.restore(path)
        assert e.value.status == 5, change
        o.close()
    data = open(path, "rb").read()
    short = tmp_path / "short.ckpt"
    short.write_bytes(data[: len(data) // 2])
    o = nsfsi.Solver(cfg)
    with pytest.raises(nsfsi.NSError) as e:
        o.restore(path)  # before initialize
    assert e.value.status == 6
    o.initialize()
    with pytest.raises(nsfsi.NSError) as e:
        o.restore(short)
    assert e.value.status == 5
    with pytest.raises(nsfsi.NSError):
        o.restore(tmp_path / "missing.ckpt")
    o.close()
    assert os.path.getsize(path) > 5 * 2 * 24 ** 3 * 8
