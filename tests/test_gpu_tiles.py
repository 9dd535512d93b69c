"""Fused-stage tile shapes and the autotuner.

Every node's arithmetic is the same whatever block computes it, so the FP64 state after a few steps is
bit-identical for every (rows per block, z-planes per block) pair; the autotuner only times launches into
scratch outputs, so a tuned run equals an untuned one bit for bit.
"""
import numpy as np
import pytest

import cases
import nsfsi

pytestmark = pytest.mark.gpu


def run(cfg, steps, **opts):
    s = nsfsi.Solver(cfg)
    s.initialize()
    for k, v in opts.items():
        s.setOption(k, v)
    s.marchTimeStep(steps)
    U = s.mesh.conserved()
    s.close()
    return U


@pytest.mark.parametrize("N", [41, 64])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tile_shape_invariance(prec, N):
    # N = 41: one x-segment; the last z-chunk of every tile height has an odd plane count (9 or 41).
    # N = 64: the last x-segment holds 2 nodes, so its tiles are half-width (two y-bands per wave), with an
    # odd y-band count at 6 and 14 rows (the last band pair has one band).
    cfg = cases.sphere48(N=N, steps=4, precision=prec)
    ref = run(cfg, 4, tile_ty=8, tile_kc=64)
    for ty in (4, 6, 14):
        for kc in (12, 16, 20, 32, 128):
            print("tile", ty, kc, flush=True)
            U = run(cfg, 4, tile_ty=ty, tile_kc=kc)
            assert np.array_equal(U, ref), (ty, kc)


def test_autotune_leaves_state_unchanged():
    cfg = cases.sphere48(N=32, steps=3)
    ref = run(cfg, 3)
    s = nsfsi.Solver(cfg)
    s.initialize()
    ty, kc = s.autotune(1)
    assert ty in (4, 6, 8, 14) and kc in (12, 16, 20, 24, 32, 40, 48, 64, 96, 128)
    s.marchTimeStep(3)
    assert np.array_equal(s.mesh.conserved(), ref)


@pytest.mark.parametrize("kahan", [1, 0])
def test_fp32_global_access_bitwise(kahan):
    """FP32 states beyond 4 GiB use per-field global addressing instead of buffer resources; forcing that
    form on a small case gives the same state bit for bit (compensated and plain combine)."""
    cfg = cases.sphere48(N=70, steps=6, precision="fp32")  # two x-segments (62 + 8 nodes)
    cfg.kahan = kahan
    ref = run(cfg, 6, tile_ty=6, tile_kc=16)
    U = run(cfg, 6, tile_ty=6, tile_kc=16, global_access=1)
    assert np.array_equal(U, ref)
