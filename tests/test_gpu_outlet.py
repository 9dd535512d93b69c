"""The outlet face applied inside the fused stage kernel.

When a row's last x-segment holds at least three nodes, the fused stage kernel writes the outlet face
(OutletBoundary::applyBoundaryCondition, src/Boundary.cpp:304-333) itself, from its own results at NI-2
and NI-3, and the edge-face pass leaves that face out. It is the same expression on the same inputs, so
the state equals, bit for bit, the one the edge-face pass gives (option fused_outlet 0).
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


# 48: one x-segment; 70: the last segment holds 8 nodes (half-width tiles); 100: 38 nodes (full-width);
# 64: 2 nodes, too few for the kernel, so both runs use the edge-face pass
@pytest.mark.parametrize("N", [48, 70, 100, 64])
@pytest.mark.parametrize("prec,kahan", [("fp64", 0), ("fp32", 1), ("fp32", 0)])
def test_fused_outlet_bitwise(N, prec, kahan):
    cfg = cases.sphere48(N=N, steps=5, precision=prec)
    cfg.kahan = kahan
    print(f"case N={N} prec={prec} kahan={kahan}", flush=True)
    ref = run(cfg, 5, fused_outlet=0, tile_ty=6, tile_kc=32)
    U = run(cfg, 5, fused_outlet=1, tile_ty=6, tile_kc=32)
    assert np.array_equal(U, ref)


def test_fused_outlet_unfused_kernel_unaffected():
    """The unfused stage pair always leaves the outlet to the edge-face pass."""
    cfg = cases.sphere48(N=48, steps=4, precision="fp64")
    ref = run(cfg, 4, stage_kernel=0, fused_outlet=0)
    U = run(cfg, 4, stage_kernel=0, fused_outlet=1)
    assert np.array_equal(U, ref)
