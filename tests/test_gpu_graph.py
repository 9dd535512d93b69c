"""HIP-graph replay of the time step equals kernel-by-kernel launches bit for bit, replay after replay,
with eager work and host synchronisation between replays."""
import numpy as np
import pytest

import cases
import nsfsi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_replays_match_eager(prec):
    cfg = cases.sphere48(N=32, steps=20, precision=prec)
    eager = nsfsi.Solver(cfg)
    eager.initialize()
    graph = nsfsi.Solver(cfg)
    graph.initialize()
    graph.setOption("graph", 1)
    for n in (1, 2, 3, 1, 4):  # odd counts exercise both buffer parities
        eager.marchTimeStep(n)
        graph.marchTimeStep(n)
        assert graph.time() == eager.time()
        assert np.array_equal(graph.mesh.conserved(), eager.mesh.conserved()), n
        graph.mesh.totals()  # eager device work between replays
    he, hg = eager.getConvergenceHistory(), graph.getConvergenceHistory()
    for k in he:
        assert np.array_equal(he[k], hg[k])
