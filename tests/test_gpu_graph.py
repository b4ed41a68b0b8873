"""Recorded launch sequences (hipGraph) of the diffusion solve against eager launches.

The solver records the launches it enqueues ahead of its first status poll
(the mixed solve's blind sweeps, the FP64 solve's residual and first chunk of
iterations) as a graph the second time it meets the same schedule and replays
it afterwards when the tissue is created with MONO_GRAPH=1; otherwise it
launches everything eagerly.  Both run the same steps, in chunks separated by
the host's own synchronising reads, and must agree bit for bit after every
chunk; the graph tissue must actually have replayed graphs.
"""
import os

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _tissue(m, precision, eager, aniso):
    if eager:
        os.environ.pop("MONO_GRAPH", None)
    else:
        os.environ["MONO_GRAPH"] = "1"
    try:
        if aniso:
            nz = 6
            t6 = m.layered_fibre_tensors(nz, -60.0, 60.0, 0.00024, 0.000045)
            return m.Tissue((48, 40, nz), model=m.MODEL_ORD2011, aniso=True, tensor=t6,
                            tensor_source=m.TENSOR_PER_LAYER, precision=precision)
        return m.Tissue((64, 64, 4), model=m.MODEL_TT2006, tensor=m.iso_tensor(0.00013), precision=precision)
    finally:
        os.environ.pop("MONO_GRAPH", None)


@pytest.mark.parametrize("aniso", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_replay_matches_eager(prec, aniso):
    import monoalg3d_c_amd as m
    precision = m.FP64 if prec == "fp64" else m.FP32_REFINE
    announce(f"graph vs eager {prec} aniso={aniso}")
    eager = _tissue(m, precision, True, aniso)
    graph = _tissue(m, precision, False, aniso)
    for chunk in range(12):
        eager.step(25)
        graph.step(25)
        ve, vg = eager.v, graph.v
        assert np.array_equal(ve, vg), (chunk, float(np.abs(ve - vg).max()))
    se, sg = eager.stats, graph.stats
    for k in ("cg_iterations_total", "inner_iterations_total", "outer_sweeps_total"):
        assert se[k] == sg[k], (k, se[k], sg[k])
    le, ce = eager.graph_counts()
    lg, cg = graph.graph_counts()
    assert le == 0 and ce == 0
    assert cg >= 1 and lg >= 100, (lg, cg)
    assert np.array_equal(eager.state, graph.state)
