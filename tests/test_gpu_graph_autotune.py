"""hipGraph-captured bounce loop and the launch autotuner: both leave results unchanged.

A captured graph is replayed several times with eager work and host synchronisation in
between, and every replay is checked against eager execution of the same steps.
"""
import numpy as np
import pytest

from raccoongpurender_amd import render_scenes as R
from tests import wf_helpers as W

pytestmark = pytest.mark.gpu


def _same(a, b):
    ca, pa = a.queue_counts()
    cb, pb = b.queue_counts()
    assert np.array_equal(ca, cb) and pa == pb
    mm = W.state_mismatches(a.state(), b.state())
    assert all(v == 0 for v in mm.values()), mm
    ra, na = a.accumulation()
    rb, nb = b.accumulation()
    assert np.array_equal(na, nb)
    assert np.allclose(ra, rb, rtol=1e-5, atol=1e-6)


def test_graph_replays_match_eager(ctx):
    sd = R.material_test_scene(48, 32)
    eager = W.gpu_lane(ctx, sd, 48 * 32 * 2)
    graph = W.gpu_lane(ctx, sd, 48 * 32 * 2)
    graph.graph_capture(4)
    for rep in range(4):
        graph.graph_launch(1)
        ctx.synchronize()
        eager.step(4)
        ctx.synchronize()
        _same(eager, graph)
    graph.graph_launch(3)
    eager.step(12)
    _same(eager, graph)


@pytest.mark.parametrize("trace_block,qmul", [(64, 1), (128, 3), (256, 16), (0, 2)])
def test_launch_shape_does_not_change_results(ctx, trace_block, qmul):
    sd = R.material_test_scene(48, 32)
    a = W.gpu_lane(ctx, sd, 48 * 32 * 2)
    b = W.gpu_lane(ctx, sd, 48 * 32 * 2)
    b.set_launch("stackless", trace_block, 32 * qmul)
    b.initialize(W.camera(sd["camera"]))
    a.step(10)
    b.step(10)
    _same(a, b)


def test_autotune_picks_a_config_and_results_match_the_oracle(ctx):
    sd = R.material_test_scene(64, 48)
    g = W.gpu_lane(ctx, sd, 64 * 48 * 4)
    ms = g.autotune(W.camera(sd["camera"]), warmup=2, steps=4)
    variant, tb, qb = g.get_launch()
    print(f"autotune: {variant} trace_block={tb} queue_blocks={qb} {ms:.4f} ms/step", flush=True)
    assert ms > 0 and tb in (0, 64, 128, 256)
    o = W.oracle_lane(sd, 64 * 48 * 4, variant=variant)
    for _ in range(6):
        g.step()
        o.step()
    cg, pg = g.queue_counts()
    co, po = o.queue_counts()
    assert np.array_equal(cg, co) and pg == po
    mm = W.state_mismatches(g.state(), o.state())
    assert all(v == 0 for v in mm.values()), mm
