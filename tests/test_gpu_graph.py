"""Captured step graphs (knp_mi355x_set_step_graph) against plain launches and the oracle.

A graph holds the launches of several steps with the step read from device memory and is replayed for the
middle of every long call. Each case runs several calls (pushed channel spikes between them, so that eager work
and synchronizations interleave with the replays) with graphs, without graphs and on the oracle: spike trains,
potentials and impact counts must be identical, and the graph must have been replayed.
"""
import sys

import numpy as np
import pytest

from knp_amd import abi, workloads
from knp_amd.backend import MI355XBackend, device_count
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert device_count() >= 1, "no HIP device visible"


def spec_builder(spec, pops=(0,)):
    def build(o):
        spec.load_into(o)
        return list(pops), None
    return build


def run_three(build, chunks, push=None, graph_steps=16, scatter=abi.SCATTER_STEPWISE, rng_seed=3,
              events_vs_oracle=True):
    """build(o) loads the network into o and returns (populations to record, push projection or None); push:
    the projection's channel count (5 % of them pushed before every call). events_vs_oracle: also compare the
    impact counts with the oracle's (a dense input projection counts its impacts a whole time block at a time,
    so only the two device runs are compared then)."""
    runs = {}
    for name in ("graph", "plain", "oracle"):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name != "oracle":
            o.set_scatter_mode(scatter)
            o.set_step_graph(graph_steps if name == "graph" else 0)
        pops, pq = build(o)
        total = sum(chunks)
        for p in pops:
            o.record_spikes(p, total)
        rng = np.random.default_rng(rng_seed)
        for c in chunks:
            if pq is not None:
                idx = np.nonzero(rng.random(push) < 0.05)[0].astype(np.uint32)
                o.push_spikes(pq, idx)
            o.step(c)
        runs[name] = ([o.read_spike_record(p) for p in pops], [o.potentials(p) for p in pops], o.events)
        if name == "graph":
            g, r = o.step_graph()
            say(f"  graph of {g} steps replayed {r} times")
            assert g > 0 and r > 0
        if name == "plain":
            assert o.step_graph() == (0, 0)
    for k in range(len(runs["oracle"][0])):
        so = runs["oracle"][0][k]
        assert so.sum() > 20, k
        for name in ("graph", "plain"):
            assert np.array_equal(runs[name][0][k], so), (name, k, np.argwhere(runs[name][0][k] != so)[:5].tolist())
            assert np.allclose(runs[name][1][k], runs["oracle"][1][k], rtol=1e-6, atol=1e-30), (name, k)
    assert runs["graph"][2] == runs["plain"][2]
    if events_vs_oracle:
        assert runs["plain"][2] == runs["oracle"][2]
    return runs


def test_graph_resource_stdp():
    """The STDP configuration's network, one instance (one launch per step: not tiled), with training."""
    say("case: step graphs, STDP network")
    spec = workloads.stdp_learning(n_images=20, n_in=144, n_out=96)
    run_three(spec_builder(spec), [1, 60, 3, 100, 37])


def test_graph_tiled_stepwise():
    """Every population tiled, one fused launch per step (phase 1 of the next step inside it), pushes."""
    import test_gpu_parity as tp

    say("case: step graphs, tiled populations, one launch per step")
    run_three(lambda o: tp.all_tiled_network(o, 8), [1, 50, 2, 70, 1, 40], push=1200)


def test_graph_dense_time_blocks():
    """Uneven instances with an FP8 dense input projection (time block 32): one graph per time block."""
    say("case: step graphs, dense input projection in time blocks of 32")
    spec = workloads.uneven_inference(n_instances=5, n_min=20, n_max=150, n_in=50, time_block=32, p_in=0.4,
                                      threshold=0.6, w_inh=0.4)
    run_three(spec_builder(spec), [1, 100, 5, 90], events_vs_oracle=False)


def test_graph_mfma_expansion():
    """Projections expanded on MFMA (two-part BF16) in a graph, and the no-allocation step through it."""
    say("case: step graphs, projections expanded on MFMA")
    spec = workloads.batched_stdp(batch=3, n_images=20, dense="skip", split=True, n_in=144, n_out=96)
    runs = {}
    for name in ("graph", "plain"):
        g = MI355XBackend(0)
        g.set_step_graph(16 if name == "graph" else 0)
        spec.load_into(g)
        g._init()
        g.record_spikes(0, 200)
        g.step(50)
        g.step_noalloc(100)
        g.step(50)
        runs[name] = (g.read_spike_record(0), g.potentials(0), g.events, g.step_graph())
    assert runs["graph"][3][1] > 0
    assert runs["graph"][0].sum() > 20
    assert np.array_equal(runs["graph"][0], runs["plain"][0])
    assert np.array_equal(runs["graph"][1], runs["plain"][1])
    assert runs["graph"][2] == runs["plain"][2]
