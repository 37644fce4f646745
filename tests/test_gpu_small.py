"""GPU tests of the single-workgroup schedule (k_small_run): networks whose neurons all fit on one workgroup's
threads run whole multi-step calls inside one workgroup, rings and spiking rows in LDS.

Each case runs the same network with that schedule (AUTO), with one launch per step (STEPWISE, the tile
kernels) and on the oracle, over calls of 1..70 steps with channel spikes pushed before every call: spike
trains, last spikes and impact counts identical, potentials after every call identical between the device
schedules and within 1e-6 relative of the oracle's double potentials.
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


def full_network(o, seed):
    """Two FULL-state BLIFAT populations (A 400 with a refractory period, postsynaptic traces and a potential
    floor; B 300 with threshold increments) with recurrent and cross projections (delays 1..16, 20 % or 40 %
    INHIBITORY_CURRENT), a pushed channel projection into A and a Bernoulli source into B. Weights on a 2^-8
    grid, so FP32 potentials are exact."""
    rng = np.random.default_rng(seed)
    na, nb = 400, 300
    a = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(
        absolute_refractory_period=2, postsynaptic_trace_decay=0.5, postsynaptic_trace_increment=1.0,
        min_potential=-2.0), na))
    # the dynamic threshold only grows by 0.25 (no decay): the device keeps it in FP32, exact on this grid
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(
        threshold_decay=1.0, threshold_increment=0.25, absolute_refractory_period=1), nb))
    ident = np.arange(na, dtype=np.uint32)
    push = o.add_delta_projection(abi.KNP_NO_POPULATION, na, a, ident, ident, np.full(na, 1.0, np.float32),
                                  np.ones(na, np.uint32))

    def proj(pre_pop, n_pre, post_pop, n_post, S, wlo, whi, dmax, inh):
        o.add_delta_projection(pre_pop, n_pre, post_pop, rng.integers(0, n_pre, S).astype(np.uint32),
                               rng.integers(0, n_post, S).astype(np.uint32),
                               (rng.integers(wlo, whi, S) / 256.0).astype(np.float32),
                               rng.integers(1, dmax + 1, S).astype(np.uint32),
                               rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[1 - inh, inh]).astype(np.uint8))

    proj(a, na, a, na, 20_000, 4, 96, 16, 0.2)
    proj(a, na, b, nb, 12_000, 16, 128, 6, 0.1)
    proj(b, nb, a, na, 2_000, 32, 200, 3, 0.4)
    nch = 300
    src = o.add_delta_projection(abi.KNP_NO_POPULATION, nch, b, np.arange(nch, dtype=np.uint32),
                                 np.arange(nch, dtype=np.uint32), np.full(nch, 1.0, np.float32),
                                 np.ones(nch, np.uint32))
    o.add_bernoulli_source(src, np.full(nch, workloads.bernoulli_threshold(0.03), np.uint32), seed=9)
    return (a, b), push, na


def lean_network(o, seed):
    """The small plain network's shape (1,000 default BLIFAT, 100,000 recurrent synapses with delays 1..5, a
    Bernoulli source) plus a pushed channel projection."""
    spec = workloads.plain_small(seed=seed)
    spec.load_into(o)
    n = 1000
    ident = np.arange(n, dtype=np.uint32)
    push = o.add_delta_projection(abi.KNP_NO_POPULATION, n, 0, ident, ident, np.full(n, 1.0, np.float32),
                                  np.ones(n, np.uint32))
    return (0,), push, n


@pytest.mark.parametrize("build", [full_network, lean_network], ids=["full-state", "default-blifat"])
def test_single_workgroup_matches_tiles_and_oracle(build):
    say(f"case: single-workgroup schedule, {build.__name__}, calls of 1..70 steps with pushes")
    chunks = [1, 9, 70, 2, 1, 33, 4, 50]
    steps = sum(chunks)
    runs = {}
    modes = {"stepwise": abi.SCATTER_STEPWISE, "tiles": abi.SCATTER_TILES}
    for name in ("single", "tiles", "stepwise", "oracle"):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name in modes:
            o.set_scatter_mode(modes[name])
        pops, push, n_push = build(o, 11)
        if name == "single":
            assert o.schedule() == abi.SCHEDULE_SINGLE_WORKGROUP
        if name == "stepwise":
            assert o.schedule() == abi.SCHEDULE_STEPWISE
        if name == "tiles":
            assert o.schedule() == abi.SCHEDULE_PERSISTENT
        for p in pops:
            o.record_spikes(p, steps)
        rng = np.random.default_rng(4)
        mids = []
        for c in chunks:
            idx = np.nonzero(rng.random(n_push) < 0.05)[0].astype(np.uint32)
            o.push_spikes(push, np.concatenate([idx, idx[:3]]))  # a few indexes twice
            o.step(c)
            mids.append(([o.potentials(p).copy() for p in pops], [o.last_spikes(p).copy() for p in pops]))
        runs[name] = ([o.read_spike_record(p) for p in pops], mids, o.events)
    devs = ("single", "tiles", "stepwise")
    for k in range(len(runs["oracle"][0])):
        so = runs["oracle"][0][k]
        for name in devs:
            say(f"  {name} population {k}: first differing (step, neuron) "
                f"{np.argwhere(runs[name][0][k] != so)[:3].tolist()}")
    for k in range(len(runs["oracle"][0])):
        so = runs["oracle"][0][k]
        assert so.sum() > 100, k
        for name in devs:
            assert np.array_equal(runs[name][0][k], so), (name, k, np.argwhere(runs[name][0][k] != so)[:5].tolist())
    # potentials after every call: the device schedules bit for bit alike; against the oracle's double
    # potentials within 1e-6 relative (FP32 state; the full-state network's values are exact on its grid)
    for j in range(len(chunks)):
        for k in range(len(runs["oracle"][0])):
            vo = np.asarray(runs["oracle"][1][j][0][k], np.float64)
            for name in devs:
                vg = np.asarray(runs[name][1][j][0][k], np.float64)
                assert np.array_equal(vg, np.asarray(runs["stepwise"][1][j][0][k], np.float64)), (name, j, k)
                assert np.max(np.abs(vg - vo) / np.maximum(np.abs(vo), 1e-30)) <= 1e-6, (name, j, k)
                assert np.array_equal(runs[name][1][j][1][k], runs["oracle"][1][j][1][k]), (name, j, k)
    say(f"  events: {[runs[x][2] for x in devs + ('oracle',)]}")
    assert runs["single"][2] == runs["tiles"][2] == runs["stepwise"][2] == runs["oracle"][2]


def test_too_large_for_one_workgroup_stays_persistent():
    """1,200 + 600 neurons do not fit one workgroup's threads: the cooperative persistent schedule runs."""
    import test_gpu_parity as tp

    say("case: schedule of a network larger than one workgroup")
    g = MI355XBackend(0)
    tp.all_tiled_network(g, 8)
    assert g.schedule() == abi.SCHEDULE_PERSISTENT
