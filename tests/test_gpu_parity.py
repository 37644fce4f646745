"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle on the same seeded inputs.

Bar (integer/index work and default BLIFAT): spike trains bit-exact, SynapticImpact counts equal.
Floating point: FP32 device state against the oracle's double state within the tolerances written in each
test (the bar for potentials: 1e-6 relative).
"""
import sys

import numpy as np
import pytest

from knp_amd import abi, workloads
from knp_amd.backend import BackendError, MI355XBackend, device_count
from oracle.pyoracle import Oracle

import netfile
from networks import lif_network

pytestmark = pytest.mark.gpu

POT_RTOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert device_count() >= 1, "no HIP device visible"


# Synaptic-resource training runs a workgroup per candidate in populations up to a size (the default limit
# covers every network here) and a wave per candidate in larger ones; limit 0 forces the wave form.
TRAIN_FORMS = {"workgroup": None, "wave": 0}


def device(train="workgroup"):
    g = MI355XBackend(0)
    if TRAIN_FORMS[train] is not None:
        g.set_train_workgroup_limit(TRAIN_FORMS[train])
    return g


def both(spec, train="workgroup"):
    return spec.load_into(device(train)), spec.load_into(Oracle("exact"))


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30))) if a.size else 0.0


def smallest(o, proj_kind="delta"):
    pop = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 1))
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1])
    o.add_delta_projection(pop, 1, pop, [0], [0], [1.1], [6])
    return pop, inp


def run_smallest(o, pop, inp, steps=20):
    res = []
    for step in range(steps):
        if step % 5 == 0:
            o.push_spikes(inp, [0])
        o.step()
        if len(o.last_spikes(pop)):
            res.append(step)
    return res


EXPECTED = [1, 6, 7, 11, 12, 13, 16, 17, 18, 19]


def test_smallest_network():
    say("case: SmallestNetwork")
    g = MI355XBackend(0)
    pop, inp = smallest(g)
    assert run_smallest(g, pop, inp) == EXPECTED
    assert g.get_step() == 20


def test_additive_stdp_network():
    say("case: AdditiveSTDPBLIFATNetwork")
    res = {}
    for name, o in (("gpu", MI355XBackend(0)), ("oracle", Oracle("exact"))):
        pop = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 1))
        inp = o.add_additive_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1], [2.0], [2.0])
        loop = o.add_additive_stdp_projection(pop, 1, pop, [0], [0], [1.0], [6], [1.0], [1.0],
                                              stdp_pops=[pop], stdp_modes=[abi.STDP_AND_SPIKE])
        res[name] = (run_smallest(o, pop, inp), o.weights(loop)[0], o.weights(inp)[0])
    assert res["gpu"][0] == EXPECTED
    assert res["gpu"][2] == 1.0
    assert res["gpu"][1] == pytest.approx(res["oracle"][1], rel=1e-6)


def resource_smallest(o, model="blifat"):
    """ResourceSTDPBLIFATNetwork / ResourceSTDPAltAILIFNetwork (single_thread_cpu_test.cpp:179-358)."""
    kw = dict(synaptic_resource_threshold=1.0, free_synaptic_resource=2.0, isi_max=0)
    if model == "altai":
        pop = o.add_resource_altai_lif_population(abi.neuron_array(abi.resource_altai_lif_defaults(**kw), 1))
    else:
        pop = o.add_resource_blifat_population(abi.neuron_array(abi.resource_blifat_defaults(**kw), 1))
    r_in = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=1.0, w_max=2.0, d_u=0.1), 1)
    r_loop = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=1.0, w_max=2.0), 1)
    inp = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1], r_in)
    loop = o.add_resource_stdp_projection(pop, 1, pop, [0], [0], [1.0], [6], r_loop)
    o.start_learning()
    return pop, inp, loop


@pytest.mark.parametrize("model", ["blifat", "altai"])
def test_resource_stdp_network(model):
    say(f"case: ResourceSTDP network, {model} neurons")
    g, o = MI355XBackend(0), Oracle("exact")
    gp, gi, gl = resource_smallest(g, model)
    op, oi, ol = resource_smallest(o, model)
    assert run_smallest(g, gp, gi) == EXPECTED
    assert run_smallest(o, op, oi) == EXPECTED
    for proj_g, proj_o in ((gi, oi), (gl, ol)):
        assert np.array_equal(g.weights(proj_g), o.weights(proj_o).astype(np.float32))
        rg, ro = g.resource_rules(proj_g)[0], o.resource_rules(proj_o)[0]
        assert (rg.synaptic_resource, rg.last_spike_step, rg.had_hebbian_update, rg.has_contributed) == \
            (ro.synaptic_resource, ro.last_spike_step, ro.had_hebbian_update, ro.has_contributed)
    read = "resource_altai_lif_neurons" if model == "altai" else "resource_neurons"
    ng, no = getattr(g, read)(gp)[0], getattr(o, read)(op)[0]
    for f in ("free_synaptic_resource", "stability", "isi_status", "last_step", "last_spike_step",
              "first_isi_spike"):
        assert getattr(ng, f) == getattr(no, f), f
    if model == "altai":
        for f in ("potential", "activity_time", "additional_threshold"):
            assert getattr(ng.base, f) == getattr(no.base, f), f
    assert g.weights(gl)[0] != 1.0


@pytest.mark.parametrize("train", list(TRAIN_FORMS))
def test_resource_stdp_hand_traces_on_device(train):
    """The hand-computed synaptic-resource STDP traces (tests/resource_traces.py) on the device: every
    expected resource, weight, free resource, stability, ISI field and flag, bit for bit."""
    import resource_traces

    say(f"case: synaptic-resource STDP hand traces on the device, a {train} per candidate")
    resource_traces.run_all_traces(lambda: device(train))


@pytest.mark.parametrize("train", list(TRAIN_FORMS))
def test_resource_free_negative_zero(train):
    """The free-resource replay skips zero-d_u additions except when free is -0 (-0 + +0 = +0 in the
    reference's float order): a neuron loaded with free = -0.0 and d_u = 0 must follow the oracle bit for bit."""
    import struct
    say("case: free synaptic resource starting at -0.0 with d_u = 0")
    runs = []
    for o in (device(train), Oracle("exact")):
        cell = abi.resource_blifat_defaults(free_synaptic_resource=-0.0, isi_max=3, d_h=0.25)
        pop = o.add_resource_blifat_population(abi.neuron_array(cell, 1))
        rules = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=1.0, w_max=2.0, d_u=0.0,
                                                          dopamine_plasticity_period=2), 3)
        inp = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 3, pop, [0, 1, 2], [0, 0, 0], [1.0] * 3, [1] * 3,
                                             rules)
        o.start_learning()
        bits = []
        for t in range(12):
            if t % 4 == 0:
                o.push_spikes(inp, [0, 2])
            o.step()
            bits.append(struct.pack("<f", o.resource_neurons(pop)[0].free_synaptic_resource))
        runs.append(bits)
    assert runs[0] == runs[1]
    assert runs[1][0] == struct.pack("<f", -0.0) and runs[1][-1] != struct.pack("<f", -0.0)


def resource_altai_network(o, seed, varied=False):
    """Synaptic-resource AltAI-LIF population (256 neurons, leak -1, threshold 4) trained by a plastic
    channel projection (64 channels, Bernoulli source), with lateral inhibition and a dopamine channel.
    varied: w_min, w_max, d_u, the dopamine plasticity period and the delay differ from synapse to synapse
    (else every synapse carries the same ones)."""
    rng = np.random.default_rng(seed)
    n, nch, S = 256, 64, 3000
    base = abi.altai_lif_defaults(activation_threshold=4, potential_leak=-1)
    cell = abi.resource_altai_lif_defaults(base, synaptic_resource_threshold=1.5, free_synaptic_resource=1.0,
                                           isi_max=4, d_h=0.25, dopamine_plasticity_time=4,
                                           stability_change_parameter=0.5, stability_change_at_isi=0.125)
    pop = o.add_resource_altai_lif_population(abi.neuron_array(cell, n))
    rules = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=0.5, w_max=3.0, d_u=0.0625,
                                                      dopamine_plasticity_period=5), S)
    delays = np.ones(S, np.uint32)
    if varied:
        for i in range(S):
            rules[i].w_min = 0.5 - 0.125 * (i % 3)
            rules[i].w_max = 3.0 - 0.25 * (i % 5)
            rules[i].d_u = 0.0625 * (1 + i % 2) if i % 7 else 0.0
            rules[i].dopamine_plasticity_period = 4 + i % 3
        delays = (1 + np.arange(S) % 2).astype(np.uint32)
    pre = rng.integers(0, nch, S).astype(np.uint32)
    post = rng.integers(0, n, S).astype(np.uint32)
    inp = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, nch, pop, pre, post, np.full(S, 0.5, np.float32),
                                         delays, rules)
    o.add_bernoulli_source(inp, np.full(nch, workloads.bernoulli_threshold(0.15), np.uint32), seed=21)
    L = 4000
    o.add_delta_projection(pop, n, pop, rng.integers(0, n, L).astype(np.uint32), rng.integers(0, n, L).astype(np.uint32),
                           np.full(L, 1.0, np.float32), rng.integers(1, 3, L).astype(np.uint32),
                           np.full(L, abi.INHIBITORY_CURRENT, np.uint8))
    dop = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, np.zeros(n, np.uint32), np.arange(n, dtype=np.uint32),
                                 np.full(n, 0.5, np.float32), np.ones(n, np.uint32), np.full(n, abi.DOPAMINE, np.uint8))
    o.start_learning()
    return pop, inp, dop


@pytest.mark.parametrize("varied", [False, True], ids=["shared-rule", "varied-rules"])
@pytest.mark.parametrize("train", list(TRAIN_FORMS))
def test_resource_altai_training_matches_oracle(train, varied):
    """Synaptic-resource AltAI-LIF training (altai_stdp.h:35-149) over 300 steps with dopamine every 10
    steps: spike trains, weights and the neurons' training state identical."""
    say(f"case: resource AltAI-LIF training, 300 steps, a {train} per candidate, varied rules {varied}")
    runs = {}
    for name in ("gpu", "oracle"):
        o = Oracle("exact") if name == "oracle" else device(train)
        pop, inp, dop = resource_altai_network(o, 5, varied)
        o.record_spikes(pop, 300)
        for t in range(0, 300, 10):
            o.step(9)
            o.push_spikes(dop, [0])
            o.step(1)
        runs[name] = (o.read_spike_record(pop), o.weights(inp), o.resource_altai_lif_neurons(pop))
    sg, so = runs["gpu"][0], runs["oracle"][0]
    agr = agreement(sg, so)
    wg, wo = runs["gpu"][1], runs["oracle"][1]
    say(f"  spikes gpu={sg.sum()} oracle={so.sum()} agreement={agr:.6f} "
        f"weight max abs diff {np.max(np.abs(wg - wo)):.3e}")
    assert so.sum() > 300
    assert np.any(wo != 0.5)
    assert np.array_equal(sg, so), np.argwhere(sg != so)[:5].tolist()
    # training replays the reference's float operations in their order: weights and neuron state bit for bit
    assert np.array_equal(np.asarray(wg, np.float32), np.asarray(wo, np.float32))
    ng, no = runs["gpu"][2], runs["oracle"][2]
    for f in ("isi_status", "free_synaptic_resource", "stability", "last_spike_step", "first_isi_spike"):
        assert [getattr(x, f) for x in ng] == [getattr(x, f) for x in no], f


def test_plain_small_bit_exact():
    """Config 0 over its full 1,000 steps: spikes bit-exact, potentials within 1e-6 relative."""
    say("case: plain_small 1000 steps")
    spec = workloads.plain_small()
    g, o = both(spec)
    steps = spec.meta["steps"]
    pop = spec.meta["pop"]
    g.record_spikes(pop, steps)
    o.record_spikes(pop, steps)
    g.step(steps)
    o.step(steps)
    sg, so = g.read_spike_record(pop), o.read_spike_record(pop)
    assert sg.shape == (steps, 1000)
    assert so.sum() > 10_000
    assert np.array_equal(sg, so), np.argwhere(sg != so)[:10]
    assert rel_err(g.potentials(pop), o.potentials(pop)) <= POT_RTOL
    assert g.events == o.events
    ng, no = g.blifat_neurons(pop), o.blifat_neurons(pop)
    assert [x.n_time_steps_since_last_firing for x in ng] == [x.n_time_steps_since_last_firing for x in no]
    assert [x.total_blocking_period for x in ng] == [x.total_blocking_period for x in no]


def test_determinism():
    say("case: determinism (two identical runs)")
    outs = []
    for _ in range(2):
        spec = workloads.plain_small(seed=9, n=500, n_syn=40_000)
        g = spec.load_into(MI355XBackend(0))
        g.record_spikes(0, 300)
        g.step(300)
        outs.append((g.read_spike_record(0), g.potentials(0).tobytes(), g.events))
        g.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1:] == outs[1][1:]


def random_network(o, rng, n, S, steps, full=False, blocking=False):
    """A random recurrent network; full=True adds non-default BLIFAT parameters whose FP32 arithmetic is
    exact (powers of two, weights on a 2^-8 grid)."""
    if full:
        base = abi.blifat_defaults(potential_decay=0.5, threshold_decay=0.5, threshold_increment=0.5,
                                   inhibitory_conductance_decay=0.5, reversal_inhibitory_potential=-0.25,
                                   absolute_refractory_period=1, bursting_period=1, reflexive_weight=0.25,
                                   min_potential=-4.0)
    else:
        base = abi.blifat_defaults()
    pop = o.add_blifat_population(abi.neuron_array(base, n))
    ident = np.arange(n, dtype=np.uint32)
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, n, pop, ident, ident, np.full(n, 1.25, np.float32),
                                 np.ones(n, np.uint32))
    pre = rng.integers(0, n, S).astype(np.uint32)
    post = rng.integers(0, n, S).astype(np.uint32)
    w = (rng.integers(1, 64, S) / 256.0).astype(np.float32)
    d = rng.integers(1, 8, S).astype(np.uint32)
    types = [abi.EXCITATORY, abi.INHIBITORY_CURRENT]
    if full:
        types += [abi.INHIBITORY_CONDUCTANCE, abi.DOPAMINE]
    otype = rng.choice(types, S).astype(np.uint8)
    o.add_delta_projection(pop, n, pop, pre, post, w, d, otype)
    if blocking:
        # two blocking projections racing for the same targets with different delays
        for k in range(2):
            m = 30
            bp = rng.integers(0, n, m).astype(np.uint32)
            bq = rng.integers(0, n, m).astype(np.uint32)
            bw = (-(rng.integers(1, 4, m))).astype(np.float32)
            bd = rng.integers(1, 4, m).astype(np.uint32)
            o.add_delta_projection(pop, n, pop, bp, bq, bw, bd, np.full(m, abi.BLOCKING, np.uint8))
    return pop, inp


def drive_both(g, o, pop, inp, rng, n, steps, rate):
    g.record_spikes(pop, steps)
    o.record_spikes(pop, steps)
    for t in range(steps):
        idx = np.nonzero(rng.random(n) < rate)[0].astype(np.uint32)
        if len(idx):
            g.push_spikes(inp, idx)
            o.push_spikes(inp, idx)
        g.step()
        o.step()
    return g.read_spike_record(pop), o.read_spike_record(pop)


@pytest.mark.parametrize("seed", [0, 1])
def test_random_default_network(seed):
    say(f"case: random default network seed={seed}")
    rng = np.random.default_rng(100 + seed)
    g, o = MI355XBackend(0), Oracle("exact")
    n, S, steps = 300, 6000, 150
    state = rng.bit_generator.state
    pg, ig = random_network(g, rng, n, S, steps)
    rng.bit_generator.state = state
    po, io = random_network(o, rng, n, S, steps)
    sg, so = drive_both(g, o, pg, ig, np.random.default_rng(seed), n, steps, 0.05)
    assert so.sum() > 100
    assert np.array_equal(sg, so)
    assert np.array_equal(g.potentials(pg).astype(np.float64), o.potentials(po))
    assert g.events == o.events


@pytest.mark.parametrize("blocking", [False, True])
def test_random_full_network(blocking):
    """Non-default BLIFAT features (decays, conductance, dopamine, refractory, bursting, threshold
    increments, blocking) on exactly representable values."""
    say(f"case: random full-feature network blocking={blocking}")
    rng = np.random.default_rng(7)
    g, o = MI355XBackend(0), Oracle("exact")
    n, S, steps = 200, 3000, 120
    state = rng.bit_generator.state
    pg, ig = random_network(g, rng, n, S, steps, full=True, blocking=blocking)
    rng.bit_generator.state = state
    po, io = random_network(o, rng, n, S, steps, full=True, blocking=blocking)
    sg, so = drive_both(g, o, pg, ig, np.random.default_rng(3), n, steps, 0.08)
    assert so.sum() > 50
    assert np.array_equal(sg, so), np.argwhere(sg != so)[:10]
    ng, no = g.blifat_neurons(pg), o.blifat_neurons(po)
    for f in ("potential", "dynamic_threshold", "inhibitory_conductance", "dopamine_value"):
        assert rel_err([getattr(x, f) for x in ng], [getattr(x, f) for x in no]) <= POT_RTOL, f
    for f in ("n_time_steps_since_last_firing", "bursting_phase", "total_blocking_period"):
        assert [getattr(x, f) for x in ng] == [getattr(x, f) for x in no], f
    assert g.events == o.events


def mixed_network(o, seed):
    """Populations on both summation paths: A (default BLIFAT, delays up to 20) only receives plain Delta
    EXCITATORY / INHIBITORY_CURRENT impacts, so it is tiled: long recurrent rows (tile directories), short rows
    from B (read whole), host pushes and a Bernoulli source over scattered targets; B (dopamine input) and
    C (a delay of 40) stay on the atomic path."""
    rng = np.random.default_rng(seed)
    na, nb, nc = 1500, 700, 300
    a = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), na))
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), nb))
    c = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), nc))
    ident = np.arange(na, dtype=np.uint32)
    push = o.add_delta_projection(abi.KNP_NO_POPULATION, na, a, ident, ident, np.full(na, 1.0, np.float32),
                                  np.ones(na, np.uint32))
    S = 60_000
    o.add_delta_projection(a, na, a, rng.integers(0, na, S).astype(np.uint32), rng.integers(0, na, S).astype(np.uint32),
                           (rng.integers(1, 48, S) / 256.0).astype(np.float32), rng.integers(1, 21, S).astype(np.uint32),
                           rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[0.8, 0.2]).astype(np.uint8))
    S2 = 1400
    o.add_delta_projection(b, nb, a, rng.integers(0, nb, S2).astype(np.uint32), rng.integers(0, na, S2).astype(np.uint32),
                           (rng.integers(1, 64, S2) / 128.0).astype(np.float32), rng.integers(1, 6, S2).astype(np.uint32))
    S3 = 40_000
    o.add_delta_projection(a, na, b, rng.integers(0, na, S3).astype(np.uint32), rng.integers(0, nb, S3).astype(np.uint32),
                           (rng.integers(64, 160, S3) / 128.0).astype(np.float32), rng.integers(1, 6, S3).astype(np.uint32),
                           rng.choice([abi.EXCITATORY, abi.DOPAMINE], S3, p=[0.7, 0.3]).astype(np.uint8))
    S4 = 6000
    o.add_delta_projection(a, na, c, rng.integers(0, na, S4).astype(np.uint32), rng.integers(0, nc, S4).astype(np.uint32),
                           (rng.integers(1, 64, S4) / 128.0).astype(np.float32), rng.integers(30, 41, S4).astype(np.uint32))
    nch = 900
    src = o.add_delta_projection(abi.KNP_NO_POPULATION, nch, a, np.arange(nch, dtype=np.uint32),
                                 rng.integers(0, na, nch).astype(np.uint32), np.full(nch, 0.75, np.float32),
                                 rng.integers(1, 4, nch).astype(np.uint32))
    o.add_bernoulli_source(src, np.full(nch, workloads.bernoulli_threshold(0.03), np.uint32), seed=11)
    return (a, b, c), push


def test_tiled_scatter_matches_atomic_and_oracle():
    say("case: tile-owner scatter vs atomic scatter vs oracle (mixed network)")
    steps = 200
    runs = {}
    for name in ("tiled", "atomic", "oracle"):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name == "atomic":
            o.set_scatter_mode(abi.SCATTER_ATOMIC)
        pops, push = mixed_network(o, 5)
        if name == "tiled":
            assert [o.tiled(p) for p in pops] == [True, False, False]
        if name == "atomic":
            assert not any(o.tiled(p) for p in pops)
        for p in pops:
            o.record_spikes(p, steps)
        rng = np.random.default_rng(21)
        for t in range(steps):
            idx = np.nonzero(rng.random(1500) < 0.02)[0].astype(np.uint32)
            if len(idx):
                o.push_spikes(push, idx)
            o.step()
        runs[name] = ([o.read_spike_record(p) for p in pops], [o.potentials(p) for p in pops], o.events)
        if name == "tiled":
            parts, cells = o.tile_counters()
            say(f"  tiled path: {parts} row parts, {cells} ring cells")
            assert parts > 0 and cells > 0
    for k in range(3):
        so = runs["oracle"][0][k]
        assert so.sum() > 50, k
        for name in ("tiled", "atomic"):
            assert np.array_equal(runs[name][0][k], so), (name, k)
            assert rel_err(runs[name][1][k], runs["oracle"][1][k]) <= POT_RTOL, (name, k)
    assert runs["tiled"][2] == runs["atomic"][2] == runs["oracle"][2]


def all_tiled_network(o, seed):
    """Every population tiled (only plain Delta EXCITATORY / INHIBITORY_CURRENT input, delays <= 24): the
    device runs phase 1 of step t + 1 inside the tile kernel of step t. A (default BLIFAT) and B (refractory
    period, postsynaptic traces, potential floor: the FULL state path) with recurrent, cross and channel
    projections and a Bernoulli source. Potentials stay sums of weights on a 2^-8 grid, so FP32 is exact."""
    rng = np.random.default_rng(seed)
    na, nb = 1200, 600
    a = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), na))
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(
        absolute_refractory_period=2, postsynaptic_trace_decay=0.5, postsynaptic_trace_increment=1.0,
        min_potential=-2.0), nb))
    ident = np.arange(na, dtype=np.uint32)
    push = o.add_delta_projection(abi.KNP_NO_POPULATION, na, a, ident, ident, np.full(na, 1.0, np.float32),
                                  np.ones(na, np.uint32))
    def proj(pre_pop, n_pre, post_pop, n_post, S, wlo, whi, dmax, inh):
        o.add_delta_projection(pre_pop, n_pre, post_pop, rng.integers(0, n_pre, S).astype(np.uint32),
                               rng.integers(0, n_post, S).astype(np.uint32),
                               (rng.integers(wlo, whi, S) / 256.0).astype(np.float32),
                               rng.integers(1, dmax + 1, S).astype(np.uint32),
                               rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[1 - inh, inh]).astype(np.uint8))
    proj(a, na, a, na, 48_000, 4, 96, 24, 0.2)
    proj(a, na, b, nb, 30_000, 16, 128, 6, 0.1)
    proj(b, nb, a, na, 1_000, 32, 200, 3, 0.5)
    nch = 600
    src = o.add_delta_projection(abi.KNP_NO_POPULATION, nch, b, np.arange(nch, dtype=np.uint32),
                                 np.arange(nch, dtype=np.uint32) % nb, np.full(nch, 1.0, np.float32),
                                 np.ones(nch, np.uint32))
    o.add_bernoulli_source(src, np.full(nch, workloads.bernoulli_threshold(0.02), np.uint32), seed=5)
    return (a, b), push


def test_fused_steps_match_oracle():
    """All populations tiled: persistent multi-step launches (auto), one fused launch per step (stepwise) and
    the atomic path against the oracle, over calls of 1..40 steps with pushed spikes between calls."""
    say("case: fused tile steps (all populations tiled), calls of 1..40 steps with pushes between calls")
    chunks = [1, 7, 40, 2, 1, 25, 3, 31, 1, 9]
    steps = sum(chunks)
    runs = {}
    modes = {"persistent": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE, "atomic": abi.SCATTER_ATOMIC}
    for name in ("persistent", "stepwise", "atomic", "oracle"):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name != "oracle":
            o.set_scatter_mode(modes[name])
            o.enable_timing(True)
        pops, push = all_tiled_network(o, 8)
        if name in ("persistent", "stepwise"):
            assert all(o.tiled(p) for p in pops)
        for p in pops:
            o.record_spikes(p, steps)
        rng = np.random.default_rng(4)
        mids = []
        for c in chunks:
            idx = np.nonzero(rng.random(1200) < 0.05)[0].astype(np.uint32)
            o.push_spikes(push, idx)
            o.step(c)
            mids.append((o.potentials(pops[0]).copy(), o.last_spikes(pops[1]).copy()))
        runs[name] = ([o.read_spike_record(p) for p in pops], mids, o.events)
        if name == "persistent":
            assert o.kernel_time(abi.KERNEL_TILE_RUN)[1] > 0  # the persistent launch ran
        if name == "stepwise":
            assert o.kernel_time(abi.KERNEL_TILE_RUN)[1] == 0
    devs = ("persistent", "stepwise", "atomic")
    for k in range(2):
        so = runs["oracle"][0][k]
        assert so.sum() > 100, k
        for name in devs:
            assert np.array_equal(runs[name][0][k], so), (name, k)
    for j in range(len(chunks)):
        for name in devs:
            assert rel_err(runs[name][1][j][0], runs["oracle"][1][j][0]) <= POT_RTOL, (name, j)
            assert np.array_equal(runs[name][1][j][1], runs["oracle"][1][j][1]), (name, j)
    assert runs["persistent"][2] == runs["stepwise"][2] == runs["atomic"][2] == runs["oracle"][2]


def test_persistent_wait_ends_with_an_error():
    """A persistent launch in which one workgroup withholds its spike counts (testing hook, a small bound on
    every wait): the waits end, the call reports the failure instead of hanging, and a new network on the
    same device steps normally afterwards."""
    say("case: persistent launch with a withheld segment (bounded waits end with an error)")
    g = MI355XBackend(0)
    pops, push = all_tiled_network(g, 8)
    g.step(3)
    assert g.tiled(pops[0])
    g.set_wait_test(0, 4096)
    with pytest.raises(BackendError, match="timed out"):
        g.step(20)
    g.close()
    o, h = MI355XBackend(0), Oracle("exact")
    for x in (o, h):
        all_tiled_network(x, 8)
        x.record_spikes(0, 12)
        x.step(12)
    assert np.array_equal(o.read_spike_record(0), h.read_spike_record(0))


def test_persistent_wait_error_survives_later_launches():
    """One call split into several persistent launches (a dense projection starts a new launch at every
    32-step time block); a workgroup withholds its spike counts only before step 10, so only the first launch
    gives up a wait: the later launches must not hide the failure."""
    say("case: a wait given up in the first of several persistent launches of one call")
    g = MI355XBackend(0)
    pops, push = all_tiled_network(g, 8)
    n = g.pop_sizes[pops[0]]
    w = abi.bf16_encode(np.full((n, 16), 2.0 ** -5, np.float32))
    g.add_dense_input_projection(pops[0], [0], [n], 16, abi.DENSE_BF16, w, None,
                                 workloads.bernoulli_threshold(0.1), seed=3, time_block=32)
    g.step(3)
    assert g.tiled(pops[0]) and g.dense_time_block(0) == 32
    g.set_wait_test(0, 4096, hold_until=10)
    with pytest.raises(BackendError, match="timed out"):
        g.step(100)
    g.close()


def heavy_parts_network(o, seed):
    """Every population tiled, with the persistent kernel's rarer paths in use: more than 512 spiking rows per
    tile and step (a 2,048-channel source at p = 0.3 read whole by every tile: several gather passes), row
    parts of 300 records (200 of delay 1) and of 270 records of delay 1 (more than 255: the whole part is
    added before phase 1), long parts on both sides of phase 1, a ranged directory (rows of A spanning one tile of the
    16-tile population B), pushed channel messages, source blocks run by the workgroups of both
    populations' tiles, and a sparse B -> A projection. Weights on a 2^-8 grid keep FP32 potentials exact."""
    rng = np.random.default_rng(seed)
    na, nb = 1024, 4096
    a = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), na))
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(min_potential=-4.0), nb))
    nch = 2048
    src = o.add_delta_projection(abi.KNP_NO_POPULATION, nch, a, np.arange(nch, dtype=np.uint32),
                                 (np.arange(nch, dtype=np.uint32) * 7) % na, np.full(nch, 0.375, np.float32),
                                 np.ones(nch, np.uint32))
    o.add_bernoulli_source(src, np.full(nch, workloads.bernoulli_threshold(0.3), np.uint32), seed=11)
    ident = np.arange(na, dtype=np.uint32)
    push = o.add_delta_projection(abi.KNP_NO_POPULATION, na, a, ident, ident, np.full(na, 1.0, np.float32),
                                  np.ones(na, np.uint32))
    # A -> A: every row 300 synapses into tile 0 (200 of delay 1, 100 longer: long parts on both sides of
    # phase 1) and 270 of delay 1 into tile 1 (more than 255: the whole part before phase 1), half inhibitory
    for tile, n_d1, per in ((0, 200, 300), (1, 270, 270)):
        pre = np.repeat(np.arange(na, dtype=np.uint32), per)
        post = (tile * 256 + rng.integers(0, 256, na * per)).astype(np.uint32)
        delay = np.tile(np.concatenate([np.ones(n_d1), rng.integers(2, 6, per - n_d1)]), na).astype(np.uint32)
        w = (rng.integers(1, 3, na * per) / 256.0).astype(np.float32)
        ty = rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], na * per).astype(np.uint8)
        o.add_delta_projection(a, na, a, pre, post, w, delay, ty)
    # A -> A: random, 64 per row, delays 1..8
    S = na * 64
    o.add_delta_projection(a, na, a, rng.integers(0, na, S).astype(np.uint32), rng.integers(0, na, S).astype(np.uint32),
                           (rng.integers(4, 64, S) / 256.0).astype(np.float32), rng.integers(1, 9, S).astype(np.uint32),
                           rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[0.7, 0.3]).astype(np.uint8))
    # A -> B: row i into tile i % 16 of B only (ranged directory), 48 per row, delays 1..3
    S = na * 48
    pre = np.repeat(np.arange(na, dtype=np.uint32), 48)
    post = ((pre % 16) * 256 + rng.integers(0, 256, S)).astype(np.uint32)
    o.add_delta_projection(a, na, b, pre, post, (rng.integers(128, 320, S) / 256.0).astype(np.float32),
                           rng.integers(1, 4, S).astype(np.uint32),
                           rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[0.8, 0.2]).astype(np.uint8))
    # B -> A: 4 per row, so that every workgroup of A also waits for every workgroup of B (the persistent
    # launch needs each segment's readers to be waited for by its writer)
    S = nb * 4
    o.add_delta_projection(b, nb, a, np.repeat(np.arange(nb, dtype=np.uint32), 4), rng.integers(0, na, S).astype(np.uint32),
                           (rng.integers(1, 4, S) / 256.0).astype(np.float32), rng.integers(1, 6, S).astype(np.uint32),
                           rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S).astype(np.uint8))
    return (a, b), push


def test_persistent_rare_paths_match_oracle():
    """heavy_parts_network: persistent launches, one launch per step and atomics against the oracle."""
    say("case: persistent kernel rare paths (several passes, parts of 300 records, ranged directory)")
    chunks = [1, 12, 3, 20]
    steps = sum(chunks)
    runs = {}
    modes = {"persistent": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE, "atomic": abi.SCATTER_ATOMIC}
    for name in ("persistent", "stepwise", "atomic", "oracle"):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name != "oracle":
            o.set_scatter_mode(modes[name])
            o.enable_timing(True)
        pops, push = heavy_parts_network(o, 21)
        if name in ("persistent", "stepwise"):
            assert all(o.tiled(p) for p in pops)
        for p in pops:
            o.record_spikes(p, steps)
        rng = np.random.default_rng(9)
        pots = []
        for c in chunks:
            o.push_spikes(push, np.nonzero(rng.random(1024) < 0.1)[0].astype(np.uint32))
            o.step(c)
            pots.append([o.potentials(p).copy() for p in pops])
        runs[name] = ([o.read_spike_record(p) for p in pops], pots, o.events)
        if name == "persistent":
            assert o.kernel_time(abi.KERNEL_TILE_RUN)[1] > 0
    for k in range(2):
        so = runs["oracle"][0][k]
        assert so.sum() > 1000, k
        for name in ("persistent", "stepwise", "atomic"):
            assert np.array_equal(runs[name][0][k], so), (name, k, np.argwhere(runs[name][0][k] != so)[:5])
    for j in range(len(chunks)):
        for k in range(2):
            for name in ("persistent", "stepwise", "atomic"):
                assert rel_err(runs[name][1][j][k], runs["oracle"][1][j][k]) <= POT_RTOL, (name, j, k)
    assert runs["persistent"][2] == runs["stepwise"][2] == runs["atomic"][2] == runs["oracle"][2]


def feed_forward_network(o, back):
    """Population A (512 default BLIFAT) driven by a Bernoulli source, A -> B (512) with delays 1..3 and, with
    `back`, B -> A. Without it no workgroup of A reads a segment that B's workgroups write (weights on a 2^-8
    grid: every sum exact)."""
    rng = np.random.default_rng(5)
    n, S = 512, 20_000
    a = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), n))
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), n))
    ident = np.arange(n, dtype=np.uint32)
    src = o.add_delta_projection(abi.KNP_NO_POPULATION, n, a, ident, ident, np.full(n, 1.0, np.float32),
                                 np.ones(n, np.uint32))
    o.add_bernoulli_source(src, np.full(n, workloads.bernoulli_threshold(0.05), np.uint32), seed=3)
    for pre, post, lo, hi in ((a, b, 64, 192),) + (((b, a, 16, 64),) if back else ()):
        o.add_delta_projection(pre, n, post, rng.integers(0, n, S).astype(np.uint32),
                               rng.integers(0, n, S).astype(np.uint32), (rng.integers(lo, hi, S) / 256.0).astype(np.float32),
                               rng.integers(1, 4, S).astype(np.uint32))
    return a, b


@pytest.mark.parametrize("back", [False, True])
def test_persistent_launch_needs_mutual_waits(back):
    """A persistent launch only runs networks where every workgroup whose segment another reads also reads a
    segment of that reader (else a writer could overwrite step t's segment before a reader has taken it): the
    feed-forward network runs stepwise, the one with a back projection persistent; both equal the oracle."""
    say(f"case: persistent eligibility, back projection {back}")
    steps = 300
    g, o = MI355XBackend(0), Oracle("exact")
    g.set_scatter_mode(abi.SCATTER_TILES)  # (not the one-workgroup schedule, which runs every step in order)
    for x in (g, o):
        a, b = feed_forward_network(x, back)
        x.record_spikes(a, steps)
        x.record_spikes(b, steps)
    assert g.schedule() == (abi.SCHEDULE_PERSISTENT if back else abi.SCHEDULE_STEPWISE)
    g.step(steps)
    o.step(steps)
    for p in (a, b):
        so = o.read_spike_record(p)
        assert so.sum() > 200, p
        assert np.array_equal(g.read_spike_record(p), so), p
    assert g.events == o.events
    g.close()


def agreement(a, b):
    inter = np.logical_and(a, b).sum()
    union = np.logical_or(a, b).sum()
    return 1.0 - (union - inter) / max(1, union)


@pytest.mark.parametrize("train", list(TRAIN_FORMS))
def test_stdp_learning_short(train):
    """Config 2 shape, 100 images (2,000 steps): FP32 device training against the KNP-exact oracle, spikes
    and weights identical."""
    say(f"case: stdp_learning 100 images, a {train} per candidate")
    spec = workloads.stdp_learning(n_images=100)
    g, o = both(spec, train)
    steps = 2000
    g.record_spikes(0, steps)
    o.record_spikes(0, steps)
    g.step(steps)
    o.step(steps)
    sg, so = g.read_spike_record(0), o.read_spike_record(0)
    agr = agreement(sg, so)
    wg, wo = g.weights(0), o.weights(0)
    first = np.argwhere(sg != so)
    say(f"  spikes gpu={sg.sum()} oracle={so.sum()} agreement={agr:.6f} weight max abs diff "
        f"{np.max(np.abs(wg - wo)):.3e} first differing (step, neuron) {first[:1].tolist()}")
    assert so.sum() > 1000
    # the device replays the reference's float order of training: identical spikes and weights
    assert np.array_equal(sg, so)
    assert np.array_equal(np.asarray(wg, np.float32), np.asarray(wo, np.float32))


@pytest.mark.parametrize("train", list(TRAIN_FORMS))
@pytest.mark.parametrize("rows", ["shared", "mixed"])
def test_stdp_loaded_last_spike_steps(train, rows):
    """Synapses loaded with last_spike_step_ values that fall inside the first steps' contribution windows:
    shared by every synapse of a presynaptic row (the row record starts from them), or differing inside some
    rows (each synapse keeps its own value until its row spikes). Config-2 shape at 200 inputs x 256 neurons,
    600 steps: spikes, weights and every rule's state identical to the oracle."""
    import ctypes

    say(f"case: stdp_learning with loaded last spike steps ({rows} by row), a {train} per candidate")
    spec = workloads.stdp_learning(n_images=30, n_in=200, n_out=256)
    op = next(x for x in spec.ops if x[0] == "add_resource_stdp_projection")
    pre, rules = op[1][3], op[1][7]
    raw = np.frombuffer(rules, dtype=np.uint8).reshape(len(pre), ctypes.sizeof(abi.ResourceRule))
    lss = 40 + (pre.astype(np.uint64) * 7) % 29
    if rows == "mixed":
        lss[::5] += 3
    off = abi.ResourceRule.last_spike_step.offset
    raw[:, off:off + 8] = lss.view(np.uint8).reshape(-1, 8)
    g, o = both(spec, train)
    steps = 600
    for x in (g, o):
        x.record_spikes(0, steps)
        x.step(steps)
    sg, so = g.read_spike_record(0), o.read_spike_record(0)
    say(f"  spikes gpu={sg.sum()} oracle={so.sum()}")
    assert so.sum() > 100
    assert np.array_equal(sg, so)
    R = abi.ResourceRule
    dt = np.dtype({"names": ["r", "lss", "hh", "hc"], "formats": [np.float32, np.uint64, np.uint8, np.uint8],
                   "offsets": [R.synaptic_resource.offset, R.last_spike_step.offset, R.had_hebbian_update.offset,
                               R.has_contributed.offset], "itemsize": ctypes.sizeof(R)})
    rg, ro = (np.frombuffer(x.resource_rules(0), dtype=dt) for x in (g, o))
    for f in dt.names:
        assert np.array_equal(rg[f], ro[f]), f
    assert np.array_equal(np.asarray(g.weights(0), np.float32), np.asarray(o.weights(0), np.float32))
    g.close()


@pytest.mark.parametrize("train", list(TRAIN_FORMS))
def test_stdp_thresholds_renormalization_stability(train):
    """Config-2 shape at 200 inputs x 256 neurons with the neuron rules the configuration leaves at their
    defaults switched on: the weight sum raises the threshold (synapse_sum_threshold_coefficient_ 0.02), the
    free resource is renormalized (threshold 1.5, drain coefficient 3) and stability changes (0.05 by dopamine,
    0.02 at ISI starts). 600 steps: spikes, weights, rule states and neuron states identical to the oracle."""
    import ctypes

    say(f"case: stdp_learning with threshold coefficient, renormalization and stability, a {train} per candidate")
    spec = workloads.stdp_learning(n_images=30, n_in=200, n_out=256)
    neurons = next(x for x in spec.ops if x[0] == "add_resource_blifat_population")[1][0]
    for nr in neurons:
        nr.synapse_sum_threshold_coefficient = 0.02
        nr.synaptic_resource_threshold = 1.5
        nr.resource_drain_coefficient = 3
        nr.stability_change_parameter = 0.05
        nr.stability_change_at_isi = 0.02
    g, o = both(spec, train)
    steps = 600
    for x in (g, o):
        x.record_spikes(0, steps)
        x.step(steps)
    sg, so = g.read_spike_record(0), o.read_spike_record(0)
    say(f"  spikes gpu={sg.sum()} oracle={so.sum()}")
    assert so.sum() > 100
    assert np.array_equal(sg, so)
    assert np.array_equal(np.asarray(g.weights(0), np.float32), np.asarray(o.weights(0), np.float32))
    R = abi.ResourceRule
    dt = np.dtype({"names": ["r", "hh", "hc"], "formats": [np.float32, np.uint8, np.uint8],
                   "offsets": [R.synaptic_resource.offset, R.had_hebbian_update.offset, R.has_contributed.offset],
                   "itemsize": ctypes.sizeof(R)})
    rg, ro = (np.frombuffer(x.resource_rules(0), dtype=dt) for x in (g, o))
    for f in dt.names:
        assert np.array_equal(rg[f], ro[f]), f
    ng, no = g.resource_neurons(0), o.resource_neurons(0)
    for f in ("free_synaptic_resource", "stability", "isi_status", "last_step", "last_spike_step", "first_isi_spike"):
        assert [getattr(x, f) for x in ng] == [getattr(x, f) for x in no], f
    assert [x.base.additional_threshold for x in ng] == [np.float32(x.base.additional_threshold) for x in no]
    g.close()


def test_plain_gpu_scale_prefix():
    """Config 1 at full size (100,000 neurons, 1e8 synapses), first 40 steps: spikes bit-exact."""
    say("case: plain_gpu_scale 40 steps (full size)")
    spec = workloads.plain_gpu_scale()
    g = spec.load_into(MI355XBackend(0))
    steps = 40
    g.record_spikes(0, steps)
    g.step(steps)
    sg = g.read_spike_record(0)
    ev_g = g.events
    vg = g.potentials(0)
    g.close()
    o = spec.load_into(Oracle("exact"))
    o.record_spikes(0, steps)
    o.step(steps)
    so = o.read_spike_record(0)
    say(f"  spikes {sg.sum()} events {ev_g}")
    assert so.sum() > 5000
    assert np.array_equal(sg, so)
    assert ev_g == o.events
    assert rel_err(vg, o.potentials(0)) <= POT_RTOL


def test_errors_are_reported():
    say("case: error paths")
    g = MI355XBackend(0)
    pop = g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 4))
    with pytest.raises(BackendError, match="delay"):
        g.add_delta_projection(pop, 4, pop, [0], [1], [0.5], [0])
    with pytest.raises(BackendError, match="out of range"):
        g.add_delta_projection(pop, 4, pop, [9], [1], [0.5], [1])
    with pytest.raises(BackendError, match="Unhandled synapse type"):
        g.add_delta_projection(pop, 4, pop, [0], [1], [0.5], [1], [7])
    q = g.add_delta_projection(pop, 4, pop, [0], [1], [0.5], [1])
    with pytest.raises(BackendError, match="out of the presynaptic range"):
        g.push_spikes(q, [4])
    g.step(2)
    with pytest.raises(BackendError, match="before the first step"):
        g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 4))
    h = MI355XBackend(0)
    h.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 1000))
    h.set_memory_cap(1024)
    with pytest.raises(BackendError, match="cap"):
        h.step(1)


def test_lif_population_matches_oracle():
    """LIFNeuron populations (lif_impl.h:31-84) on the persistent, stepwise and atomic paths against the
    oracle: spikes, potentials and refractory counters identical; EXCITATORY / INHIBITORY_CURRENT only."""
    say("case: LIF + BLIFAT network, persistent / stepwise / atomic, calls of 1..30 steps")
    chunks = [1, 5, 30, 2, 17, 1, 24]
    steps = sum(chunks)
    runs = {}
    modes = {"persistent": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE, "atomic": abi.SCATTER_ATOMIC}
    for name in ("persistent", "stepwise", "atomic", "oracle"):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name != "oracle":
            o.set_scatter_mode(modes[name])
        pops, push = lif_network(o, 21)
        if name in ("persistent", "stepwise"):
            assert all(o.tiled(p) for p in pops)
        for p in pops:
            o.record_spikes(p, steps)
        rng = np.random.default_rng(6)
        mids = []
        for c in chunks:
            o.push_spikes(push, np.nonzero(rng.random(1500) < 0.08)[0].astype(np.uint32))
            o.step(c)
            lf = o.lif_neurons(pops[0])
            mids.append((np.array([x.potential for x in lf]), np.array([x.refract_counter for x in lf])))
        runs[name] = ([o.read_spike_record(p) for p in pops], mids, o.events)
    so = runs["oracle"][0]
    assert so[0].sum() > 200 and so[1].sum() > 50
    for name in ("persistent", "stepwise", "atomic"):
        for k in range(2):
            assert np.array_equal(runs[name][0][k], so[k]), (name, k)
        for j in range(len(chunks)):
            assert np.array_equal(runs[name][1][j][0], runs["oracle"][1][j][0]), (name, j)
            assert np.array_equal(runs[name][1][j][1], runs["oracle"][1][j][1]), (name, j)
        assert runs[name][2] == runs["oracle"][2]
    g = MI355XBackend(0)
    lp = g.add_lif_population(abi.neuron_array(abi.lif_defaults(), 4))
    with pytest.raises(BackendError, match="Unhandled synapse type"):
        g.add_delta_projection(abi.KNP_NO_POPULATION, 4, lp, [0], [1], [0.5], [1], [abi.DOPAMINE])
    with pytest.raises(BackendError, match="not a BLIFAT population"):
        g.blifat_neurons(lp)


def altai_network(o, seed, dopamine):
    """AltAI-LIF population A with per-neuron thresholds, leaks, negative thresholds, resets, activity times
    and flag sets, next to a default BLIFAT population B: recurrent A->A (integer weights, delays up to 8),
    A->B, B->A, host pushes into A, a Bernoulli source into B; dopamine=True adds a DOPAMINE projection A->A
    (A is then summed by atomics). Weights on a 2^-2 grid (2^-8 into B) keep every FP32 sum exact."""
    rng = np.random.default_rng(seed)
    na, nb = 1200, 400
    cells = abi.neuron_array(abi.altai_lif_defaults(), na)
    for i in range(na):
        c = cells[i]
        c.activation_threshold = int(rng.integers(4, 16))
        c.additional_threshold = float(rng.choice([0.0, 0.5]))
        c.negative_activation_threshold = int(rng.integers(6, 30))
        c.potential_leak = int(rng.integers(-2, 3))
        c.potential_reset_value = int(rng.integers(0, 3))
        c.is_diff, c.is_reset, c.leak_rev, c.saturate = (int(b) for b in rng.integers(0, 2, 4))
        c.do_not_save = int(rng.random() < 0.05)
        c.activity_time = int(rng.choice([0, 0, 0, 2, -2, 5]))
        c.potential = float(rng.integers(-4, 8))
    a = o.add_altai_lif_population(cells)
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), nb))
    ident = np.arange(na, dtype=np.uint32)
    push = o.add_delta_projection(abi.KNP_NO_POPULATION, na, a, ident, ident, np.full(na, 6.0, np.float32),
                                  np.ones(na, np.uint32))
    def proj(pre_pop, n_pre, post_pop, n_post, S, w, dmax, types, p):
        o.add_delta_projection(pre_pop, n_pre, post_pop, rng.integers(0, n_pre, S).astype(np.uint32),
                               rng.integers(0, n_post, S).astype(np.uint32), w(S).astype(np.float32),
                               rng.integers(1, dmax + 1, S).astype(np.uint32), rng.choice(types, S, p=p).astype(np.uint8))
    EI = [abi.EXCITATORY, abi.INHIBITORY_CURRENT]
    proj(a, na, a, na, 12_000, lambda S: rng.integers(1, 12, S) / 4.0, 8, EI, [0.55, 0.45])
    proj(a, na, b, nb, 6_000, lambda S: rng.integers(8, 64, S) / 256.0, 4, EI, [0.7, 0.3])
    proj(b, nb, a, na, 3_000, lambda S: rng.integers(1, 20, S) / 4.0, 3, EI, [0.5, 0.5])
    if dopamine:
        proj(a, na, a, na, 2_000, lambda S: rng.integers(1, 8, S) / 4.0, 2, [abi.DOPAMINE], [1.0])
    src = o.add_delta_projection(abi.KNP_NO_POPULATION, nb, b, np.arange(nb, dtype=np.uint32),
                                 np.arange(nb, dtype=np.uint32), np.full(nb, 1.0, np.float32), np.ones(nb, np.uint32))
    o.add_bernoulli_source(src, np.full(nb, workloads.bernoulli_threshold(0.03), np.uint32), seed=13)
    return (a, b), push


@pytest.mark.parametrize("dopamine", [False, True])
def test_altai_lif_population_matches_oracle(dopamine):
    """AltAI-LIF populations (altai_impl.h:31-150) against the oracle: spikes, potentials, pre-impact
    potentials, activity times and dopamine identical, on every summation path the network allows."""
    say(f"case: AltAI-LIF + BLIFAT network dopamine={dopamine}, calls of 1..25 steps")
    chunks = [1, 4, 25, 2, 13, 1, 20]
    steps = sum(chunks)
    names = ("auto", "atomic") if dopamine else ("auto", "stepwise", "atomic")
    modes = {"auto": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE, "atomic": abi.SCATTER_ATOMIC}
    fields = ("potential", "pre_impact_potential", "activity_time", "dopamine_value")
    runs = {}
    for name in names + ("oracle",):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name != "oracle":
            o.set_scatter_mode(modes[name])
        pops, push = altai_network(o, 31, dopamine)
        if name in ("auto", "stepwise"):
            assert o.tiled(pops[0]) == (not dopamine)
        for p in pops:
            o.record_spikes(p, steps)
        rng = np.random.default_rng(8)
        mids = []
        for c in chunks:
            o.push_spikes(push, np.nonzero(rng.random(1200) < 0.06)[0].astype(np.uint32))
            o.step(c)
            x = o.altai_lif_neurons(pops[0])
            mids.append({f: np.array([getattr(v, f) for v in x]) for f in fields})
        runs[name] = ([o.read_spike_record(p) for p in pops], mids, o.events)
    so = runs["oracle"][0]
    assert so[0].sum() > 200 and so[1].sum() > 50
    for name in names:
        for k in range(2):
            assert np.array_equal(runs[name][0][k], so[k]), (name, k)
        for j in range(len(chunks)):
            for f in fields:
                assert np.array_equal(runs[name][1][j][f], runs["oracle"][1][j][f]), (name, j, f)
        assert runs[name][2] == runs["oracle"][2]
    if not dopamine:
        g = MI355XBackend(0)
        ap = g.add_altai_lif_population(abi.neuron_array(abi.altai_lif_defaults(), 4))
        with pytest.raises(BackendError, match="BLOCKING"):
            g.add_delta_projection(abi.KNP_NO_POPULATION, 4, ap, [0], [1], [2.0], [1], [abi.BLOCKING])
        with pytest.raises(BackendError, match="must be a float"):
            g.add_altai_lif_population(abi.neuron_array(abi.altai_lif_defaults(is_diff=1, additional_threshold=0.1), 2))


def test_network_file_round_trip(tmp_path):
    """save_network / load_network (network_io.h:41,51): the file holds the state the device reports, and a
    network loaded from it evolves on the device exactly as the oracle built from the same file does."""
    say("case: network file round trip (LIF + BLIFAT; trained resource AltAI-LIF)")
    g = MI355XBackend(0)
    pops, push = lif_network(g, 21)
    rng = np.random.default_rng(3)
    for _ in range(5):
        g.push_spikes(push, np.nonzero(rng.random(1500) < 0.08)[0].astype(np.uint32))
        g.step(6)
    path = str(tmp_path / "lif.knpnet")
    g.save_network(path)
    fpops, fprojs = netfile.parse(path)
    assert bytes(fpops[0][1]) == bytes(g.lif_neurons(pops[0]))
    assert bytes(fpops[1][1]) == bytes(g.blifat_neurons(pops[1]))
    for q, p in enumerate(fprojs):
        assert np.array_equal(p["weight"], g.weights(q)), q
    g2, o = MI355XBackend(0), Oracle("exact")
    g2.load_network(path)
    netfile.load_into(o, path)
    assert bytes(g2.lif_neurons(0)) == bytes(fpops[0][1])
    steps = 40
    for x in (g2, o):
        x.record_spikes(0, steps)
        x.record_spikes(1, steps)
    for _ in range(steps):
        idx = np.nonzero(rng.random(1500) < 0.08)[0].astype(np.uint32)
        for x in (g2, o):
            x.push_spikes(0, idx)
            x.step()
    for k in range(2):
        sg, so = g2.read_spike_record(k), o.read_spike_record(k)
        assert so.sum() > 50 and np.array_equal(sg, so), k
    assert bytes(g2.lif_neurons(0)) == bytes(o.lif_neurons(0))

    # a trained synaptic-resource network: rules, weights and neuron state survive the file
    g = MI355XBackend(0)
    pop, inp, dop = resource_altai_network(g, 5)
    for _ in range(6):
        g.step(9)
        g.push_spikes(dop, [0])
        g.step(1)
    path = str(tmp_path / "res.knpnet")
    g.save_network(path)
    fpops, fprojs = netfile.parse(path)
    assert bytes(fpops[0][1]) == bytes(g.resource_altai_lif_neurons(pop))
    assert bytes(fprojs[inp]["rules"]) == bytes(g.resource_rules(inp))
    assert np.any(fprojs[inp]["weight"] != 0.5) and fprojs[inp]["locked"] == 0
    g3 = MI355XBackend(0)
    g3.load_network(path)
    assert bytes(g3.resource_altai_lif_neurons(0)) == bytes(fpops[0][1])
    assert bytes(g3.resource_rules(inp)) == bytes(fprojs[inp]["rules"])
    assert np.array_equal(g3.weights(inp), fprojs[inp]["weight"])
    with pytest.raises(BackendError, match="empty handle"):
        g3.load_network(path)

    # additive STDP: time constants, STDP populations and processing types survive the file, and the loaded
    # network (spike-time histories start empty, as after a fresh load in the reference) matches the oracle
    g = MI355XBackend(0)
    pop = g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 1))
    inp = g.add_additive_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1], [2.0], [2.0])
    loop = g.add_additive_stdp_projection(pop, 1, pop, [0], [0], [1.0], [6], [1.0], [1.0],
                                          stdp_pops=[pop], stdp_modes=[abi.STDP_AND_SPIKE])
    run_smallest(g, pop, inp)
    assert g.weights(loop)[0] != 1.0
    path = str(tmp_path / "add.knpnet")
    g.save_network(path)
    fpops, fprojs = netfile.parse(path)
    assert fprojs[loop]["kind"] == abi.PROJ_ADDITIVE and list(fprojs[loop]["stdp_pops"]) == [pop]
    assert list(fprojs[loop]["stdp_modes"]) == [abi.STDP_AND_SPIKE] and fprojs[inp]["tau_plus"][0] == 2.0
    assert fprojs[loop]["weight"][0] == g.weights(loop)[0]
    g4, o = MI355XBackend(0), Oracle("exact")
    g4.load_network(path)
    netfile.load_into(o, path)
    assert run_smallest(g4, 0, inp) == run_smallest(o, 0, inp)
    assert g4.weights(loop)[0] == pytest.approx(o.weights(loop)[0], rel=1e-6)


def test_batched_instances_match_standalone():
    """Instances composed into one device network (config 3 layout, small batch) evolve exactly as the same
    networks run alone on the oracle: composition changes launch shapes, not results."""
    say("case: batched composition, 3 instances x 100 steps")
    B, steps = 3, 100
    spec = workloads.batched_stdp(batch=B, n_images=5)
    g = spec.load_into(MI355XBackend(0))
    g.record_spikes(0, steps)
    g.step(steps)
    sg = g.read_spike_record(0)
    wg = g.weights(0)
    for b in range(B):
        part = workloads.stdp_learning(seed=30 + b, n_images=5, source_seed=30 * 7919 + 17, instance=b)
        o = part.load_into(Oracle("exact"))
        o.record_spikes(0, steps)
        o.step(steps)
        so = o.read_spike_record(0)
        assert so.sum() > 0
        assert agreement(sg[:, b * 1024:(b + 1) * 1024], so) >= 0.999
        n = part.proj_sizes[0]
        assert np.max(np.abs(wg[b * n:(b + 1) * n] - o.weights(0))) <= 1e-6


@pytest.mark.parametrize("fmt,time_block", [(abi.DENSE_FP8_E4M3, 64), (abi.DENSE_BF16, 32), (abi.DENSE_FP8_E4M3, 0)])
def test_dense_input_projection(fmt, time_block):
    """Dense feed-forward projection on MFMA (uneven instances, n_in not a multiple of 16) against the same
    projection stated as a sparse Delta projection on the oracle: spikes bit-exact, impacts counted equal."""
    say(f"case: dense input projection fmt={fmt} T={time_block}")
    spec = workloads.uneven_inference(n_instances=5, n_min=20, n_max=150, n_in=50, fmt=fmt, time_block=time_block,
                                      p_in=0.4, threshold=0.6, w_inh=0.4)
    steps = 150
    g, o = both(spec)
    g.record_spikes(0, steps)
    o.record_spikes(0, steps)
    g.step(steps)
    o.step(steps)
    sg, so = g.read_spike_record(0), o.read_spike_record(0)
    assert so.sum() > 100
    assert np.array_equal(sg, so), np.argwhere(sg != so)[:10]
    assert rel_err(g.potentials(0), o.potentials(0)) <= POT_RTOL
    if time_block:
        # whole time blocks were expanded on the device: compare impact counts at a block boundary
        g.step((-steps) % time_block)
        o.step((-steps) % time_block)
        assert g.events == o.events
    assert g.dense_time_block(0) % 32 == 0


def test_dense_input_projection_at_shape():
    """The uneven configuration's shape on a reduced batch: 16 instances of 200..5,000 neurons (log-uniform),
    784 inputs (seven chunks of 8 K-steps in the MFMA kernel), the autotuned time block, FP8 weights: spikes and
    potentials against the oracle bit for bit, impact counts equal at a time-block boundary."""
    say("case: dense input projection at the configuration's shape, 16 instances")
    spec = workloads.uneven_inference(n_instances=16, steps=150)
    assert spec.meta["n_in"] == 784
    g, o = both(spec)
    steps = 150
    g.record_spikes(0, steps)
    o.record_spikes(0, steps)
    g.step(steps)
    o.step(steps)
    sg, so = g.read_spike_record(0), o.read_spike_record(0)
    say(f"  neurons {spec.pop_sizes[0]}, time block {g.dense_time_block(0)}, spikes {int(so.sum())}")
    assert so.sum() > 1000
    assert np.array_equal(sg, so), np.argwhere(sg != so)[:10]
    assert rel_err(g.potentials(0), o.potentials(0)) <= POT_RTOL
    T = g.dense_time_block(0)
    g.step((-steps) % T)
    o.step((-steps) % T)
    assert g.events == o.events


def test_inexact_weights_are_counted():
    """Weights off the accumulator's grid are counted at load: a float weight of 1e-12 has mantissa bits below
    2^-52 (the fraction bits of a population whose incoming sums stay below 1); weights on a 2^-8 grid do not."""
    say("case: weights off the exact accumulator's grid")
    g = MI355XBackend(0)
    pop = g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 4))
    g.add_delta_projection(abi.KNP_NO_POPULATION, 4, pop, [0, 1, 2, 3], [0, 1, 2, 3],
                           np.array([0.25, 1e-12, 0.5, 3 / 256], np.float32), [1, 1, 1, 1])
    assert g.accumulator_bits(pop) == 52
    assert g.inexact_weights(pop) == 1
    g.close()
    spec = workloads.plain_small(seed=11, n=300, n_syn=20_000)
    g = spec.load_into(MI355XBackend(0))
    say(f"  plain_small: {g.inexact_weights(0)} inexact of {spec.n_synapses} (frac bits {g.accumulator_bits(0)})")
    g.close()

