"""Pins the CPU oracle (oracle/knp_oracle.c) before it is trusted as the parity reference.

Checks, in order:
  * the reference's own known-answer tests for this path
    (knp/tests/backend/single_thread_cpu_test.cpp:54-264): spike steps of SmallestNetwork,
    AdditiveSTDPBLIFATNetwork and ResourceSTDPBLIFATNetwork, and that their weights change;
  * an exact hand trace of the additive-STDP weight (additive_stdp.h:50-76, 173-196, including the
    unsigned-wrap quirk that makes negative spike-time differences contribute 0);
  * total-resource conservation of synaptic-resource STDP (every transfer in blifat_stdp.h:35-155 moves
    resource between a synapse and the neuron's free pool; renormalization with drain 0 spreads all of it);
  * the closed form of default (memoryless) BLIFAT networks against an independent numpy simulation;
  * hand-computed single-neuron traces for decay, refractory period, conductance, blocking, threshold
    increments and bursting (blifat_impl.h:31-168);
  * hand-computed LIF traces (leak, strict threshold, reset, refractory counter: lif_impl.h:31-84) and a
    random LIF network against an independent numpy simulation of the same rules;
  * the reference's AltAI-LIF known answers (knp/tests/framework/altai_lif_test.cpp:103-196: reversed and
    plain leak, saturation, spike generation, single-neuron impacts) plus hand traces of is_diff, negative
    resets, rounding, do_not_save and activity time (altai_impl.h:31-150), a random AltAI-LIF network
    against an independent numpy simulation, and ResourceSTDPAltAILIFNetwork
    (single_thread_cpu_test.cpp:267-358);
  * the Bernoulli channel hash against an independent numpy implementation.
"""
import numpy as np
import pytest

from knp_amd import abi, workloads
from oracle.pyoracle import Oracle, bernoulli_hash

POLICIES = ["exact", "f64"]


def smallest_network(o, neuron=None, loop_w=1.1):
    """SmallestNetwork (single_thread_cpu_test.cpp:54-92; generators.h:48-57)."""
    pop = o.add_blifat_population(abi.neuron_array(neuron or abi.blifat_defaults(), 1))
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1])
    o.add_delta_projection(pop, 1, pop, [0], [0], [loop_w], [6])
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


@pytest.mark.parametrize("policy", POLICIES)
def test_smallest_network_kat(oracle_built, policy):
    o = Oracle(policy)
    pop, inp = smallest_network(o)
    assert run_smallest(o, pop, inp) == EXPECTED


@pytest.mark.parametrize("policy", POLICIES)
def test_additive_stdp_kat_and_hand_trace(oracle_built, policy):
    """AdditiveSTDPBLIFATNetwork (single_thread_cpu_test.cpp:95-176)."""
    o = Oracle(policy)
    pop = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 1))
    inp = o.add_additive_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1], [2.0], [2.0])
    loop = o.add_additive_stdp_projection(pop, 1, pop, [0], [0], [1.0], [6], [1.0], [1.0],
                                          stdp_pops=[pop], stdp_modes=[abi.STDP_AND_SPIKE])
    assert run_smallest(o, pop, inp) == EXPECTED
    w = o.weights(loop)[0]
    assert w != 1.0
    assert o.weights(inp)[0] == 1.0
    # Hand trace: both histories of the loop synapse receive every population spike (cap tau+ + tau- = 2);
    # when both hold 2 times the weight grows by sum_pre sum_post stdp_w(t_post - t_pre), where
    # t_post < t_pre wraps to a huge positive float and contributes exp(-huge) = 0.
    f32 = np.float32
    dt = f32 if policy == "exact" else np.float64
    expw = dt(1.0)
    for a, b in [(1, 6), (7, 11), (12, 13), (16, 17), (18, 19)]:
        wj = dt(0)
        for tf in (a, b):
            for tn in (a, b):
                diff = dt(float((tn - tf) % 2**64))
                wj = dt(wj + (dt(np.exp(-diff / dt(1))) if diff > 0 else dt(np.exp(diff / dt(1)))))
        expw = dt(expw + wj)
    assert w == pytest.approx(float(expw), rel=1e-6)
    assert abs(float(expw) - (11 + np.exp(-5) + np.exp(-4) + 3 * np.exp(-1))) < 1e-5


def resource_smallest(o):
    """ResourceSTDPBLIFATNetwork (single_thread_cpu_test.cpp:179-264)."""
    neuron = abi.resource_blifat_defaults(synaptic_resource_threshold=1.0, free_synaptic_resource=2.0, isi_max=0)
    pop = o.add_resource_blifat_population(abi.neuron_array(neuron, 1))
    r_in = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=1.0, w_max=2.0, d_u=0.1), 1)
    r_loop = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=1.0, w_max=2.0), 1)
    inp = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1], r_in)
    loop = o.add_resource_stdp_projection(pop, 1, pop, [0], [0], [1.0], [6], r_loop)
    o.start_learning()
    return pop, inp, loop


@pytest.mark.parametrize("policy", POLICIES)
def test_resource_stdp_kat(oracle_built, policy):
    o = Oracle(policy)
    pop, inp, loop = resource_smallest(o)
    assert run_smallest(o, pop, inp) == EXPECTED
    assert o.weights(loop)[0] != 1.0


@pytest.mark.parametrize("policy", POLICIES)
def test_resource_stdp_altai_kat(oracle_built, policy):
    """ResourceSTDPAltAILIFNetwork (single_thread_cpu_test.cpp:267-358): the same network and expected spike
    steps as the BLIFAT variant, on default AltAI-LIF neurons; the loop weight changes."""
    o = Oracle(policy)
    neuron = abi.resource_altai_lif_defaults(synaptic_resource_threshold=1.0, free_synaptic_resource=2.0, isi_max=0)
    pop = o.add_resource_altai_lif_population(abi.neuron_array(neuron, 1))
    r_in = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=1.0, w_max=2.0, d_u=0.1), 1)
    r_loop = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.0, w_min=1.0, w_max=2.0), 1)
    inp = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1], r_in)
    loop = o.add_resource_stdp_projection(pop, 1, pop, [0], [0], [1.0], [6], r_loop)
    o.start_learning()
    assert run_smallest(o, pop, inp) == EXPECTED
    assert o.weights(loop)[0] != 1.0
    x = o.resource_altai_lif_neurons(pop)[0]
    # 0 -> unlimited (INT64_MAX) at step 0, then one step less of activity per step (altai_impl.h:103-114)
    assert x.base.activity_time == abi.INT64_MAX - 19 and x.isi_status != abi.ISI_NOT_IN_PERIOD


def test_resource_conservation(oracle_built):
    """free + sum of synaptic resources is invariant under resource-STDP training with drain 0."""
    o = Oracle("f64")
    pop, inp, loop = resource_smallest(o)
    total0 = 2.0
    for step in range(40):
        if step % 5 == 0:
            o.push_spikes(inp, [0])
        o.step()
        n = o.resource_neurons(pop)[0]
        total = n.free_synaptic_resource + o.resources(inp)[0] + o.resources(loop)[0]
        assert total == pytest.approx(total0, abs=1e-6)


def test_locked_projection_does_not_train(oracle_built):
    """Projections are locked by default (projection.h:432): without start_learning nothing trains."""
    o = Oracle("exact")
    neuron = abi.resource_blifat_defaults(free_synaptic_resource=2.0)
    pop = o.add_resource_blifat_population(abi.neuron_array(neuron, 1))
    r = abi.rule_array(abi.resource_rule_defaults(synaptic_resource=0.5, w_min=0.0, w_max=2.0), 1)
    inp = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.5], [1], r)
    for step in range(10):
        o.push_spikes(inp, [0])
        o.step()
    assert o.weights(inp)[0] == np.float32(1.5)
    assert o.resources(inp)[0] == np.float32(0.5)


def memoryless_reference(n, steps, syn, inputs):
    """Independent simulation of default BLIFAT networks: V_t = sum of impacts arriving at t, spike iff
    t > 0 and V_t >= 1 (activation threshold), V reset to 0, clamp at -1e9. Impacts from a spike at step s
    through a synapse of delay d arrive at step s + d."""
    pre, post, w, d, sign = syn
    in_idx, in_w = inputs  # per step: list of channel indexes (one-to-one weight in_w, delay 1)
    spikes = np.zeros((steps, n), dtype=bool)
    V = np.zeros((steps, n))
    for t in range(steps):
        acc = np.zeros(n)
        if t >= 1:
            for c in in_idx[t - 1]:
                acc[c] += in_w
        for dd in np.unique(d):
            if t - dd < 0:
                continue
            m = (d == dd) & spikes[t - dd][pre]
            np.add.at(acc, post[m], sign[m] * w[m])
        s = (acc >= 1.0) & (t > 0)
        spikes[t] = s
        V[t] = np.where(s, 0.0, np.maximum(acc, -1e9))
    return spikes, V


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_memoryless_closed_form(oracle_built, seed):
    rng = np.random.default_rng(seed)
    n, S, steps = 40, 600, 60
    pre = rng.integers(0, n, S)
    post = rng.integers(0, n, S)
    # weights on a 2^-12 grid keep every double sum exact, so the decision is order-free on both sides
    w = (rng.integers(1, 400, S) / 4096.0).astype(np.float32)
    d = rng.integers(1, 6, S)
    otype = (rng.random(S) < 0.3).astype(np.uint8)  # 30% INHIBITORY_CURRENT
    sign = np.where(otype == 1, -1.0, 1.0)
    in_idx = [np.unique(rng.integers(0, n, rng.integers(0, 6))).astype(np.uint32) for _ in range(steps)]
    o = Oracle("exact")
    pop = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), n))
    ident = np.arange(n, dtype=np.uint32)
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, n, pop, ident, ident, np.ones(n, np.float32),
                                 np.ones(n, np.uint32))
    o.add_delta_projection(pop, n, pop, pre, post, w, d, otype)
    o.record_spikes(pop, steps)
    Vs = []
    for t in range(steps):
        if len(in_idx[t]):
            o.push_spikes(inp, in_idx[t])
        o.step()
        Vs.append(o.potentials(pop))
    got = o.read_spike_record(pop)
    ref, Vref = memoryless_reference(n, steps, (pre, post, w.astype(np.float64), d, sign), (in_idx, 1.0))
    assert got.sum() > 0
    assert np.array_equal(got, ref)
    assert np.array_equal(np.array(Vs), Vref)
    expected_events = sum(len(x) for x in in_idx) + sum(
        np.bincount(pre, minlength=n)[np.nonzero(ref[t])[0]].sum() for t in range(steps))
    assert o.events == expected_events


def single_neuron(o, neuron, extra_types=()):
    """One neuron with one channel-fed projection per impact type (weights given per push)."""
    pop = o.add_blifat_population(abi.neuron_array(neuron, 1))
    return pop


def drive(o, pop, plan, steps, types=(abi.EXCITATORY,)):
    """plan: {step: [(type, weight), ...]} delivered with delay 1 (applied at step + 1)."""
    projs = {}
    # one projection per (type, weight) pair, loaded in plan order
    keys = []
    for t in sorted(plan):
        for k in plan[t]:
            if k not in keys:
                keys.append(k)
    for k in keys:
        projs[k] = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [k[1]], [1], [k[0]])
    out = []
    for t in range(steps):
        for k in plan.get(t, []):
            o.push_spikes(projs[k], [0])
        o.step()
        n = o.blifat_neurons(pop)[0]
        out.append((len(o.last_spikes(pop)) > 0, n.potential, n))
    return out


def test_hand_potential_decay(oracle_built):
    o = Oracle("exact")
    pop = single_neuron(o, abi.blifat_defaults(potential_decay=0.5))
    tr = drive(o, pop, {0: [(0, 0.625)], 1: [(0, 0.625)], 2: [(0, 0.625)]}, 5)
    assert [x[0] for x in tr] == [False, False, False, True, False]
    assert [x[1] for x in tr] == [0.0, 0.625, 0.9375, 0.0, 0.0]  # 0.3125 + 0.625 = 0.9375; then >= 1 -> reset


def test_hand_refractory(oracle_built):
    o = Oracle("exact")
    pop = single_neuron(o, abi.blifat_defaults(absolute_refractory_period=2))
    tr = drive(o, pop, {t: [(0, 1.0)] for t in range(12)}, 12)
    assert [t for t, x in enumerate(tr) if x[0]] == [3, 6, 9]


def test_hand_conductance(oracle_built):
    o = Oracle("exact")
    pop = single_neuron(o, abi.blifat_defaults(inhibitory_conductance_decay=0.5))
    tr = drive(o, pop, {0: [(0, 1.25), (abi.INHIBITORY_CONDUCTANCE, 0.5)], 1: [(0, 1.5)]}, 4)
    # step 1: V = 1.25, g = 0.5 -> V -= (V + 0.3) * g = 0.475 (no spike);
    # step 2: g = 0.25 (decay 0.5), V = 1.5 -> 1.5 - 1.8 * 0.25 = 1.05 -> spike
    v1 = 1.25
    v1 -= (v1 - (-0.3)) * 0.5
    assert [x[0] for x in tr] == [False, False, True, False]
    assert tr[1][1] == v1
    assert tr[1][2].inhibitory_conductance == 0.5
    # conductance >= 1 pins the potential to the reversal potential
    o2 = Oracle("exact")
    pop2 = single_neuron(o2, abi.blifat_defaults())
    tr2 = drive(o2, pop2, {0: [(0, 5.0), (abi.INHIBITORY_CONDUCTANCE, 1.5)]}, 3)
    assert tr2[1][1] == -0.3 and not tr2[1][0]


def test_hand_blocking(oracle_built):
    o = Oracle("exact")
    pop = single_neuron(o, abi.blifat_defaults())
    # a negative blocking value blocks impacts for |value| steps, then resets to INT64_MAX
    tr = drive(o, pop, {0: [(abi.BLOCKING, -2.0)], 1: [(0, 1.5)], 2: [(0, 1.5)], 3: [(0, 1.5)]}, 5)
    assert [x[0] for x in tr] == [False, False, False, True, True]
    assert tr[1][2].total_blocking_period == -1
    assert tr[2][2].total_blocking_period == abi.INT64_MAX
    assert tr[1][1] == 0.0 and tr[2][1] == 0.0  # impacts of steps 1 and 2 are undone


def test_hand_blocking_positive_and_restore(oracle_built):
    o = Oracle("exact")
    pop = single_neuron(o, abi.blifat_defaults())
    tr = drive(o, pop, {0: [(abi.BLOCKING, -1.0)], 1: [(0, 1.5)], 3: [(0, 1.5)]}, 5)
    # step 1: tbp = -1 -> potential restored to V_pre = 0 (impacts ignored), then tbp -> 0 -> INT64_MAX
    assert tr[1][2].total_blocking_period == abi.INT64_MAX and tr[1][1] == 0.0
    # step 2: impact pushed at 1 arrives; no block -> spike
    assert [x[0] for x in tr] == [False, False, True, False, True]


def test_hand_threshold_increment(oracle_built):
    o = Oracle("exact")
    pop = single_neuron(o, abi.blifat_defaults(threshold_increment=0.5, threshold_decay=0.5))
    tr = drive(o, pop, {t: [(0, 1.125)] for t in range(6)}, 6)
    # spike at 1 -> dyn 0.5; step 2 dyn 0.25 -> thr 1.25 > 1.125 no; step 3 dyn 0.125 -> 1.125 >= 1.125 spike
    # -> dyn 0.625; step 4 thr 1.3125, step 5 thr 1.15625: no
    assert [x[0] for x in tr] == [False, True, False, True, False, False]
    assert tr[1][2].dynamic_threshold == 0.5


def test_hand_bursting(oracle_built):
    o = Oracle("exact")
    pop = single_neuron(o, abi.blifat_defaults(bursting_period=1, reflexive_weight=0.75, potential_decay=1.0))
    tr = drive(o, pop, {0: [(0, 1.0)], 1: [(0, 0.25)]}, 4)
    # spike at 1 (V reset 0, phase 1); step 2: phase 1 -> 0, V += 0.75, + 0.25 impact = 1.0 -> spike again
    assert [x[0] for x in tr] == [False, True, True, False]
    assert tr[3][1] == 0.75  # step 3: phase reset to 1 by the step-2 spike -> V = 0 + 0.75


def test_resource_neuron_keeps_bursting_phase(oracle_built):
    """Resource neurons add the reflexive weight while bursting_phase_ == 1 without decrementing it
    (blifat_impl.h:62)."""
    o = Oracle("exact")
    base = abi.blifat_defaults(bursting_phase=1, reflexive_weight=0.25, potential_decay=1.0)
    pop = o.add_resource_blifat_population(abi.neuron_array(abi.resource_blifat_defaults(base), 1))
    o.step(3)
    n = o.resource_neurons(pop)[0]
    assert n.base.bursting_phase == 1
    assert n.base.potential == 0.75


def _np_hash(seed, step, ch):
    M = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    a = mix((seed + step * 0x9E3779B97F4A7C15) & M)
    b = mix(a ^ ((ch * 0xD6E8FEB86659FD93 + 0x632BE59BD9B4E019) & M))
    return b >> 32


def test_bernoulli_hash_and_rate(oracle_built):
    rng = np.random.default_rng(5)
    for _ in range(200):
        s, t, c = int(rng.integers(0, 2**63)), int(rng.integers(0, 2**40)), int(rng.integers(0, 2**32))
        assert bernoulli_hash(s, t, c) == _np_hash(s, t, c)
    thr = int(workloads.bernoulli_threshold(0.02))
    hits = sum(bernoulli_hash(99, t, c) < thr for t in range(50) for c in range(400))
    assert abs(hits / 20000 - 0.02) < 0.004


def test_source_matches_pushes(oracle_built):
    """A device-style Bernoulli source is the same as pushing its spikes from the host."""
    spec = workloads.plain_small(seed=4, n=200, n_syn=5000)
    a = spec.load_into(Oracle("exact"))
    b = Oracle("exact")
    pop = b.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 200))
    for name, args, kw in spec.ops[1:]:
        if name == "add_delta_projection":
            getattr(b, name)(*args, **kw)
    thr = int(workloads.bernoulli_threshold(0.02))
    seed = 4 * 7919 + 17
    a.record_spikes(0, 40)
    b.record_spikes(pop, 40)
    for t in range(40):
        idx = [c for c in range(200) if bernoulli_hash(seed, t, c) < thr]
        if idx:
            b.push_spikes(0, idx)
        b.step()
    a.step(40)
    assert np.array_equal(a.read_spike_record(0), b.read_spike_record(pop))
    assert a.events == b.events


def test_exact_and_f64_agree_without_plasticity(oracle_built):
    spec = workloads.plain_small(seed=3, n=300, n_syn=20000)
    a = spec.load_into(Oracle("exact"))
    b = spec.load_into(Oracle("f64"))
    for x in (a, b):
        x.record_spikes(0, 100)
        x.step(100)
    assert np.array_equal(a.read_spike_record(0), b.read_spike_record(0))
    assert np.array_equal(a.potentials(0), b.potentials(0))


# ---- LIF (lif_impl.h:31-84) ------------------------------------------------------------------------
def lif_trace(neuron, w, steps, every=1, otype=abi.EXCITATORY):
    """One LIF neuron fed by a channel that fires every `every` steps with weight w, delay 1."""
    o = Oracle("exact")
    pop = o.add_lif_population(abi.neuron_array(neuron, 1))
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [w], [1], [otype])
    spikes, pots, counters = [], [], []
    for t in range(steps):
        if t % every == 0:
            o.push_spikes(inp, [0])
        o.step()
        if len(o.last_spikes(pop)):
            spikes.append(t)
        n = o.lif_neurons(pop)[0]
        pots.append(n.potential)
        counters.append(n.refract_counter)
    return spikes, pots, counters


def test_hand_lif_leak_and_refractory(oracle_built):
    # leak 0.5, threshold 1, refractory period 2, an impact of 0.75 every step from step 1 on:
    # t1: 0.75; t2: 0.375 + 0.75 = 1.125 > 1 -> spike, V = 0, counter 2; t3, t4: no leak, impacts
    # ignored, counter 2 -> 1 -> 0; t5: 0.75; t6: spike again -> period 4
    sp, pots, cnt = lif_trace(abi.lif_defaults(leak_coefficient=0.5, refract_period=2), 0.75, 16)
    assert sp == [2, 6, 10, 14]
    assert pots[:8] == [0.0, 0.75, 0.0, 0.0, 0.0, 0.75, 0.0, 0.0]
    assert cnt[:8] == [0, 0, 2, 1, 0, 0, 2, 1]


def test_hand_lif_strict_threshold_and_reset(oracle_built):
    # no leak, 0.5 per step: V reaches exactly 1.0 at t2 without firing (V > threshold, not >=), fires at
    # t3 and resets to -0.25; then -0.25 + 0.5 + 0.5 + 0.5 = 1.25 fires at t6
    sp, pots, _ = lif_trace(abi.lif_defaults(potential_reset_value=-0.25), 0.5, 13)
    assert sp == [3, 6, 9, 12]
    assert pots[:7] == [0.0, 0.5, 1.0, -0.25, 0.25, 0.75, -0.25]


def test_hand_lif_inhibitory_and_initial_counter(oracle_built):
    # INHIBITORY_CURRENT subtracts; a neuron loaded with refract_counter 3 ignores the first impacts
    sp, pots, cnt = lif_trace(abi.lif_defaults(refract_counter=3, potential=0.5), 0.25, 7,
                              otype=abi.INHIBITORY_CURRENT)
    assert sp == []
    assert cnt[:4] == [2, 1, 0, 0]
    assert pots == [0.5, 0.5, 0.5, 0.25, 0.0, -0.25, -0.5]


def test_lif_rejects_other_synapse_types(oracle_built):
    o = Oracle("exact")
    pop = o.add_lif_population(abi.neuron_array(abi.lif_defaults(), 2))
    for ty in (abi.INHIBITORY_CONDUCTANCE, abi.DOPAMINE, abi.BLOCKING):
        with pytest.raises(RuntimeError, match="Unhandled synapse type"):
            o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [1], [0.5], [1], [ty])


def lif_numpy(n, pre, post, w, d, otype, reset, thr, period, inputs, steps):
    """Independent simulation of a recurrent LIF network (leak 1): per step, the sum of the weights of the
    synapses whose presynaptic neuron spiked `delay` steps earlier (channel inputs: one step earlier), added
    only to neurons whose refractory counter is 0."""
    sign = np.where(otype == abi.INHIBITORY_CURRENT, -1.0, 1.0)
    V = np.zeros(n)
    cnt = np.zeros(n, dtype=np.int64)
    hist = []
    out = np.zeros((steps, n), dtype=bool)
    for t in range(steps):
        s = np.zeros(n)
        if t >= 1:
            np.add.at(s, inputs[t - 1], 1.0)
        for k in range(len(pre)):
            if t - d[k] >= 0 and hist[t - d[k]][pre[k]]:
                s[post[k]] += sign[k] * w[k]
        free = cnt == 0
        V = np.where(free, V + s, V)
        fire = free & (V > thr)
        V = np.where(fire, reset, V)
        cnt = np.where(free, np.where(fire, period, 0), cnt - 1)
        hist.append(fire)
        out[t] = fire
    return out, V, cnt


def test_lif_network_matches_numpy(oracle_built):
    rng = np.random.default_rng(11)
    n, S, steps = 40, 400, 60
    pre = rng.integers(0, n, S)
    post = rng.integers(0, n, S)
    w = rng.integers(1, 64, S) / 64.0
    d = rng.integers(1, 6, S)
    otype = rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[0.7, 0.3]).astype(np.uint8)
    inputs = [np.nonzero(rng.random(n) < 0.15)[0] for _ in range(steps)]
    o = Oracle("exact")
    pop = o.add_lif_population(abi.neuron_array(abi.lif_defaults(potential_reset_value=-0.5, refract_period=2), n))
    ident = np.arange(n, dtype=np.uint32)
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, n, pop, ident, ident, np.ones(n, np.float32),
                                 np.ones(n, np.uint32))
    o.add_delta_projection(pop, n, pop, pre, post, w, d, otype)
    o.record_spikes(pop, steps)
    for t in range(steps):
        if len(inputs[t]):
            o.push_spikes(inp, inputs[t].astype(np.uint32))
        o.step()
    ref, V, cnt = lif_numpy(n, pre, post, w, d, otype, -0.5, 1.0, 2, inputs, steps)
    assert ref.sum() > 50
    assert np.array_equal(o.read_spike_record(pop), ref)
    lf = o.lif_neurons(pop)
    assert np.array_equal([x.potential for x in lf], V)
    assert np.array_equal([x.refract_counter for x in lf], cnt)


# ---- AltAI-LIF (altai_impl.h:31-150) ---------------------------------------------------------------
def run_altai_neuron(neuron, steps, impacts=(), num_neurons=1, neuron_index=0):
    """As run_altai_neuron of knp/tests/framework/altai_lif_test.cpp:60-100: impacts[t] is delivered straight
    to the neuron at step t; the potential is logged (as int16) before every step."""
    o = Oracle("exact")
    pop = o.add_altai_lif_population(abi.neuron_array(neuron, num_neurons))
    pots, spikes = [], []
    for t in range(steps):
        if t < len(impacts):
            o.inject_impacts(pop, [neuron_index], [impacts[t]])
        pots.append(int(np.int16(o.altai_lif_neurons(pop)[neuron_index].potential)))
        o.step()
        if neuron_index in o.last_spikes(pop):
            spikes.append(t)
    return pots, spikes


def test_altai_reference_leak_rev(oracle_built):
    n = abi.altai_lif_defaults(activation_threshold=200, potential_leak=-25, potential=100, leak_rev=1)
    pots, spikes = run_altai_neuron(n, 10, num_neurons=3, neuron_index=2)
    assert pots == [100, 75, 50, 25, 0, -25, 0, -25, 0, -25]
    assert spikes == []


def test_altai_reference_leak_no_rev(oracle_built):
    # negative_activation_threshold_ = -200 stored in uint16 (altai_lif_test.cpp:125)
    n = abi.altai_lif_defaults(activation_threshold=200, negative_activation_threshold=(-200) & 0xFFFF,
                               potential_leak=-25, potential=100, leak_rev=0)
    pots, spikes = run_altai_neuron(n, 10)
    assert pots == [100, 75, 50, 25, 0, -25, -50, -75, -100, -125]
    assert spikes == []


def test_altai_reference_saturate(oracle_built):
    n = abi.altai_lif_defaults(potential=100, activation_threshold=200, negative_activation_threshold=60,
                               potential_leak=-50, leak_rev=0, saturate=1)
    pots, spikes = run_altai_neuron(n, 10)
    assert pots == [100, 50, 0, -50, -60, -60, -60, -60, -60, -60]
    assert spikes == []


def test_altai_reference_spike_generation(oracle_built):
    n = abi.altai_lif_defaults(potential_leak=50, activation_threshold=110)
    pots, spikes = run_altai_neuron(n, 10)
    assert pots == [0, 50, 100, 0, 50, 100, 0, 50, 100, 0]
    assert spikes == [2, 5, 8]


def test_altai_reference_single_neuron_impact(oracle_built):
    n = abi.altai_lif_defaults(potential_leak=-30, activation_threshold=100)
    pots, spikes = run_altai_neuron(n, 10, [130, 50, 50, 50, 50, 0, 100, 80, 80, 80], num_neurons=5,
                                    neuron_index=3)
    assert pots == [0, 0, 20, 40, 60, 80, 50, 0, 50, 0]
    assert spikes == [0, 6, 8]


def test_hand_altai_diff_rounding_and_negative_reset(oracle_built):
    # is_diff without reset keeps the excess: 0 + 7.5 = 7.5 >= 5 -> spike, 2.5; next step rounds 2.5 to 3
    # (half away from zero) and adds 7.5 -> 10.5 -> spike, 5.5; then 6 + 7.5 = 13.5 -> 8.5; without input,
    # round(8.5) = 9 still fires and leaves 4, which does not
    n = abi.altai_lif_defaults(is_diff=1, is_reset=0, activation_threshold=5)
    pots, spikes = run_altai_neuron(n, 5, [7.5, 7.5, 7.5])
    assert spikes == [0, 1, 2, 3]
    assert pots == [0, 2, 5, 8, 4]  # logged as int16 before each step (2.5, 5.5, 8.5 truncate)
    o = Oracle("exact")
    pop = o.add_altai_lif_population(abi.neuron_array(n, 1))
    for w in (7.5, 7.5, 7.5):
        o.inject_impacts(pop, [0], [w])
        o.step()
    x = o.altai_lif_neurons(pop)[0]
    assert x.potential == 8.5 and x.pre_impact_potential == 6.0
    # negative side without saturation: reset gives -reset_value, diff adds the negative threshold
    for flags, expect in (({"is_reset": 1}, -2.0), ({"is_reset": 0, "is_diff": 1}, -1.0)):
        n = abi.altai_lif_defaults(saturate=0, negative_activation_threshold=10, potential_reset_value=2,
                                   activation_threshold=100, **flags)
        o = Oracle("exact")
        pop = o.add_altai_lif_population(abi.neuron_array(n, 1))
        o.inject_impacts(pop, [0], [11.0], [abi.INHIBITORY_CURRENT])
        o.step()
        assert o.altai_lif_neurons(pop)[0].potential == expect, flags


def test_hand_altai_activity_time_and_do_not_save(oracle_built):
    # activity_time_ = -2 blocks spikes for two steps (-2 -> -1 -> 0 -> unlimited): a neuron above threshold
    # every step (leak 5, threshold 1, reset) fires from the second step on
    n = abi.altai_lif_defaults(potential_leak=5, activity_time=-2)
    pots, spikes = run_altai_neuron(n, 4)
    assert spikes == [1, 2, 3]
    # do_not_save_: the potential restarts from the reset value every step, so 0.75 per step never adds up
    n = abi.altai_lif_defaults(do_not_save=1, potential_reset_value=0, activation_threshold=1)
    pots, spikes = run_altai_neuron(n, 4, [0.75] * 4)
    assert spikes == [] and pots == [0, 0, 0, 0]
    o = Oracle("exact")
    pop = o.add_altai_lif_population(abi.neuron_array(abi.altai_lif_defaults(), 2))
    o.inject_impacts(pop, [1], [0.5], [abi.DOPAMINE])
    o.step()
    x = o.altai_lif_neurons(pop)
    assert (x[0].dopamine_value, x[1].dopamine_value) == (0.0, 0.5)
    assert x[0].activity_time == abi.INT64_MAX


def test_altai_rejects_conductance_and_blocking(oracle_built):
    o = Oracle("exact")
    pop = o.add_altai_lif_population(abi.neuron_array(abi.altai_lif_defaults(), 2))
    for ty in (abi.INHIBITORY_CONDUCTANCE, abi.BLOCKING):
        with pytest.raises(RuntimeError, match="Unhandled synapse type"):
            o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [1], [0.5], [1], [ty])


def altai_numpy(n, pre, post, w, d, otype, cells, inputs, steps):
    """Independent simulation of a recurrent AltAI-LIF network with integer weights (altai_impl.h:31-150):
    per step round (or reset) the potential, add the weights of the synapses whose presynaptic neuron spiked
    `delay` steps earlier (channel inputs one step earlier), then leak, count the activity time down, fire
    and reset."""
    g = lambda f: np.array([getattr(c, f) for c in cells], dtype=np.float64)
    V = g("potential")
    act = g("activation_threshold") + g("additional_threshold")
    neg = g("negative_activation_threshold")
    leak = g("potential_leak")
    reset = g("potential_reset_value")
    flags = {f: g(f).astype(bool) for f in ("is_diff", "is_reset", "leak_rev", "saturate", "do_not_save")}
    at = np.array([c.activity_time for c in cells], dtype=np.int64)
    sign = np.where(otype == abi.INHIBITORY_CURRENT, -1.0, 1.0)
    hist, out = [], np.zeros((steps, n), dtype=bool)
    for t in range(steps):
        V = np.where(flags["do_not_save"], reset, np.sign(V) * np.floor(np.abs(V) + 0.5))
        s = np.zeros(n)
        if t >= 1:
            np.add.at(s, inputs[t - 1], 4.0)
        for k in range(len(pre)):
            if t - d[k] >= 0 and hist[t - d[k]][pre[k]]:
                s[post[k]] += sign[k] * w[k]
        V = V + s
        V = V + np.where(flags["leak_rev"] & (V < 0), -leak, leak)
        at = np.where(at > 0, at - 1, np.where(at < 0, at + 1, at))
        at = np.where(at == 0, np.iinfo(np.int64).max, at)
        above = V >= act
        fire = above & (at > 0)
        V = np.where(above & flags["is_diff"], V - act, V)
        was_reset = above & flags["is_reset"]
        V = np.where(was_reset, reset, V)
        low = (V <= -neg) & ~was_reset
        V = np.where(low & flags["saturate"], -neg,
                     np.where(low & flags["is_reset"], -reset, np.where(low & flags["is_diff"], V + neg, V)))
        hist.append(fire)
        out[t] = fire
    return out, V, at


def test_altai_network_matches_numpy(oracle_built):
    rng = np.random.default_rng(17)
    n, S, steps = 40, 300, 60
    cells = abi.neuron_array(abi.altai_lif_defaults(), n)
    for c in cells:
        c.activation_threshold = int(rng.integers(3, 9))
        c.negative_activation_threshold = int(rng.integers(4, 12))
        c.potential_leak = int(rng.integers(-1, 2))
        c.potential_reset_value = int(rng.integers(0, 2))
        c.is_diff, c.is_reset, c.leak_rev, c.saturate = (int(b) for b in rng.integers(0, 2, 4))
        c.activity_time = int(rng.choice([0, 0, 3, -2]))
    pre = rng.integers(0, n, S)
    post = rng.integers(0, n, S)
    w = rng.integers(1, 5, S).astype(np.float64)
    d = rng.integers(1, 5, S)
    otype = rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[0.65, 0.35]).astype(np.uint8)
    inputs = [np.nonzero(rng.random(n) < 0.2)[0] for _ in range(steps)]
    o = Oracle("exact")
    pop = o.add_altai_lif_population(cells)
    ident = np.arange(n, dtype=np.uint32)
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, n, pop, ident, ident, np.full(n, 4.0, np.float32),
                                 np.ones(n, np.uint32))
    o.add_delta_projection(pop, n, pop, pre, post, w, d, otype)
    o.record_spikes(pop, steps)
    for t in range(steps):
        if len(inputs[t]):
            o.push_spikes(inp, inputs[t].astype(np.uint32))
        o.step()
    ref, V, at = altai_numpy(n, pre, post, w, d, otype, cells, inputs, steps)
    assert ref.sum() > 50
    assert np.array_equal(o.read_spike_record(pop), ref)
    x = o.altai_lif_neurons(pop)
    assert np.array_equal([c.potential for c in x], V)
    assert np.array_equal([c.activity_time for c in x], at)


def test_multi_threaded_step_equals_single_thread(oracle_built):
    """The multi-threaded decomposition (cpu-multi-threaded-backend/impl/backend.cpp:77-323) gives the
    single-threaded step's spikes, potentials and event count: its chunks are merged in spike order."""
    for spec in (workloads.plain_small(seed=5, n=600, n_syn=30000),):
        a = spec.load_into(Oracle("exact"))
        b = spec.load_into(Oracle("exact", threads=4))
        for x in (a, b):
            x.record_spikes(0, 150)
            x.step(150)
        ra, rb = a.read_spike_record(0), b.read_spike_record(0)
        assert ra.sum() > 100
        assert np.array_equal(ra, rb)
        assert np.array_equal(a.potentials(0), b.potentials(0))
        assert a.events == b.events
    # several populations and projections, LIF and BLIFAT, pushed channels and a Bernoulli source
    from networks import lif_network

    runs = []
    for threads in (1, 4):
        o = Oracle("exact", threads=threads)
        pops, push = lif_network(o, 21)
        for p in pops:
            o.record_spikes(p, 40)
        rng = np.random.default_rng(6)
        for _ in range(4):
            o.push_spikes(push, np.nonzero(rng.random(1500) < 0.08)[0].astype(np.uint32))
            o.step(10)
        runs.append(([o.read_spike_record(p) for p in pops], o.potentials(pops[0]), o.potentials(pops[1]), o.events))
    assert runs[0][0][0].sum() > 100
    for k in range(2):
        assert np.array_equal(runs[0][0][k], runs[1][0][k])
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])
    assert runs[0][3] == runs[1][3]


def test_resource_stdp_hand_traces(oracle_built):
    """Synaptic-resource STDP arithmetic pinned by hand-computed traces (tests/resource_traces.py): Hebbian
    updates at stability 0 and 1, the has_contributed window with its edges and the u64 wrap, dopamine reward
    (both signs of the stability term), punishment and forced reward, renormalization with drains 0 and 2,
    a negative free resource and the weight formula with w_min != 0."""
    import resource_traces

    resource_traces.run_all_traces(lambda: Oracle("exact"))
