"""GPU parity under the multi-threaded backend's message semantics (knp_mi355x_set_message_semantics): the
device against the oracle with the same semantics (pinned in test_mt_semantics.py). Bar: spike trains
bit-exact, SynapticImpact counts equal, potentials within 1e-6 relative."""
import sys

import numpy as np
import pytest

from knp_amd import abi
from knp_amd.backend import BackendError, MI355XBackend, device_count
from oracle.pyoracle import Oracle

import networks
from networks import mt_smallest_network, multiplicity_network, precedence_network

pytestmark = pytest.mark.gpu

MT = abi.SEMANTICS_MULTI_THREADED
EXPECTED = [1, 6, 7, 11, 12, 13, 16, 17, 18, 19]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert device_count() >= 1, "no HIP device visible"


def pair():
    g, o = MI355XBackend(0), Oracle("exact")
    g.set_message_semantics(MT)
    o.set_message_semantics(MT)
    return g, o


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30))) if a.size else 0.0


@pytest.mark.parametrize("resource", [False, True])
def test_reference_mt_known_answers_on_device(resource):
    say(f"case: multi-threaded SmallestNetwork, resource={resource}")
    g, _ = pair()
    pop, inp = mt_smallest_network(g, resource)
    res = []
    for step in range(20):
        if step % 5 == 0:
            g.push_spikes(inp, [0])
        g.step()
        if len(g.last_spikes(pop)):
            res.append(step)
    assert res == EXPECTED
    g.close()


def test_multiplicity_and_first_message_on_device():
    say("case: multi-threaded repeated index, second push, population precedence")
    g, o = pair()
    for x in (g, o):
        pop, inp = multiplicity_network(x)
        x.push_spikes(inp, [0, 1, 0, 0])
        x.push_spikes(inp, [1])  # dropped
        x.step(2)
    assert list(g.last_spikes(pop)) == list(o.last_spikes(pop)) == [0]
    assert g.events == o.events == 2
    assert rel_err(g.potentials(pop), o.potentials(pop)) <= 1e-6
    g.close()
    g, o = pair()
    for x in (g, o):
        (a, b), drive, ab = precedence_network(x)
        x.push_spikes(drive, [0])
        x.step(1)
        x.push_spikes(ab, [1])
        x.step(2)
    assert list(g.potentials(b)) == list(o.potentials(b)) == [0.5, 0.0]
    assert g.events == o.events


def test_random_multi_sender_network_matches_oracle():
    """Recurrent and synaptic-resource populations fed by pushed messages with repeated indexes and second
    messages, 200 steps: spikes bit-exact, impact counts equal, the resource rules' last spike steps equal."""
    say("case: multi-threaded random network, 200 steps")
    steps = 200
    g, o = pair()
    rng = np.random.default_rng(7)
    plan = networks.mt_random_pushes(rng, steps, 600)
    for x in (g, o):
        (a, b), pa, pb = networks.mt_random_network(x, 5)
        x.record_spikes(a, steps)
        x.record_spikes(b, steps)
        for msgs in plan:
            for proj, idx in msgs:
                x.push_spikes(pa if proj == 0 else pb, idx)
            x.step(1)
    sa, sb = g.read_spike_record(a), g.read_spike_record(b)
    oa, ob = o.read_spike_record(a), o.read_spike_record(b)
    say(f"  spikes a {int(oa.sum())} b {int(ob.sum())}, events {o.events}")
    assert oa.sum() > 100
    assert np.array_equal(sa, oa) and np.array_equal(sb, ob)
    assert g.events == o.events
    assert rel_err(g.potentials(a), o.potentials(a)) <= 1e-6
    rg = np.array([r.last_spike_step for r in g.resource_rules(pb)])
    ro = np.array([r.last_spike_step for r in o.resource_rules(pb)])
    assert np.array_equal(rg, ro)
    g.close()


def test_multi_threaded_refusals():
    say("case: multi-threaded semantics refusals")
    g = MI355XBackend(0)
    g.set_message_semantics(MT)
    g.add_altai_lif_population(abi.neuron_array(abi.altai_lif_defaults(), 2))
    with pytest.raises(BackendError, match="Type is not supported"):
        g.step(1)
    g = MI355XBackend(0)
    g.set_message_semantics(MT)
    pop = g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 2))
    g.add_delta_projection(pop, 2, pop, [0], [1], [0.5], [1], [abi.BLOCKING])
    with pytest.raises(BackendError, match="BLOCKING"):
        g.step(1)
    g = MI355XBackend(0)
    g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 2))
    g.step(1)
    with pytest.raises(BackendError, match="before the first step"):
        g.set_message_semantics(MT)
