"""Pins the oracle's multi-threaded message semantics (knp/backends/cpu/cpu-multi-threaded-backend/impl/
backend.cpp:235-295, cpu-library delta_impl.h:117-175) with the reference's own known answers and hand-computed
cases:
  * SmallestNetwork and SmallestResourceNetwork (knp/tests/backend/multi_thread_cpu_test.cpp:54-141);
  * an index repeated in one message gives one impact of weight float(w * count): float(3 * float(0.1)) =
    10066330 * 2^-25 reaches a threshold that the single-threaded backend's three impacts of float(0.1)
    (40265319 * 2^-27, summed in double) do not;
  * only the first message of a step counts: a second push is dropped, and a population's message (routed
    from the backend's own endpoint, message_bus_cpu_impl.cpp:41-88) goes before a pushed one;
  * the types the multi-threaded backend does not hold, and BLOCKING synapses, are refused.
"""
import numpy as np
import pytest

from knp_amd import abi
from oracle.pyoracle import Oracle

import networks
from networks import (SUM3_DOUBLE, SUM3_FLOAT, THRESHOLD_BETWEEN, W_TENTH, mt_smallest_network, multiplicity_network,
                      precedence_network)

EXPECTED = [1, 6, 7, 11, 12, 13, 16, 17, 18, 19]
ST, MT = abi.SEMANTICS_SINGLE_THREADED, abi.SEMANTICS_MULTI_THREADED


def oracle(semantics, policy="exact"):
    o = Oracle(policy)
    o.set_message_semantics(semantics)
    return o


def run_smallest(o, pop, inp, steps=20):
    res = []
    for step in range(steps):
        if step % 5 == 0:
            o.push_spikes(inp, [0])
        o.step()
        if len(o.last_spikes(pop)):
            res.append(step)
    return res


@pytest.mark.parametrize("resource", [False, True])
@pytest.mark.parametrize("policy", ["exact", "f64"])
def test_reference_mt_known_answers(oracle_built, resource, policy):
    o = oracle(MT, policy)
    pop, inp = mt_smallest_network(o, resource)
    assert run_smallest(o, pop, inp) == EXPECTED


def test_hand_computed_constants():
    """The constants of the multiplicity case, from their integer forms."""
    assert float(W_TENTH) == 13421773 * 2.0 ** -27
    assert 3 * float(W_TENTH) == SUM3_DOUBLE  # exact in double
    assert float(np.float32(W_TENTH * np.float32(3))) == SUM3_FLOAT
    assert SUM3_DOUBLE < THRESHOLD_BETWEEN <= SUM3_FLOAT


@pytest.mark.parametrize("semantics", [ST, MT])
def test_repeated_index_is_one_weight_multiple(oracle_built, semantics):
    o = oracle(semantics)
    pop, inp = multiplicity_network(o)
    o.push_spikes(inp, [0, 1, 0, 0])
    o.step(1)  # step 0: the projection expands the message; the impacts arrive at step 1
    ev0 = o.events
    o.step(1)
    v = o.potentials(pop)
    if semantics == MT:
        assert ev0 == 2  # one impact per synapse
        assert list(o.last_spikes(pop)) == [0]  # float(3 w) = SUM3_FLOAT >= threshold
        assert v[0] == 0.0 and v[1] == float(W_TENTH)
    else:
        assert ev0 == 4  # one impact per occurrence
        assert len(o.last_spikes(pop)) == 0  # 3 w = SUM3_DOUBLE < threshold
        assert v[0] == SUM3_DOUBLE and v[1] == float(W_TENTH)


@pytest.mark.parametrize("semantics", [ST, MT])
def test_second_push_of_a_step(oracle_built, semantics):
    o = oracle(semantics)
    pop, inp = multiplicity_network(o)
    o.push_spikes(inp, [0])
    o.push_spikes(inp, [1])
    o.step(2)
    v = o.potentials(pop)
    assert v[0] == float(W_TENTH)
    assert v[1] == (0.0 if semantics == MT else float(W_TENTH))


@pytest.mark.parametrize("semantics", [ST, MT])
def test_population_message_goes_first(oracle_built, semantics):
    """A's neuron 0 fires at step 1 (driven at step 0); a message pushed into A -> B for step 1 comes after A's
    message, so the multi-threaded projection expands A's spikes only. At step 3 A is silent and the pushed
    message is the first one."""
    o = oracle(semantics)
    (a, b), drive, ab = precedence_network(o)
    o.push_spikes(drive, [0])
    o.step(1)  # step 0
    o.push_spikes(ab, [1])
    o.step(1)  # step 1: A fires neuron 0; A -> B receives A's message and the pushed one
    assert list(o.last_spikes(a)) == [0]
    o.step(1)  # step 2: B holds the impacts
    vb = o.potentials(b)
    assert vb[0] == 0.5
    assert vb[1] == (0.0 if semantics == MT else 0.5)
    o.push_spikes(ab, [1])
    o.step(2)  # steps 3 and 4: only the pushed message
    assert list(o.potentials(b)) == [0.0, 0.5]


def test_types_of_the_multi_threaded_backend_only(oracle_built):
    o = oracle(MT)
    pop = o.add_altai_lif_population(abi.neuron_array(abi.altai_lif_defaults(), 2))
    with pytest.raises(RuntimeError, match="Type is not supported"):
        o.step(1)
    o = oracle(MT)
    pop = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 2))
    o.add_delta_projection(pop, 2, pop, [0], [1], [0.5], [1], [abi.BLOCKING])
    with pytest.raises(RuntimeError, match="BLOCKING"):
        o.step(1)


def test_random_network_differs_between_semantics(oracle_built):
    """The random multi-sender network (repeated indexes, second messages) gives different spike trains under
    the two semantics: the cases the device tests compare are not degenerate."""
    recs = []
    for semantics in (ST, MT):
        o = oracle(semantics)
        (a, b), pa, pb = networks.mt_random_network(o, 5)
        o.record_spikes(a, 60)
        rng = np.random.default_rng(7)
        for msgs in networks.mt_random_pushes(rng, 60, 600):
            for proj, idx in msgs:
                o.push_spikes(pa if proj == 0 else pb, idx)
            o.step(1)
        recs.append(o.read_spike_record(a))
    assert recs[0].sum() > 100
    assert not np.array_equal(recs[0], recs[1])
