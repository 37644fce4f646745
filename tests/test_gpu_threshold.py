"""GPU parity of the spike threshold when activation_threshold_ + additional_threshold_ is not a float.

The reference decides `potential_ >= activation_threshold_ + dynamic_threshold_ + additional_threshold_` in
double (knp/backends/cpu/cpu-library/include/knp/backends/cpu-library/impl/populations/blifat/
blifat_impl.h:110-116). A float threshold would get these cases wrong:
  - T = 0.6: float(T) = 0.600000024 > T. Impacts 0.5f + 0.1f sum to 0.6000000015: the reference fires.
  - T = 0.7: float(T) = 0.699999988 < T. Impacts 0.5f + (the float below 0.2f) sum to exactly float(T): the
    reference does not fire.
  - T = 0.5 + 0.1 (activation + additional threshold, a double sum equal to 0.6) with 0.5f + 0.1f: fires.
  - T = 0.600000024 (float(0.6) as a double) with 0.5f + 0.1f: does not fire.
Three populations get the same cases: default BLIFAT (the lean kind: exact sum against ceil(T * 2^F)), a
full-state one (trace parameters set, threshold still static) and one with a dynamic threshold (increment
2^-2, decay 1/2: the float-float comparison; after a spike it fires again once the dynamic part has decayed
below 1.5e-9, 28 halvings). Every case is checked against the oracle bit for bit, on the persistent, the
one-launch-per-step and the atomic paths.
"""
import sys

import numpy as np
import pytest

from knp_amd import abi, workloads
from knp_amd.backend import MI355XBackend, device_count
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

F01 = np.float32(0.1)
BELOW_F01 = np.nextafter(F01, np.float32(0))
F02 = np.float32(0.2)
BELOW_F02 = np.nextafter(F02, np.float32(0))
# (activation_threshold_, additional_threshold_, second weight, does the reference fire)
CASES = [
    (0.6, 0.0, F01, True),
    (0.6, 0.0, BELOW_F01, False),
    (0.7, 0.0, BELOW_F02, False),
    (0.7, 0.0, F02, True),
    (0.5, 0.1, F01, True),
    (float(np.float32(0.6)), 0.0, F01, False),
]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert device_count() >= 1, "no HIP device visible"


def threshold_network(o, n=768):
    """Three populations of n neurons; neuron i has case i % len(CASES) and is driven by channel i of its
    population's Bernoulli source (p = 1/2) through two synapses of weights 0.5 and the case's second weight,
    delay 1."""
    kinds = [dict(), dict(postsynaptic_trace_decay=0.5, postsynaptic_trace_increment=1.0),
             dict(threshold_increment=0.25, threshold_decay=0.5)]
    pops = []
    for k, kw in enumerate(kinds):
        cells = abi.neuron_array(abi.blifat_defaults(**kw), n)
        for i in range(n):
            act, add, _, _ = CASES[i % len(CASES)]
            cells[i].activation_threshold = act
            cells[i].additional_threshold = add
        p = o.add_blifat_population(cells)
        ch = np.arange(n, dtype=np.uint32)
        w2 = np.array([CASES[i % len(CASES)][2] for i in range(n)], np.float32)
        pre = np.concatenate([ch, ch])
        post = np.concatenate([ch, ch])
        w = np.concatenate([np.full(n, 0.5, np.float32), w2])
        q = o.add_delta_projection(abi.KNP_NO_POPULATION, n, p, pre, post, w, np.ones(2 * n, np.uint32))
        o.add_bernoulli_source(q, np.full(n, workloads.bernoulli_threshold(0.5), np.uint32), seed=41 + k)
        pops.append(p)
    # zero-weight synapses from the first neuron of every 256-neuron segment to every segment of every
    # population: they change no potential, but every tile then reads every population segment, which the
    # persistent launch needs (each segment's writer waits for the segments of its readers)
    starts = np.arange(0, n, 256, dtype=np.uint32)
    for p in pops:
        for q2 in pops:
            pre = np.repeat(starts, len(starts))
            post = np.tile(starts + 1, len(starts))
            o.add_delta_projection(p, n, q2, pre, post, np.zeros(len(pre), np.float32), np.ones(len(pre), np.uint32))
    return pops


def test_threshold_not_a_float_matches_oracle():
    say("case: thresholds whose float differs from their double (persistent, stepwise, atomic vs oracle)")
    steps = 240
    chunks = [1, 50, 3, 90, 96]
    modes = {"persistent": abi.SCATTER_AUTO, "stepwise": abi.SCATTER_STEPWISE, "atomic": abi.SCATTER_ATOMIC}
    runs = {}
    for name in ("oracle", "persistent", "stepwise", "atomic"):
        o = Oracle("exact") if name == "oracle" else MI355XBackend(0)
        if name != "oracle":
            o.set_scatter_mode(modes[name])
            o.enable_timing(True)
        pops = threshold_network(o)
        for p in pops:
            o.record_spikes(p, steps)
        for c in chunks:
            o.step(c)
        if name in ("persistent", "stepwise"):
            assert all(o.tiled(p) for p in pops)
        if name == "persistent":
            assert o.kernel_time(abi.KERNEL_TILE_RUN)[1] > 0
        runs[name] = ([o.read_spike_record(p) for p in pops], o.events)
        say(f"  {name}: spikes per population {[int(r.sum()) for r in runs[name][0]]}")
    n = runs["oracle"][0][0].shape[1]
    for k in range(3):
        so = runs["oracle"][0][k]
        per_case = [int(so[:, c::len(CASES)].sum()) for c in range(len(CASES))]
        # the oracle (the reference's double threshold) fires exactly for the cases that should fire
        for c, (_, _, _, fires) in enumerate(CASES):
            assert (per_case[c] > 0) == fires, (k, c, per_case)
        for name in ("persistent", "stepwise", "atomic"):
            sg = runs[name][0][k]
            assert sg.shape == so.shape == (steps, n)
            assert np.array_equal(sg, so), (name, k, np.argwhere(sg != so)[:5].tolist())
    # the dynamic threshold population: a neuron fires again only after its threshold has decayed
    gaps = np.diff(np.nonzero(runs["oracle"][0][2][:, 0])[0])
    assert len(gaps) >= 3 and gaps.min() >= 28, gaps
    assert runs["persistent"][1] == runs["stepwise"][1] == runs["atomic"][1] == runs["oracle"][1]
