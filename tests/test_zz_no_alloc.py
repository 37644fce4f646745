"""knp_mi355x_step_noalloc (the allocation-free variant of knp_mi355x_step) against knp_mi355x_step.

Both run twin networks whose projections are expanded on MFMA (the kernels whose spike-count staging moves
to the library's device scratch): spike trains, potentials and impact counts must be identical. Two calls in
a row on different networks, the second smaller (the scratch keeps the first one's contents), then calls that
alternate between the two entry points on one network; sizes above the scratch and a network that is not
initialized are refused with errors.
"""
import sys

import numpy as np
import pytest

from knp_amd import abi, workloads
from knp_amd.backend import BackendError, MI355XBackend, device_count

pytestmark = pytest.mark.gpu

N_IN = 40


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert device_count() >= 1, "no HIP device visible"


def dense_network(sizes, seed):
    """Instances of default BLIFAT neurons with an all-to-all input projection from their own N_IN channels
    (weights k/256) and an all-to-all inhibitory lateral projection (1/8), both expanded on MFMA; a Bernoulli
    source (p = 0.3) drives the channels."""
    g = MI355XBackend(0)
    rng = np.random.default_rng(seed)
    sizes = np.asarray(sizes)
    begin = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    N = int(sizes.sum())
    pop = g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), N))
    pre, post, w, lp, lq = [], [], [], [], []
    for k, (b, n) in enumerate(zip(begin, sizes)):
        jj, ii = np.meshgrid(np.arange(n), np.arange(N_IN), indexing="ij")
        post.append(b + jj.ravel())
        pre.append(k * N_IN + ii.ravel())
        w.append(rng.integers(1, 65, n * N_IN) / 256.0)
        jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        lq.append(b + jj.ravel())
        lp.append(b + ii.ravel())
    pre, post, w, lp, lq = (np.concatenate(x) for x in (pre, post, w, lp, lq))
    inp = g.add_delta_projection(abi.KNP_NO_POPULATION, len(sizes) * N_IN, pop, pre, post, w.astype(np.float32),
                                 np.ones(len(pre), np.uint32))
    lat = g.add_delta_projection(pop, N, pop, lp, lq, np.full(len(lp), 0.125, np.float32), np.ones(len(lp), np.uint32),
                                 np.full(len(lp), abi.INHIBITORY_CURRENT, np.uint8))
    g.add_bernoulli_source(inp, np.full(len(sizes) * N_IN, workloads.bernoulli_threshold(0.3), np.uint32), seed=seed)
    g.set_projection_dense(inp, abi.DENSE_BF16, abi.DENSE_SKIP_EMPTY | abi.DENSE_SPLIT)
    g.set_projection_dense(lat, abi.DENSE_BF16, abi.DENSE_SKIP_EMPTY)
    g._init()
    return g, pop


def run_pair(sizes, seed, plan):
    """Twin networks; plan: list of (steps, use_noalloc_on_the_first). Returns both (record, potentials,
    events)."""
    a, pa = dense_network(sizes, seed)
    b, pb = dense_network(sizes, seed)
    total = sum(n for n, _ in plan)
    a.record_spikes(pa, total)
    b.record_spikes(pb, total)
    for n, noalloc in plan:
        if noalloc:
            a.step_noalloc(n)
        else:
            a.step(n)
        b.step(n)
    out = [(x.read_spike_record(p), x.potentials(p), x.events) for x, p in ((a, pa), (b, pb))]
    a.close()
    b.close()
    return out


def check_equal(out, label):
    (ra, va, ea), (rb, vb, eb) = out
    say(f"  {label}: spikes {int(ra.sum())} / {int(rb.sum())}, events {ea} / {eb}")
    assert rb.sum() > 50
    assert np.array_equal(ra, rb), (label, np.argwhere(ra != rb)[:5].tolist())
    assert np.array_equal(va, vb), label
    assert ea == eb, label


def test_step_noalloc_matches_step():
    say("case: knp_mi355x_step_noalloc vs knp_mi355x_step, two networks (the second smaller), then alternating")
    check_equal(run_pair([80, 33, 64, 100, 17], 5, [(60, True)]), "first call, 5 instances, 60 steps")
    check_equal(run_pair([50, 21], 6, [(30, True)]), "second call, 2 instances, 30 steps")
    check_equal(run_pair([64, 40, 9], 7, [(10, False), (15, True), (12, False), (1, True), (9, False)]),
                "alternating entry points")


def test_step_noalloc_refusals():
    say("case: knp_mi355x_step_noalloc refusals (not initialized, staging above the scratch)")
    g = MI355XBackend(0)
    g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 4))
    with pytest.raises(BackendError, match="initialized"):
        g.step_noalloc(1)
    g.close()
    # one instance of 524,300 inputs: 65,538 groups of 8, 16 bytes each, above the 1 MiB scratch
    n_in = 524_300
    g = MI355XBackend(0)
    pop = g.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 1))
    q = g.add_delta_projection(abi.KNP_NO_POPULATION, n_in, pop, np.arange(n_in, dtype=np.uint32),
                               np.zeros(n_in, np.uint32), np.full(n_in, 2.0 ** -20, np.float32), np.ones(n_in, np.uint32))
    g.set_projection_dense(q, abi.DENSE_BF16, 0)
    g._init()
    with pytest.raises(BackendError, match="scratch"):
        g.step_noalloc(1)
    g.step(2)  # the allocating entry point still runs it
    g.close()
