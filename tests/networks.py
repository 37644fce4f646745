"""Network builders shared by the parity tests (they load into the device backend or the oracle alike)."""
import numpy as np

from knp_amd import abi, workloads


def lif_network(o, seed):
    """LIF population A (leak 1, reset -0.5, refractory period 2, some neurons loaded refractory) with a
    default BLIFAT population B: recurrent A->A (delays up to 12), A->B, B->A, host pushes into A and a
    Bernoulli source into B. Only plain Delta EXCITATORY / INHIBITORY_CURRENT input, so both are tiled;
    weights on a 2^-8 grid keep every FP32 sum exact."""
    rng = np.random.default_rng(seed)
    na, nb = 1500, 500
    cells = abi.neuron_array(abi.lif_defaults(potential_reset_value=-0.5, refract_period=2), na)
    for i in range(0, na, 7):
        cells[i].refract_counter = 1 + i % 3
        cells[i].potential = 0.25
    a = o.add_lif_population(cells)
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), nb))
    ident = np.arange(na, dtype=np.uint32)
    push = o.add_delta_projection(abi.KNP_NO_POPULATION, na, a, ident, ident, np.full(na, 0.75, np.float32),
                                  np.ones(na, np.uint32))
    def proj(pre_pop, n_pre, post_pop, n_post, S, wlo, whi, dmax, inh):
        o.add_delta_projection(pre_pop, n_pre, post_pop, rng.integers(0, n_pre, S).astype(np.uint32),
                               rng.integers(0, n_post, S).astype(np.uint32),
                               (rng.integers(wlo, whi, S) / 256.0).astype(np.float32),
                               rng.integers(1, dmax + 1, S).astype(np.uint32),
                               rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[1 - inh, inh]).astype(np.uint8))
    proj(a, na, a, na, 60_000, 4, 64, 12, 0.25)
    proj(a, na, b, nb, 20_000, 16, 128, 5, 0.1)
    proj(b, nb, a, na, 2_000, 32, 160, 3, 0.5)
    src = o.add_delta_projection(abi.KNP_NO_POPULATION, nb, b, np.arange(nb, dtype=np.uint32),
                                 np.arange(nb, dtype=np.uint32), np.full(nb, 1.0, np.float32), np.ones(nb, np.uint32))
    o.add_bernoulli_source(src, np.full(nb, workloads.bernoulli_threshold(0.03), np.uint32), seed=9)
    return (a, b), push


# ---- networks of the multi-threaded backend's message semantics (cpu-multi-threaded-backend/impl/backend.cpp)

def mt_smallest_network(o, resource=False):
    """SmallestNetwork and SmallestResourceNetwork of knp/tests/backend/multi_thread_cpu_test.cpp:54-141
    (generators knp/tests/common/generators.h:48-100): one neuron, an input projection (weight 1.0 or, for the
    resource form, 1.5, delay 1) and a delay-6 loop (1.1 or 1.5); resource projections stay locked."""
    if resource:
        pop = o.add_resource_blifat_population(abi.neuron_array(abi.resource_blifat_defaults(), 1))
        rule = abi.resource_rule_defaults(w_max=2.0, w_min=1.0, synaptic_resource=1.0, dopamine_plasticity_period=5)
        inp = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.5], [1], abi.rule_array(rule, 1))
        o.add_resource_stdp_projection(pop, 1, pop, [0], [0], [1.5], [6], abi.rule_array(rule, 1))
    else:
        pop = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 1))
        inp = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1])
        o.add_delta_projection(pop, 1, pop, [0], [0], [1.1], [6])
    return pop, inp


# float(0.1) = 13421773 * 2^-27. Three impacts of it sum (in double) to 40265319 * 2^-27 = 0.30000000447034836;
# one impact of float(3 * float(0.1)) is 40265319 * 2^-27 rounded to 24 bits: 10066330 * 2^-25 =
# 0.30000001192092896. A threshold between the two separates them.
W_TENTH = np.float32(0.1)
SUM3_DOUBLE = 40265319 * 2.0 ** -27
SUM3_FLOAT = 10066330 * 2.0 ** -25
THRESHOLD_BETWEEN = 0.3000000080


def multiplicity_network(o):
    """Two default BLIFAT neurons, threshold between 3 * float(0.1) (double) and float(3 * float(0.1)), fed
    one to one by a channel projection of weight float(0.1), delay 1."""
    pop = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(activation_threshold=THRESHOLD_BETWEEN), 2))
    inp = o.add_delta_projection(abi.KNP_NO_POPULATION, 2, pop, [0, 1], [0, 1], np.array([W_TENTH] * 2), [1, 1])
    return pop, inp


def precedence_network(o):
    """Population A (2 default BLIFAT neurons) driven one to one by channel projection `drive` (weight 1,
    delay 1); projection A -> B (2 neurons, one to one, weight 0.5, delay 1) that also takes pushed messages."""
    a = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 2))
    b = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), 2))
    drive = o.add_delta_projection(abi.KNP_NO_POPULATION, 2, a, [0, 1], [0, 1], [1.0, 1.0], [1, 1])
    ab = o.add_delta_projection(a, 2, b, [0, 1], [0, 1], [0.5, 0.5], [1, 1])
    return (a, b), drive, ab


def mt_random_network(o, seed):
    """A recurrent default-BLIFAT population with a synaptic-resource population behind it, both fed by pushed
    messages with repeated indexes (weights on a 2^-8 grid: every sum exact)."""
    rng = np.random.default_rng(seed)
    n, m = 600, 300
    a = o.add_blifat_population(abi.neuron_array(abi.blifat_defaults(), n))
    b = o.add_resource_blifat_population(abi.neuron_array(abi.resource_blifat_defaults(), m))
    ident = np.arange(n, dtype=np.uint32)
    push_a = o.add_delta_projection(abi.KNP_NO_POPULATION, n, a, ident, ident, np.full(n, 0.375, np.float32),
                                    np.ones(n, np.uint32))
    S = 30_000
    o.add_delta_projection(a, n, a, rng.integers(0, n, S).astype(np.uint32), rng.integers(0, n, S).astype(np.uint32),
                           (rng.integers(1, 12, S) / 256.0).astype(np.float32), rng.integers(1, 6, S).astype(np.uint32),
                           rng.choice([abi.EXCITATORY, abi.INHIBITORY_CURRENT], S, p=[0.8, 0.2]).astype(np.uint8))
    S2 = 8_000
    rule = abi.resource_rule_defaults(w_max=0.5, w_min=0.0, synaptic_resource=1.0)
    push_b = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 50, b, rng.integers(0, 50, S2).astype(np.uint32),
                                            rng.integers(0, m, S2).astype(np.uint32),
                                            (rng.integers(8, 64, S2) / 256.0).astype(np.float32),
                                            rng.integers(1, 4, S2).astype(np.uint32), abi.rule_array(rule, S2))
    o.add_delta_projection(a, n, b, rng.integers(0, n, S2).astype(np.uint32), rng.integers(0, m, S2).astype(np.uint32),
                           (rng.integers(16, 96, S2) / 256.0).astype(np.float32), np.ones(S2, np.uint32))
    return (a, b), push_a, push_b


def mt_random_pushes(rng, steps, n_a, n_b=50):
    """Per step: pushes into the two channel projections, some with repeated indexes (up to 3 occurrences),
    some steps with a second message (dropped under the multi-threaded semantics)."""
    out = []
    for _ in range(steps):
        msgs = []
        for proj, n in ((0, n_a), (1, n_b)):
            for _k in range(int(rng.integers(0, 3))):
                idx = rng.choice(n, int(rng.integers(0, n // 8 + 1)), replace=False).astype(np.uint32)
                rep = idx[: len(idx) // 4]
                msgs.append((proj, np.concatenate([idx, rep, rep[:len(rep) // 2]])))
        out.append(msgs)
    return out
