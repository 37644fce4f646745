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
