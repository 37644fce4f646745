"""Reader of the network files written by knp_mi355x_save_network (test helper).

Layout (little-endian): magic "KNPNET01", u32 version, u32 populations, u32 projections; per population
u32 model (0 BLIFAT, 1 resource BLIFAT, 2 LIF, 3 AltAI-LIF, 4 resource AltAI-LIF), u32 count, u32 struct
size, then the C ABI structs; per projection u32 kind (0 Delta, 1 synaptic resource, 2 additive), u32
presynaptic population, u32 presynaptic size, u32 postsynaptic population, u64 synapses, u32 locked, then
pre u32[], post u32[], weight f32[], delay u32[], output type u8[]; resource: u32 rule size and the rules;
additive: tau_plus f32[], tau_minus f32[], u32 count, STDP populations u32[], processing types u8[].
"""
import ctypes as C

import numpy as np

from knp_amd import abi

MODELS = {0: ("add_blifat_population", abi.BLIFATNeuron),
          1: ("add_resource_blifat_population", abi.ResourceBLIFATNeuron),
          2: ("add_lif_population", abi.LIFNeuron),
          3: ("add_altai_lif_population", abi.AltAILIFNeuron),
          4: ("add_resource_altai_lif_population", abi.ResourceAltAILIFNeuron)}


def parse(path):
    data = open(path, "rb").read()
    pos = 0

    def take(dtype, count=1):
        nonlocal pos
        a = np.frombuffer(data, dtype=dtype, count=count, offset=pos)
        pos += a.nbytes
        return a

    assert data[:8] == b"KNPNET01"
    pos = 8
    version, n_pops, n_projs = (int(x) for x in take(np.uint32, 3))
    assert version == 1
    pops, projs = [], []
    for _ in range(n_pops):
        model, n, size = (int(x) for x in take(np.uint32, 3))
        method, struct = MODELS[model]
        assert size == C.sizeof(struct)
        arr = (struct * n).from_buffer_copy(data[pos:pos + n * size]) if n else (struct * 0)()
        pos += n * size
        pops.append((method, arr))
    for _ in range(n_projs):
        kind, pre_pop, n_pre, post_pop = (int(x) for x in take(np.uint32, 4))
        n = int(take(np.uint64)[0])
        locked = int(take(np.uint32)[0])
        p = dict(kind=kind, pre_pop=pre_pop, n_pre=n_pre, post_pop=post_pop, locked=locked,
                 pre=take(np.uint32, n), post=take(np.uint32, n), weight=take(np.float32, n),
                 delay=take(np.uint32, n), otype=take(np.uint8, n))
        if kind == abi.PROJ_RESOURCE:
            size = int(take(np.uint32)[0])
            assert size == C.sizeof(abi.ResourceRule)
            p["rules"] = (abi.ResourceRule * n).from_buffer_copy(data[pos:pos + n * size])
            pos += n * size
        elif kind == abi.PROJ_ADDITIVE:
            p["tau_plus"], p["tau_minus"] = take(np.float32, n), take(np.float32, n)
            k = int(take(np.uint32)[0])
            p["stdp_pops"], p["stdp_modes"] = take(np.uint32, k), take(np.uint8, k)
        projs.append(p)
    assert pos == len(data)
    return pops, projs


def load_into(o, path):
    """Add a network file's populations and projections to `o` (the oracle or the device backend)."""
    pops, projs = parse(path)
    for method, arr in pops:
        getattr(o, method)(arr)
    for q, p in enumerate(projs):
        args = (p["pre_pop"], p["n_pre"], p["post_pop"], p["pre"], p["post"], p["weight"], p["delay"])
        if p["kind"] == abi.PROJ_DELTA:
            o.add_delta_projection(*args, p["otype"])
        elif p["kind"] == abi.PROJ_RESOURCE:
            o.add_resource_stdp_projection(*args, p["rules"], p["otype"])
        else:
            o.add_additive_stdp_projection(*args, p["tau_plus"], p["tau_minus"], p["stdp_pops"], p["stdp_modes"],
                                           p["otype"])
        o.set_locked(q, p["locked"])
    return pops, projs
