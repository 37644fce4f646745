"""The network file layout (knp_mi355x_save_network / load_network) read back on the CPU: a file written by
hand in the documented layout is parsed by the test reader and by the package's size reader, and the oracle
loaded from it runs the reference's SmallestNetwork known answer (single_thread_cpu_test.cpp:54-92)."""
import ctypes as C

import numpy as np

from knp_amd import abi
from knp_amd.backend import read_network_sizes
from oracle.pyoracle import Oracle

import netfile


def smallest_network_file(path):
    """SmallestNetwork: one default BLIFAT neuron, a channel projection (w 1, d 1) and a loop (w 1.1, d 6)."""
    out = [b"KNPNET01", np.array([1, 1, 2], np.uint32).tobytes()]
    neuron = abi.blifat_defaults()
    out += [np.array([0, 1, C.sizeof(neuron)], np.uint32).tobytes(), bytes(neuron)]
    for pre_pop, w, d in ((abi.KNP_NO_POPULATION, 1.0, 1), (0, 1.1, 6)):
        out += [np.array([abi.PROJ_DELTA, pre_pop, 1, 0], np.uint32).tobytes(), np.array([1], np.uint64).tobytes(),
                np.array([1], np.uint32).tobytes(), np.array([0], np.uint32).tobytes(), np.array([0], np.uint32).tobytes(),
                np.array([w], np.float32).tobytes(), np.array([d], np.uint32).tobytes(),
                np.array([abi.EXCITATORY], np.uint8).tobytes()]
    open(path, "wb").write(b"".join(out))


def test_network_file_layout(tmp_path, oracle_built):
    path = str(tmp_path / "smallest.knpnet")
    smallest_network_file(path)
    pops, projs = netfile.parse(path)
    assert [p[0] for p in pops] == ["add_blifat_population"] and len(projs) == 2
    assert read_network_sizes(path) == ([1], [1, 1])
    o = Oracle("exact")
    netfile.load_into(o, path)
    res = []
    for step in range(20):
        if step % 5 == 0:
            o.push_spikes(0, [0])
        o.step()
        if len(o.last_spikes(0)):
            res.append(step)
    assert res == [1, 6, 7, 11, 12, 13, 16, 17, 18, 19]
