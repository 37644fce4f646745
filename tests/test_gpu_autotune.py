"""GPU tests of the step tuner (knp_mi355x_autotune): tuning from any state leaves that state as it was
(every choice it tries gives identical results and the state is restored after each trial), so a tuned
network continues exactly as the oracle; the choice is cached per network shape."""
import os
import sys

import numpy as np
import pytest

from knp_amd import abi, workloads
from knp_amd.backend import MI355XBackend, device_count
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert device_count() >= 1, "no HIP device visible"


@pytest.fixture()
def tune_cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("KNP_MI355X_TUNE_CACHE", str(path))
    return path


@pytest.mark.parametrize("scatter", [abi.SCATTER_AUTO, abi.SCATTER_TILES, abi.SCATTER_STEPWISE])
def test_tuning_mid_run_keeps_the_state(tune_cache, scatter):
    """The small plain network (one workgroup, persistent and stepwise schedules all possible): 100 steps, a
    pushed message, tuning, 200 more steps; spikes, impact counts and potentials equal the oracle's."""
    say(f"case: step tuner on plain_small, scatter mode {scatter}")
    spec = workloads.plain_small(seed=4, n=1000, n_syn=100_000)
    g = MI355XBackend(0)
    g.set_scatter_mode(scatter)
    spec.load_into(g)
    o = spec.load_into(Oracle("exact"))
    for x in (g, o):
        x.record_spikes(0, 300)
        x.step(100)
        x.push_spikes(spec.meta.get("input_proj", 0), [3, 5, 7])
    choice = g.autotune(64)
    say(f"  chosen {choice}")
    assert g.get_step() == 100
    g.step(200)
    o.step(200)
    assert np.array_equal(g.read_spike_record(0), o.read_spike_record(0))
    assert g.events == o.events
    vg, vo = g.potentials(0).astype(np.float64), o.potentials(0)
    assert np.max(np.abs(vg - vo) / np.maximum(np.abs(vo), 1e-30)) <= 1e-6
    assert tune_cache.read_text().startswith("step ")
    # a network of the same shape takes the cached choice
    h = MI355XBackend(0)
    h.set_scatter_mode(scatter)
    spec.load_into(h)
    assert h.autotune(64) == choice
    g.close()
    h.close()


def test_tuning_a_training_network(tune_cache):
    """Synaptic-resource training (stepwise launches: graph length and training form are tried): tuned after 40
    steps, then 200 steps; spikes and weights equal the oracle's bit for bit."""
    say("case: step tuner on the STDP network")
    spec = workloads.stdp_learning(n_images=20)
    g = spec.load_into(MI355XBackend(0))
    o = spec.load_into(Oracle("exact"))
    for x in (g, o):
        x.record_spikes(0, 240)
        x.step(40)
    choice = g.autotune(64)
    say(f"  chosen {choice}")
    assert choice["schedule"] == abi.SCHEDULE_STEPWISE
    g.step(200)
    o.step(200)
    assert np.array_equal(g.read_spike_record(0), o.read_spike_record(0))
    assert np.array_equal(np.asarray(g.weights(0), np.float32), np.asarray(o.weights(0), np.float32))
    g.close()
