"""The C ABI library loads, exports every entry point include/knp_mi355x.h declares, and its struct layout
matches the ctypes mirror. No compute calls: these run without a GPU."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from knp_amd import abi
from knp_amd.backend import LIB_PATH, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "knp_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(knp_\w+)\s*\(", text)))


def test_library_present():
    assert os.path.exists(LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_declared_function():
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_exports_nothing_undeclared():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = sorted({l.split()[-1] for l in out.splitlines() if " T knp" in l})
    assert exported == declared_functions()


def test_defaults_match_reference_values():
    lib = load_library()
    n = abi.BLIFATNeuron()
    lib.knp_blifat_defaults(C.byref(n))
    assert bytes(n) == bytes(abi.blifat_defaults())
    assert n.n_time_steps_since_last_firing == 2**64 - 1  # blifat.h:53
    assert n.activation_threshold == 1.0 and n.reversal_inhibitory_potential == -0.3
    assert n.min_potential == -1e9 and n.total_blocking_period == 2**63 - 1
    r = abi.ResourceBLIFATNeuron()
    lib.knp_resource_blifat_defaults(C.byref(r))
    ref = abi.resource_blifat_defaults()
    for f, _ in abi.ResourceBLIFATNeuron._fields_[1:]:
        assert getattr(r, f) == getattr(ref, f), f
    assert r.synaptic_resource_threshold == np.finfo(np.float32).max
    rr = abi.ResourceRule()
    lib.knp_resource_rule_defaults(C.byref(rr))
    assert (rr.w_min, rr.w_max, rr.d_u, rr.synaptic_resource) == (0.0, 1.0, 0.0, 0.0)
    lf = abi.LIFNeuron()
    lib.knp_lif_defaults(C.byref(lf))
    assert bytes(lf) == bytes(abi.lif_defaults())
    # lif.h:43-67: potential 0, reset 0, threshold 1, leak 1, counter 0, period 0
    assert (lf.potential, lf.potential_reset_value, lf.activation_threshold, lf.leak_coefficient) == (0, 0, 1, 1)
    assert (lf.refract_counter, lf.refract_period) == (0, 0)
    al = abi.AltAILIFNeuron()
    lib.knp_altai_lif_defaults(C.byref(al))
    assert bytes(al) == bytes(abi.altai_lif_defaults())
    # altai_lif.h:43-122
    assert (al.is_diff, al.is_reset, al.leak_rev, al.saturate, al.do_not_save) == (0, 1, 1, 1, 0)
    assert (al.activation_threshold, al.negative_activation_threshold, al.potential_leak) == (1, 30000, 0)
    assert al.activity_time == 0 and al.additional_threshold == 0.0
    ra = abi.ResourceAltAILIFNeuron()
    lib.knp_resource_altai_lif_defaults(C.byref(ra))
    assert bytes(ra.base) == bytes(al)
    ref = abi.resource_altai_lif_defaults()
    for f, _ in abi.ResourceAltAILIFNeuron._fields_[1:]:
        assert getattr(ra, f) == getattr(ref, f) == getattr(r, f), f


def test_bernoulli_hash_matches_oracle(oracle_built):
    from oracle.pyoracle import bernoulli_hash

    lib = load_library()
    rng = np.random.default_rng(0)
    for _ in range(100):
        s, t, c = int(rng.integers(0, 2**63)), int(rng.integers(0, 2**40)), int(rng.integers(0, 2**32))
        assert lib.knp_bernoulli_hash(s, t, c) == bernoulli_hash(s, t, c)


def test_struct_layout_matches_header():
    """sizeof/offsetof of the header structs, compiled with gcc, equal the ctypes mirror."""
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "knp_mi355x.h"
#define P(T, F) printf(#T "." #F " %zu\n", offsetof(T, F));
int main(void) {
  printf("knp_blifat_neuron %zu\n", sizeof(knp_blifat_neuron));
  printf("knp_resource_blifat_neuron %zu\n", sizeof(knp_resource_blifat_neuron));
  printf("knp_resource_rule %zu\n", sizeof(knp_resource_rule));
  printf("knp_lif_neuron %zu\n", sizeof(knp_lif_neuron));
  P(knp_lif_neuron, refract_counter) P(knp_lif_neuron, refract_period)
  printf("knp_altai_lif_neuron %zu\n", sizeof(knp_altai_lif_neuron));
  printf("knp_resource_altai_lif_neuron %zu\n", sizeof(knp_resource_altai_lif_neuron));
  P(knp_resource_altai_lif_neuron, d_h) P(knp_resource_altai_lif_neuron, is_being_forced)
  P(knp_altai_lif_neuron, potential) P(knp_altai_lif_neuron, potential_reset_value)
  P(knp_altai_lif_neuron, dopamine_value) P(knp_altai_lif_neuron, activity_time)
  P(knp_blifat_neuron, potential) P(knp_blifat_neuron, total_blocking_period)
  P(knp_blifat_neuron, random_number_generator_state)
  P(knp_resource_blifat_neuron, d_h) P(knp_resource_blifat_neuron, is_being_forced)
  P(knp_resource_rule, last_spike_step) P(knp_resource_rule, has_contributed)
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(src, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, src], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    vals = dict(l.rsplit(" ", 1) for l in out.strip().splitlines())
    assert int(vals["knp_blifat_neuron"]) == C.sizeof(abi.BLIFATNeuron)
    assert int(vals["knp_resource_blifat_neuron"]) == C.sizeof(abi.ResourceBLIFATNeuron)
    assert int(vals["knp_resource_rule"]) == C.sizeof(abi.ResourceRule)
    assert int(vals["knp_lif_neuron"]) == C.sizeof(abi.LIFNeuron)
    assert int(vals["knp_lif_neuron.refract_counter"]) == abi.LIFNeuron.refract_counter.offset
    assert int(vals["knp_lif_neuron.refract_period"]) == abi.LIFNeuron.refract_period.offset
    assert int(vals["knp_altai_lif_neuron"]) == C.sizeof(abi.AltAILIFNeuron)
    for f in ("potential", "potential_reset_value", "dopamine_value", "activity_time"):
        assert int(vals["knp_altai_lif_neuron." + f]) == getattr(abi.AltAILIFNeuron, f).offset, f
    assert int(vals["knp_resource_altai_lif_neuron"]) == C.sizeof(abi.ResourceAltAILIFNeuron)
    for f in ("d_h", "is_being_forced"):
        assert int(vals["knp_resource_altai_lif_neuron." + f]) == getattr(abi.ResourceAltAILIFNeuron, f).offset, f
    assert int(vals["knp_blifat_neuron.potential"]) == abi.BLIFATNeuron.potential.offset
    assert int(vals["knp_blifat_neuron.total_blocking_period"]) == abi.BLIFATNeuron.total_blocking_period.offset
    assert int(vals["knp_blifat_neuron.random_number_generator_state"]) == \
        abi.BLIFATNeuron.random_number_generator_state.offset
    assert int(vals["knp_resource_blifat_neuron.d_h"]) == abi.ResourceBLIFATNeuron.d_h.offset
    assert int(vals["knp_resource_blifat_neuron.is_being_forced"]) == abi.ResourceBLIFATNeuron.is_being_forced.offset
    assert int(vals["knp_resource_rule.last_spike_step"]) == abi.ResourceRule.last_spike_step.offset
    assert int(vals["knp_resource_rule.has_contributed"]) == abi.ResourceRule.has_contributed.offset


def test_create_without_device_fails_loudly():
    """Without a visible HIP device the backend refuses to construct (no CPU fallback)."""
    from knp_amd.backend import BackendError, MI355XBackend, device_count

    if device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(BackendError):
        MI355XBackend(0)
