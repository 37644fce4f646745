"""The resource-STDP hand traces (tests/resource_traces.py) pin the oracle: each test here plants one
plausible mistake in a copy of the oracle's source (a dropped term, a wrong branch, a missing wrap, a wrong
sign), builds it with the oracle's own flags, and checks that the traces reject it, while the unmodified
copy passes them. Runs on the CPU."""
import os
import shutil
import subprocess

import pytest

import resource_traces
from oracle.pyoracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (name, text in oracle/knp_oracle.c, replacement)
MUTATIONS = [
    ("renormalization without the drain coefficient",
     "(wreal)(uint64_t)(k + x->resource_drain_coefficient)", "(wreal)(uint64_t)(k)"),
    ("stability branches swapped",
     "if (x->is_being_forced || dopamine < 0) {", "if (!(x->is_being_forced || dopamine < 0)) {"),
    ("has_contributed window without the u64 wrap",
     "if (b > e) { uint64_t t = b; b = e; e = t; }", ""),
    ("Hebbian increment without the stability factor",
     "const wreal d_h = x->d_h * stability_factor(x->stability);", "const wreal d_h = x->d_h;"),
    ("stability factor with the wrong sign",
     "wreal v = (wreal)pow(2.0, -(double)stability);", "wreal v = (wreal)pow(2.0, (double)stability);"),
    ("weight recompute without w_min",
     "s[i].proj->w[s[i].syn] = r->w_min + weight_diff * syn_w / (weight_diff + syn_w);",
     "s[i].proj->w[s[i].syn] = weight_diff * syn_w / (weight_diff + syn_w);"),
    ("free resource not credited with d_u",
     "x->free_synaptic_resource += r->d_u;", ""),
    ("dopamine time window off by one",
     "if (step - x->last_spike_step <= (uint64_t)x->dopamine_plasticity_time && r->has_contributed) {",
     "if (step - x->last_spike_step < (uint64_t)x->dopamine_plasticity_time && r->has_contributed) {"),
]


def build_copy(tmp_path, old=None, new=None):
    (tmp_path / "oracle").mkdir()
    (tmp_path / "include").mkdir()
    shutil.copy(os.path.join(ROOT, "include", "knp_mi355x.h"), tmp_path / "include")
    src = open(os.path.join(ROOT, "oracle", "knp_oracle.c")).read()
    if old is not None:
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    (tmp_path / "oracle" / "knp_oracle.c").write_text(src)
    lib = tmp_path / "oracle" / "liboracle_planted.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-std=c11",
                    "-DKO_WREAL=float", "-o", str(lib), str(tmp_path / "oracle" / "knp_oracle.c"), "-lm"],
                   check=True)
    return str(lib)


def test_unmodified_copy_passes(tmp_path):
    lib = build_copy(tmp_path)
    resource_traces.run_all_traces(lambda: Oracle("exact", lib_path=lib))


@pytest.mark.parametrize("name,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_is_caught(tmp_path, name, old, new):
    lib = build_copy(tmp_path, old, new)
    with pytest.raises(AssertionError):
        resource_traces.run_all_traces(lambda: Oracle("exact", lib_path=lib))
