"""ctypes binding of the CPU oracle (oracle/knp_oracle.c). TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this module; the product path
(knp_amd) never does. The oracle restates the semantics of the reference single-threaded CPU backend
(knp/backends/cpu/cpu-single-threaded-backend/impl/backend.cpp:89-141); see knp_oracle.c for the
file-by-file citations.
"""
import ctypes as C
import os

import numpy as np

from knp_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    import subprocess

    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def _load(policy, path=None):
    """The oracle library of a precision policy; path: a build of the oracle elsewhere (the tests that plant
    mistakes in a copy of the oracle source load their builds this way)."""
    key = path or policy
    if key in _LIBS:
        return _LIBS[key]
    if path is None:
        name = {"exact": "liboracle_exact.so", "f64": "liboracle_f64.so", "mt": "liboracle_mt.so",
                "release": "liboracle_release.so"}[policy]
        path = os.path.join(_HERE, "lib", name)
        if not os.path.exists(path):
            build()
    lib = C.CDLL(path)
    P = C.c_void_p
    u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int
    pu32, pf32, pu8, pf64 = C.POINTER(u32), C.POINTER(C.c_float), C.POINTER(C.c_uint8), C.POINTER(C.c_double)
    sig = {
        "ko_create": (P, []),
        "ko_destroy": (None, [P]),
        "ko_last_error": (C.c_char_p, [P]),
        "ko_set_threads": (i32, [P, i32]),
        "ko_set_message_semantics": (i32, [P, i32]),
        "ko_add_blifat_population": (i32, [P, C.POINTER(abi.BLIFATNeuron), u32]),
        "ko_add_resource_blifat_population": (i32, [P, C.POINTER(abi.ResourceBLIFATNeuron), u32]),
        "ko_add_lif_population": (i32, [P, C.POINTER(abi.LIFNeuron), u32]),
        "ko_add_altai_lif_population": (i32, [P, C.POINTER(abi.AltAILIFNeuron), u32]),
        "ko_add_resource_altai_lif_population": (i32, [P, C.POINTER(abi.ResourceAltAILIFNeuron), u32]),
        "ko_inject_impacts": (i32, [P, u32, pu32, pf32, pu8, u32]),
        "ko_add_delta_projection": (i32, [P, u32, u32, u32, pu32, pu32, pf32, pu32, pu8, u64]),
        "ko_add_resource_stdp_projection": (
            i32, [P, u32, u32, u32, pu32, pu32, pf32, pu32, pu8, C.POINTER(abi.ResourceRule), u64]),
        "ko_add_additive_stdp_projection": (
            i32, [P, u32, u32, u32, pu32, pu32, pf32, pu32, pu8, pf32, pf32, u64, pu32, pu8, u32]),
        "ko_set_locked": (i32, [P, u32, i32]),
        "ko_record_spikes": (i32, [P, u32, u64]),
        "ko_read_spike_record": (i32, [P, u32, pu32, u64, C.POINTER(u64)]),
        "ko_push_spikes": (i32, [P, u32, pu32, u32]),
        "ko_add_bernoulli_source": (i32, [P, u32, u32, u32, pu32, u32, u32, u32, u32, u64, u32]),
        "ko_step": (i32, [P, u64]),
        "ko_get_step": (u64, [P]),
        "ko_get_event_count": (u64, [P]),
        "ko_read_potentials": (i32, [P, u32, pf64, u32]),
        "ko_read_last_spikes": (i32, [P, u32, pu32, u32, C.POINTER(u32)]),
        "ko_read_blifat_neurons": (i32, [P, u32, C.POINTER(abi.BLIFATNeuron), u32]),
        "ko_read_resource_neurons": (i32, [P, u32, C.POINTER(abi.ResourceBLIFATNeuron), u32]),
        "ko_read_lif_neurons": (i32, [P, u32, C.POINTER(abi.LIFNeuron), u32]),
        "ko_read_altai_lif_neurons": (i32, [P, u32, C.POINTER(abi.AltAILIFNeuron), u32]),
        "ko_read_resource_altai_lif_neurons": (i32, [P, u32, C.POINTER(abi.ResourceAltAILIFNeuron), u32]),
        "ko_read_weights": (i32, [P, u32, pf64, u64]),
        "ko_read_resources": (i32, [P, u32, pf64, u64]),
        "ko_read_resource_rules": (i32, [P, u32, C.POINTER(abi.ResourceRule), u64]),
        "ko_bernoulli_hash": (u32, [u64, u64, u32]),
        "ko_wreal_bytes": (i32, []),
    }
    for name_, (res, args) in sig.items():
        f = getattr(lib, name_)
        f.restype = res
        f.argtypes = args
    _LIBS[key] = lib
    return lib


def bernoulli_hash(seed, step, channel):
    return _load("exact").ko_bernoulli_hash(seed, step, channel)


class Oracle:
    """One network instance of the restated single-threaded CPU backend. threads > 1: the multi-threaded
    backend's decomposition of the same step (exact policy only; same results as one thread)."""

    def __init__(self, policy="exact", threads=1, lib_path=None):
        if threads > 1:
            if policy != "exact":
                raise ValueError("the multi-threaded step is built for the exact policy only")
            policy = "mt"
        self.lib = _load(policy, lib_path)
        self.net = self.lib.ko_create()
        if threads > 1:
            self.lib.ko_set_threads(self.net, int(threads))
        self.pop_sizes = []
        self.proj_sizes = []
        self._rec_cap = {}

    def close(self):
        if self.net:
            self.lib.ko_destroy(self.net)
            self.net = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise RuntimeError("oracle: " + self.lib.ko_last_error(self.net).decode())
        return rc

    # ---- loading ------------------------------------------------------------------------------
    def add_blifat_population(self, neurons):
        n = len(neurons)
        pid = self._chk(self.lib.ko_add_blifat_population(self.net, neurons, n))
        self.pop_sizes.append(n)
        return pid

    def add_resource_blifat_population(self, neurons):
        n = len(neurons)
        pid = self._chk(self.lib.ko_add_resource_blifat_population(self.net, neurons, n))
        self.pop_sizes.append(n)
        return pid

    def add_lif_population(self, neurons):
        n = len(neurons)
        pid = self._chk(self.lib.ko_add_lif_population(self.net, neurons, n))
        self.pop_sizes.append(n)
        return pid

    def add_altai_lif_population(self, neurons):
        n = len(neurons)
        pid = self._chk(self.lib.ko_add_altai_lif_population(self.net, neurons, n))
        self.pop_sizes.append(n)
        return pid

    def add_resource_altai_lif_population(self, neurons):
        n = len(neurons)
        pid = self._chk(self.lib.ko_add_resource_altai_lif_population(self.net, neurons, n))
        self.pop_sizes.append(n)
        return pid

    def inject_impacts(self, pop, neurons, weights, otype=None):
        """Impacts delivered straight to `pop` at the next step (a test hook with no projection)."""
        neurons = np.ascontiguousarray(neurons, dtype=np.uint32)
        weights = np.ascontiguousarray(weights, dtype=np.float32)
        otype = None if otype is None else np.ascontiguousarray(otype, dtype=np.uint8)
        self._chk(self.lib.ko_inject_impacts(self.net, pop, abi.u32p(neurons), abi.f32p(weights),
                                             abi.u8p(otype), len(neurons)))

    @staticmethod
    def _syn(pre, post, weight, delay, otype):
        pre = np.ascontiguousarray(pre, dtype=np.uint32)
        post = np.ascontiguousarray(post, dtype=np.uint32)
        weight = np.ascontiguousarray(weight, dtype=np.float32)
        delay = np.ascontiguousarray(delay, dtype=np.uint32)
        otype = None if otype is None else np.ascontiguousarray(otype, dtype=np.uint8)
        return pre, post, weight, delay, otype

    def add_delta_projection(self, pre_pop, n_pre, post_pop, pre, post, weight, delay, otype=None):
        pre, post, weight, delay, otype = self._syn(pre, post, weight, delay, otype)
        q = self._chk(self.lib.ko_add_delta_projection(
            self.net, pre_pop, n_pre, post_pop, abi.u32p(pre), abi.u32p(post), abi.f32p(weight), abi.u32p(delay),
            abi.u8p(otype), len(pre)))
        self.proj_sizes.append(len(pre))
        return q

    def add_resource_stdp_projection(self, pre_pop, n_pre, post_pop, pre, post, weight, delay, rules, otype=None):
        pre, post, weight, delay, otype = self._syn(pre, post, weight, delay, otype)
        q = self._chk(self.lib.ko_add_resource_stdp_projection(
            self.net, pre_pop, n_pre, post_pop, abi.u32p(pre), abi.u32p(post), abi.f32p(weight), abi.u32p(delay),
            abi.u8p(otype), rules, len(pre)))
        self.proj_sizes.append(len(pre))
        return q

    def add_additive_stdp_projection(self, pre_pop, n_pre, post_pop, pre, post, weight, delay, tau_plus, tau_minus,
                                     stdp_pops=(), stdp_modes=(), otype=None):
        pre, post, weight, delay, otype = self._syn(pre, post, weight, delay, otype)
        tp = np.ascontiguousarray(tau_plus, dtype=np.float32)
        tm = np.ascontiguousarray(tau_minus, dtype=np.float32)
        sp = np.ascontiguousarray(stdp_pops, dtype=np.uint32)
        sm = np.ascontiguousarray(stdp_modes, dtype=np.uint8)
        q = self._chk(self.lib.ko_add_additive_stdp_projection(
            self.net, pre_pop, n_pre, post_pop, abi.u32p(pre), abi.u32p(post), abi.f32p(weight), abi.u32p(delay),
            abi.u8p(otype), abi.f32p(tp), abi.f32p(tm), len(pre), abi.u32p(sp), abi.u8p(sm), len(sp)))
        self.proj_sizes.append(len(pre))
        return q

    def set_locked(self, proj, locked):
        self._chk(self.lib.ko_set_locked(self.net, proj, 1 if locked else 0))

    def start_learning(self):
        for q in range(len(self.proj_sizes)):
            self.set_locked(q, False)

    def set_message_semantics(self, semantics):
        """abi.SEMANTICS_SINGLE_THREADED or abi.SEMANTICS_MULTI_THREADED (the multi-threaded backend's projections:
        the first message of a step only, repeated indexes as weight multiples)."""
        self._chk(self.lib.ko_set_message_semantics(self.net, int(semantics)))

    # ---- inputs / stepping --------------------------------------------------------------------
    def push_spikes(self, proj, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        self._chk(self.lib.ko_push_spikes(self.net, proj, abi.u32p(idx), len(idx)))

    def add_bernoulli_source(self, proj, thresholds, steps_per_pattern=1, active_lo=0, active_hi=None, seed=0,
                             first_index=0, hash_offset=0):
        thr = np.ascontiguousarray(np.atleast_2d(thresholds), dtype=np.uint32)
        n_patterns, n_ch = thr.shape
        hi = steps_per_pattern - 1 if active_hi is None else active_hi
        self._chk(self.lib.ko_add_bernoulli_source(self.net, proj, first_index, n_ch, abi.u32p(thr), n_patterns,
                                                   steps_per_pattern, active_lo, hi, seed, hash_offset))

    def add_dense_input_projection(self, post_pop, post_begin, post_count, n_in, fmt, weights, row_scale,
                                   threshold, seed, hash_offset=0, time_block=0):
        """The dense projection stated as what it stands for: a Projection<DeltaSynapse> from channel block k
        (indexes k * n_in + i) to every neuron of instance k with weight row_scale[j] * w(j, i), EXCITATORY,
        delay 1, fed by one Bernoulli source over all channels."""
        pb = np.asarray(post_begin, dtype=np.int64)
        pc = np.asarray(post_count, dtype=np.int64)
        K = len(pb)
        rows = int(pc.sum())
        if fmt == "f32":  # the unquantized weights (rows x n_in float32), for agreement measurements
            w = np.asarray(weights, dtype=np.float64).reshape(rows, n_in)
        elif fmt == abi.DENSE_FP8_E4M3:
            w = abi.e4m3_decode(np.asarray(weights).reshape(rows, n_in))
        else:
            w = abi.bf16_decode(np.asarray(weights).reshape(rows, n_in)).astype(np.float64)
        sc = np.ones(rows) if row_scale is None else np.asarray(row_scale, dtype=np.float64)
        wf = (w * sc[:, None]).astype(np.float32)
        pre, post, wl = [], [], []
        r0 = 0
        for k in range(K):
            jj, ii = np.meshgrid(np.arange(pc[k]), np.arange(n_in), indexing="ij")
            post.append((pb[k] + jj).ravel())
            pre.append((k * n_in + ii).ravel())
            wl.append(wf[r0:r0 + pc[k]].ravel())
            r0 += pc[k]
        pre = np.concatenate(pre)
        post = np.concatenate(post)
        wl = np.concatenate(wl)
        q = self.add_delta_projection(abi.KNP_NO_POPULATION, K * n_in, post_pop, pre, post, wl,
                                      np.ones(len(pre), np.uint32))
        self.add_bernoulli_source(q, np.full(K * n_in, threshold, np.uint32), seed=seed, hash_offset=hash_offset)
        return q

    def step(self, n=1):
        self._chk(self.lib.ko_step(self.net, n))

    @property
    def current_step(self):
        return self.lib.ko_get_step(self.net)

    @property
    def events(self):
        return self.lib.ko_get_event_count(self.net)

    # ---- outputs ------------------------------------------------------------------------------
    def record_spikes(self, pop, capacity_steps):
        self._chk(self.lib.ko_record_spikes(self.net, pop, capacity_steps))
        self._rec_cap[pop] = capacity_steps

    def read_spike_record(self, pop):
        n = self.pop_sizes[pop]
        words_per = (n + 31) // 32
        cap = self._rec_cap.get(pop, 0)
        buf = np.zeros(words_per * max(cap, 1), dtype=np.uint32)
        steps = C.c_uint64(0)
        rc = self.lib.ko_read_spike_record(self.net, pop, abi.u32p(buf), buf.size, C.byref(steps))
        if rc < 0:
            raise RuntimeError("oracle: spike record larger than the read buffer")
        return abi.unpack_bitmaps(buf, steps.value, n)

    def last_spikes(self, pop):
        n = self.pop_sizes[pop]
        buf = np.zeros(max(n, 1), dtype=np.uint32)
        cnt = C.c_uint32(0)
        self._chk(self.lib.ko_read_last_spikes(self.net, pop, abi.u32p(buf), n, C.byref(cnt)))
        return buf[: cnt.value].copy()

    def potentials(self, pop):
        n = self.pop_sizes[pop]
        out = np.zeros(n, dtype=np.float64)
        self._chk(self.lib.ko_read_potentials(self.net, pop, abi.f64p(out), n))
        return out

    def blifat_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.BLIFATNeuron * n)()
        self._chk(self.lib.ko_read_blifat_neurons(self.net, pop, out, n))
        return out

    def resource_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.ResourceBLIFATNeuron * n)()
        self._chk(self.lib.ko_read_resource_neurons(self.net, pop, out, n))
        return out

    def lif_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.LIFNeuron * n)()
        self._chk(self.lib.ko_read_lif_neurons(self.net, pop, out, n))
        return out

    def altai_lif_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.AltAILIFNeuron * n)()
        self._chk(self.lib.ko_read_altai_lif_neurons(self.net, pop, out, n))
        return out

    def resource_altai_lif_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.ResourceAltAILIFNeuron * n)()
        self._chk(self.lib.ko_read_resource_altai_lif_neurons(self.net, pop, out, n))
        return out

    def weights(self, proj):
        n = self.proj_sizes[proj]
        out = np.zeros(n, dtype=np.float64)
        self._chk(self.lib.ko_read_weights(self.net, proj, abi.f64p(out), n))
        return out

    def resources(self, proj):
        n = self.proj_sizes[proj]
        out = np.zeros(n, dtype=np.float64)
        self._chk(self.lib.ko_read_resources(self.net, proj, abi.f64p(out), n))
        return out

    def resource_rules(self, proj):
        n = self.proj_sizes[proj]
        out = (abi.ResourceRule * n)()
        self._chk(self.lib.ko_read_resource_rules(self.net, proj, out, n))
        return out
