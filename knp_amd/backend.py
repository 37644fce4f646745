"""Python binding of the MI355X backend library (knp_amd/libknp_mi355x.so, C ABI in include/knp_mi355x.h).

`MI355XBackend` mirrors the loading / stepping surface of KNP's backends
(knp/core-library/include/knp/core/backend.h:55-457 and SingleThreadedCPUBackend,
knp/backends/cpu/cpu-single-threaded-backend/include/knp/backends/cpu-single-threaded/backend.h):
populations and projections are added in load order, `start_learning` / `stop_learning` unlock / lock
projections, `_step` advances one network step and `get_step` returns the step counter. Errors raised by the
library surface as `BackendError` (a RuntimeError) carrying the library's message.

There is no fallback: if the library is missing or no HIP device is present, construction raises.
"""
import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libknp_mi355x.so")
_LIB = None
# Networks still open when the interpreter exits are destroyed by an exit hook, while the HIP runtime is still
# up; a finalizer that runs later (interpreter teardown) makes no HIP call.
_OPEN = weakref.WeakSet()


@atexit.register
def _close_open_networks():
    for g in list(_OPEN):
        g.close()


class BackendError(RuntimeError):
    pass


def _declare(lib):
    P = C.c_void_p
    u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int
    pu32, pf32, pu8 = C.POINTER(u32), C.POINTER(C.c_float), C.POINTER(C.c_uint8)
    sig = {
        "knp_mi355x_version": (C.c_char_p, []),
        "knp_mi355x_last_error": (C.c_char_p, []),
        "knp_mi355x_device_count": (i32, [C.POINTER(i32)]),
        "knp_mi355x_device_name": (i32, [i32, C.c_char_p, C.c_size_t]),
        "knp_mi355x_create": (i32, [i32, C.POINTER(P)]),
        "knp_mi355x_destroy": (i32, [P]),
        "knp_mi355x_set_memory_cap": (i32, [P, u64]),
        "knp_mi355x_set_scatter_mode": (i32, [P, i32]),
        "knp_mi355x_set_message_semantics": (i32, [P, i32]),
        "knp_mi355x_autotune": (i32, [P, u32]),
        "knp_mi355x_get_tuning": (i32, [P, C.POINTER(i32), C.POINTER(i32), pu32, pu32, pu32]),
        "knp_mi355x_set_wait_test": (i32, [P, u32, u32, u64]),
        "knp_mi355x_set_train_workgroup_limit": (i32, [P, u32]),
        "knp_mi355x_get_tiled": (i32, [P, u32, C.POINTER(i32)]),
        "knp_mi355x_get_schedule": (i32, [P, C.POINTER(i32)]),
        "knp_mi355x_get_tile_counters": (i32, [P, C.POINTER(u64), C.POINTER(u64)]),
        "knp_mi355x_add_blifat_population": (i32, [P, C.POINTER(abi.BLIFATNeuron), u32, pu32]),
        "knp_mi355x_add_resource_blifat_population": (i32, [P, C.POINTER(abi.ResourceBLIFATNeuron), u32, pu32]),
        "knp_mi355x_add_lif_population": (i32, [P, C.POINTER(abi.LIFNeuron), u32, pu32]),
        "knp_mi355x_add_altai_lif_population": (i32, [P, C.POINTER(abi.AltAILIFNeuron), u32, pu32]),
        "knp_mi355x_add_resource_altai_lif_population": (i32, [P, C.POINTER(abi.ResourceAltAILIFNeuron), u32, pu32]),
        "knp_mi355x_add_delta_projection": (i32, [P, u32, u32, u32, pu32, pu32, pf32, pu32, pu8, u64, pu32]),
        "knp_mi355x_add_resource_stdp_projection": (
            i32, [P, u32, u32, u32, pu32, pu32, pf32, pu32, pu8, C.POINTER(abi.ResourceRule), u64, pu32]),
        "knp_mi355x_add_additive_stdp_projection": (
            i32, [P, u32, u32, u32, pu32, pu32, pf32, pu32, pu8, pf32, pf32, u64, pu32, pu8, u32, pu32]),
        "knp_mi355x_set_locked": (i32, [P, u32, i32]),
        "knp_mi355x_start_learning": (i32, [P]),
        "knp_mi355x_stop_learning": (i32, [P]),
        "knp_mi355x_init": (i32, [P]),
        "knp_mi355x_push_spikes": (i32, [P, u32, pu32, u32]),
        "knp_mi355x_add_bernoulli_source": (i32, [P, u32, u32, u32, pu32, u32, u32, u32, u32, u64, u32]),
        "knp_mi355x_add_dense_input_projection": (
            i32, [P, u32, u32, pu32, pu32, u32, i32, C.c_void_p, pf32, u32, u64, u32, u32, pu32]),
        "knp_mi355x_get_dense_time_block": (i32, [P, u32, pu32]),
        "knp_mi355x_set_projection_dense": (i32, [P, u32, i32, u32]),
        "knp_mi355x_read_dense_weights": (i32, [P, u32, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), u64]),
        "knp_mi355x_get_dense_weight_bytes": (i32, [P, C.POINTER(u64), C.POINTER(u64)]),
        "knp_mi355x_step": (i32, [P, u64]),
        "knp_mi355x_step_noalloc": (i32, [P, u64]),
        "knp_mi355x_set_step_graph": (i32, [P, u32]),
        "knp_mi355x_get_inexact_weights": (i32, [P, u32, C.POINTER(u64)]),
        "knp_mi355x_get_step_graph": (i32, [P, pu32, C.POINTER(u64)]),
        "knp_mi355x_get_step": (i32, [P, C.POINTER(u64)]),
        "knp_mi355x_synchronize": (i32, [P]),
        "knp_mi355x_record_spikes": (i32, [P, u32, u64]),
        "knp_mi355x_read_spike_record": (i32, [P, u32, pu32, u64, C.POINTER(u64)]),
        "knp_mi355x_read_last_spikes": (i32, [P, u32, pu32, u32, C.POINTER(u32)]),
        "knp_mi355x_read_potentials": (i32, [P, u32, pf32, u32]),
        "knp_mi355x_read_blifat_neurons": (i32, [P, u32, C.POINTER(abi.BLIFATNeuron), u32]),
        "knp_mi355x_read_resource_neurons": (i32, [P, u32, C.POINTER(abi.ResourceBLIFATNeuron), u32]),
        "knp_mi355x_read_lif_neurons": (i32, [P, u32, C.POINTER(abi.LIFNeuron), u32]),
        "knp_mi355x_read_altai_lif_neurons": (i32, [P, u32, C.POINTER(abi.AltAILIFNeuron), u32]),
        "knp_mi355x_read_resource_altai_lif_neurons": (i32, [P, u32, C.POINTER(abi.ResourceAltAILIFNeuron), u32]),
        "knp_mi355x_read_weights": (i32, [P, u32, pf32, u64]),
        "knp_mi355x_read_resource_rules": (i32, [P, u32, C.POINTER(abi.ResourceRule), u64]),
        "knp_mi355x_get_event_count": (i32, [P, C.POINTER(u64)]),
        "knp_mi355x_save_network": (i32, [P, C.c_char_p]),
        "knp_mi355x_load_network": (i32, [P, C.c_char_p]),
        "knp_mi355x_get_counters": (i32, [P, C.POINTER(u64), C.POINTER(u64)]),
        "knp_mi355x_get_accumulator_bits": (i32, [P, u32, C.POINTER(C.c_int32)]),
        "knp_mi355x_enable_timing": (i32, [P, i32]),
        "knp_mi355x_kernel_time": (i32, [P, i32, C.POINTER(C.c_double), C.POINTER(u64)]),
        "knp_blifat_defaults": (None, [C.POINTER(abi.BLIFATNeuron)]),
        "knp_resource_blifat_defaults": (None, [C.POINTER(abi.ResourceBLIFATNeuron)]),
        "knp_lif_defaults": (None, [C.POINTER(abi.LIFNeuron)]),
        "knp_altai_lif_defaults": (None, [C.POINTER(abi.AltAILIFNeuron)]),
        "knp_resource_altai_lif_defaults": (None, [C.POINTER(abi.ResourceAltAILIFNeuron)]),
        "knp_resource_rule_defaults": (None, [C.POINTER(abi.ResourceRule)]),
        "knp_bernoulli_hash": (u32, [u64, u64, u32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load_library(path=None):
    """Load the built library (raises if it has not been built)."""
    global _LIB
    if path is None:
        path = LIB_PATH
    if _LIB is None:
        if not os.path.exists(path):
            raise BackendError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        _LIB = _declare(C.CDLL(path))
    return _LIB


def device_count():
    lib = load_library()
    n = C.c_int(0)
    if lib.knp_mi355x_device_count(C.byref(n)) != 0:
        raise BackendError(lib.knp_mi355x_last_error().decode())
    return n.value


def device_name(dev=0):
    lib = load_library()
    buf = C.create_string_buffer(256)
    if lib.knp_mi355x_device_name(dev, buf, 256) != 0:
        raise BackendError(lib.knp_mi355x_last_error().decode())
    return buf.value.decode()


def read_network_sizes(path):
    """Population sizes and synapse counts of a network file, in load order (the layout written by
    knp_mi355x_save_network: an 8-byte magic and version / population / projection counts, then per
    population a model tag, count and struct size before its structs, per projection a fixed header before
    its synapse arrays)."""
    with open(path, "rb") as f:
        head = f.read(20)
        n_pops, n_projs = np.frombuffer(head[12:20], dtype=np.uint32)
        pops, projs = [], []
        for _ in range(int(n_pops)):
            tag, n, size = np.frombuffer(f.read(12), dtype=np.uint32)
            pops.append(int(n))
            f.seek(int(n) * int(size), 1)
        for _ in range(int(n_projs)):
            kind, _pre, _npre, _post = np.frombuffer(f.read(16), dtype=np.uint32)
            n = int(np.frombuffer(f.read(8), dtype=np.uint64)[0])
            f.seek(4 + n * (4 + 4 + 4 + 4 + 1), 1)
            if kind == abi.PROJ_RESOURCE:
                size = int(np.frombuffer(f.read(4), dtype=np.uint32)[0])
                f.seek(n * size, 1)
            elif kind == abi.PROJ_ADDITIVE:
                f.seek(8 * n, 1)
                n_stdp = int(np.frombuffer(f.read(4), dtype=np.uint32)[0])
                f.seek(5 * n_stdp, 1)
            projs.append(n)
    return pops, projs


class MI355XBackend:
    """A network instance on one MI355X (gfx950) device."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.net = C.c_void_p()
        self._chk(self.lib.knp_mi355x_create(device, C.byref(self.net)))
        _OPEN.add(self)
        self.pop_sizes = []
        self.proj_sizes = []
        self._rec_cap = {}

    # ---- lifetime -----------------------------------------------------------------------------
    def close(self):
        if self.net:
            self.lib.knp_mi355x_destroy(self.net)
            self.net = C.c_void_p()
        _OPEN.discard(self)

    def __del__(self):
        if sys.is_finalizing():
            return  # the exit hook has closed it; no HIP call after the runtime's teardown
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise BackendError(f"knp_mi355x error {rc}: {self.lib.knp_mi355x_last_error().decode()}")
        return rc

    def set_memory_cap(self, nbytes):
        self._chk(self.lib.knp_mi355x_set_memory_cap(self.net, nbytes))

    # ---- loading ------------------------------------------------------------------------------
    def add_blifat_population(self, neurons):
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_blifat_population(self.net, neurons, len(neurons), C.byref(out)))
        self.pop_sizes.append(len(neurons))
        return out.value

    def add_resource_blifat_population(self, neurons):
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_resource_blifat_population(self.net, neurons, len(neurons), C.byref(out)))
        self.pop_sizes.append(len(neurons))
        return out.value

    def add_lif_population(self, neurons):
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_lif_population(self.net, neurons, len(neurons), C.byref(out)))
        self.pop_sizes.append(len(neurons))
        return out.value

    def add_altai_lif_population(self, neurons):
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_altai_lif_population(self.net, neurons, len(neurons), C.byref(out)))
        self.pop_sizes.append(len(neurons))
        return out.value

    def add_resource_altai_lif_population(self, neurons):
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_resource_altai_lif_population(self.net, neurons, len(neurons),
                                                                        C.byref(out)))
        self.pop_sizes.append(len(neurons))
        return out.value

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
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_delta_projection(
            self.net, pre_pop, n_pre, post_pop, abi.u32p(pre), abi.u32p(post), abi.f32p(weight), abi.u32p(delay),
            abi.u8p(otype), len(pre), C.byref(out)))
        self.proj_sizes.append(len(pre))
        return out.value

    def add_resource_stdp_projection(self, pre_pop, n_pre, post_pop, pre, post, weight, delay, rules, otype=None):
        pre, post, weight, delay, otype = self._syn(pre, post, weight, delay, otype)
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_resource_stdp_projection(
            self.net, pre_pop, n_pre, post_pop, abi.u32p(pre), abi.u32p(post), abi.f32p(weight), abi.u32p(delay),
            abi.u8p(otype), rules, len(pre), C.byref(out)))
        self.proj_sizes.append(len(pre))
        return out.value

    def add_additive_stdp_projection(self, pre_pop, n_pre, post_pop, pre, post, weight, delay, tau_plus, tau_minus,
                                     stdp_pops=(), stdp_modes=(), otype=None):
        pre, post, weight, delay, otype = self._syn(pre, post, weight, delay, otype)
        tp = np.ascontiguousarray(tau_plus, dtype=np.float32)
        tm = np.ascontiguousarray(tau_minus, dtype=np.float32)
        sp = np.ascontiguousarray(stdp_pops, dtype=np.uint32)
        sm = np.ascontiguousarray(stdp_modes, dtype=np.uint8)
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_additive_stdp_projection(
            self.net, pre_pop, n_pre, post_pop, abi.u32p(pre), abi.u32p(post), abi.f32p(weight), abi.u32p(delay),
            abi.u8p(otype), abi.f32p(tp), abi.f32p(tm), len(pre), abi.u32p(sp), abi.u8p(sm), len(sp), C.byref(out)))
        self.proj_sizes.append(len(pre))
        return out.value

    def set_locked(self, proj, locked):
        self._chk(self.lib.knp_mi355x_set_locked(self.net, proj, 1 if locked else 0))

    def start_learning(self):
        self._chk(self.lib.knp_mi355x_start_learning(self.net))

    def stop_learning(self):
        self._chk(self.lib.knp_mi355x_stop_learning(self.net))

    def _init(self):
        self._chk(self.lib.knp_mi355x_init(self.net))

    # ---- inputs / stepping --------------------------------------------------------------------
    def push_spikes(self, proj, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        self._chk(self.lib.knp_mi355x_push_spikes(self.net, proj, abi.u32p(idx), len(idx)))

    def add_bernoulli_source(self, proj, thresholds, steps_per_pattern=1, active_lo=0, active_hi=None, seed=0,
                             first_index=0, hash_offset=0):
        thr = np.ascontiguousarray(np.atleast_2d(thresholds), dtype=np.uint32)
        n_patterns, n_ch = thr.shape
        hi = steps_per_pattern - 1 if active_hi is None else active_hi
        self._chk(self.lib.knp_mi355x_add_bernoulli_source(self.net, proj, first_index, n_ch, abi.u32p(thr),
                                                           n_patterns, steps_per_pattern, active_lo, hi, seed,
                                                           hash_offset))

    def add_dense_input_projection(self, post_pop, post_begin, post_count, n_in, fmt, weights, row_scale,
                                   threshold, seed, hash_offset=0, time_block=0):
        """Dense feed-forward projection from per-instance Bernoulli channel blocks (see the C ABI)."""
        pb = np.ascontiguousarray(post_begin, dtype=np.uint32)
        pc = np.ascontiguousarray(post_count, dtype=np.uint32)
        w = np.ascontiguousarray(weights)
        assert w.dtype == (np.uint8 if fmt == abi.DENSE_FP8_E4M3 else np.uint16)
        sc = None if row_scale is None else np.ascontiguousarray(row_scale, dtype=np.float32)
        out = C.c_uint32()
        self._chk(self.lib.knp_mi355x_add_dense_input_projection(
            self.net, post_pop, len(pb), abi.u32p(pb), abi.u32p(pc), n_in, fmt, w.ctypes.data_as(C.c_void_p),
            abi.f32p(sc), int(threshold), seed, hash_offset, time_block, C.byref(out)))
        return out.value

    def set_projection_dense(self, proj, fmt=abi.DENSE_BF16, flags=0):
        """Expand projection `proj` on MFMA (spike counts x a BF16 copy of its weights; see the C ABI)."""
        self._chk(self.lib.knp_mi355x_set_projection_dense(self.net, proj, fmt, flags))

    def dense_weights(self, proj, low=False):
        """The BF16 weight copy of a projection expanded on MFMA (bit patterns, synapse-index order); with
        low=True, (hi, lo) of a split copy."""
        n = self.proj_sizes[proj]
        out = np.zeros(max(n, 1), dtype=np.uint16)
        lo = np.zeros(max(n, 1), dtype=np.uint16)
        p16 = C.POINTER(C.c_uint16)
        self._chk(self.lib.knp_mi355x_read_dense_weights(self.net, proj, out.ctypes.data_as(p16),
                                                         lo.ctypes.data_as(p16) if low else None, n))
        return (out[:n], lo[:n]) if low else out[:n]

    def dense_weight_bytes(self):
        """(bytes read so far by the skipping MFMA expansions, bytes per step of the full ones)."""
        a, b = C.c_uint64(), C.c_uint64()
        self._chk(self.lib.knp_mi355x_get_dense_weight_bytes(self.net, C.byref(a), C.byref(b)))
        return a.value, b.value

    def dense_time_block(self, dense=0):
        t = C.c_uint32()
        self._chk(self.lib.knp_mi355x_get_dense_time_block(self.net, dense, C.byref(t)))
        return t.value

    def set_step_graph(self, steps):
        """Steps per captured HIP graph of one-launch-per-step runs (0: no graphs)."""
        self._chk(self.lib.knp_mi355x_set_step_graph(self.net, steps))

    def step_graph(self):
        """(steps of the captured graph or 0, replays so far)."""
        g, r = C.c_uint32(), C.c_uint64()
        self._chk(self.lib.knp_mi355x_get_step_graph(self.net, C.byref(g), C.byref(r)))
        return g.value, r.value

    def step_noalloc(self, n=1):
        """knp_mi355x_step_noalloc: steps an initialized network without allocating device memory."""
        self._chk(self.lib.knp_mi355x_step_noalloc(self.net, n))

    def step(self, n=1):
        self._chk(self.lib.knp_mi355x_step(self.net, n))

    _step = step

    def synchronize(self):
        self._chk(self.lib.knp_mi355x_synchronize(self.net))

    def get_step(self):
        s = C.c_uint64()
        self._chk(self.lib.knp_mi355x_get_step(self.net, C.byref(s)))
        return s.value

    current_step = property(get_step)

    @property
    def events(self):
        e = C.c_uint64()
        self._chk(self.lib.knp_mi355x_get_event_count(self.net, C.byref(e)))
        return e.value

    def counters(self):
        e, r = C.c_uint64(), C.c_uint64()
        self._chk(self.lib.knp_mi355x_get_counters(self.net, C.byref(e), C.byref(r)))
        return e.value, r.value

    def set_scatter_mode(self, mode):
        """abi.SCATTER_AUTO (tiled populations where eligible, persistent multi-step launches when every
        population is tiled), abi.SCATTER_STEPWISE (tiled, one launch per step) or abi.SCATTER_ATOMIC;
        before the first step."""
        self._chk(self.lib.knp_mi355x_set_scatter_mode(self.net, mode))

    def autotune(self, trial_steps=256):
        """Times the execution choices of multi-step calls from the current state (restored after each trial) and
        keeps the fastest; returns tuning()."""
        self._chk(self.lib.knp_mi355x_autotune(self.net, int(trial_steps)))
        return self.tuning()

    def tuning(self):
        """{"schedule", "wide_tiles", "graph_steps", "train_workgroup_limit", "scatter_grid"}: the execution
        choices in force."""
        s, w, g, t, sg = C.c_int(), C.c_int(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._chk(self.lib.knp_mi355x_get_tuning(self.net, C.byref(s), C.byref(w), C.byref(g), C.byref(t), C.byref(sg)))
        return {"schedule": s.value, "wide_tiles": w.value, "graph_steps": g.value, "train_workgroup_limit": t.value,
                "scatter_grid": sg.value}

    def set_message_semantics(self, semantics):
        """abi.SEMANTICS_SINGLE_THREADED (default) or abi.SEMANTICS_MULTI_THREADED: the multi-threaded CPU
        backend's projections (only the first SpikeMessage of a step, repeated indexes as weight multiples);
        before the first step."""
        self._chk(self.lib.knp_mi355x_set_message_semantics(self.net, semantics))

    def set_train_workgroup_limit(self, max_neurons):
        """Synaptic-resource populations of at most max_neurons neurons train with a workgroup per candidate,
        larger ones with a wave per candidate (identical results; see knp_mi355x_set_train_workgroup_limit)."""
        self._chk(self.lib.knp_mi355x_set_train_workgroup_limit(self.net, max_neurons))

    def set_wait_test(self, hold_block, spin_limit, hold_until=2**64 - 1):
        """Testing hook: a persistent workgroup that withholds its spike counts at steps before hold_until, and
        a small bound on every wait, so that a step call must end with an error (see knp_mi355x_set_wait_test)."""
        self._chk(self.lib.knp_mi355x_set_wait_test(self.net, hold_block, spin_limit, hold_until))

    def schedule(self):
        """How multi-step calls run: abi.SCHEDULE_ATOMIC, _STEPWISE, _PERSISTENT or _SINGLE_WORKGROUP."""
        x = C.c_int()
        self._chk(self.lib.knp_mi355x_get_schedule(self.net, C.byref(x)))
        return x.value

    def tiled(self, pop):
        t = C.c_int()
        self._chk(self.lib.knp_mi355x_get_tiled(self.net, pop, C.byref(t)))
        return bool(t.value)

    def tile_counters(self):
        """(row parts read, ring cells updated) by the tile-owner scatter so far."""
        a, b = C.c_uint64(), C.c_uint64()
        self._chk(self.lib.knp_mi355x_get_tile_counters(self.net, C.byref(a), C.byref(b)))
        return a.value, b.value

    def inexact_weights(self, pop):
        """Loaded weights into `pop` off its accumulator's 2^-frac_bits grid (rounded as impacts)."""
        n = C.c_uint64()
        self._chk(self.lib.knp_mi355x_get_inexact_weights(self.net, pop, C.byref(n)))
        return n.value

    def accumulator_bits(self, pop):
        b = C.c_int32()
        self._chk(self.lib.knp_mi355x_get_accumulator_bits(self.net, pop, C.byref(b)))
        return b.value

    # ---- outputs ------------------------------------------------------------------------------
    def record_spikes(self, pop, capacity_steps):
        self._chk(self.lib.knp_mi355x_record_spikes(self.net, pop, capacity_steps))
        self._rec_cap[pop] = capacity_steps

    def read_spike_record(self, pop):
        n = self.pop_sizes[pop]
        w = (n + 31) // 32
        cap = self._rec_cap.get(pop, 0)
        buf = np.zeros(max(w * cap, 1), dtype=np.uint32)
        steps = C.c_uint64()
        self._chk(self.lib.knp_mi355x_read_spike_record(self.net, pop, abi.u32p(buf), buf.size, C.byref(steps)))
        return abi.unpack_bitmaps(buf, steps.value, n)

    def last_spikes(self, pop):
        n = self.pop_sizes[pop]
        buf = np.zeros(max(n, 1), dtype=np.uint32)
        cnt = C.c_uint32()
        self._chk(self.lib.knp_mi355x_read_last_spikes(self.net, pop, abi.u32p(buf), n, C.byref(cnt)))
        return buf[: cnt.value].copy()

    def potentials(self, pop):
        n = self.pop_sizes[pop]
        out = np.zeros(n, dtype=np.float32)
        self._chk(self.lib.knp_mi355x_read_potentials(self.net, pop, abi.f32p(out), n))
        return out

    def blifat_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.BLIFATNeuron * n)()
        self._chk(self.lib.knp_mi355x_read_blifat_neurons(self.net, pop, out, n))
        return out

    def resource_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.ResourceBLIFATNeuron * n)()
        self._chk(self.lib.knp_mi355x_read_resource_neurons(self.net, pop, out, n))
        return out

    def lif_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.LIFNeuron * n)()
        self._chk(self.lib.knp_mi355x_read_lif_neurons(self.net, pop, out, n))
        return out

    def altai_lif_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.AltAILIFNeuron * n)()
        self._chk(self.lib.knp_mi355x_read_altai_lif_neurons(self.net, pop, out, n))
        return out

    def resource_altai_lif_neurons(self, pop):
        n = self.pop_sizes[pop]
        out = (abi.ResourceAltAILIFNeuron * n)()
        self._chk(self.lib.knp_mi355x_read_resource_altai_lif_neurons(self.net, pop, out, n))
        return out

    def weights(self, proj):
        n = self.proj_sizes[proj]
        out = np.zeros(n, dtype=np.float32)
        self._chk(self.lib.knp_mi355x_read_weights(self.net, proj, abi.f32p(out), n))
        return out

    def resource_rules(self, proj):
        n = self.proj_sizes[proj]
        out = (abi.ResourceRule * n)()
        self._chk(self.lib.knp_mi355x_read_resource_rules(self.net, proj, out, n))
        return out

    def resources(self, proj):
        r = self.resource_rules(proj)
        return np.array([x.synaptic_resource for x in r], dtype=np.float32)

    # ---- kernel timing ------------------------------------------------------------------------
    def save_network(self, path):
        """Write every population and projection with its current state to `path`."""
        self._chk(self.lib.knp_mi355x_save_network(self.net, os.fsencode(path)))

    def load_network(self, path):
        """Add the populations and projections of a network file to this (empty) backend."""
        self._chk(self.lib.knp_mi355x_load_network(self.net, os.fsencode(path)))
        self.pop_sizes, self.proj_sizes = read_network_sizes(path)

    def enable_timing(self, on=True):
        self._chk(self.lib.knp_mi355x_enable_timing(self.net, 1 if on else 0))

    def kernel_time(self, kernel_id):
        avg = C.c_double()
        n = C.c_uint64()
        self._chk(self.lib.knp_mi355x_kernel_time(self.net, kernel_id, C.byref(avg), C.byref(n)))
        return avg.value, n.value


KERNEL_IDS = {"neuron": abi.KERNEL_NEURON, "scatter": abi.KERNEL_SCATTER, "train": abi.KERNEL_TRAIN,
              "dense": abi.KERNEL_DENSE, "tile_scatter": abi.KERNEL_TILE_SCATTER, "tile_run": abi.KERNEL_TILE_RUN,
              "dense_expand": abi.KERNEL_DENSE_EXPAND}
