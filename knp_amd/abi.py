"""ctypes mirror of include/knp_mi355x.h (the C ABI of the MI355X backend).

The structs follow the reference trait structs field by field:
  knp_blifat_neuron           neuron_parameters<BLIFATNeuron>      knp/neuron-traits-library/include/knp/neuron-traits/blifat.h:207-343
  knp_resource_blifat_neuron  neuron_parameters<SynapticResourceSTDPNeuron<BLIFATNeuron>>
                                                                   knp/neuron-traits-library/include/knp/neuron-traits/stdp_synaptic_resource_rule.h:91-192
  knp_resource_rule           STDPSynapticResourceRule<DeltaSynapse>
                                                                   knp/synapse-traits-library/include/knp/synapse-traits/stdp_synaptic_resource_rule.h:43-100
  knp_lif_neuron              neuron_parameters<LIFNeuron>         knp/neuron-traits-library/include/knp/neuron-traits/lif.h:72-117
  knp_altai_lif_neuron        neuron_parameters<AltAILIF>          knp/neuron-traits-library/include/knp/neuron-traits/altai_lif.h:128-354
"""
import ctypes as C

import numpy as np

KNP_OK = 0
KNP_NO_POPULATION = 0xFFFFFFFF

# OutputType, knp/synapse-traits-library/include/knp/synapse-traits/output_types.h:36-62
EXCITATORY = 0
INHIBITORY_CURRENT = 1
INHIBITORY_CONDUCTANCE = 2
DOPAMINE = 3
BLOCKING = 4

# ISIPeriodType
ISI_IS_FORCED = 0
ISI_PERIOD_STARTED = 1
ISI_PERIOD_CONTINUED = 2
ISI_NOT_IN_PERIOD = 3

# ProcessingType of additive STDP projections
STDP_ONLY = 0
STDP_AND_SPIKE = 1

KERNEL_NEURON = 0
KERNEL_SCATTER = 1
KERNEL_TRAIN = 2
KERNEL_DENSE = 3
KERNEL_TILE_SCATTER = 4
KERNEL_TILE_RUN = 5
KERNEL_DENSE_EXPAND = 6

# projection kinds in a network file
PROJ_DELTA = 0
PROJ_RESOURCE = 1
PROJ_ADDITIVE = 2

# knp_scatter_mode
SCATTER_AUTO = 0
SCATTER_ATOMIC = 1
SCATTER_STEPWISE = 2
SCATTER_TILES = 3
# enum knp_message_semantics
SEMANTICS_SINGLE_THREADED = 0
SEMANTICS_MULTI_THREADED = 1
SCHEDULE_ATOMIC, SCHEDULE_STEPWISE, SCHEDULE_PERSISTENT, SCHEDULE_SINGLE_WORKGROUP = 0, 1, 2, 3

# knp_dense_format
DENSE_FP8_E4M3 = 0
DENSE_BF16 = 1
DENSE_SKIP_EMPTY = 1  # knp_mi355x_set_projection_dense flags
DENSE_SPLIT = 2

SIZE_MAX = 0xFFFFFFFFFFFFFFFF
INT64_MAX = 0x7FFFFFFFFFFFFFFF
FLT_MAX = float(np.finfo(np.float32).max)


class BLIFATNeuron(C.Structure):
    _fields_ = [
        ("n_time_steps_since_last_firing", C.c_uint64),
        ("activation_threshold", C.c_double),
        ("additional_threshold", C.c_double),
        ("dynamic_threshold", C.c_double),
        ("threshold_decay", C.c_double),
        ("threshold_increment", C.c_double),
        ("postsynaptic_trace", C.c_double),
        ("postsynaptic_trace_decay", C.c_double),
        ("postsynaptic_trace_increment", C.c_double),
        ("inhibitory_conductance", C.c_double),
        ("inhibitory_conductance_decay", C.c_double),
        ("potential", C.c_double),
        ("pre_impact_potential", C.c_double),
        ("potential_decay", C.c_double),
        ("bursting_phase", C.c_uint32),
        ("bursting_period", C.c_uint32),
        ("reflexive_weight", C.c_double),
        ("reversal_inhibitory_potential", C.c_double),
        ("absolute_refractory_period", C.c_uint32),
        ("potential_reset_value", C.c_double),
        ("min_potential", C.c_double),
        ("total_blocking_period", C.c_int64),
        ("dopamine_value", C.c_double),
        ("stochastic_stimulation", C.c_double),
        ("random_number_generator_state", C.c_uint32),
    ]


class ResourceBLIFATNeuron(C.Structure):
    _fields_ = [
        ("base", BLIFATNeuron),
        ("synapse_sum_threshold_coefficient", C.c_float),
        ("dopamine_plasticity_time", C.c_uint32),
        ("free_synaptic_resource", C.c_float),
        ("synaptic_resource_threshold", C.c_float),
        ("resource_drain_coefficient", C.c_uint32),
        ("stability", C.c_float),
        ("stability_change_parameter", C.c_float),
        ("stability_change_at_isi", C.c_float),
        ("isi_max", C.c_uint32),
        ("d_h", C.c_float),
        ("isi_status", C.c_int32),
        ("last_step", C.c_uint64),
        ("last_spike_step", C.c_uint64),
        ("first_isi_spike", C.c_uint64),
        ("is_being_forced", C.c_uint8),
    ]


class LIFNeuron(C.Structure):
    """neuron_parameters<LIFNeuron> (lif.h:72-117)."""
    _fields_ = [
        ("potential", C.c_float),
        ("potential_reset_value", C.c_float),
        ("activation_threshold", C.c_float),
        ("leak_coefficient", C.c_float),
        ("refract_counter", C.c_uint32),
        ("refract_period", C.c_uint32),
    ]


class AltAILIFNeuron(C.Structure):
    """neuron_parameters<AltAILIF> (altai_lif.h:128-354); the bool flags are bytes."""
    _fields_ = [
        ("is_diff", C.c_uint8),
        ("is_reset", C.c_uint8),
        ("leak_rev", C.c_uint8),
        ("saturate", C.c_uint8),
        ("do_not_save", C.c_uint8),
        ("potential", C.c_float),
        ("pre_impact_potential", C.c_float),
        ("activation_threshold", C.c_uint16),
        ("negative_activation_threshold", C.c_uint16),
        ("potential_leak", C.c_int16),
        ("potential_reset_value", C.c_uint16),
        ("dopamine_value", C.c_double),
        ("additional_threshold", C.c_double),
        ("activity_time", C.c_int64),
    ]


class ResourceAltAILIFNeuron(C.Structure):
    """neuron_parameters<SynapticResourceSTDPAltAILIFNeuron>: AltAI-LIF fields, then the resource fields."""
    _fields_ = [("base", AltAILIFNeuron)] + ResourceBLIFATNeuron._fields_[1:]


class ResourceRule(C.Structure):
    _fields_ = [
        ("synaptic_resource", C.c_float),
        ("w_min", C.c_float),
        ("w_max", C.c_float),
        ("d_u", C.c_float),
        ("dopamine_plasticity_period", C.c_uint32),
        ("last_spike_step", C.c_uint64),
        ("had_hebbian_update", C.c_uint8),
        ("has_contributed", C.c_uint8),
    ]


def blifat_defaults(**overrides):
    """default_values<BLIFATNeuron> (blifat.h:47-199) as a BLIFATNeuron struct."""
    n = BLIFATNeuron()
    n.n_time_steps_since_last_firing = SIZE_MAX
    n.activation_threshold = 1.0
    n.reversal_inhibitory_potential = -0.3
    n.min_potential = -1.0e9
    n.total_blocking_period = INT64_MAX
    for k, v in overrides.items():
        setattr(n, k, v)
    return n


def lif_defaults(**overrides):
    """default_values<LIFNeuron> (lif.h:43-67) as a LIFNeuron struct."""
    n = LIFNeuron()
    n.activation_threshold = 1.0
    n.leak_coefficient = 1.0
    for k, v in overrides.items():
        setattr(n, k, v)
    return n


def altai_lif_defaults(**overrides):
    """default_values<AltAILIF> (altai_lif.h:43-122) as an AltAILIFNeuron struct."""
    n = AltAILIFNeuron()
    n.is_reset = 1
    n.leak_rev = 1
    n.saturate = 1
    n.activation_threshold = 1
    n.negative_activation_threshold = 30000
    for k, v in overrides.items():
        setattr(n, k, v)
    return n


def resource_blifat_defaults(base=None, **overrides):
    """neuron_parameters<SynapticResourceSTDPNeuron<BLIFATNeuron>> defaults
    (stdp_synaptic_resource_rule.h:91-192)."""
    r = ResourceBLIFATNeuron()
    r.base = base if base is not None else blifat_defaults()
    r.synapse_sum_threshold_coefficient = 0.0
    r.dopamine_plasticity_time = 1
    r.free_synaptic_resource = 1.0
    r.synaptic_resource_threshold = FLT_MAX
    r.resource_drain_coefficient = 0
    r.stability = 0.0
    r.stability_change_parameter = 0.0
    r.stability_change_at_isi = 0.0
    r.isi_max = 1
    r.d_h = 1.0
    r.isi_status = ISI_NOT_IN_PERIOD
    r.last_step = 0
    r.last_spike_step = 0
    r.first_isi_spike = 0
    r.is_being_forced = 0
    for k, v in overrides.items():
        setattr(r, k, v)
    return r


def resource_altai_lif_defaults(base=None, **overrides):
    """neuron_parameters<SynapticResourceSTDPAltAILIFNeuron> defaults (resource fields as for BLIFAT)."""
    r = ResourceAltAILIFNeuron()
    r.base = base if base is not None else altai_lif_defaults()
    ref = resource_blifat_defaults()
    for f, _ in ResourceBLIFATNeuron._fields_[1:]:
        setattr(r, f, getattr(ref, f))
    for k, v in overrides.items():
        setattr(r, k, v)
    return r


def resource_rule_defaults(**overrides):
    """STDPSynapticResourceRule defaults (synapse-traits stdp_synaptic_resource_rule.h:43-100)."""
    r = ResourceRule()
    r.synaptic_resource = 0.0
    r.w_min = 0.0
    r.w_max = 1.0
    r.d_u = 0.0
    r.dopamine_plasticity_period = 0
    r.last_spike_step = 0
    r.had_hebbian_update = 0
    r.has_contributed = 0
    for k, v in overrides.items():
        setattr(r, k, v)
    return r


def neuron_array(proto, n):
    """A ctypes array of n copies of `proto` (one of the neuron structs above)."""
    arr = (type(proto) * n)()
    if n:
        view = np.frombuffer(arr, dtype=np.uint8).reshape(n, C.sizeof(proto))
        view[:] = np.frombuffer(bytes(proto), dtype=np.uint8)
    return arr


def rule_array(proto, n):
    return neuron_array(proto, n)


def u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32)) if a is not None else None


def f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def u8p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8)) if a is not None else None


def unpack_bitmaps(words, n_steps, n_neurons):
    """Spike bitmaps (ceil(n/32) uint32 words per step, bit i = neuron i) -> bool array [steps, n]."""
    w = (n_neurons + 31) // 32
    a = np.asarray(words[: n_steps * w], dtype=np.uint32).reshape(n_steps, w)
    bits = np.unpackbits(a.view(np.uint8).reshape(n_steps, w * 4), axis=1, bitorder="little")
    return bits[:, :n_neurons].astype(bool)


# ---- FP8 E4M3 (OCP "E4M3FN": bias 7, no infinities, max 448, NaN = S.1111.111) and BF16 codecs ---------
def _e4m3_table():
    vals = np.zeros(128, dtype=np.float64)
    for b in range(128):
        e, m = (b >> 3) & 15, b & 7
        vals[b] = np.ldexp(m / 8.0, -6) if e == 0 else np.ldexp(1.0 + m / 8.0, e - 7)
    vals[127] = np.nan
    return vals


E4M3_VALUES = _e4m3_table()  # code -> value for the 127 non-negative finite codes (index 127 is NaN)


def e4m3_decode(codes):
    codes = np.asarray(codes, dtype=np.uint8)
    mag = E4M3_VALUES[codes & 0x7F]
    return np.where(codes & 0x80, -mag, mag)


def e4m3_encode(x, min_magnitude=0.0):
    """Round to the nearest e4m3 value (ties to the even code), saturating at +-448. Nonzero codes below
    `min_magnitude` are not produced (such inputs round to 0 or to the smallest allowed code)."""
    x = np.asarray(x, dtype=np.float64)
    a = np.minimum(np.abs(x), 448.0)
    codes = np.array([0] + [c for c in range(1, 127) if E4M3_VALUES[c] >= min_magnitude])
    grid = E4M3_VALUES[codes]
    hi = np.clip(np.searchsorted(grid, a, side="left"), 0, len(grid) - 1)
    lo = np.maximum(hi - 1, 0)
    d_lo = a - grid[lo]
    d_hi = grid[hi] - a
    pick_hi = (d_hi < d_lo) | ((d_hi == d_lo) & (codes[hi] % 2 == 0))
    code = np.where(pick_hi, codes[hi], codes[lo]).astype(np.uint8)
    return np.where((x < 0) & (code != 0), code | 0x80, code).astype(np.uint8)


def bf16_encode(x):
    """float32 -> bf16 bits, round to nearest even."""
    f = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((f + 0x7FFF + ((f >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def bf16_decode(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)
