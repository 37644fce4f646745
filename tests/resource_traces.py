"""Hand-computed traces of synaptic-resource STDP on one or two neurons with one to five synapses.

Every expected number below is derived by hand from the reference's formulas, step by step, in the comments;
none comes from running the oracle or the device. Where a value is not a short binary fraction it is written
as the float32 rounding of the exact rational or of the float operations the reference performs, in the
reference's order (IEEE float arithmetic rounds each operation once, so these are what the reference computes).

Reference (paths relative to knp/backends/cpu/cpu-library/include/knp/backends/cpu-library/impl/populations/):
  blifat/blifat_stdp.h:35-103   process_spiking_neurons_impl: ISI update, has_contributed window (with the
                                u64 wrap of step - period and the swap in is_point_in_interval), Hebbian
                                r += d_h * min((float)pow(2, -stability), 1) once per ISI, free -= ...
  blifat/blifat_stdp.h:106-155  do_dopamine_plasticity_impl: resource change (float)(D * factor) for
                                contributing synapses while step - last_spike_step <= dopamine_plasticity_time;
                                stability -= D * param (forced reward or punishment, clamped at 0), else
                                stability += param * D * max(2 - |step - first_isi - isi_max| / isi_max, -1)
                                with the u64 difference converted to double
  training/stdp.h:40-53         weight = w_min + (w_max - w_min) * max(r, 0) / ((w_max - w_min) + max(r, 0))
  training/stdp.h:88-126        update_isi (last_step - step compared unsigned)
  training/stdp.h:152-186       renormalize_resource: out of ISI (step - last_step > isi_max, or forced) and
                                |free| >= threshold: r += free / (synapses + drain), free = 0
Step order (cpu-single-threaded-backend/impl/backend.cpp:89-141): a channel spike pushed before step t is
expanded by its projection at step t (last_spike_step_ = t, synaptic_resource_stdp.h:36-62) and, with
delay 1, reaches the neuron at step t + 1.
"""
import numpy as np

from knp_amd import abi

f32 = np.float32


def run(o, plan, steps):
    """plan: {step: [(projection, [channel indexes]), ...]} pushed before that step."""
    for t in range(steps):
        for q, idx in plan.get(t, []):
            o.push_spikes(q, idx)
        o.step()


def resource_pop(o, cells):
    arr = abi.neuron_array(cells[0], len(cells))
    for i, c in enumerate(cells):
        arr[i] = c
    return o.add_resource_blifat_population(arr)


def rules(n, **kw):
    return abi.rule_array(abi.resource_rule_defaults(**kw), n)


# ---------------------------------------------------------------------------------------------------------
# H: Hebbian updates at stability 0 (factor 1) and 1 (factor 1/2), twice in a row.
# Two neurons, d_h = 0.25, isi_max = 3, free = 1. n0: stability 0, stability_change_at_isi 0. n1: stability 2,
# stability_change_at_isi 1. Synapses (channels 0, 1 -> each neuron): r = 1, w_min 0, w_max 1, so
# w = 1 * 1 / (1 + 1) = 0.5 and two of them reach the threshold 1.0; period 2, d_u 0, delay 1.
# Channels 0 and 1 spike at steps 4 and 6, both neurons spike at steps 5 and 7.
#   step 5, n0: not_in_period -> period_started, first_isi 5; stability -= 0; window [3, 5] holds 4 + 1 - 1
#     = 4: both synapses contributed; d_h = 0.25 * min(2^0, 1) = 0.25: r = 1.25, free 1 -> 0.75 -> 0.5;
#     w = 1.25 / 2.25 = 5/9.
#   step 5, n1: stability 2 - 1 = 1; d_h = 0.25 * 2^-1 = 0.125: r = 1.125, free 0.75; w = 1.125/2.125 = 9/17.
#   step 7: last_step - step = 5 - 7 wraps, not < isi_max: the status stays period_started (first_isi stays 5),
#     so had_hebbian_update is cleared and the stability change applies again; window [5, 7] holds 6.
#     n0: r = 1.5, free 0; w = 1.5 / 2.5 = 0.6.   n1: stability 0, d_h 0.25: r = 1.375, free 0.25;
#     w = 1.375 / 2.375 = 11/19.
def build_hebbian(o):
    c0 = abi.resource_blifat_defaults(d_h=0.25, isi_max=3, free_synaptic_resource=1.0)
    c1 = abi.resource_blifat_defaults(d_h=0.25, isi_max=3, free_synaptic_resource=1.0, stability=2.0,
                                      stability_change_at_isi=1.0)
    pop = resource_pop(o, [c0, c1])
    q = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 2, pop, [0, 1, 0, 1], [0, 0, 1, 1], [0.5] * 4,
                                       [1] * 4, rules(4, synaptic_resource=1.0, w_min=0.0, w_max=1.0,
                                                      dopamine_plasticity_period=2))
    o.start_learning()
    return pop, q


HEBBIAN_PLAN = lambda q: {4: [(q, [0, 1])], 6: [(q, [0, 1])]}  # noqa: E731
HEBBIAN_AFTER_6 = {  # state after steps 0..5
    "free": [0.5, 0.75], "stability": [0.0, 1.0], "isi_status": [abi.ISI_PERIOD_STARTED] * 2,
    "last_step": [5, 5], "last_spike_step": [5, 5], "first_isi_spike": [5, 5],
    "r": [1.25, 1.25, 1.125, 1.125], "w": [f32(5 / 9)] * 2 + [f32(9 / 17)] * 2,
    "had_hebbian": [1, 1, 1, 1], "has_contributed": [1, 1, 1, 1], "rule_last_spike": [4, 4, 4, 4],
}
HEBBIAN_AFTER_8 = {  # state after steps 0..7
    "free": [0.0, 0.25], "stability": [0.0, 0.0], "isi_status": [abi.ISI_PERIOD_STARTED] * 2,
    "last_step": [7, 7], "last_spike_step": [7, 7], "first_isi_spike": [5, 5],
    "r": [1.5, 1.5, 1.375, 1.375], "w": [f32(0.6)] * 2 + [f32(11 / 19)] * 2,
    "had_hebbian": [1, 1, 1, 1], "has_contributed": [1, 1, 1, 1], "rule_last_spike": [6, 6, 6, 6],
}
HEBBIAN_SPIKES = {0: [5, 7], 1: [5, 7]}


# ---------------------------------------------------------------------------------------------------------
# W: the has_contributed window, its edges and the u64 wrap of step - period.
# One neuron, d_h = 0.5, isi_max 3, free 2. Trainable synapses (weights 0: r 0, w_min 0, w_max 1):
#   A ch0, B ch1, C ch2: delay 1, period 2;  D ch3: delay 1, period 10;  E ch4: delay 3, period 10.
# A locked synaptic-resource projection (ch0 -> neuron, weight 1, not forcing, not trained) makes the neuron
# spike one step after its channel spikes: at steps 3 and 9.
# Channel spikes: E at 1, D at 2, trigger at 2; then A at 5, B at 7, C at 8, trigger at 8.
#   step 3: D and E have step - period = 3 - 10, which wraps to 2^64 - 7; the interval is swapped to
#     [3, 2^64 - 7]. D's point 2 + 1 - 1 = 2 is outside it, E's point 1 + 3 - 1 = 3 inside: only E
#     contributed (A, B, C: [1, 3] does not hold their point 0). E: r = 0 + 0.5 = 0.5, w = 0.5 / 1.5 = 1/3;
#     free 2 - 0.5 = 1.5.
#   step 9: window [7, 9] for A, B, C: A (5) no, B (7, the lower edge) yes, C (8) yes; D, E: [9, 2^64 - 1],
#     neither. Status stays period_started (3 - 9 wraps), had_hebbian_update is cleared for all five, B and C
#     get 0.5 each: r 0.5, w 1/3, free 0.5. E keeps r 0.5 but has_contributed is now false.
def build_window(o):
    cell = abi.resource_blifat_defaults(d_h=0.5, isi_max=3, free_synaptic_resource=2.0)
    pop = resource_pop(o, [cell])
    rl = rules(5, synaptic_resource=0.0, w_min=0.0, w_max=1.0, dopamine_plasticity_period=2)
    rl[3].dopamine_plasticity_period = 10
    rl[4].dopamine_plasticity_period = 10
    q = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 5, pop, [0, 1, 2, 3, 4], [0] * 5, [0.0] * 5,
                                       [1, 1, 1, 1, 3], rl)
    trig = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [1.0], [1],
                                          rules(1, synaptic_resource=1.0, w_min=1.0, w_max=2.0))
    o.set_locked(q, False)
    return pop, q, trig


WINDOW_PLAN = lambda q, trig: {1: [(q, [4])], 2: [(q, [3]), (trig, [0])], 5: [(q, [0])], 7: [(q, [1])],  # noqa: E731
                               8: [(q, [2]), (trig, [0])]}
WINDOW_AFTER_4 = {"free": [1.5], "r": [0, 0, 0, 0, 0.5], "w": [0, 0, 0, 0, f32(1 / 3)],
                  "has_contributed": [0, 0, 0, 0, 1], "had_hebbian": [0, 0, 0, 0, 1],
                  "isi_status": [abi.ISI_PERIOD_STARTED], "first_isi_spike": [3]}
WINDOW_AFTER_10 = {"free": [0.5], "r": [0, 0.5, 0.5, 0, 0.5], "w": [0, f32(1 / 3), f32(1 / 3), 0, f32(1 / 3)],
                   "has_contributed": [0, 1, 1, 0, 0], "had_hebbian": [0, 1, 1, 0, 0],
                   "rule_last_spike": [5, 7, 8, 2, 1], "isi_status": [abi.ISI_PERIOD_STARTED],
                   "first_isi_spike": [3], "last_spike_step": [9]}
WINDOW_SPIKES = {0: [3, 9]}


# ---------------------------------------------------------------------------------------------------------
# D: dopamine reward (not forced, with both signs of the stability term), punishment, and forced reward.
# One neuron: d_h 0.25, isi_max 4, dopamine_plasticity_time 10, stability 0, stability_change_parameter
# 0.5, free 1. Trainable synapses ch0, ch1: r 1, w 0.5, period 2. Plain Delta projections from one channel:
# "dop" DOPAMINE +0.5, "pun" DOPAMINE -0.5, "force" EXCITATORY 0.25 (plain Delta impacts are forcing; 0.25
# alone stays below threshold). All delay 1.
#   step 5 (ch0, ch1 at 4): spike, Hebbian as H/n0: r 1.25, free 0.5, first_isi 5, last_spike_step 5.
#   step 10 (dop at 9): D = 0.5; 10 - 5 <= 10 and both contributed: change = (float)(0.5 * 1) = 0.5:
#     r 1.75, free 0.5 - 0.5 - 0.5 = -0.5, w = 1.75 / 2.75 = 7/11. Not forced, D > 0:
#     |10 - 5 - 4| = 1, 2 - 1/4 = 1.75: stability 0 + 0.5 * 0.5 * 1.75 = 0.4375.
#   step 11 (dop at 10): factor = (float)2^-0.4375; change = (float)(0.5 * factor) = factor / 2 (exact);
#     r = 1.75 + change, free = (-0.5 - change) - change; |11 - 5 - 4| = 2: 2 - 0.5 = 1.5:
#     stability 0.4375 + 0.375 = 0.8125.
#   step 12 (pun at 11): D = -0.5 < 0, status period_started (not forced): change = (float)(-0.5 * factor),
#     factor = (float)2^-0.8125; stability = 0.8125 - (-0.5 * 0.5) = 1.0625.
#   step 13 (dop and force at 12): forced reward: factor (float)2^-1.0625, change = factor / 2;
#     stability = 1.0625 - 0.5 * 0.5 = 0.8125.
#   step 14 (dop at 13): not forced, factor (float)2^-0.8125, change = factor / 2;
#     |14 - 5 - 4| = 5, 2 - 5/4 = 0.75: stability 0.8125 + 0.25 * 0.75 = 1.
#   step 16 (pun at 15): 16 - 5 > 10, no resource change; stability 1 - (-0.5 * 0.5) = 1.25.
#   step 20 (dop at 19): no resource change; |20 - 5 - 4| = 11, 2 - 11/4 = -0.75:
#     stability 1.25 + 0.25 * -0.75 = 1.0625.
def build_dopamine(o):
    cell = abi.resource_blifat_defaults(d_h=0.25, isi_max=4, dopamine_plasticity_time=10,
                                        stability_change_parameter=0.5, free_synaptic_resource=1.0)
    pop = resource_pop(o, [cell])
    q = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 2, pop, [0, 1], [0, 0], [0.5, 0.5], [1, 1],
                                       rules(2, synaptic_resource=1.0, w_min=0.0, w_max=1.0,
                                             dopamine_plasticity_period=2))
    dop = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [0.5], [1], [abi.DOPAMINE])
    pun = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [-0.5], [1], [abi.DOPAMINE])
    force = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [0.25], [1], [abi.EXCITATORY])
    o.set_locked(q, False)
    return pop, q, dop, pun, force


DOPAMINE_PLAN = lambda q, dop, pun, force: {  # noqa: E731
    4: [(q, [0, 1])], 9: [(dop, [0])], 10: [(dop, [0])], 11: [(pun, [0])], 12: [(dop, [0]), (force, [0])],
    13: [(dop, [0])], 15: [(pun, [0])], 19: [(dop, [0])]}


def dopamine_expected():
    """(step count, expected state) checkpoints of trace D; the float operations in the reference's order."""
    out = []
    r, free = f32(1.25), f32(0.5)
    # step 10
    r, free = r + f32(0.5), (free - f32(0.5)) - f32(0.5)
    assert (r, free) == (f32(1.75), f32(-0.5))
    out.append((11, {"stability": [0.4375], "r": [1.75, 1.75], "free": [-0.5], "w": [f32(7 / 11)] * 2}))
    # step 11
    c = f32(f32(2.0 ** -0.4375) / 2)
    r, free = f32(r + c), f32(f32(free - c) - c)
    out.append((12, {"stability": [0.8125], "r": [r, r], "free": [free]}))
    # step 12: punishment
    c = f32(-0.5 * float(f32(2.0 ** -0.8125)))
    r, free = f32(r + c), f32(f32(free - c) - c)
    out.append((13, {"stability": [1.0625], "r": [r, r], "free": [free]}))
    # step 13: forced reward
    c = f32(f32(2.0 ** -1.0625) / 2)
    r, free = f32(r + c), f32(f32(free - c) - c)
    out.append((14, {"stability": [0.8125], "r": [r, r], "free": [free]}))
    # step 14: reward, 2 - 5/4
    c = f32(f32(2.0 ** -0.8125) / 2)
    r, free = f32(r + c), f32(f32(free - c) - c)
    out.append((15, {"stability": [1.0], "r": [r, r], "free": [free]}))
    # step 16: punishment, no resource change (16 - 5 > 10): stability 1 + 0.25
    out.append((17, {"stability": [1.25], "r": [r, r], "free": [free]}))
    # step 20: reward, no resource change (20 - 5 > 10), 2 - 11/4 = -0.75
    out.append((21, {"stability": [1.0625], "r": [r, r], "free": [free]}))
    return out


DOPAMINE_SPIKES = {0: [5]}


# ---------------------------------------------------------------------------------------------------------
# R: renormalization, drains 0 and 2, a negative free resource, and the weight formula with w_min != 0.
# Three neurons, synaptic_resource_threshold 1, isi_max 2, status not_in_period, last_step 0, no spikes:
#   n0: free 1.5, drain 0;  n1: free 1.25, drain 2;  n2: free -1.5, drain 0.
# Three synapses into each (channels 0..2), r 0; n0: w_min 0, w_max 2; n1: w_min 0.25, w_max 1.25; n2: w_min 0,
# w_max 1.
# Steps 0..2: step - last_step <= 2, in the ISI window: nothing. Step 3:
#   n0: 1.5 / (3 + 0) = 0.5 per synapse, w = 2 * 0.5 / (2 + 0.5) = 0.4, free 0.
#   n1: 1.25 / (3 + 2) = 0.25, w = 0.25 + 1 * 0.25 / 1.25 = 0.25 + (float)0.2 (then rounded), free 0.
#   n2: -1.5 / 3 = -0.5, w = w_min + 1 * max(-0.5, 0) / (...) = 0, free 0.
def build_renorm(o):
    cells = [abi.resource_blifat_defaults(synaptic_resource_threshold=1.0, isi_max=2, free_synaptic_resource=fr,
                                          resource_drain_coefficient=dr)
             for fr, dr in ((1.5, 0), (1.25, 2), (-1.5, 0))]
    pop = resource_pop(o, cells)
    pre = np.tile(np.arange(3), 3)
    post = np.repeat(np.arange(3), 3)
    rl = rules(9, synaptic_resource=0.0, w_min=0.0, w_max=1.0)
    for k in range(3):
        rl[k].w_max = 2.0
    for k in range(3, 6):
        rl[k].w_min, rl[k].w_max = 0.25, 1.25
    w0 = [0.0] * 3 + [0.25] * 3 + [0.0] * 3
    q = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 3, pop, pre, post, w0, [1] * 9, rl)
    o.set_locked(q, False)
    return pop, q


RENORM_AFTER_3 = {"free": [1.5, 1.25, -1.5], "r": [0.0] * 9}
RENORM_AFTER_4 = {"free": [0.0, 0.0, 0.0], "r": [0.5] * 3 + [0.25] * 3 + [-0.5] * 3,
                  "w": [f32(0.4)] * 3 + [f32(f32(0.25) + f32(0.2))] * 3 + [0.0] * 3}


# ---------------------------------------------------------------------------------------------------------
# U: a non-zero d_u, and dopamine at the edge of dopamine_plasticity_time.
# One neuron: d_h 0.25, isi_max 4, dopamine_plasticity_time 5, stability 0, stability_change_parameter 0
# (the stability stays 0, factor 1), free 1. Trainable synapses ch0, ch1: r 1, w 0.5, period 2, d_u 0.125.
# "dop": plain Delta DOPAMINE +0.5 from one channel. All delay 1.
#   step 5 (ch0, ch1 at 4): spike; not_in_period -> period_started, first_isi 5; window [3, 5] holds 4: both
#     contributed. Per synapse in order: r -= d_u, free += d_u, then r += d_h, free -= d_h:
#     syn 0: r 1 - 0.125 + 0.25 = 1.125, free 1 + 0.125 - 0.25 = 0.875;
#     syn 1: r 1.125, free 0.875 + 0.125 - 0.25 = 0.75.  w = 1.125 / 2.125 = 9/17.
#   step 10 (dop at 9): 10 - 5 = 5 <= 5: change (float)(0.5 * 1) = 0.5 for both: r 1.625,
#     free 0.75 - 0.5 - 0.5 = -0.25, w = 1.625 / 2.625 = 13/21; stability += 0 * ... = 0.
#   step 11 (dop at 10): 11 - 5 = 6 > 5: no resource change; weights recalculated unchanged.
def build_du(o):
    cell = abi.resource_blifat_defaults(d_h=0.25, isi_max=4, dopamine_plasticity_time=5,
                                        stability_change_parameter=0.0, free_synaptic_resource=1.0)
    pop = resource_pop(o, [cell])
    q = o.add_resource_stdp_projection(abi.KNP_NO_POPULATION, 2, pop, [0, 1], [0, 0], [0.5, 0.5], [1, 1],
                                       rules(2, synaptic_resource=1.0, w_min=0.0, w_max=1.0, d_u=0.125,
                                             dopamine_plasticity_period=2))
    dop = o.add_delta_projection(abi.KNP_NO_POPULATION, 1, pop, [0], [0], [0.5], [1], [abi.DOPAMINE])
    o.set_locked(q, False)
    return pop, q, dop


DU_PLAN = lambda q, dop: {4: [(q, [0, 1])], 9: [(dop, [0])], 10: [(dop, [0])]}  # noqa: E731
DU_AFTER_6 = {"r": [1.125, 1.125], "free": [0.75], "w": [f32(9 / 17)] * 2, "stability": [0.0],
              "isi_status": [abi.ISI_PERIOD_STARTED], "first_isi_spike": [5], "last_spike_step": [5],
              "has_contributed": [1, 1], "had_hebbian": [1, 1]}
DU_AFTER_11 = {"r": [1.625, 1.625], "free": [-0.25], "w": [f32(13 / 21)] * 2, "stability": [0.0]}
DU_AFTER_12 = DU_AFTER_11
DU_SPIKES = {0: [5]}


# ---------------------------------------------------------------------------------------------------------
def state(o, pop, q):
    n = o.resource_neurons(pop)
    rl = o.resource_rules(q)
    return {
        "free": [x.free_synaptic_resource for x in n], "stability": [x.stability for x in n],
        "isi_status": [x.isi_status for x in n], "last_step": [x.last_step for x in n],
        "last_spike_step": [x.last_spike_step for x in n], "first_isi_spike": [x.first_isi_spike for x in n],
        "r": [x.synaptic_resource for x in rl], "w": list(np.asarray(o.weights(q), np.float32)),
        "had_hebbian": [x.had_hebbian_update for x in rl], "has_contributed": [x.has_contributed for x in rl],
        "rule_last_spike": [x.last_spike_step for x in rl],
    }


def check(o, pop, q, expected, where):
    got = state(o, pop, q)
    for k, v in expected.items():
        g = got[k]
        if k in ("free", "stability", "r", "w"):
            assert np.array_equal(np.asarray(g, np.float32), np.asarray(v, np.float32)), (where, k, g, v)
        else:
            assert list(g) == list(v), (where, k, g, v)


def spike_steps(o, pop, n_neurons, steps):
    rec = o.read_spike_record(pop)
    return {i: [int(t) for t in np.nonzero(rec[:steps, i])[0]] for i in range(n_neurons)}


def run_all_traces(make, dense_flags=None):
    """Runs the five traces on fresh networks from make() (an oracle or a device backend) and checks every
    expected value. dense_flags (device only): expand the trained projection on MFMA with these flags; the
    window trace (its synapses have two delays, which the expansion does not take) is then left out."""
    def dense(o, q):
        if dense_flags is not None:
            o.set_projection_dense(q, abi.DENSE_BF16, dense_flags)

    # H
    o = make()
    pop, q = build_hebbian(o)
    dense(o, q)
    o.record_spikes(pop, 8)
    run(o, HEBBIAN_PLAN(q), 6)
    check(o, pop, q, HEBBIAN_AFTER_6, "hebbian after 6 steps")
    run(o, {0: HEBBIAN_PLAN(q)[6]}, 2)
    check(o, pop, q, HEBBIAN_AFTER_8, "hebbian after 8 steps")
    assert spike_steps(o, pop, 2, 8) == HEBBIAN_SPIKES
    # W
    o = make()
    pop, q, trig = build_window(o) if dense_flags is None else (None, None, None)
    if pop is not None:
        o.record_spikes(pop, 10)
        plan = WINDOW_PLAN(q, trig)
        run(o, plan, 4)
        check(o, pop, q, WINDOW_AFTER_4, "window after 4 steps")
        run(o, {t - 4: v for t, v in plan.items() if t >= 4}, 6)
        check(o, pop, q, WINDOW_AFTER_10, "window after 10 steps")
        assert spike_steps(o, pop, 1, 10) == WINDOW_SPIKES
    # D
    o = make()
    pop, q, dop, pun, force = build_dopamine(o)
    dense(o, q)
    o.record_spikes(pop, 21)
    plan = DOPAMINE_PLAN(q, dop, pun, force)
    done = 0
    for upto, exp in dopamine_expected():
        run(o, {t - done: v for t, v in plan.items() if done <= t < upto}, upto - done)
        done = upto
        check(o, pop, q, exp, f"dopamine after {upto} steps")
    assert spike_steps(o, pop, 1, 21) == DOPAMINE_SPIKES
    # U
    o = make()
    pop, q, dop = build_du(o)
    dense(o, q)
    o.record_spikes(pop, 12)
    plan = DU_PLAN(q, dop)
    run(o, plan, 6)
    check(o, pop, q, DU_AFTER_6, "d_u after 6 steps")
    run(o, {t - 6: v for t, v in plan.items() if t >= 6}, 5)
    check(o, pop, q, DU_AFTER_11, "d_u: dopamine at the edge of the window")
    run(o, {}, 1)
    check(o, pop, q, DU_AFTER_12, "d_u: dopamine past the window")
    assert spike_steps(o, pop, 1, 12) == DU_SPIKES
    # R
    o = make()
    pop, q = build_renorm(o)
    dense(o, q)
    run(o, {}, 3)
    check(o, pop, q, RENORM_AFTER_3, "renormalization after 3 steps")
    run(o, {}, 1)
    check(o, pop, q, RENORM_AFTER_4, "renormalization after 4 steps")
