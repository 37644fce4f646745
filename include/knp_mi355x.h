/*
 * knp_mi355x.h - C ABI of the MI355X (gfx950) backend for the KNP spiking-network step.
 *
 * This is the drop-in boundary between a KNP-style C++ backend object (knp::core::Backend,
 * reference knp/core-library/include/knp/core/backend.h:55-457) and the HIP kernels that run the
 * per-step update of the reference's single-threaded CPU backend
 * (knp/backends/cpu/cpu-single-threaded-backend/impl/backend.cpp:89-141).
 *
 * Conventions
 *   - Opaque handle, plain pointers and sizes, no C++ or torch types.
 *   - Every function returns 0 (KNP_OK) on success or a negative KNP_ERR_* code; the message of the
 *     last failure on the calling thread is returned by knp_mi355x_last_error().
 *   - Host buffers passed in are copied before the call returns; the caller keeps ownership.
 *   - One host thread drives one handle at a time; handles are independent.
 *
 * Every entry point below names the reference interface it replaces.
 */
#ifndef KNP_MI355X_H
#define KNP_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------------------------------ */
#define KNP_OK 0
#define KNP_ERR_INVALID_ARGUMENT (-1) /* bad index, size or parameter */
#define KNP_ERR_UNSUPPORTED (-2)      /* "Type is not supported." (knp/meta-library variant_helpers.h:67) */
#define KNP_ERR_HIP (-3)              /* a HIP runtime call failed */
#define KNP_ERR_OUT_OF_MEMORY (-4)    /* device or host allocation failed, or the allocation cap was hit */
#define KNP_ERR_STATE (-5)            /* call not allowed in the current state (e.g. load after start) */
#define KNP_ERR_RANGE (-6)            /* exact accumulator or buffer capacity exceeded */

#define KNP_NO_POPULATION 0xFFFFFFFFu /* presynaptic side of a channel-fed projection (null UID) */

/* OutputType, knp/synapse-traits-library/include/knp/synapse-traits/output_types.h:36-62 */
enum knp_output_type {
    KNP_EXCITATORY = 0,
    KNP_INHIBITORY_CURRENT = 1,
    KNP_INHIBITORY_CONDUCTANCE = 2,
    KNP_DOPAMINE = 3,
    KNP_BLOCKING = 4
};

/* ISIPeriodType, knp/neuron-traits-library/include/knp/neuron-traits/stdp_synaptic_resource_rule.h */
enum knp_isi_status {
    KNP_ISI_IS_FORCED = 0,
    KNP_ISI_PERIOD_STARTED = 1,
    KNP_ISI_PERIOD_CONTINUED = 2,
    KNP_ISI_NOT_IN_PERIOD = 3
};

/* Projection::SharedSynapseParameters::ProcessingType, knp/core-library/include/knp/core/projection.h:139-158 */
enum knp_stdp_processing { KNP_STDP_ONLY = 0, KNP_STDP_AND_SPIKE = 1 };

/* neuron_parameters<BLIFATNeuron>, knp/neuron-traits-library/include/knp/neuron-traits/blifat.h:207-343.
 * Same fields, types and meaning; knp_blifat_defaults() returns default_values<BLIFATNeuron>. */
typedef struct knp_blifat_neuron {
    uint64_t n_time_steps_since_last_firing;
    double activation_threshold;
    double additional_threshold;
    double dynamic_threshold;
    double threshold_decay;
    double threshold_increment;
    double postsynaptic_trace;
    double postsynaptic_trace_decay;
    double postsynaptic_trace_increment;
    double inhibitory_conductance;
    double inhibitory_conductance_decay;
    double potential;
    double pre_impact_potential;
    double potential_decay;
    uint32_t bursting_phase;
    uint32_t bursting_period;
    double reflexive_weight;
    double reversal_inhibitory_potential;
    uint32_t absolute_refractory_period;
    double potential_reset_value;
    double min_potential;
    int64_t total_blocking_period;
    double dopamine_value;
    double stochastic_stimulation;
    uint32_t random_number_generator_state;
} knp_blifat_neuron;

/* neuron_parameters<SynapticResourceSTDPNeuron<BLIFATNeuron>>,
 * knp/neuron-traits-library/include/knp/neuron-traits/stdp_synaptic_resource_rule.h:91-192 */
typedef struct knp_resource_blifat_neuron {
    knp_blifat_neuron base;
    float synapse_sum_threshold_coefficient;
    uint32_t dopamine_plasticity_time;
    float free_synaptic_resource;
    float synaptic_resource_threshold;
    uint32_t resource_drain_coefficient;
    float stability;
    float stability_change_parameter;
    float stability_change_at_isi;
    uint32_t isi_max;
    float d_h;
    int32_t isi_status; /* enum knp_isi_status */
    uint64_t last_step;
    uint64_t last_spike_step;
    uint64_t first_isi_spike;
    uint8_t is_being_forced;
} knp_resource_blifat_neuron;

/* STDPSynapticResourceRule<DeltaSynapse>,
 * knp/synapse-traits-library/include/knp/synapse-traits/stdp_synaptic_resource_rule.h:43-100 */
typedef struct knp_resource_rule {
    float synaptic_resource;
    float w_min;
    float w_max;
    float d_u;
    uint32_t dopamine_plasticity_period;
    uint64_t last_spike_step;
    uint8_t had_hebbian_update;
    uint8_t has_contributed;
} knp_resource_rule;

/* neuron_parameters<LIFNeuron>, knp/neuron-traits-library/include/knp/neuron-traits/lif.h:72-117.
 * Same fields, types and meaning; knp_lif_defaults() returns default_values<LIFNeuron> (lif.h:43-67). */
typedef struct knp_lif_neuron {
    float potential;
    float potential_reset_value;
    float activation_threshold;
    float leak_coefficient;
    uint32_t refract_counter;
    uint32_t refract_period;
} knp_lif_neuron;

/* neuron_parameters<AltAILIF>, knp/neuron-traits-library/include/knp/neuron-traits/altai_lif.h:128-354.
 * Same fields, order and meaning (the five bool flags as bytes); knp_altai_lif_defaults() returns
 * default_values<AltAILIF> (altai_lif.h:43-122). */
typedef struct knp_altai_lif_neuron {
    uint8_t is_diff;
    uint8_t is_reset;
    uint8_t leak_rev;
    uint8_t saturate;
    uint8_t do_not_save;
    float potential;
    float pre_impact_potential;
    uint16_t activation_threshold;
    uint16_t negative_activation_threshold;
    int16_t potential_leak;
    uint16_t potential_reset_value;
    double dopamine_value;
    double additional_threshold;
    int64_t activity_time;
} knp_altai_lif_neuron;

/* neuron_parameters<SynapticResourceSTDPNeuron<AltAILIF>> (SynapticResourceSTDPAltAILIFNeuron,
 * stdp_synaptic_resource_rule.h:91-192 over altai_lif.h): the AltAI-LIF fields, then the same resource fields
 * as knp_resource_blifat_neuron. */
typedef struct knp_resource_altai_lif_neuron {
    knp_altai_lif_neuron base;
    float synapse_sum_threshold_coefficient;
    uint32_t dopamine_plasticity_time;
    float free_synaptic_resource;
    float synaptic_resource_threshold;
    uint32_t resource_drain_coefficient;
    float stability;
    float stability_change_parameter;
    float stability_change_at_isi;
    uint32_t isi_max;
    float d_h;
    int32_t isi_status; /* enum knp_isi_status */
    uint64_t last_step;
    uint64_t last_spike_step;
    uint64_t first_isi_spike;
    uint8_t is_being_forced;
} knp_resource_altai_lif_neuron;

/* Defaults of the reference trait structs. */
void knp_blifat_defaults(knp_blifat_neuron *out);
void knp_lif_defaults(knp_lif_neuron *out);
void knp_altai_lif_defaults(knp_altai_lif_neuron *out);
void knp_resource_blifat_defaults(knp_resource_blifat_neuron *out);
void knp_resource_altai_lif_defaults(knp_resource_altai_lif_neuron *out);
void knp_resource_rule_defaults(knp_resource_rule *out);

/* ---- library / devices ------------------------------------------------------------------------ */
typedef struct knp_net knp_net;

const char *knp_mi355x_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *knp_mi355x_last_error(void);
/* Backend::get_devices(), knp/core-library/include/knp/core/backend.h:172 (HIP device list). */
int knp_mi355x_device_count(int *count);
int knp_mi355x_device_name(int device, char *buf, size_t cap);

/* Backend construction: SingleThreadedCPUBackend::create(),
 * knp/backends/cpu/cpu-single-threaded-backend/impl/backend.cpp:48-52. */
int knp_mi355x_create(int device, knp_net **out);
int knp_mi355x_destroy(knp_net *net);
/* Caps the device memory the handle may allocate (0 = no cap); exceeding it gives KNP_ERR_OUT_OF_MEMORY. */
int knp_mi355x_set_memory_cap(knp_net *net, uint64_t bytes);
/* How impacts are summed (no reference counterpart: an execution choice with identical results).
 * AUTO: populations that only receive plain Delta EXCITATORY / INHIBITORY_CURRENT impacts with delays up to
 * 32 are tiled (each 256-neuron tile owned by one workgroup, sums in LDS, no global atomics); the rest use
 * per-impact 64-bit atomics. When every population is tiled and all tiles fit on the device at once, a
 * multi-step call runs as one persistent launch (ring kept in LDS; each step's spike lists carry a step
 * stamp that the readers wait for).
 * A network whose neurons all fit on one workgroup's threads then runs its multi-step calls in one workgroup
 * instead (rings and the step's spiking rows in LDS: no device-scope stores or polls).
 * STEPWISE: as AUTO but one launch per step. ATOMIC: atomics everywhere. TILES: as AUTO without the
 * one-workgroup schedule. Set before the first step. */
enum knp_scatter_mode { KNP_SCATTER_AUTO = 0, KNP_SCATTER_ATOMIC = 1, KNP_SCATTER_STEPWISE = 2, KNP_SCATTER_TILES = 3 };
int knp_mi355x_set_scatter_mode(knp_net *net, int mode);
/* Which reference backend's projection messages the network follows. SINGLE_THREADED (default):
 * SingleThreadedCPUBackend, whose projections expand every SpikeMessage of the step, each occurrence of an
 * index once (delta_impl.h:54-114). MULTI_THREADED: MultiThreadedCPUBackend::calculate_projections
 * (knp/backends/cpu/cpu-multi-threaded-backend/impl/backend.cpp:250-295, delta_impl.h:117-175): a projection
 * expands only the first SpikeMessage of the step. Its bus routes the population messages of the step and the
 * channel messages sent before it in one routing, the backend's endpoint first (message_bus_cpu_impl.cpp:41-88),
 * so the first message is its presynaptic population's if that population spiked, else the first one pushed
 * with knp_mi355x_push_spikes, else that of its first device source; later pushes of the step are dropped. An
 * index occurring c times in that message gives one impact of weight float(weight * c) per synapse. The
 * network may hold only the multi-threaded backend's types (BLIFAT, synaptic-resource BLIFAT and LIF
 * populations; Delta and synaptic-resource STDP Delta projections, presynaptic ranges below 2^24), no BLOCKING
 * synapses (that backend's last-impact order follows its threads' completion order) and no MFMA expansions;
 * impacts are summed on the atomic path. The swapped presynaptic / postsynaptic UIDs of that backend's
 * SynapticImpactMessages (delta_impl.h:164-169) concern message headers only, which this library does not
 * build. Set before the first step. */
enum knp_message_semantics { KNP_SEMANTICS_SINGLE_THREADED = 0, KNP_SEMANTICS_MULTI_THREADED = 1 };
int knp_mi355x_set_message_semantics(knp_net *net, int semantics);
/* Testing hook (no reference counterpart): in persistent launches, workgroup hold_block (~0u: none) does
 * not publish its spike counts at steps before hold_until (UINT64_MAX: every step), and every wait gives up
 * after spin_limit polls (at least 64; default 2^20). A launch whose workgroups cannot all proceed then ends,
 * later launches of the same call return at once, and the call returns KNP_ERR_HIP instead of hanging, which
 * is what the hook lets a test check. The network state is undefined after such an error. */
int knp_mi355x_set_wait_test(knp_net *net, uint32_t hold_block, uint32_t spin_limit, uint64_t hold_until);
/* Execution choice of synaptic-resource training (no reference counterpart, identical results): populations
 * of at most max_neurons neurons spread each training candidate's synapses over a workgroup (latency: the few
 * candidates of a step cannot fill the device otherwise); larger ones give each candidate a wave (throughput:
 * batches of instances, whose candidates fill it). Default 16384. Takes effect at the next step call. */
int knp_mi355x_set_train_workgroup_limit(knp_net *net, uint32_t max_neurons);
/* 1 if the population's impacts are summed by tiles (valid after the first step or an explicit init). */
int knp_mi355x_get_tiled(knp_net *net, uint32_t pop, int *tiled);
/* How multi-step calls run (no reference counterpart; initializes the network): ATOMIC one launch per step with
 * atomic impact sums, STEPWISE one or more launches per step (tile kernels and/or atomics; captured step graphs
 * where enabled), PERSISTENT one cooperative launch of every tile for many steps, SINGLE_WORKGROUP the whole
 * network in one workgroup for many steps (every neuron on one workgroup's threads, rings in LDS). */
enum knp_schedule { KNP_SCHEDULE_ATOMIC = 0, KNP_SCHEDULE_STEPWISE = 1, KNP_SCHEDULE_PERSISTENT = 2,
                    KNP_SCHEDULE_SINGLE_WORKGROUP = 3 };
int knp_mi355x_get_schedule(knp_net *net, int *schedule);
/* Step tuner (no reference counterpart: every choice gives identical results). Picks the execution choices of
 * multi-step calls by timing trial_steps steps (16..65536) of each candidate from the network's current state,
 * restoring that state (every device allocation, the step, pending pushes) after each: the schedule among
 * those the network allows (one workgroup, persistent, stepwise launches), then for stepwise launches the tile
 * kernel's occupancy (6 or 2 waves per SIMD), the steps per captured step graph (0, 8, 32, 128; for networks
 * with dense input projections, off or their time block), the training form (a workgroup or a wave per
 * candidate) and the atomic scatter's workgroups (its default, a quarter, a sixteenth, a sixty-fourth). The choice is cached per network shape in the tune cache (KNP_MI355X_TUNE_CACHE, else
 * ~/.knp_mi355x_tune.txt) and reused without trials. Needs device memory for one copy of the network's device
 * state. Initializes the network; pending pushes stay pending. */
int knp_mi355x_autotune(knp_net *net, uint32_t trial_steps);
/* The execution choices in force: schedule (enum knp_schedule), the 6-waves tile kernel (0/1), steps per
 * captured step graph, the population size up to which training runs a workgroup per candidate, and the
 * atomic scatter's workgroups. */
int knp_mi355x_get_tuning(knp_net *net, int *schedule, int *wide_tiles, uint32_t *graph_steps, uint32_t *train_limit,
                          uint32_t *scatter_grid);

/* ---- loading (Backend::load_all_populations / load_all_projections, backend.h:137,146;
 *      SingleThreadedCPUBackend::load_populations / load_projections, backend.cpp:144-170) ---------- */
/* Populations are numbered in load order (the order the reference updates them in). */
int knp_mi355x_add_blifat_population(knp_net *net, const knp_blifat_neuron *neurons, uint32_t n, uint32_t *out_pop);
int knp_mi355x_add_resource_blifat_population(
    knp_net *net, const knp_resource_blifat_neuron *neurons, uint32_t n, uint32_t *out_pop);
/* Population<LIFNeuron> (step: lif_impl.h:31-84). It accepts EXCITATORY and INHIBITORY_CURRENT impacts only:
 * the reference throws "Unhandled synapse type." when another type reaches it (lif_impl.h:62-64); here the
 * projection carrying it is rejected, with the same message, when it is added. */
int knp_mi355x_add_lif_population(knp_net *net, const knp_lif_neuron *neurons, uint32_t n, uint32_t *out_pop);
/* Population<AltAILIF> (step: altai_impl.h:31-150). Accepts EXCITATORY, INHIBITORY_CURRENT and DOPAMINE
 * impacts. BLOCKING impacts (altai_impl.h:72-79: a fold over the step's impacts in delivery order whose
 * result depends on that order) and INHIBITORY_CONDUCTANCE (the reference throws) are rejected when the
 * projection is added. With is_diff_ set, activation_threshold_ + additional_threshold_ must be a float. */
int knp_mi355x_add_altai_lif_population(
    knp_net *net, const knp_altai_lif_neuron *neurons, uint32_t n, uint32_t *out_pop);
/* Population<SynapticResourceSTDPAltAILIFNeuron> (SingleThreadedCPUBackend::calculate_population,
 * backend.cpp:244-251; training altai_stdp.h:35-149, the same rules as the BLIFAT variant). The AltAI-LIF
 * restrictions above apply; the threshold is activation_threshold_ + the trained additional threshold. */
int knp_mi355x_add_resource_altai_lif_population(
    knp_net *net, const knp_resource_altai_lif_neuron *neurons, uint32_t n, uint32_t *out_pop);

/* Projection<DeltaSynapse>. Synapse i of the reference container is (pre[i], post[i], weight[i], delay[i],
 * output_type[i]); output_type may be NULL (all KNP_EXCITATORY). pre_pop is a population index or
 * KNP_NO_POPULATION for a projection fed only by channels (knp_mi355x_push_spikes / sources); n_pre is the
 * size of the presynaptic index range. Projections are numbered in load order. */
int knp_mi355x_add_delta_projection(
    knp_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop, const uint32_t *pre, const uint32_t *post,
    const float *weight, const uint32_t *delay, const uint8_t *output_type, uint64_t n_syn, uint32_t *out_proj);

/* Projection<SynapticResourceSTDPDeltaSynapse>: as above plus one rule per synapse. Locked by default
 * (knp/core-library/include/knp/core/projection.h:432); see knp_mi355x_set_locked. */
int knp_mi355x_add_resource_stdp_projection(
    knp_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop, const uint32_t *pre, const uint32_t *post,
    const float *weight, const uint32_t *delay, const uint8_t *output_type, const knp_resource_rule *rules,
    uint64_t n_syn, uint32_t *out_proj);

/* Projection<AdditiveSTDPDeltaSynapse> with rule {tau_plus, tau_minus} per synapse and the shared
 * stdp_populations_ map (population index -> enum knp_stdp_processing). */
int knp_mi355x_add_additive_stdp_projection(
    knp_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop, const uint32_t *pre, const uint32_t *post,
    const float *weight, const uint32_t *delay, const uint8_t *output_type, const float *tau_plus,
    const float *tau_minus, uint64_t n_syn, const uint32_t *stdp_pops, const uint8_t *stdp_modes,
    uint32_t n_stdp_pops, uint32_t *out_proj);

/* Projection::lock_weights / unlock_weights (projection.h:415-426). */
int knp_mi355x_set_locked(knp_net *net, uint32_t proj, int locked);
/* Backend::start_learning / stop_learning (SingleThreadedCPUBackend, backend.h:324-332): unlock/lock all. */
int knp_mi355x_start_learning(knp_net *net);
int knp_mi355x_stop_learning(knp_net *net);

/* Backend::_init (backend.cpp:207-214 + cpu-library init.h:93-110): uploads the network and builds the
 * device routing. Called implicitly by the first step; loading afterwards is KNP_ERR_STATE. */
int knp_mi355x_init(knp_net *net);

/* ---- inputs (InputChannel::send -> SpikeMessage to a projection, base-framework io/input_channel.h:99-118) */
/* One SpikeMessage with these presynaptic indexes, delivered to projection `proj` in the next step. */
int knp_mi355x_push_spikes(knp_net *net, uint32_t proj, const uint32_t *idx, uint32_t n);
/* A device-side Bernoulli channel group feeding presynaptic indexes first_index .. first_index+n_channels-1
 * of `proj` every step: channel c spikes at step t iff
 * knp_bernoulli_hash(seed, t, hash_offset + c) < thresholds[pattern * n_channels + c], where
 * pattern = (t / steps_per_pattern) % n_patterns, and only when lo <= t % steps_per_pattern <= hi.
 * (Stands in for an InputChannel whose generator the host would otherwise run every step.) */
int knp_mi355x_add_bernoulli_source(
    knp_net *net, uint32_t proj, uint32_t first_index, uint32_t n_channels, const uint32_t *thresholds,
    uint32_t n_patterns, uint32_t steps_per_pattern, uint32_t active_lo, uint32_t active_hi, uint64_t seed,
    uint32_t hash_offset);

/* A dense feed-forward Projection<DeltaSynapse> from per-instance blocks of device Bernoulli channels into
 * many independent network instances at once (inference: the weights never change). Instance k owns
 * postsynaptic neurons post_begin[k] .. post_begin[k] + post_count[k] - 1 of `post_pop` and n_in channels;
 * channel i of instance k spikes at step t iff knp_bernoulli_hash(seed, t, hash_offset + k * n_in + i) <
 * threshold. Each spike is a SynapticImpact of weight row_scale[j] * w(j, i) on every neuron j of its
 * instance, EXCITATORY, delay 1 - the same as the equivalent sparse projection - but the impacts are
 * computed time_block steps at a time as (binary spikes) x (weights) on MFMA with FP32 accumulation.
 * weights: row-major [sum(post_count)][n_in] in `format`; row_scale may be NULL (all 1). Power-of-two
 * scales keep the dequantized weights exact. time_block: a multiple of 32, or 0 to let the autotuner pick
 * (cached in $KNP_MI355X_TUNE_CACHE). One dense projection per population; not for resource populations.
 * (No counterpart class in the reference: it is the batched form of its DeltaSynapse projection.) */
enum knp_dense_format { KNP_DENSE_FP8_E4M3 = 0, KNP_DENSE_BF16 = 1 };
int knp_mi355x_add_dense_input_projection(
    knp_net *net, uint32_t post_pop, uint32_t n_instances, const uint32_t *post_begin, const uint32_t *post_count,
    uint32_t n_in, int format, const void *weights, const float *row_scale, uint32_t threshold, uint64_t seed,
    uint32_t hash_offset, uint32_t time_block, uint32_t *out_dense);
/* The time block the dense projection runs with (after init). */
int knp_mi355x_get_dense_time_block(knp_net *net, uint32_t dense, uint32_t *time_block);

/* Expand a loaded Projection<DeltaSynapse> or Projection<SynapticResourceSTDPDeltaSynapse> on the matrix cores
 * instead of by the scatter: each step, the spike counts of its presynaptic indexes are multiplied by a BF16
 * copy of its weights with FP32 accumulation (the batched dense projection of many instances: no counterpart
 * class in the reference, which expands the same projection spike by spike in calculate_projection_impl,
 * knp/backends/cpu/cpu-library/include/knp/backends/cpu-library/impl/projections/delta/delta_impl.h:54-114).
 * The synapses must form dense blocks: one delay, one output type (EXCITATORY or INHIBITORY_CURRENT), and
 * every postsynaptic neuron with synapses takes each presynaptic index of one range [lo, lo + n) exactly once,
 * n the same for all; consecutive neurons with the same range form an instance, and instances do not share
 * presynaptic indexes (an all-to-all projection, or a block-diagonal batch of them). Synaptic-resource STDP
 * still trains the FP32 weights (read by knp_mi355x_read_weights) and refreshes the BF16 copy. The impacts
 * then differ from the reference's by the rounding of the weights to BF16 (8 significant bits) and of the
 * FP32 accumulation. flags (or-ed): KNP_DENSE_SKIP_EMPTY multiplies only the groups of 8 presynaptic indexes that hold
 * a spike (an exact shortcut: the others contribute 0). format: KNP_DENSE_BF16. Set before the first step;
 * the postsynaptic population is then summed on the atomic path. */
#define KNP_DENSE_SKIP_EMPTY 1u
/* KNP_DENSE_SPLIT: each weight is kept as two BF16 values, hi = bf16(w) and lo = bf16(w - hi) (16 significant
 * bits), and multiplied twice; the impacts then differ from the reference's by about 2^-17 of a weight. */
#define KNP_DENSE_SPLIT 2u
int knp_mi355x_set_projection_dense(knp_net *net, uint32_t proj, int format, uint32_t flags);
/* Weight bytes the MFMA expansions read: so far by the KNP_DENSE_SKIP_EMPTY ones (a group of 8 presynaptic
 * indexes with a spike reads 16 bytes per postsynaptic neuron of its instance, padded to 32), and per step by
 * the others (every group). */
int knp_mi355x_get_dense_weight_bytes(knp_net *net, uint64_t *skip_bytes, uint64_t *full_bytes_per_step);
/* The BF16 weight copy of a projection expanded on MFMA, in synapse-index order (bf16 bit patterns); out_lo
 * (may be NULL) receives the low parts of a KNP_DENSE_SPLIT copy (0 otherwise). */
int knp_mi355x_read_dense_weights(knp_net *net, uint32_t proj, uint16_t *out, uint16_t *out_lo, uint64_t n_syn);

/* ---- stepping (Backend::_step, backend.cpp:89-141) ------------------------------------------------ */
int knp_mi355x_step(knp_net *net, uint64_t n_steps);
/* No counterpart in the reference: knp_mi355x_step for callers that cannot allocate device memory (code under
 * HIP stream capture, serving loops, harnesses that forbid hipMalloc). The network must be initialized already
 * (knp_mi355x_init allocates its state; calling this first is KNP_ERR_STATE). The spike counts that the MFMA
 * expansions (knp_mi355x_set_projection_dense) stage between their two kernels go to a file-scope device
 * scratch of 1 MiB instead of the projections' own buffers: 16 bytes per group of 8 presynaptic indexes of
 * every such projection, at most 1 MiB in all (more is KNP_ERR_RANGE). No kernel timing events are created.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU (threads, streams) race.
 * Results equal knp_mi355x_step's. */
int knp_mi355x_step_noalloc(knp_net *net, uint64_t n_steps);
/* Captured step graphs (no reference counterpart: an execution choice with identical results). Where a call
 * runs one launch per step (populations that are not all tiled, or tiles that do not all fit on the device at
 * once), the launches of `steps` consecutive steps are captured once into a HIP graph, whose kernels read the
 * step from device memory, and replayed for the middle of every longer call (the first and last steps of a
 * call run as plain launches; pushed channel spikes are delivered at the first). With dense input projections
 * a graph holds one time block. 0 disables graphs; the default is 32. Kernel timing turns them off. */
int knp_mi355x_set_step_graph(knp_net *net, uint32_t steps);
/* The steps of the captured graph (0: none yet) and the number of replays so far. */
int knp_mi355x_get_step_graph(knp_net *net, uint32_t *steps, uint64_t *replays);
/* Backend::get_step (backend.h:290). */
int knp_mi355x_get_step(knp_net *net, uint64_t *step);
int knp_mi355x_synchronize(knp_net *net);

/* ---- outputs ------------------------------------------------------------------------------------- */
/* Observe a population: its spikes of every following step are kept on device as bitmaps (ceil(n/32)
 * words per step, bit i = neuron i spiked) for up to capacity_steps steps. */
int knp_mi355x_record_spikes(knp_net *net, uint32_t pop, uint64_t capacity_steps);
/* Copies the recorded bitmaps (oldest first) and the number of recorded steps, then clears the record. */
int knp_mi355x_read_spike_record(knp_net *net, uint32_t pop, uint32_t *words, uint64_t n_words, uint64_t *n_steps);
/* SpikeMessage of the last step for `pop`: ascending neuron indexes (populations_old.h:110-118). */
int knp_mi355x_read_last_spikes(knp_net *net, uint32_t pop, uint32_t *idx, uint32_t cap, uint32_t *n);
/* get_network_data (backend.h:420; ST get_network.cpp:28-93): state of every neuron / synapse,
 * in load order and synapse-index order. */
int knp_mi355x_read_potentials(knp_net *net, uint32_t pop, float *out, uint32_t n);
int knp_mi355x_read_blifat_neurons(knp_net *net, uint32_t pop, knp_blifat_neuron *out, uint32_t n);
int knp_mi355x_read_resource_neurons(knp_net *net, uint32_t pop, knp_resource_blifat_neuron *out, uint32_t n);
int knp_mi355x_read_lif_neurons(knp_net *net, uint32_t pop, knp_lif_neuron *out, uint32_t n);
int knp_mi355x_read_altai_lif_neurons(knp_net *net, uint32_t pop, knp_altai_lif_neuron *out, uint32_t n);
int knp_mi355x_read_resource_altai_lif_neurons(knp_net *net, uint32_t pop, knp_resource_altai_lif_neuron *out,
                                               uint32_t n);
int knp_mi355x_read_weights(knp_net *net, uint32_t proj, float *out, uint64_t n_syn);
int knp_mi355x_read_resource_rules(knp_net *net, uint32_t proj, knp_resource_rule *out, uint64_t n_syn);
/* Network file: save_network / load_network of knp/framework-library/include/knp/framework/sonata/
 * network_io.h:41,51 (SONATA over HDF5 in the reference; HDF5 is absent here). The file holds every
 * population (neuron model and the C ABI struct of each neuron with its current state) and projection
 * (kind, presynaptic / postsynaptic populations, synapses with current weights, synaptic-resource rules
 * or additive-STDP time constants, learning lock), in load order. As in the reference, impacts still in
 * flight, spike records, channel sources and dense input projections are not part of the network file.
 * load_network adds the file's populations and projections to a handle that has none yet; when it fails,
 * the handle may hold part of the file's network and should be destroyed. */
int knp_mi355x_save_network(knp_net *net, const char *path);
int knp_mi355x_load_network(knp_net *net, const char *path);
/* Number of SynapticImpacts created so far (sum over steps of spikes x out-degree over all projections). */
int knp_mi355x_get_event_count(knp_net *net, uint64_t *events);
/* Counters of the spike -> impact expansion so far: impacts (as above) and presynaptic rows expanded
 * (one per spike per projection that receives it). */
int knp_mi355x_get_counters(knp_net *net, uint64_t *events, uint64_t *rows);
/* Fixed-point fraction bits of the exact impact accumulator of a population (chosen at init). */
int knp_mi355x_get_accumulator_bits(knp_net *net, uint32_t pop, int32_t *frac_bits);
/* Loaded synapse weights into a population that are not multiples of 2^-frac_bits (float weights below
 * 2^(23 - frac_bits) with low mantissa bits): each of their impacts is rounded to the accumulator's grid, so
 * sums that include them are exact only to 2^-(frac_bits + 1) per impact. 0 means every sum is exact. */
int knp_mi355x_get_inexact_weights(knp_net *net, uint32_t pop, uint64_t *count);
/* Work of the tiled path so far: row parts read (a spiking row's synapses in one tile) and ring cells
 * updated in HBM (non-zero tile sums). */
int knp_mi355x_get_tile_counters(knp_net *net, uint64_t *row_parts, uint64_t *ring_cells);

/* ---- kernel timing (HIP events on the library's own stream) -------------------------------------- */
enum knp_kernel_id {
    KNP_KERNEL_NEURON = 0, /* fused population step (+ spike compaction) */
    KNP_KERNEL_SCATTER = 1, /* spike -> impact scatter into the delay ring */
    KNP_KERNEL_TRAIN = 2,   /* synaptic-resource STDP */
    KNP_KERNEL_DENSE = 3,   /* dense input projection on MFMA (spike bits + product), per time block */
    KNP_KERNEL_TILE_SCATTER = 4, /* tile-owner scatter (tiled populations), one step per launch */
    KNP_KERNEL_TILE_RUN = 5,     /* persistent tile launch: several steps per launch */
    KNP_KERNEL_DENSE_EXPAND = 6, /* projections expanded on MFMA: spike counts + product, per step */
    KNP_KERNEL_COUNT = 7
};
int knp_mi355x_enable_timing(knp_net *net, int enable);
/* Average duration (ms) and launch count of a kernel class since timing was enabled. */
int knp_mi355x_kernel_time(knp_net *net, int kernel_id, double *avg_ms, uint64_t *launches);

/* ---- shared helper: the counter-based hash behind Bernoulli sources (identical on host and device) */
uint32_t knp_bernoulli_hash(uint64_t seed, uint64_t step, uint32_t channel);

#ifdef __cplusplus
}
#endif
#endif /* KNP_MI355X_H */
