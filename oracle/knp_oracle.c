/*
 * knp_oracle.c - CPU restatement of the KNP single-threaded backend step (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the oracle the HIP backend is checked against. Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it; the product path never does.
 *
 * It restates, with its own data structures, the semantics of:
 *   - SingleThreadedCPUBackend::_step          knp/backends/cpu/cpu-single-threaded-backend/impl/backend.cpp:89-141
 *   - calculate_any_population / resource pop   knp/backends/cpu/cpu-library/include/knp/backends/cpu-library/populations_old.h:104-217
 *   - BLIFAT pre / impact / post                cpu-library/impl/populations/blifat/blifat_impl.h:31-168
 *   - LIF pre / impact / post                   cpu-library/impl/populations/lif/lif_impl.h:31-84
 *   - AltAI-LIF pre / impact / post             cpu-library/impl/populations/altai/altai_impl.h:31-150
 *   - resource STDP training                    cpu-library/impl/populations/blifat/blifat_stdp.h:35-172
 *                                               cpu-library/impl/populations/training/stdp.h:40-186
 *   - delta projection + delay queue            cpu-library/projections.h:58-84, impl/projections/delta/delta_impl.h:54-114
 *   - projection STDP hooks                     impl/projections/training/stdp/{stdp_dispatcher.h,synaptic_resource_stdp.h,additive_stdp.h}
 *   - message delivery order of the CPU bus     knp/core-library/impl/message_bus_cpu_impl/message_bus_cpu_impl.cpp:41-88,
 *                                               knp/core-library/impl/message_endpoint.cpp:151-241
 *   - the multi-threaded backend's decomposition knp/backends/cpu/cpu-multi-threaded-backend/impl/backend.cpp:77-323
 *     (ko_set_threads, built with OpenMP): neuron pre / post states in parallel parts, impacts applied per
 *     population in delivery order, projection impacts built from parallel chunks of the spike list. Thread
 *     buffers are merged in chunk order, so impact order, sums and spike order equal the one-thread step
 *     (the reference's MT merge order depends on thread timing; with exact sums its spike sets agree).
 *
 * Precision policy (compile-time): KO_WREAL is the type of every value the reference keeps in `float`
 * (weights, impacts, synaptic resources, the float fields of resource neurons).
 *   KO_WREAL=float  -> "KNP-exact": identical types to the reference (BLIFAT state in double).
 *   KO_WREAL=double -> "FP64 oracle": every real in double.
 * Build strict IEEE (-O2 -ffp-contract=off, no fast-math).
 *
 * Fixed choices where the reference leaves order to Boost's hashed index (projection.h:450-460):
 * synapses of one presynaptic (or postsynaptic) neuron are visited in ascending synapse index.
 * A projection onto a LIF population with synapses other than EXCITATORY / INHIBITORY_CURRENT is rejected
 * when it is added; the reference throws "Unhandled synapse type." once such an impact is delivered
 * (lif_impl.h:62-64). Onto AltAI-LIF populations, INHIBITORY_CONDUCTANCE (the reference throws) and BLOCKING
 * (order-dependent, altai_impl.h:72-79; not supported by the device) are rejected the same way.
 */
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/knp_mi355x.h"

#ifndef KO_WREAL
#define KO_WREAL float
#endif
typedef KO_WREAL wreal;

#define KO_PROJ_DELTA 0
#define KO_PROJ_RESOURCE 1
#define KO_PROJ_ADDITIVE 2

/* ---------------------------------------------------------------------------------------------- */
typedef struct {
    wreal synapse_sum_threshold_coefficient;
    uint32_t dopamine_plasticity_time;
    wreal free_synaptic_resource;
    wreal synaptic_resource_threshold;
    uint32_t resource_drain_coefficient;
    wreal stability;
    wreal stability_change_parameter;
    wreal stability_change_at_isi;
    uint32_t isi_max;
    wreal d_h;
    int isi_status;
    uint64_t last_step, last_spike_step, first_isi_spike;
    int is_being_forced;
} ko_res_ext;

typedef struct {
    wreal synaptic_resource, w_min, w_max, d_u;
    uint32_t dopamine_plasticity_period;
    uint64_t last_spike_step;
    int had_hebbian_update, has_contributed;
} ko_rule;

typedef struct {
    uint64_t syn;
    wreal w;
    int type;
    uint32_t pre, post;
} ko_impact;

typedef struct {
    ko_impact *v;
    size_t n, cap;
    int exists;
} ko_msg;

typedef struct {
    uint32_t *idx;
    uint32_t n;
    uint64_t send_time;
    int from_pop; /* population index of the sender, -1 for a channel */
} ko_spike_msg;

/* neuron_parameters<LIFNeuron> (lif.h:72-117); the reference keeps the reals in float */
typedef struct {
    wreal potential, potential_reset_value, activation_threshold, leak_coefficient;
    uint32_t refract_counter, refract_period;
} ko_lif;

/* neuron_parameters<AltAILIF> (altai_lif.h:128-354); reals as the reference keeps them */
typedef struct {
    int is_diff, is_reset, leak_rev, saturate, do_not_save;
    wreal potential, pre_impact_potential;
    uint16_t activation_threshold, negative_activation_threshold;
    int16_t potential_leak;
    uint16_t potential_reset_value;
    double dopamine_value, additional_threshold;
    int64_t activity_time;
} ko_altai;

typedef struct {
    int resource;
    int lif, altai;
    uint32_t n;
    ko_lif *nl;            /* LIF populations */
    ko_altai *na;          /* AltAI-LIF populations */
    ko_impact *injected;   /* impacts delivered straight to the population at the next step */
    size_t n_injected;
    knp_blifat_neuron *nb; /* BLIFAT fields, double as in the reference */
    ko_res_ext *rx;        /* resource fields (resource populations only) */
    int record;
    uint32_t *rec;         /* bitmaps per recorded step */
    uint64_t rec_steps, rec_cap;
    uint32_t *last_spikes;
    uint32_t n_last;
} ko_pop;

typedef struct {
    uint32_t first, n_channels, n_patterns, steps_per_pattern, lo, hi, hash_offset;
    uint32_t *thr;
    uint64_t seed;
    uint32_t proj;
} ko_source;

typedef struct {
    int kind, locked;
    uint32_t pre_pop, n_pre, post_pop;
    uint64_t n;
    uint32_t *pre, *post, *delay;
    uint8_t *type;
    wreal *w;
    ko_rule *rule;                   /* resource */
    wreal *tau_plus, *tau_minus;     /* additive */
    uint64_t **pre_times, **post_times;
    uint32_t *n_pre_times, *n_post_times, *cap_times;
    uint32_t *stdp_pops;
    uint8_t *stdp_modes;
    uint32_t n_stdp;
    uint64_t *row_ptr, *row_syn;     /* by presynaptic, ascending synapse index */
    uint64_t *col_ptr, *col_syn;     /* by postsynaptic, ascending synapse index */
    uint32_t max_delay;
    ko_msg *queue;                   /* ring indexed by key % max_delay */
    ko_msg outbox;                   /* message sent this step, delivered next step */
    int has_outbox;
    ko_spike_msg *inbox;             /* channel messages for the next step */
    uint32_t n_inbox, cap_inbox;
} ko_proj;

typedef struct ko_net {
    ko_pop *pops;
    uint32_t n_pops;
    ko_proj *projs;
    uint32_t n_projs;
    ko_source *srcs;
    uint32_t n_srcs;
    uint64_t step;
    uint64_t events;
    int initialized;
    int threads;  /* > 1: the multi-threaded step (OpenMP builds) */
    int semantics; /* enum knp_message_semantics: whose projection messages the step follows */
    char err[256];
} ko_net;

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s ? s : 1);
    if (!p) {
        fprintf(stderr, "knp_oracle: out of memory\n");
        abort();
    }
    return p;
}

static void msg_push(ko_msg *m, const ko_impact *imp) {
    if (m->n == m->cap) {
        m->cap = m->cap ? 2 * m->cap : 16;
        m->v = (ko_impact *)realloc(m->v, m->cap * sizeof(ko_impact));
        if (!m->v) abort();
    }
    m->v[m->n++] = *imp;
    m->exists = 1;
}

/* ---- the counter-based hash of Bernoulli sources (the same definition the device uses) ----------- */
static uint64_t ko_mix64(uint64_t z) {
    /* splitmix64 finalizer (Steele, Lea, Flood 2014) */
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint32_t ko_bernoulli_hash(uint64_t seed, uint64_t step, uint32_t channel) {
    uint64_t a = ko_mix64(seed + step * 0x9E3779B97F4A7C15ULL);
    uint64_t b = ko_mix64(a ^ ((uint64_t)channel * 0xD6E8FEB86659FD93ULL + 0x632BE59BD9B4E019ULL));
    return (uint32_t)(b >> 32);
}

/* ---- construction ---------------------------------------------------------------------------- */
ko_net *ko_create(void) { return (ko_net *)xcalloc(1, sizeof(ko_net)); }

void ko_destroy(ko_net *net) {
    if (!net) return;
    for (uint32_t i = 0; i < net->n_pops; ++i) {
        free(net->pops[i].nb);
        free(net->pops[i].nl);
        free(net->pops[i].na);
        free(net->pops[i].injected);
        free(net->pops[i].rx);
        free(net->pops[i].rec);
        free(net->pops[i].last_spikes);
    }
    for (uint32_t q = 0; q < net->n_projs; ++q) {
        ko_proj *p = &net->projs[q];
        free(p->pre); free(p->post); free(p->delay); free(p->type); free(p->w); free(p->rule);
        free(p->tau_plus); free(p->tau_minus);
        if (p->pre_times) {
            for (uint64_t s = 0; s < p->n; ++s) { free(p->pre_times[s]); free(p->post_times[s]); }
        }
        free(p->pre_times); free(p->post_times); free(p->n_pre_times); free(p->n_post_times); free(p->cap_times);
        free(p->stdp_pops); free(p->stdp_modes);
        free(p->row_ptr); free(p->row_syn); free(p->col_ptr); free(p->col_syn);
        if (p->queue) { for (uint32_t k = 0; k < p->max_delay; ++k) free(p->queue[k].v); }
        free(p->queue); free(p->outbox.v);
        for (uint32_t m = 0; m < p->n_inbox; ++m) free(p->inbox[m].idx);
        free(p->inbox);
    }
    for (uint32_t s = 0; s < net->n_srcs; ++s) free(net->srcs[s].thr);
    free(net->pops); free(net->projs); free(net->srcs);
    free(net);
}

const char *ko_last_error(ko_net *net) { return net->err; }

/* Threads of the step (1: the single-threaded backend; more: the multi-threaded decomposition, effective in
 * OpenMP builds). */
int ko_set_threads(ko_net *net, int threads) {
    net->threads = threads > 1 ? threads : 1;
    return 0;
}

int ko_set_message_semantics(ko_net *net, int semantics) {
    if (semantics != KNP_SEMANTICS_SINGLE_THREADED && semantics != KNP_SEMANTICS_MULTI_THREADED) return -1;
    net->semantics = semantics;
    return 0;
}

/* Under the multi-threaded backend's semantics the network holds that backend's types only
 * (cpu-multi-threaded-backend/include/knp/backends/cpu-multi-threaded/backend.h:80-88) and no BLOCKING
 * synapses (their last-impact order follows that backend's thread completion order). */
static int check_multi_threaded(ko_net *net) {
    for (uint32_t k = 0; k < net->n_pops; ++k)
        if (net->pops[k].altai) {
            snprintf(net->err, sizeof net->err, "Type is not supported.");
            return -1;
        }
    for (uint32_t q = 0; q < net->n_projs; ++q) {
        const ko_proj *p = &net->projs[q];
        if (p->kind == KO_PROJ_ADDITIVE) {
            snprintf(net->err, sizeof net->err, "Type is not supported.");
            return -1;
        }
        for (uint64_t s = 0; s < p->n; ++s)
            if (p->type[s] == KNP_BLOCKING) {
                snprintf(net->err, sizeof net->err, "BLOCKING synapses under the multi-threaded semantics");
                return -1;
            }
    }
    return 0;
}

static int add_pop(ko_net *net, uint32_t n, int resource) {
    net->pops = (ko_pop *)realloc(net->pops, (net->n_pops + 1) * sizeof(ko_pop));
    ko_pop *p = &net->pops[net->n_pops];
    memset(p, 0, sizeof(*p));
    p->n = n;
    p->resource = resource;
    p->nb = (knp_blifat_neuron *)xcalloc(n, sizeof(knp_blifat_neuron));
    p->last_spikes = (uint32_t *)xcalloc(n, sizeof(uint32_t));
    if (resource) p->rx = (ko_res_ext *)xcalloc(n, sizeof(ko_res_ext));
    return (int)(net->n_pops++);
}

int ko_add_blifat_population(ko_net *net, const knp_blifat_neuron *neurons, uint32_t n) {
    int id = add_pop(net, n, 0);
    memcpy(net->pops[id].nb, neurons, (size_t)n * sizeof(knp_blifat_neuron));
    return id;
}

int ko_add_lif_population(ko_net *net, const knp_lif_neuron *neurons, uint32_t n) {
    int id = add_pop(net, n, 0);
    ko_pop *p = &net->pops[id];
    p->lif = 1;
    p->nl = (ko_lif *)xcalloc(n, sizeof(ko_lif));
    for (uint32_t i = 0; i < n; ++i) {
        ko_lif *x = &p->nl[i];
        x->potential = neurons[i].potential;
        x->potential_reset_value = neurons[i].potential_reset_value;
        x->activation_threshold = neurons[i].activation_threshold;
        x->leak_coefficient = neurons[i].leak_coefficient;
        x->refract_counter = neurons[i].refract_counter;
        x->refract_period = neurons[i].refract_period;
    }
    return id;
}

int ko_add_altai_lif_population(ko_net *net, const knp_altai_lif_neuron *neurons, uint32_t n) {
    int id = add_pop(net, n, 0);
    ko_pop *p = &net->pops[id];
    p->altai = 1;
    p->na = (ko_altai *)xcalloc(n, sizeof(ko_altai));
    for (uint32_t i = 0; i < n; ++i) {
        const knp_altai_lif_neuron *s = &neurons[i];
        ko_altai *x = &p->na[i];
        x->is_diff = s->is_diff != 0;
        x->is_reset = s->is_reset != 0;
        x->leak_rev = s->leak_rev != 0;
        x->saturate = s->saturate != 0;
        x->do_not_save = s->do_not_save != 0;
        x->potential = s->potential;
        x->pre_impact_potential = s->pre_impact_potential;
        x->activation_threshold = s->activation_threshold;
        x->negative_activation_threshold = s->negative_activation_threshold;
        x->potential_leak = s->potential_leak;
        x->potential_reset_value = s->potential_reset_value;
        x->dopamine_value = s->dopamine_value;
        x->additional_threshold = s->additional_threshold;
        x->activity_time = s->activity_time;
    }
    return id;
}

/* Impacts (neuron, weight, type) delivered to `pop` at the next step, before the projections' messages: the
 * reference's population tests send SynapticImpactMessages straight to a population
 * (knp/tests/framework/altai_lif_test.cpp:60-100). Test hook; no projection is involved. */
int ko_inject_impacts(ko_net *net, uint32_t pop, const uint32_t *neuron, const float *w, const uint8_t *type,
                      uint32_t n) {
    if (pop >= net->n_pops) return -1;
    ko_pop *P = &net->pops[pop];
    P->injected = (ko_impact *)realloc(P->injected, (P->n_injected + n) * sizeof(ko_impact));
    for (uint32_t k = 0; k < n; ++k) {
        if (neuron[k] >= P->n) return -1;
        ko_impact *imp = &P->injected[P->n_injected++];
        memset(imp, 0, sizeof(*imp));
        imp->w = (wreal)w[k];
        imp->type = type ? type[k] : KNP_EXCITATORY;
        imp->post = neuron[k];
    }
    return 0;
}

/* the resource fields of a knp_resource_*_neuron (everything after `base`) into a ko_res_ext */
#define KO_SET_RES_EXT(x, s)                                                     \
    do {                                                                         \
        (x)->synapse_sum_threshold_coefficient = (s)->synapse_sum_threshold_coefficient; \
        (x)->dopamine_plasticity_time = (s)->dopamine_plasticity_time;           \
        (x)->free_synaptic_resource = (s)->free_synaptic_resource;               \
        (x)->synaptic_resource_threshold = (s)->synaptic_resource_threshold;     \
        (x)->resource_drain_coefficient = (s)->resource_drain_coefficient;       \
        (x)->stability = (s)->stability;                                         \
        (x)->stability_change_parameter = (s)->stability_change_parameter;       \
        (x)->stability_change_at_isi = (s)->stability_change_at_isi;             \
        (x)->isi_max = (s)->isi_max;                                             \
        (x)->d_h = (s)->d_h;                                                     \
        (x)->isi_status = (s)->isi_status;                                       \
        (x)->last_step = (s)->last_step;                                         \
        (x)->last_spike_step = (s)->last_spike_step;                             \
        (x)->first_isi_spike = (s)->first_isi_spike;                             \
        (x)->is_being_forced = (s)->is_being_forced;                             \
    } while (0)

int ko_add_resource_altai_lif_population(ko_net *net, const knp_resource_altai_lif_neuron *neurons, uint32_t n) {
    knp_altai_lif_neuron *base = (knp_altai_lif_neuron *)xcalloc(n ? n : 1, sizeof(knp_altai_lif_neuron));
    for (uint32_t i = 0; i < n; ++i) base[i] = neurons[i].base;
    int id = ko_add_altai_lif_population(net, base, n);
    free(base);
    ko_pop *p = &net->pops[id];
    p->resource = 1;
    p->rx = (ko_res_ext *)xcalloc(n, sizeof(ko_res_ext));
    for (uint32_t i = 0; i < n; ++i) KO_SET_RES_EXT(&p->rx[i], &neurons[i]);
    return id;
}

int ko_add_resource_blifat_population(ko_net *net, const knp_resource_blifat_neuron *neurons, uint32_t n) {
    int id = add_pop(net, n, 1);
    ko_pop *p = &net->pops[id];
    for (uint32_t i = 0; i < n; ++i) {
        const knp_resource_blifat_neuron *s = &neurons[i];
        p->nb[i] = s->base;
        KO_SET_RES_EXT(&p->rx[i], s);
    }
    return id;
}

static int add_proj(ko_net *net, int kind, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop, const uint32_t *pre,
                    const uint32_t *post, const float *weight, const uint32_t *delay, const uint8_t *type, uint64_t n) {
    if (post_pop >= net->n_pops) { snprintf(net->err, sizeof net->err, "bad post population"); return -1; }
    if (pre_pop != KNP_NO_POPULATION && pre_pop >= net->n_pops) {
        snprintf(net->err, sizeof net->err, "bad pre population");
        return -1;
    }
    uint32_t n_post = net->pops[post_pop].n;
    if (pre_pop != KNP_NO_POPULATION) n_pre = net->pops[pre_pop].n;
    net->projs = (ko_proj *)realloc(net->projs, (net->n_projs + 1) * sizeof(ko_proj));
    ko_proj *p = &net->projs[net->n_projs];
    memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->locked = 1; /* projection.h:432 */
    p->pre_pop = pre_pop;
    p->n_pre = n_pre;
    p->post_pop = post_pop;
    p->n = n;
    p->pre = (uint32_t *)xcalloc(n, 4);
    p->post = (uint32_t *)xcalloc(n, 4);
    p->delay = (uint32_t *)xcalloc(n, 4);
    p->type = (uint8_t *)xcalloc(n, 1);
    p->w = (wreal *)xcalloc(n, sizeof(wreal));
    p->max_delay = 1;
    for (uint64_t s = 0; s < n; ++s) {
        if (pre[s] >= n_pre || post[s] >= n_post) { snprintf(net->err, sizeof net->err, "synapse index out of range"); return -1; }
        if (delay[s] == 0) { snprintf(net->err, sizeof net->err, "delay 0 is not supported"); return -1; }
        p->pre[s] = pre[s];
        p->post[s] = post[s];
        p->w[s] = (wreal)weight[s];
        p->delay[s] = delay[s];
        p->type[s] = type ? type[s] : KNP_EXCITATORY;
        const int t = p->type[s];
        const ko_pop *Pp = &net->pops[post_pop];
        if (t > KNP_BLOCKING || (Pp->lif && t > KNP_INHIBITORY_CURRENT) ||
            (Pp->altai && (t == KNP_INHIBITORY_CONDUCTANCE || t == KNP_BLOCKING))) {
            snprintf(net->err, sizeof net->err, "Unhandled synapse type.");
            return -1;
        }
        if (delay[s] > p->max_delay) p->max_delay = delay[s];
    }
    /* by-presynaptic and by-postsynaptic indexes (counting sort keeps ascending synapse index) */
    p->row_ptr = (uint64_t *)xcalloc((size_t)n_pre + 1, 8);
    p->row_syn = (uint64_t *)xcalloc(n, 8);
    p->col_ptr = (uint64_t *)xcalloc((size_t)n_post + 1, 8);
    p->col_syn = (uint64_t *)xcalloc(n, 8);
    for (uint64_t s = 0; s < n; ++s) { p->row_ptr[p->pre[s] + 1]++; p->col_ptr[p->post[s] + 1]++; }
    for (uint32_t i = 0; i < n_pre; ++i) p->row_ptr[i + 1] += p->row_ptr[i];
    for (uint32_t i = 0; i < n_post; ++i) p->col_ptr[i + 1] += p->col_ptr[i];
    uint64_t *rf = (uint64_t *)xcalloc((size_t)n_pre + 1, 8), *cf = (uint64_t *)xcalloc((size_t)n_post + 1, 8);
    memcpy(rf, p->row_ptr, ((size_t)n_pre + 1) * 8);
    memcpy(cf, p->col_ptr, ((size_t)n_post + 1) * 8);
    for (uint64_t s = 0; s < n; ++s) { p->row_syn[rf[p->pre[s]]++] = s; p->col_syn[cf[p->post[s]]++] = s; }
    free(rf); free(cf);
    p->queue = (ko_msg *)xcalloc(p->max_delay, sizeof(ko_msg));
    return (int)(net->n_projs++);
}

int ko_add_delta_projection(ko_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop, const uint32_t *pre,
                            const uint32_t *post, const float *weight, const uint32_t *delay, const uint8_t *type,
                            uint64_t n) {
    return add_proj(net, KO_PROJ_DELTA, pre_pop, n_pre, post_pop, pre, post, weight, delay, type, n);
}

int ko_add_resource_stdp_projection(ko_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop,
                                    const uint32_t *pre, const uint32_t *post, const float *weight,
                                    const uint32_t *delay, const uint8_t *type, const knp_resource_rule *rules,
                                    uint64_t n) {
    int id = add_proj(net, KO_PROJ_RESOURCE, pre_pop, n_pre, post_pop, pre, post, weight, delay, type, n);
    if (id < 0) return id;
    if (!net->pops[post_pop].resource) {
        snprintf(net->err, sizeof net->err, "resource STDP projection must target a resource population");
        net->n_projs--;
        return -1;
    }
    ko_proj *p = &net->projs[id];
    p->rule = (ko_rule *)xcalloc(n, sizeof(ko_rule));
    for (uint64_t s = 0; s < n; ++s) {
        p->rule[s].synaptic_resource = rules[s].synaptic_resource;
        p->rule[s].w_min = rules[s].w_min;
        p->rule[s].w_max = rules[s].w_max;
        p->rule[s].d_u = rules[s].d_u;
        p->rule[s].dopamine_plasticity_period = rules[s].dopamine_plasticity_period;
        p->rule[s].last_spike_step = rules[s].last_spike_step;
        p->rule[s].had_hebbian_update = rules[s].had_hebbian_update;
        p->rule[s].has_contributed = rules[s].has_contributed;
    }
    return id;
}

int ko_add_additive_stdp_projection(ko_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop,
                                    const uint32_t *pre, const uint32_t *post, const float *weight,
                                    const uint32_t *delay, const uint8_t *type, const float *tau_plus,
                                    const float *tau_minus, uint64_t n, const uint32_t *stdp_pops,
                                    const uint8_t *stdp_modes, uint32_t n_stdp) {
    int id = add_proj(net, KO_PROJ_ADDITIVE, pre_pop, n_pre, post_pop, pre, post, weight, delay, type, n);
    if (id < 0) return id;
    ko_proj *p = &net->projs[id];
    p->tau_plus = (wreal *)xcalloc(n, sizeof(wreal));
    p->tau_minus = (wreal *)xcalloc(n, sizeof(wreal));
    p->pre_times = (uint64_t **)xcalloc(n, sizeof(uint64_t *));
    p->post_times = (uint64_t **)xcalloc(n, sizeof(uint64_t *));
    p->n_pre_times = (uint32_t *)xcalloc(n, 4);
    p->n_post_times = (uint32_t *)xcalloc(n, 4);
    p->cap_times = (uint32_t *)xcalloc(n, 4);
    for (uint64_t s = 0; s < n; ++s) {
        p->tau_plus[s] = tau_plus[s];
        p->tau_minus[s] = tau_minus[s];
        /* the list only grows while size() < tau_minus_ + tau_plus_ (additive_stdp.h:108) */
        double lim = (double)tau_minus[s] + (double)tau_plus[s];
        uint32_t cap = lim <= 0 ? 0 : (uint32_t)ceil(lim);
        p->cap_times[s] = cap;
        p->pre_times[s] = (uint64_t *)xcalloc(cap ? cap : 1, 8);
        p->post_times[s] = (uint64_t *)xcalloc(cap ? cap : 1, 8);
    }
    p->n_stdp = n_stdp;
    p->stdp_pops = (uint32_t *)xcalloc(n_stdp, 4);
    p->stdp_modes = (uint8_t *)xcalloc(n_stdp, 1);
    for (uint32_t k = 0; k < n_stdp; ++k) { p->stdp_pops[k] = stdp_pops[k]; p->stdp_modes[k] = stdp_modes[k]; }
    return id;
}

int ko_set_locked(ko_net *net, uint32_t proj, int locked) {
    if (proj >= net->n_projs) return -1;
    net->projs[proj].locked = locked ? 1 : 0;
    return 0;
}

int ko_record_spikes(ko_net *net, uint32_t pop, uint64_t cap) {
    if (pop >= net->n_pops) return -1;
    ko_pop *p = &net->pops[pop];
    uint32_t words = (p->n + 31) / 32;
    free(p->rec);
    p->rec = (uint32_t *)xcalloc((size_t)(cap ? cap : 1) * words, 4);
    p->rec_cap = cap;
    p->rec_steps = 0;
    p->record = 1;
    return 0;
}

int ko_read_spike_record(ko_net *net, uint32_t pop, uint32_t *words, uint64_t n_words, uint64_t *n_steps) {
    if (pop >= net->n_pops) return -1;
    ko_pop *p = &net->pops[pop];
    uint64_t w = (p->n + 31) / 32;
    uint64_t need = w * p->rec_steps;
    if (n_words < need) return -1;
    memcpy(words, p->rec, need * 4);
    *n_steps = p->rec_steps;
    p->rec_steps = 0;
    return 0;
}

int ko_push_spikes(ko_net *net, uint32_t proj, const uint32_t *idx, uint32_t n) {
    if (proj >= net->n_projs) return -1;
    ko_proj *p = &net->projs[proj];
    for (uint32_t i = 0; i < n; ++i)
        if (idx[i] >= p->n_pre) { snprintf(net->err, sizeof net->err, "spike index out of range"); return -1; }
    if (p->n_inbox == p->cap_inbox) {
        p->cap_inbox = p->cap_inbox ? 2 * p->cap_inbox : 4;
        p->inbox = (ko_spike_msg *)realloc(p->inbox, p->cap_inbox * sizeof(ko_spike_msg));
    }
    ko_spike_msg *m = &p->inbox[p->n_inbox++];
    m->idx = (uint32_t *)xcalloc(n, 4);
    memcpy(m->idx, idx, (size_t)n * 4);
    m->n = n;
    m->send_time = net->step;
    m->from_pop = -1;
    return 0;
}

int ko_add_bernoulli_source(ko_net *net, uint32_t proj, uint32_t first, uint32_t n_channels, const uint32_t *thr,
                            uint32_t n_patterns, uint32_t steps_per_pattern, uint32_t lo, uint32_t hi, uint64_t seed,
                            uint32_t hash_offset) {
    if (proj >= net->n_projs || n_patterns == 0 || steps_per_pattern == 0) return -1;
    if ((uint64_t)first + n_channels > net->projs[proj].n_pre) return -1;
    net->srcs = (ko_source *)realloc(net->srcs, (net->n_srcs + 1) * sizeof(ko_source));
    ko_source *s = &net->srcs[net->n_srcs++];
    s->first = first;
    s->n_channels = n_channels;
    s->n_patterns = n_patterns;
    s->steps_per_pattern = steps_per_pattern;
    s->lo = lo;
    s->hi = hi;
    s->seed = seed;
    s->hash_offset = hash_offset;
    s->proj = proj;
    s->thr = (uint32_t *)xcalloc((size_t)n_channels * n_patterns, 4);
    memcpy(s->thr, thr, (size_t)n_channels * n_patterns * 4);
    return 0;
}

/* ---- BLIFAT (blifat_impl.h:31-168) ----------------------------------------------------------- */
static void blifat_pre(knp_blifat_neuron *n) {
    ++n->n_time_steps_since_last_firing;
    n->dynamic_threshold *= n->threshold_decay;
    n->postsynaptic_trace *= n->postsynaptic_trace_decay;
    n->inhibitory_conductance *= n->inhibitory_conductance_decay;
    n->potential *= n->potential_decay;
    if (1 == n->bursting_phase) {
        --n->bursting_phase;
        n->potential += n->reflexive_weight;
    }
    n->pre_impact_potential = n->potential;
}

static void resource_pre(knp_blifat_neuron *n, ko_res_ext *x) {
    ++n->n_time_steps_since_last_firing;
    n->dynamic_threshold *= n->threshold_decay;
    n->postsynaptic_trace *= n->postsynaptic_trace_decay;
    n->inhibitory_conductance *= n->inhibitory_conductance_decay;
    n->dopamine_value = 0.0;
    x->is_being_forced = 0;
    n->potential *= n->potential_decay;
    if (1 == n->bursting_phase) n->potential += n->reflexive_weight; /* no decrement: blifat_impl.h:62 */
    n->pre_impact_potential = n->potential;
}

static void blifat_impact(knp_blifat_neuron *n, const ko_impact *imp) {
    switch (imp->type) {
        case KNP_EXCITATORY: n->potential += imp->w; break;
        case KNP_INHIBITORY_CURRENT: n->potential -= imp->w; break;
        case KNP_INHIBITORY_CONDUCTANCE: n->inhibitory_conductance += imp->w; break;
        case KNP_DOPAMINE: n->dopamine_value += imp->w; break;
        case KNP_BLOCKING: n->total_blocking_period = (int64_t)imp->w; break;
        default: break; /* rejected at load */
    }
}

static int blifat_post(knp_blifat_neuron *n) {
    int spike = 0;
    if (n->total_blocking_period <= 0) n->potential = n->pre_impact_potential;
    if (n->total_blocking_period < 0) {
        ++n->total_blocking_period;
        if (n->total_blocking_period == 0) n->total_blocking_period = INT64_MAX;
    } else if (n->total_blocking_period > 0) {
        n->total_blocking_period -= 1;
    }
    if (n->inhibitory_conductance < 1.0)
        n->potential -= (n->potential - n->reversal_inhibitory_potential) * n->inhibitory_conductance;
    else
        n->potential = n->reversal_inhibitory_potential;
    spike = (n->n_time_steps_since_last_firing > n->absolute_refractory_period) &&
            (n->potential >= n->activation_threshold + n->dynamic_threshold + n->additional_threshold);
    if (spike) {
        n->dynamic_threshold += n->threshold_increment;
        n->postsynaptic_trace += n->postsynaptic_trace_increment;
        n->potential = n->potential_reset_value;
        n->bursting_phase = n->bursting_period;
        n->n_time_steps_since_last_firing = 0;
    }
    if (n->potential < n->min_potential) n->potential = n->min_potential;
    return spike;
}

/* ---- LIF (lif_impl.h:31-84) ------------------------------------------------------------------- */
static void lif_pre(ko_lif *n) {
    if (0 == n->refract_counter) n->potential *= n->leak_coefficient;
}

static void lif_impact(ko_lif *n, const ko_impact *imp) {
    if (0 != n->refract_counter) return; /* impacts are ignored while refractory */
    if (imp->type == KNP_EXCITATORY) n->potential += imp->w;
    else if (imp->type == KNP_INHIBITORY_CURRENT) n->potential -= imp->w;
}

static int lif_post(ko_lif *n) {
    if (0 == n->refract_counter) {
        if (n->potential > n->activation_threshold) {
            n->potential = n->potential_reset_value;
            n->refract_counter = n->refract_period;
            return 1;
        }
    } else {
        --n->refract_counter;
    }
    return 0;
}

/* ---- AltAI-LIF (altai_impl.h:31-150) ---------------------------------------------------------- */
static void altai_pre(ko_altai *n) {
    n->potential = n->do_not_save ? (wreal)n->potential_reset_value : (wreal)round(n->potential);
    n->pre_impact_potential = n->potential;
}

static void altai_impact(ko_altai *n, const ko_impact *imp) {
    if (imp->type == KNP_EXCITATORY) n->potential += imp->w;
    else if (imp->type == KNP_INHIBITORY_CURRENT) n->potential -= imp->w;
    else if (imp->type == KNP_DOPAMINE) n->dopamine_value += imp->w;
}

static int altai_post(ko_altai *n) {
    const int sign = (n->leak_rev && n->potential < 0) ? -1 : 1;
    n->potential += (wreal)(n->potential_leak * sign);
    int spiked = 0;
    if (n->activity_time > 0) --n->activity_time;
    else if (n->activity_time < 0) ++n->activity_time;
    if (n->activity_time == 0) n->activity_time = INT64_MAX;
    int was_reset = 0;
    const double thr = (double)n->activation_threshold + n->additional_threshold;
    if ((double)n->potential >= thr) {
        if (n->activity_time > 0) spiked = 1;
        if (n->is_diff) n->potential = (wreal)((double)n->potential - thr);
        if (n->is_reset) {
            n->potential = (wreal)n->potential_reset_value;
            was_reset = 1;
        }
    }
    if (n->potential <= -(wreal)n->negative_activation_threshold && !was_reset) {
        if (n->saturate) n->potential = -(wreal)n->negative_activation_threshold;
        else if (n->is_reset) n->potential = (wreal)(-(int)n->potential_reset_value);
        else if (n->is_diff) n->potential += (wreal)n->negative_activation_threshold;
    }
    return spiked;
}

/* ---- resource STDP training (blifat_stdp.h, training/stdp.h) --------------------------------- */
typedef struct {
    ko_proj *proj;
    uint64_t syn;
} ko_synref;

static size_t connected_synapses(ko_net *net, uint32_t pop, uint32_t neuron, ko_synref **out, size_t *cap) {
    /* get_all_connected_synapses over working projections: unlocked resource projections targeting pop,
     * in load order (populations_old.h:57-83; training/stdp.h:63-78) */
    size_t k = 0;
    for (uint32_t q = 0; q < net->n_projs; ++q) {
        ko_proj *p = &net->projs[q];
        if (p->kind != KO_PROJ_RESOURCE || p->locked || p->post_pop != pop) continue;
        for (uint64_t j = p->col_ptr[neuron]; j < p->col_ptr[neuron + 1]; ++j) {
            if (k == *cap) {
                *cap = *cap ? 2 * *cap : 64;
                *out = (ko_synref *)realloc(*out, *cap * sizeof(ko_synref));
            }
            (*out)[k].proj = p;
            (*out)[k].syn = p->col_syn[j];
            ++k;
        }
    }
    return k;
}

static void recalc_weights(ko_synref *s, size_t k) {
    for (size_t i = 0; i < k; ++i) {
        ko_rule *r = &s[i].proj->rule[s[i].syn];
        wreal syn_w = r->synaptic_resource > (wreal)0 ? r->synaptic_resource : (wreal)0;
        wreal weight_diff = r->w_max - r->w_min;
        s[i].proj->w[s[i].syn] = r->w_min + weight_diff * syn_w / (weight_diff + syn_w);
    }
}

/* additional_threshold_ and dopamine_value_ of neuron i of a resource population (BLIFAT or AltAI-LIF base);
 * the training rules read nothing else of the base neuron (blifat_stdp.h, altai_stdp.h) */
static double *res_add_thr(ko_pop *P, uint32_t i) {
    return P->altai ? &P->na[i].additional_threshold : &P->nb[i].additional_threshold;
}
static double res_dop(const ko_pop *P, uint32_t i) {
    return P->altai ? P->na[i].dopamine_value : P->nb[i].dopamine_value;
}

static int update_isi(double dopamine, ko_res_ext *x, uint64_t step) {
    if (x->is_being_forced) {
        x->isi_status = KNP_ISI_IS_FORCED;
        return x->isi_status;
    }
    switch (x->isi_status) {
        case KNP_ISI_NOT_IN_PERIOD:
        case KNP_ISI_IS_FORCED:
            x->isi_status = KNP_ISI_PERIOD_STARTED;
            x->first_isi_spike = step;
            break;
        case KNP_ISI_PERIOD_STARTED:
            if (x->last_step - step < (uint64_t)x->isi_max) x->isi_status = KNP_ISI_PERIOD_CONTINUED;
            break;
        case KNP_ISI_PERIOD_CONTINUED:
            if (x->last_step - step >= (uint64_t)x->isi_max || dopamine != 0) {
                x->isi_status = KNP_ISI_PERIOD_STARTED;
                x->first_isi_spike = step;
            }
            break;
        default: break;
    }
    x->last_step = step;
    return x->isi_status;
}

static int point_in_interval(uint64_t b, uint64_t e, uint64_t p) {
    if (b > e) { uint64_t t = b; b = e; e = t; }
    return b <= p && p <= e;
}

static wreal stability_factor(wreal stability) {
    /* std::min(static_cast<float>(std::pow(2, -stability)), 1.F) */
    wreal v = (wreal)pow(2.0, -(double)stability);
    return (wreal)1 < v ? (wreal)1 : v; /* std::min(a, b) = (b < a) ? b : a */
}

static void train_population(ko_net *net, uint32_t pop, const uint32_t *spikes, uint32_t n_spikes, uint64_t step) {
    ko_pop *P = &net->pops[pop];
    ko_synref *syn = NULL;
    size_t cap = 0;
    /* process_spiking_neurons_impl (blifat_stdp.h:35-103) */
    for (uint32_t si = 0; si < n_spikes; ++si) {
        uint32_t idx = spikes[si];
        size_t k = connected_synapses(net, pop, idx, &syn, &cap);
        double *add_thr = res_add_thr(P, idx);
        ko_res_ext *x = &P->rx[idx];
        x->last_spike_step = step;
        update_isi(res_dop(P, idx), x, step);
        if (x->isi_status == KNP_ISI_PERIOD_STARTED) x->stability -= x->stability_change_at_isi;
        *add_thr = 0.0;
        for (size_t i = 0; i < k; ++i) {
            ko_rule *r = &syn[i].proj->rule[syn[i].syn];
            wreal w = syn[i].proj->w[syn[i].syn];
            *add_thr += w * (wreal)(w > 0);
            int had_spike = point_in_interval(step - r->dopamine_plasticity_period, step,
                                              r->last_spike_step + syn[i].proj->delay[syn[i].syn] - 1);
            if (x->isi_status != KNP_ISI_PERIOD_CONTINUED || had_spike) r->has_contributed = had_spike;
        }
        *add_thr *= x->synapse_sum_threshold_coefficient;
        if (x->isi_status != KNP_ISI_PERIOD_CONTINUED)
            for (size_t i = 0; i < k; ++i) syn[i].proj->rule[syn[i].syn].had_hebbian_update = 0;
        if (x->isi_status != KNP_ISI_IS_FORCED) {
            for (size_t i = 0; i < k; ++i) {
                ko_rule *r = &syn[i].proj->rule[syn[i].syn];
                r->synaptic_resource -= r->d_u;
                x->free_synaptic_resource += r->d_u;
                if (r->has_contributed && !r->had_hebbian_update) {
                    const wreal d_h = x->d_h * stability_factor(x->stability);
                    r->synaptic_resource += d_h;
                    x->free_synaptic_resource -= d_h;
                    r->had_hebbian_update = 1;
                }
            }
        }
        recalc_weights(syn, k);
    }
    /* do_dopamine_plasticity_impl (blifat_stdp.h:106-155) */
    for (uint32_t idx = 0; idx < P->n; ++idx) {
        const double dopamine = res_dop(P, idx);
        ko_res_ext *x = &P->rx[idx];
        if (dopamine > 0.0 || (dopamine < 0.0 && x->isi_status != KNP_ISI_IS_FORCED)) {
            size_t k = connected_synapses(net, pop, idx, &syn, &cap);
            for (size_t i = 0; i < k; ++i) {
                ko_rule *r = &syn[i].proj->rule[syn[i].syn];
                if (step - x->last_spike_step <= (uint64_t)x->dopamine_plasticity_time && r->has_contributed) {
                    wreal resource_change = (wreal)(dopamine * stability_factor(x->stability));
                    r->synaptic_resource += resource_change;
                    x->free_synaptic_resource -= resource_change;
                }
            }
            if (x->is_being_forced || dopamine < 0) {
                x->stability = (wreal)(x->stability - dopamine * x->stability_change_parameter);
                x->stability = x->stability < (wreal)0 ? (wreal)0 : x->stability; /* std::max(a, b) = (a < b) ? b : a */
            } else {
                const double dopamine_constant = 2.0;
                const double difference = fabs((double)(step - x->first_isi_spike - (uint64_t)x->isi_max));
                double f = dopamine_constant - difference / (double)x->isi_max;
                if (f < -1.0) f = -1.0;
                x->stability = (wreal)(x->stability + x->stability_change_parameter * dopamine * f);
            }
            recalc_weights(syn, k);
        }
    }
    /* renormalize_resource (training/stdp.h:152-186) */
    for (uint32_t idx = 0; idx < P->n; ++idx) {
        ko_res_ext *x = &P->rx[idx];
        if (step - x->last_step <= (uint64_t)x->isi_max && x->isi_status != KNP_ISI_IS_FORCED) continue;
        wreal f = x->free_synaptic_resource;
        if ((f < 0 ? -f : f) < x->synaptic_resource_threshold) continue;
        size_t k = connected_synapses(net, pop, idx, &syn, &cap);
        wreal add = x->free_synaptic_resource / (wreal)(uint64_t)(k + x->resource_drain_coefficient);
        for (size_t i = 0; i < k; ++i) syn[i].proj->rule[syn[i].syn].synaptic_resource += add;
        x->free_synaptic_resource = 0;
        recalc_weights(syn, k);
    }
    free(syn);
}

/* ---- projections (delta_impl.h:54-114, additive_stdp.h) -------------------------------------- */
static wreal stdp_w(wreal diff, wreal tau_plus, wreal tau_minus) {
    /* STDPFormula::stdp_w with a_plus = a_minus = 1 (additive_stdp.h:50-55, 187) */
#if defined(KO_WREAL_IS_DOUBLE)
    return diff > 0 ? exp(-diff / tau_plus) : exp(diff / tau_minus);
#else
    return diff > 0 ? expf(-diff / tau_plus) : expf(diff / tau_minus);
#endif
}

static void additive_init_projection(ko_net *net, ko_proj *p, ko_spike_msg *msgs, uint32_t n_msgs) {
    (void)net;
    for (uint32_t m = 0; m < n_msgs; ++m) {
        if (msgs[m].from_pop < 0) continue;
        int mode = -1;
        for (uint32_t k = 0; k < p->n_stdp; ++k)
            if (p->stdp_pops[k] == (uint32_t)msgs[m].from_pop) mode = p->stdp_modes[k];
        if (mode < 0) continue;
        for (int list = 0; list < 2; ++list) {
            /* list 0: postsynaptic history (STDPOnly or STDPAndSpike); list 1: presynaptic (STDPAndSpike) */
            if (list == 1 && mode != KNP_STDP_AND_SPIKE) continue;
            for (uint32_t i = 0; i < msgs[m].n; ++i) {
                uint32_t nidx = msgs[m].idx[i];
                if (nidx >= net->pops[p->post_pop].n) continue;
                for (uint64_t j = p->col_ptr[nidx]; j < p->col_ptr[nidx + 1]; ++j) {
                    uint64_t s = p->col_syn[j];
                    uint32_t *cnt = list == 0 ? &p->n_post_times[s] : &p->n_pre_times[s];
                    uint64_t *arr = list == 0 ? p->post_times[s] : p->pre_times[s];
                    if ((double)*cnt < (double)p->tau_minus[s] + (double)p->tau_plus[s] && *cnt < p->cap_times[s])
                        arr[(*cnt)++] = msgs[m].send_time;
                }
            }
        }
        if (mode == KNP_STDP_ONLY) msgs[m].n = 0;
    }
}

static void additive_modify_weights(ko_proj *p) {
    for (uint64_t s = 0; s < p->n; ++s) {
        wreal period = p->tau_plus[s] + p->tau_minus[s];
        if ((wreal)p->n_pre_times[s] >= period && (wreal)p->n_post_times[s] >= period) {
            wreal w_j = 0;
            for (uint32_t a = 0; a < p->n_pre_times[s]; ++a)
                for (uint32_t b = 0; b < p->n_post_times[s]; ++b)
                    w_j += stdp_w((wreal)(p->post_times[s][b] - p->pre_times[s][a]), p->tau_plus[s], p->tau_minus[s]);
            p->w[s] += w_j;
            p->n_pre_times[s] = 0;
            p->n_post_times[s] = 0;
        }
    }
}

/* Impacts of the spikes [lo, hi) of the flattened message list into per-delay-slot buffers. */
static uint64_t build_impacts(const ko_proj *p, const uint32_t *flat, size_t lo, size_t hi, ko_msg *slots,
                              uint64_t step) {
    uint64_t ev = 0;
    for (size_t i = lo; i < hi; ++i) {
        const uint32_t idx = flat[i];
        if (idx >= p->n_pre) continue;
        for (uint64_t j = p->row_ptr[idx]; j < p->row_ptr[idx + 1]; ++j) {
            const uint64_t s = p->row_syn[j];
            const uint64_t future_step = (uint64_t)p->delay[s] + step - 1;
            const ko_impact imp = {s, p->w[s], p->type[s], p->pre[s], p->post[s]};
            msg_push(&slots[future_step % p->max_delay], &imp);
            ev++;
        }
    }
    return ev;
}

/* The multi-threaded form of the spike loop below for plain Delta projections: contiguous chunks of the
 * spike list build impacts into thread buffers, merged into the queue slots in chunk order. */
static void calculate_projection_mt(ko_net *net, ko_proj *p, ko_spike_msg *msgs, uint32_t n_msgs, uint64_t step) {
    size_t total = 0;
    for (uint32_t m = 0; m < n_msgs; ++m) total += msgs[m].n;
    uint32_t *flat = (uint32_t *)xcalloc(total ? total : 1, 4);
    for (uint32_t m = 0, k = 0; m < n_msgs; ++m)
        for (uint32_t i = 0; i < msgs[m].n; ++i) flat[k++] = msgs[m].idx[i];
    const int T = net->threads;
    ko_msg *buf = (ko_msg *)xcalloc((size_t)T * p->max_delay, sizeof(ko_msg));
    uint64_t ev = 0;
#pragma omp parallel num_threads(T) reduction(+ : ev)
    {
        int th = 0, nth = 1;
#ifdef _OPENMP
        th = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        const size_t lo = total * (size_t)th / (size_t)nth, hi = total * (size_t)(th + 1) / (size_t)nth;
        ev += build_impacts(p, flat, lo, hi, buf + (size_t)th * p->max_delay, step);
    }
    for (uint32_t d = 0; d < p->max_delay; ++d)
        for (int th = 0; th < T; ++th) {
            ko_msg *b = &buf[(size_t)th * p->max_delay + d];
            for (size_t k = 0; k < b->n; ++k) msg_push(&p->queue[d], &b->v[k]);
            free(b->v);
        }
    free(buf);
    free(flat);
    net->events += ev;
}

static void send_due(ko_proj *p, uint64_t step);

static void calculate_projection(ko_net *net, ko_proj *p, ko_spike_msg *msgs, uint32_t n_msgs, uint64_t step) {
    if (p->kind == KO_PROJ_ADDITIVE) additive_init_projection(net, p, msgs, n_msgs);
    if (net->threads > 1 && p->kind == KO_PROJ_DELTA) {
        calculate_projection_mt(net, p, msgs, n_msgs, step);
        n_msgs = 0;  /* impacts already built */
    }
    for (uint32_t m = 0; m < n_msgs; ++m) {
        for (uint32_t i = 0; i < msgs[m].n; ++i) {
            uint32_t idx = msgs[m].idx[i];
            if (idx >= p->n_pre) continue;
            for (uint64_t j = p->row_ptr[idx]; j < p->row_ptr[idx + 1]; ++j) {
                uint64_t s = p->row_syn[j];
                if (p->kind == KO_PROJ_RESOURCE) p->rule[s].last_spike_step = step; /* init_synapse */
                uint64_t future_step = (uint64_t)p->delay[s] + step - 1;
                ko_impact imp = {s, p->w[s], p->type[s], p->pre[s], p->post[s]};
                msg_push(&p->queue[future_step % p->max_delay], &imp);
                net->events++;
            }
        }
    }
    if (p->kind == KO_PROJ_ADDITIVE) additive_modify_weights(p);
    send_due(p, step);
}

/* MultiThreadedCPUBackend::calculate_projections (cpu-multi-threaded-backend/impl/backend.cpp:250-295) over
 * calculate_projection_multithreaded_impl (delta_impl.h:117-175): only message m, its indexes counted
 * (convert_spikes, backend.cpp:235-247); every synapse whose presynaptic index occurs c times gives one impact
 * of weight weight * c, synapses in index order (that backend merges its parts in thread completion order,
 * which with exact sums and no BLOCKING changes no result). */
static void calculate_projection_first_message(ko_net *net, ko_proj *p, const ko_spike_msg *m, uint64_t step) {
    uint32_t *count = (uint32_t *)xcalloc(p->n_pre ? p->n_pre : 1, 4);
    for (uint32_t i = 0; i < m->n; ++i)
        if (m->idx[i] < p->n_pre) count[m->idx[i]]++;
    for (uint64_t s = 0; s < p->n; ++s) {
        const uint32_t c = count[p->pre[s]];
        if (!c) continue;
        if (p->kind == KO_PROJ_RESOURCE) p->rule[s].last_spike_step = step; /* delta_impl.h:137-140 */
        const uint64_t future_step = (uint64_t)p->delay[s] + step - 1;
        const ko_impact imp = {s, (wreal)(p->w[s] * (wreal)c), p->type[s], p->pre[s], p->post[s]};
        msg_push(&p->queue[future_step % p->max_delay], &imp);
        net->events++;
    }
    free(count);
    send_due(p, step);
}

/* The projection's message of this step from its delay queue, sent (cpu-library projections.h:76-84;
 * cpu-multi-threaded-backend/impl/backend.cpp:284-294). */
static void send_due(ko_proj *p, uint64_t step) {
    ko_msg *out = &p->queue[step % p->max_delay];
    p->has_outbox = 0;
    if (out->exists) {
        /* send and erase: swap the slot into the outbox */
        ko_msg tmp = p->outbox;
        p->outbox = *out;
        *out = tmp;
        out->n = 0;
        out->exists = 0;
        p->has_outbox = 1;
    }
}

/* ---- the step (backend.cpp:89-141) ----------------------------------------------------------- */
static void population_step(ko_net *net, uint32_t pop, uint64_t step) {
    ko_pop *P = &net->pops[pop];
    const int T = net->threads > 1 ? net->threads : 1;
    (void)T;
#pragma omp parallel for num_threads(T) schedule(static) if (T > 1)
    for (uint32_t i = 0; i < P->n; ++i) {
        if (P->lif) lif_pre(&P->nl[i]);
        else if (P->altai) {
            altai_pre(&P->na[i]);
            if (P->resource) { /* altai_impl.h:43-52 */
                P->na[i].dopamine_value = 0.0;
                P->rx[i].is_being_forced = 0;
            }
        }
        else if (P->resource) resource_pre(&P->nb[i], &P->rx[i]);
        else blifat_pre(&P->nb[i]);
    }
    /* impacts injected by a test, then impact messages in delivery order: projections that sent last step,
     * in load order */
    for (size_t k = 0; k < P->n_injected; ++k) {
        const ko_impact *imp = &P->injected[k];
        if (P->lif) lif_impact(&P->nl[imp->post], imp);
        else if (P->altai) altai_impact(&P->na[imp->post], imp);
        else blifat_impact(&P->nb[imp->post], imp);
    }
    P->n_injected = 0;
    for (uint32_t q = 0; q < net->n_projs; ++q) {
        ko_proj *p = &net->projs[q];
        if (p->post_pop != pop || !p->has_outbox) continue;
        int is_forcing = p->kind == KO_PROJ_DELTA; /* stdp_dispatcher.h:54-58 */
        for (size_t k = 0; k < p->outbox.n; ++k) {
            const ko_impact *imp = &p->outbox.v[k];
            if (P->lif) {
                lif_impact(&P->nl[imp->post], imp);
                continue;
            }
            if (P->altai) {
                altai_impact(&P->na[imp->post], imp);
                if (P->resource && imp->type == KNP_EXCITATORY) P->rx[imp->post].is_being_forced |= is_forcing;
                continue;
            }
            knp_blifat_neuron *n = &P->nb[imp->post];
            blifat_impact(n, imp);
            if (P->resource && imp->type == KNP_EXCITATORY) P->rx[imp->post].is_being_forced |= is_forcing;
        }
    }
    P->n_last = 0;
    if (T > 1) {
        /* post-impact states in parallel parts, the spike message in ascending order */
        uint8_t *fired = (uint8_t *)xcalloc(P->n ? P->n : 1, 1);
#pragma omp parallel for num_threads(T) schedule(static)
        for (uint32_t i = 0; i < P->n; ++i)
            fired[i] = (uint8_t)(P->lif ? lif_post(&P->nl[i]) : P->altai ? altai_post(&P->na[i]) : blifat_post(&P->nb[i]));
        for (uint32_t i = 0; i < P->n; ++i)
            if (fired[i]) P->last_spikes[P->n_last++] = i;
        free(fired);
    } else {
        for (uint32_t i = 0; i < P->n; ++i)
            if (P->lif ? lif_post(&P->nl[i]) : P->altai ? altai_post(&P->na[i]) : blifat_post(&P->nb[i]))
                P->last_spikes[P->n_last++] = i;
    }
    if (P->resource) train_population(net, pop, P->last_spikes, P->n_last, step);
    if (P->record && P->rec_steps < P->rec_cap) {
        uint32_t words = (P->n + 31) / 32;
        uint32_t *row = P->rec + P->rec_steps * words;
        memset(row, 0, (size_t)words * 4);
        for (uint32_t k = 0; k < P->n_last; ++k) row[P->last_spikes[k] / 32] |= 1u << (P->last_spikes[k] % 32);
        P->rec_steps++;
    }
}

static void generate_sources(ko_net *net, uint64_t step) {
    for (uint32_t s = 0; s < net->n_srcs; ++s) {
        ko_source *src = &net->srcs[s];
        uint32_t phase = (uint32_t)(step % src->steps_per_pattern);
        if (phase < src->lo || phase > src->hi) continue;
        uint32_t pattern = (uint32_t)((step / src->steps_per_pattern) % src->n_patterns);
        const uint32_t *thr = src->thr + (size_t)pattern * src->n_channels;
        uint32_t *idx = (uint32_t *)xcalloc(src->n_channels, 4);
        uint32_t n = 0;
        for (uint32_t c = 0; c < src->n_channels; ++c)
            if (ko_bernoulli_hash(src->seed, step, src->hash_offset + c) < thr[c]) idx[n++] = src->first + c;
        if (n) ko_push_spikes(net, src->proj, idx, n);
        free(idx);
    }
}

int ko_step(ko_net *net, uint64_t n_steps) {
    if (net->semantics == KNP_SEMANTICS_MULTI_THREADED && check_multi_threaded(net)) return -1;
    for (uint64_t it = 0; it < n_steps; ++it) {
        uint64_t step = net->step;
        generate_sources(net, step);
        for (uint32_t pop = 0; pop < net->n_pops; ++pop) population_step(net, pop, step);
        for (uint32_t q = 0; q < net->n_projs; ++q) {
            ko_proj *p = &net->projs[q];
            /* messages for this projection: channel messages first (routed at the start of the step),
             * then population spike messages in population load order (routed after the population loop) */
            uint32_t n_pop_msgs = 0;
            ko_spike_msg pm[64];
            for (uint32_t pop = 0; pop < net->n_pops && n_pop_msgs < 64; ++pop) {
                int subscribed = (p->pre_pop == pop);
                for (uint32_t k = 0; k < p->n_stdp; ++k) subscribed |= (p->stdp_pops[k] == pop);
                if (!subscribed || net->pops[pop].n_last == 0) continue;
                pm[n_pop_msgs].idx = net->pops[pop].last_spikes;
                pm[n_pop_msgs].n = net->pops[pop].n_last;
                pm[n_pop_msgs].send_time = step;
                pm[n_pop_msgs].from_pop = (int)pop;
                n_pop_msgs++;
            }
            uint32_t total = p->n_inbox + n_pop_msgs;
            ko_spike_msg *all = (ko_spike_msg *)xcalloc(total ? total : 1, sizeof(ko_spike_msg));
            for (uint32_t m = 0; m < p->n_inbox; ++m) all[m] = p->inbox[m];
            for (uint32_t m = 0; m < n_pop_msgs; ++m) all[p->n_inbox + m] = pm[m];
            if (net->semantics == KNP_SEMANTICS_MULTI_THREADED) {
                /* one routing carries the step's population messages and the channel messages sent before
                 * the step, the backend's endpoint first (message_bus_cpu_impl.cpp:41-88); the projection
                 * expands only the first (cpu-multi-threaded-backend/impl/backend.cpp:267) */
                const ko_spike_msg *first = n_pop_msgs ? &pm[0] : p->n_inbox ? &p->inbox[0] : NULL;
                if (first) calculate_projection_first_message(net, p, first, step);
                else send_due(p, step);
            } else {
                /* `all` holds copies of the message headers: additive STDPOnly empties only this view */
                calculate_projection(net, p, all, total, step);
            }
            for (uint32_t m = 0; m < p->n_inbox; ++m) free(p->inbox[m].idx);
            p->n_inbox = 0;
            free(all);
        }
        net->step++;
    }
    return 0;
}

/* ---- state export ---------------------------------------------------------------------------- */
uint64_t ko_get_step(ko_net *net) { return net->step; }
uint64_t ko_get_event_count(ko_net *net) { return net->events; }

int ko_read_potentials(ko_net *net, uint32_t pop, double *out, uint32_t n) {
    if (pop >= net->n_pops || n > net->pops[pop].n) return -1;
    const ko_pop *P = &net->pops[pop];
    for (uint32_t i = 0; i < n; ++i)
        out[i] = P->lif ? (double)P->nl[i].potential : P->altai ? (double)P->na[i].potential : P->nb[i].potential;
    return 0;
}

int ko_read_lif_neurons(ko_net *net, uint32_t pop, knp_lif_neuron *out, uint32_t n) {
    if (pop >= net->n_pops || n > net->pops[pop].n || !net->pops[pop].lif) return -1;
    const ko_lif *x = net->pops[pop].nl;
    for (uint32_t i = 0; i < n; ++i) {
        out[i].potential = (float)x[i].potential;
        out[i].potential_reset_value = (float)x[i].potential_reset_value;
        out[i].activation_threshold = (float)x[i].activation_threshold;
        out[i].leak_coefficient = (float)x[i].leak_coefficient;
        out[i].refract_counter = x[i].refract_counter;
        out[i].refract_period = x[i].refract_period;
    }
    return 0;
}

int ko_read_last_spikes(ko_net *net, uint32_t pop, uint32_t *idx, uint32_t cap, uint32_t *n) {
    if (pop >= net->n_pops) return -1;
    ko_pop *P = &net->pops[pop];
    uint32_t k = P->n_last < cap ? P->n_last : cap;
    memcpy(idx, P->last_spikes, (size_t)k * 4);
    *n = P->n_last;
    return 0;
}

int ko_read_altai_lif_neurons(ko_net *net, uint32_t pop, knp_altai_lif_neuron *out, uint32_t n) {
    if (pop >= net->n_pops || n > net->pops[pop].n || !net->pops[pop].altai) return -1;
    const ko_altai *x = net->pops[pop].na;
    for (uint32_t i = 0; i < n; ++i) {
        knp_altai_lif_neuron *o = &out[i];
        memset(o, 0, sizeof(*o));
        o->is_diff = (uint8_t)x[i].is_diff;
        o->is_reset = (uint8_t)x[i].is_reset;
        o->leak_rev = (uint8_t)x[i].leak_rev;
        o->saturate = (uint8_t)x[i].saturate;
        o->do_not_save = (uint8_t)x[i].do_not_save;
        o->potential = (float)x[i].potential;
        o->pre_impact_potential = (float)x[i].pre_impact_potential;
        o->activation_threshold = x[i].activation_threshold;
        o->negative_activation_threshold = x[i].negative_activation_threshold;
        o->potential_leak = x[i].potential_leak;
        o->potential_reset_value = x[i].potential_reset_value;
        o->dopamine_value = x[i].dopamine_value;
        o->additional_threshold = x[i].additional_threshold;
        o->activity_time = x[i].activity_time;
    }
    return 0;
}

int ko_read_resource_altai_lif_neurons(ko_net *net, uint32_t pop, knp_resource_altai_lif_neuron *out, uint32_t n) {
    if (pop >= net->n_pops || n > net->pops[pop].n || !net->pops[pop].altai || !net->pops[pop].resource) return -1;
    knp_altai_lif_neuron *base = (knp_altai_lif_neuron *)xcalloc(n ? n : 1, sizeof(knp_altai_lif_neuron));
    ko_read_altai_lif_neurons(net, pop, base, n);
    const ko_res_ext *rx = net->pops[pop].rx;
    for (uint32_t i = 0; i < n; ++i) {
        knp_resource_altai_lif_neuron *o = &out[i];
        const ko_res_ext *x = &rx[i];
        memset(o, 0, sizeof(*o));
        o->base = base[i];
        o->synapse_sum_threshold_coefficient = (float)x->synapse_sum_threshold_coefficient;
        o->dopamine_plasticity_time = x->dopamine_plasticity_time;
        o->free_synaptic_resource = (float)x->free_synaptic_resource;
        o->synaptic_resource_threshold = (float)x->synaptic_resource_threshold;
        o->resource_drain_coefficient = x->resource_drain_coefficient;
        o->stability = (float)x->stability;
        o->stability_change_parameter = (float)x->stability_change_parameter;
        o->stability_change_at_isi = (float)x->stability_change_at_isi;
        o->isi_max = x->isi_max;
        o->d_h = (float)x->d_h;
        o->isi_status = x->isi_status;
        o->last_step = x->last_step;
        o->last_spike_step = x->last_spike_step;
        o->first_isi_spike = x->first_isi_spike;
        o->is_being_forced = (uint8_t)x->is_being_forced;
    }
    free(base);
    return 0;
}

int ko_read_blifat_neurons(ko_net *net, uint32_t pop, knp_blifat_neuron *out, uint32_t n) {
    if (pop >= net->n_pops || n > net->pops[pop].n || net->pops[pop].lif || net->pops[pop].altai) return -1;
    memcpy(out, net->pops[pop].nb, (size_t)n * sizeof(knp_blifat_neuron));
    return 0;
}

int ko_read_resource_neurons(ko_net *net, uint32_t pop, knp_resource_blifat_neuron *out, uint32_t n) {
    if (pop >= net->n_pops || n > net->pops[pop].n || !net->pops[pop].resource || net->pops[pop].altai) return -1;
    ko_pop *P = &net->pops[pop];
    for (uint32_t i = 0; i < n; ++i) {
        out[i].base = P->nb[i];
        ko_res_ext *x = &P->rx[i];
        out[i].synapse_sum_threshold_coefficient = (float)x->synapse_sum_threshold_coefficient;
        out[i].dopamine_plasticity_time = x->dopamine_plasticity_time;
        out[i].free_synaptic_resource = (float)x->free_synaptic_resource;
        out[i].synaptic_resource_threshold = (float)x->synaptic_resource_threshold;
        out[i].resource_drain_coefficient = x->resource_drain_coefficient;
        out[i].stability = (float)x->stability;
        out[i].stability_change_parameter = (float)x->stability_change_parameter;
        out[i].stability_change_at_isi = (float)x->stability_change_at_isi;
        out[i].isi_max = x->isi_max;
        out[i].d_h = (float)x->d_h;
        out[i].isi_status = x->isi_status;
        out[i].last_step = x->last_step;
        out[i].last_spike_step = x->last_spike_step;
        out[i].first_isi_spike = x->first_isi_spike;
        out[i].is_being_forced = (uint8_t)x->is_being_forced;
    }
    return 0;
}

/* Weights as double (exact for the float policy). */
int ko_read_weights(ko_net *net, uint32_t proj, double *out, uint64_t n) {
    if (proj >= net->n_projs || n > net->projs[proj].n) return -1;
    for (uint64_t s = 0; s < n; ++s) out[s] = (double)net->projs[proj].w[s];
    return 0;
}

int ko_read_resources(ko_net *net, uint32_t proj, double *out, uint64_t n) {
    if (proj >= net->n_projs || n > net->projs[proj].n || !net->projs[proj].rule) return -1;
    for (uint64_t s = 0; s < n; ++s) out[s] = (double)net->projs[proj].rule[s].synaptic_resource;
    return 0;
}

int ko_read_resource_rules(ko_net *net, uint32_t proj, knp_resource_rule *out, uint64_t n) {
    if (proj >= net->n_projs || n > net->projs[proj].n || !net->projs[proj].rule) return -1;
    for (uint64_t s = 0; s < n; ++s) {
        ko_rule *r = &net->projs[proj].rule[s];
        out[s].synaptic_resource = (float)r->synaptic_resource;
        out[s].w_min = (float)r->w_min;
        out[s].w_max = (float)r->w_max;
        out[s].d_u = (float)r->d_u;
        out[s].dopamine_plasticity_period = r->dopamine_plasticity_period;
        out[s].last_spike_step = r->last_spike_step;
        out[s].had_hebbian_update = (uint8_t)r->had_hebbian_update;
        out[s].has_contributed = (uint8_t)r->has_contributed;
    }
    return 0;
}

int ko_wreal_bytes(void) { return (int)sizeof(wreal); }
