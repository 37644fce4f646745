// step_kernels.h - HIP kernels (gfx950, wave64) of one KNP network step.
//
// One step of the reference single-threaded backend (knp/backends/cpu/cpu-single-threaded-backend/
// impl/backend.cpp:89-141) becomes, on the device:
//   k_phase1   every population's fused BLIFAT update (pre-impact, impact, post-impact, threshold) for all
//              populations at once, spike compaction (wave ballot + one atomic per wave), spike bitmaps,
//              training candidates, plus the device Bernoulli channel sources;
//   k_train    synaptic-resource STDP of one population (only for its candidate neurons);
//   k_additive_append / k_additive_modify   additive STDP histories and weight updates;
//   k_scatter  every projection's spike -> impact expansion: one wave per spiking presynaptic row, impacts
//              added into per-population delay rings of exact int64 fixed-point accumulators.
// Integer accumulation makes the result independent of the order in which impacts arrive, so the
// spike trains do not depend on thread scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

#include "device_tables.h"
#include "float_repeat.h"

namespace knp_mi355x {

// Pointers read from the network tables are generic (flat) to the compiler; flat accesses also count on the
// LDS counter, so every LDS wait and workgroup barrier would wait for them. Accesses through gp() are
// global-memory instructions.
#define KNP_GLOBAL __attribute__((address_space(1)))
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ KNP_GLOBAL T *gp(T *p) {
    return (KNP_GLOBAL T *)p;
}

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    // splitmix64 finalizer (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint32_t bernoulli_hash(uint64_t seed, uint64_t step, uint32_t ch) {
    uint64_t a = mix64(seed + step * 0x9E3779B97F4A7C15ULL);
    uint64_t b = mix64(a ^ ((uint64_t)ch * 0xD6E8FEB86659FD93ULL + 0x632BE59BD9B4E019ULL));
    return (uint32_t)(b >> 32);
}

// A kernel's step argument: an absolute step, or kStepDev + d for "the device's current step + d" (the form the
// launches of a captured graph take, so that one graph serves every replay; k_step_advance moves the step).
constexpr uint64_t kStepDev = 1ull << 63;
__device__ __forceinline__ uint64_t resolve_step(const NetDev *__restrict__ net, uint64_t t) {
    return t >= kStepDev ? (uint64_t)*gp(net->step_ptr) + (t - kStepDev) : t;
}

__device__ __forceinline__ float pf(const ParamF &p, uint32_t i) { return p.a ? gp(p.a)[i] : p.v; }
__device__ __forceinline__ uint32_t pu(const ParamU &p, uint32_t i) { return p.a ? gp(p.a)[i] : p.v; }

// n_time_steps_since_last_firing_ kept in 32 bits: 0xFFFFFFFF stands for SIZE_MAX (wraps to 0 on the next
// increment, blifat_impl.h:34 with the default of blifat.h:53); counts saturate at 0xFFFFFFFE.
__device__ __forceinline__ uint32_t ts_inc(uint32_t x) {
    return x == 0xFFFFFFFFu ? 0u : (x >= 0xFFFFFFFDu ? 0xFFFFFFFEu : x + 1u);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ int64_t pi64(const ParamI64 &p, uint32_t i) { return p.a ? gp(p.a)[i] : p.v; }

// s + e = x + y exactly (Knuth's TwoSum; the library is built without FMA contraction or reassociation).
__device__ __forceinline__ void two_sum(float x, float y, float &s, float &e) {
    s = x + y;
    const float bb = s - x;
    e = (x - (s - bb)) + (y - bb);
}

// A threshold with a dynamic part, hi + lo + x1 + x2 + x2_lo, as a float-float pair s + l (about 48 bits).
__device__ __forceinline__ void threshold_ff(float hi, float lo, float x1, float x2, float x2_lo, float &s, float &l) {
    float s1, e1, e2;
    two_sum(hi, x1, s1, e1);
    two_sum(s1, x2, s, e2);
    l = lo + (e1 + e2) + x2_lo;
}

// d as a float pair hi + lo (about 48 bits).
__device__ __forceinline__ void split_double(double d, float &hi, float &lo) {
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

// Does an exact accumulator sum a (on the 2^-frac_bits grid, |a| < 2^60) reach the threshold s + l?
// scale = 2^frac_bits. The threshold times scale is split into integer parts and fractions, so the comparison
// is made in int64 except for the fractions' sum.
__device__ __forceinline__ bool fixed_ge_ff(int64_t a, float s, float l, float scale) {
    const float H = s * scale, L = l * scale;
    if (!(fabsf(H) < 2.0e18f)) return H < 0.f;  // beyond any sum the accumulator holds
    const float Hf = floorf(H), Lf = floorf(L);
    const int64_t d = a - (int64_t)Hf - (int64_t)Lf;
    if (d >= 2) return true;
    if (d < 0) return false;
    return (float)d >= (H - Hf) + (L - Lf);  // the fractions sum to [0, 2)
}

// Does a float potential V reach the threshold s + l?
__device__ __forceinline__ bool float_ge_ff(float V, float s, float l) {
    float r, er, d, ed;
    two_sum(V, -s, r, er);  // r + er = V - s
    two_sum(r, -l, d, ed);  // d + ed = r - l
    return d >= -(ed + er);
}

__device__ __forceinline__ float fixed_to_float(int64_t a, float inv_scale) {
    return __ll2float_rn((long long)a) * inv_scale;
}

// Stores and loads of data that another workgroup reads or wrote within the same persistent launch:
// device-scope relaxed atomics, i.e. written through and read past this XCD's L2 (the L2s of the XCDs are
// not coherent with each other inside a kernel). COHERENT = false: plain accesses (kernel boundaries order
// them).
template <bool COHERENT>
__device__ __forceinline__ void st_u32(uint32_t *p, uint32_t v) {
    if (COHERENT) __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *gp(p) = v;
}
template <bool COHERENT>
__device__ __forceinline__ uint32_t ld_u32(const uint32_t *p) {
    if (COHERENT) return __hip_atomic_load(gp(const_cast<uint32_t *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *gp(p);
}

// Word 0 of a population or source segment: the count (at most 256) in the low kStampShift bits; entries
// from word 1. In a persistent launch, where the readers of a segment run in other workgroups while it is
// written, the writer stores word 0 only once the entries have reached memory, with the step's stamp above
// the count (the step modulo 2^20 - 1, plus 1: never 0, the stamp of a plain count, and different for steps
// t and t - 2, which share the segment), so a reader that sees the stamp of step t also reads complete
// entries. Channel segments hold the host's plain counts.
constexpr uint32_t kStampShift = 12, kCountMask = (1u << kStampShift) - 1u;
__device__ __forceinline__ uint32_t list_stamp(uint64_t t) { return (uint32_t)(t % 0xFFFFFull) + 1u; }

// Bounded waits of a persistent launch: after kSpinLimit polls (NetDev::spin_limit), or once another
// workgroup has given up, a wait ends with a failure, raises the abort flag (word kBarAbort of grid_bar) and
// the launch stops (the host reports the failure), so no wait lasts forever.
constexpr uint32_t kSpinLimit = 1u << 20;
constexpr uint32_t kBarAbort = 0, kBarWords = 32;  // words of grid_bar: the abort flag

// Ordered compaction of a workgroup's 256 candidates into its own segment (count in word 0, indexes from
// word 1): ascending order, no global atomics. NT = threads of the workgroup (threads from 256 on carry no
// candidate).
// Returns (in thread 0) the rows published: the segment's count times the projections that read it.
// vp_pre: the segment's 4 virtual positions if the caller holds them (else they are loaded).
template <uint32_t NT = kBlock, bool COHERENT = false>
__device__ __forceinline__ uint32_t block_segment_write(const NetDev *__restrict__ net, uint64_t t, bool pred,
                                                        uint32_t value, uint32_t *segment, uint32_t seg,
                                                        const uint32_t *vp_pre = nullptr) {
    __shared__ uint32_t wcnt[NT / 64];
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t off = 0, total = 0;
    for (uint32_t k = 0; k < NT / 64; ++k) {
        if (k < w) off += wcnt[k];
        total += wcnt[k];
    }
    if (pred) st_u32<COHERENT>(&segment[1 + off + __popcll(mask & ((1ull << lane) - 1ull))], value);
    if (COHERENT) {
        __builtin_amdgcn_s_waitcnt(0);  // the entries have reached memory before the stamped count
        __syncthreads();
    }
    uint32_t rows = 0;
    if (threadIdx.x < 4) {
        const size_t par = (size_t)(t & 1);
        if (threadIdx.x == 0) {
            st_u32<COHERENT>(&segment[0], COHERENT ? total | list_stamp(t) << kStampShift : total);
            st_u32<COHERENT>(&net->seg_count[par * net->n_seg + seg], total);
        }
        uint32_t vps[4];
        for (uint32_t k = 0; k < 4; ++k) vps[k] = vp_pre ? vp_pre[k] : gp(net->seg_vpos)[4 * seg + k];
        const uint32_t vp = vps[threadIdx.x & 3];
        if (vp != 0xFFFFFFFFu) st_u32<COHERENT>(&net->vcount[par * net->n_vseg + vp], total);
        if (threadIdx.x == 0)
            for (uint32_t k = 0; k < 4; ++k) rows += vps[k] != 0xFFFFFFFFu ? total : 0u;
    }
    __syncthreads();
    return rows;
}

// Append the workgroup's values with pred set to a list: one atomic on the list's counter per workgroup (a
// counter shared by every workgroup of a large population is otherwise hit by each of its waves). The
// entries of a workgroup are contiguous; the order of workgroups is whatever the atomics give. Every thread
// of the workgroup calls it.
template <uint32_t NT>
__device__ __forceinline__ void block_append(bool pred, uint32_t value, uint32_t *counter, uint32_t *list) {
    __shared__ uint32_t wcnt[NT / 64];
    __shared__ uint32_t base;
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t off = 0, total = 0;
    for (uint32_t k = 0; k < NT / 64; ++k) {
        if (k < w) off += wcnt[k];
        total += wcnt[k];
    }
    if (threadIdx.x == 0 && total) base = atomicAdd((uint32_t *)gp(counter), total);
    __syncthreads();
    if (pred) gp(list)[base + off + __popcll(mask & ((1ull << lane) - 1ull))] = value;
    __syncthreads();  // wcnt and base are free for the next call
}

// Write the wave's 64-neuron spike mask into the bitmap words (and the record), bit i = neuron i.
__device__ __forceinline__ void wave_bitmap(uint64_t mask, uint32_t first_neuron, const PopDev &P, uint64_t t) {
    uint32_t lane = lane_id();
    uint32_t w0 = first_neuron >> 5;
    if (lane < 2 && w0 + lane < P.words) {
        uint32_t word = lane == 0 ? (uint32_t)mask : (uint32_t)(mask >> 32);
        gp(P.bitmap)[w0 + lane] = word;
        if (P.rec && t >= P.rec_t0 && t - P.rec_t0 < P.rec_cap) gp(P.rec)[(t - P.rec_t0) * P.words + w0 + lane] = word;
    }
}

// Decode a BLOCKING key (see k_scatter) into the blocking period the reference assigns:
// total_blocking_period_ = static_cast<int64_t>(impact_value_) (blifat_impl.h:86-88).
__device__ __forceinline__ int64_t blocking_value(const NetDev *net, uint64_t key) {
    uint64_t k = key - 1;
    uint32_t bid = (uint32_t)(k >> 57);
    uint32_t pre = (uint32_t)((k >> 17) & 0x1FFFFFFFull);
    uint32_t pos = (uint32_t)(k & 0x1FFFFull);
    const ProjDev &Q = net->projs[net->blk_proj[bid]];
    float w = Q.w_csr[Q.row_ptr[pre] + pos];
    return (int64_t)w;
}

// ---- phase 1: populations + sources ----------------------------------------------------------------
// NT = threads of the workgroup (threads from 256 on have no neuron). FUSED: the exact sum of the step's
// impacts is passed in (a_in) by the tile kernel instead of being read from the ring.
// The table entry of population p, or the caller's LDS copy of it (COPY: known at compile time, so that
// the compiler addresses the copy as LDS).
template <bool COPY, typename T>
__device__ __forceinline__ const T &entry_of(const T *table, uint32_t i, const T *copy) {
    if constexpr (COPY) return *copy;
    else return table[i];
}

// Synaptic-resource neurons after their update: store the step's dopamine and forcing for training and
// return whether the neuron is a training candidate: it spiked, has dopamine, or is due for resource
// renormalization (training/stdp.h:152-186; the test only depends on state that training leaves unchanged
// for other neurons).
__device__ __forceinline__ bool resource_candidate(const PopDev &P, uint32_t i, uint64_t t, bool spike,
                                                   float dop_now, int64_t dop_fix, uint32_t forced) {
    const ResDev &R = P.res;
    gp(R.dop_now)[i] = dop_now;  // dopamine_value_ is reset every step for resource neurons
    gp(R.dop_fix)[i] = dop_fix;
    gp(R.forced_now)[i] = (uint8_t)(forced != 0);
    bool renorm = false;
    const float fr = gp(R.free_res)[i];
    if (fabsf(fr) >= pf(R.res_thr, i)) {
        const uint8_t st = gp(R.isi_status)[i];
        const bool in_isi = (t - gp(R.last_step)[i] <= (uint64_t)pu(R.isi_max, i)) && st != 0;
        renorm = !in_isi;
    }
    return spike || dop_fix != 0 || renorm;
}

// The BLIFAT update of neuron i of population P at step t (blifat_impl.h:31-168): returns the spike and sets
// cand (synaptic-resource neurons: a training candidate). FUSED: a_in is the exact sum of the step's impacts
// (else it is read from the ring and the slot cleared). KEEP (default BLIFAT only): the potential and the
// last spike step go to the caller's registers (V_keep, ls_keep), which store them once a run of steps is
// done; the default BLIFAT step reads neither.
template <uint32_t KIND, bool FUSED, bool KEEP = false>
__device__ __forceinline__ bool population_update(const NetDev *__restrict__ net, const PopDev &P, uint32_t i,
                                                  bool active, uint64_t t, int64_t a_in, bool &cand,
                                                  float *V_keep = nullptr, uint32_t *ls_keep = nullptr) {
    static_assert(!KEEP || KIND == POP_LEAN, "only the default BLIFAT step keeps its state in registers");
    const uint32_t slot = (uint32_t)(t % P.D);
    const size_t ai = (size_t)slot * P.n + i;
    bool spike = false;
    cand = false;
    if (active) {
        int64_t a = a_in;
        if (!FUSED && P.acc_v) {
            a = gp(P.acc_v)[ai];
            gp(P.acc_v)[ai] = 0;
        }
        if (P.ff) a += __float2ll_rn(gp(P.ff)[(size_t)(t % (2ull * P.ff_T)) * P.n + i] * P.scale);
        if (KIND == POP_LEAN) {
            // default BLIFAT (blifat.h defaults): V_pre = 0, no decays, no refractory period, no blocking.
            const bool can = !(P.lean_ts_max && t == P.t_first);  // step 0 rule (wrap of SIZE_MAX)
            float V = fixed_to_float(a, P.inv_scale);
            spike = can && a >= pi64(P.thr_fix, i);
            if (spike) {
                V = pf(P.V_reset, i);
                if (KEEP) *ls_keep = (uint32_t)t;
                else gp(P.last_spike)[i] = (uint32_t)t;
            }
            const float vmin = pf(P.V_min, i);
            if (V < vmin) V = vmin;
            if (KEEP) *V_keep = V;
            else gp(P.V)[i] = V;
        } else if (P.model == MODEL_ALTAI) {
            // AltAI-LIF (altai_impl.h:31-150). pre: the potential is rounded (or reset when do_not_save_);
            // impacts; post: signed leak, activity-time countdown (0 means unlimited activity), spike when
            // the potential reaches the threshold while active, then the positive and negative resets.
            // Synaptic-resource AltAI-LIF neurons (altai_impl.h:43-52, 85-94) also reset their dopamine
            // every step and note forcing, and their threshold adds the trained additional threshold.
            const uint32_t fl = pu(P.alt_flags, i);
            float V = (fl & kAltDoNotSave) ? pf(P.V_reset, i) : roundf(gp(P.V)[i]);
            gp(P.Vpre)[i] = V;
            V += fixed_to_float(a, P.inv_scale);
            float dop_now = 0.f;
            int64_t dop_fix = 0;
            if (P.acc_d) {
                dop_fix = gp(P.acc_d)[ai];
                gp(P.acc_d)[ai] = 0;
                dop_now = fixed_to_float(dop_fix, P.inv_scale);
                if (KIND != POP_RESOURCE && dop_fix) gp(P.dop)[i] += dop_now;
            }
            uint32_t forced = 0;
            if (KIND == POP_RESOURCE && P.acc_f) {
                forced = gp(P.acc_f)[ai];
                gp(P.acc_f)[ai] = 0;
            }
            const float leak = pf(P.V_decay, i);
            V += ((fl & kAltLeakRev) && V < 0.f) ? -leak : leak;
            int64_t at = gp(P.tbp)[i];
            if (at > 0) --at;
            else if (at < 0) ++at;
            if (at == 0) at = INT64_MAX;
            bool was_reset = false;
            // plain: the smallest float not below activation + additional threshold (exact for V >= T);
            // resource: activation_threshold_ + the trained threshold, as a float-float sum
            float ts = pf(P.act, i), tl = 0.f;
            if (KIND == POP_RESOURCE) {
                float add, add_lo;
                split_double(gp(P.res.add_thr)[i], add, add_lo);
                two_sum(ts, add, ts, tl);
                tl += add_lo;
            }
            if (KIND == POP_RESOURCE ? float_ge_ff(V, ts, tl) : V >= ts) {
                spike = at > 0;
                if (fl & kAltDiff) {
                    // V - T rounded once: (V - ts) as r + er exactly, then r + (er - tl)
                    float r, er;
                    two_sum(V, -ts, r, er);
                    V = r + (er - tl);
                }
                if (fl & kAltReset) {
                    V = pf(P.V_reset, i);
                    was_reset = true;
                }
            }
            const float neg = pf(P.V_min, i);  // -negative_activation_threshold_
            if (V <= neg && !was_reset) {
                if (fl & kAltSaturate) V = neg;
                else if (fl & kAltReset) V = -pf(P.V_reset, i);
                else if (fl & kAltDiff) V -= neg;
            }
            gp(P.V)[i] = V;
            gp(P.tbp)[i] = at;
            if (KIND == POP_RESOURCE) cand = resource_candidate(P, i, t, spike, dop_now, dop_fix, forced);
        } else if (P.model == MODEL_LIF) {
            // LIFNeuron (lif_impl.h:31-84): while the refractory counter is non-zero the potential neither
            // leaks nor takes impacts and the counter counts down; else leak, add the step's impacts and fire
            // strictly above the threshold.
            uint32_t rc = gp(P.n_ts)[i];
            float V = gp(P.V)[i];
            if (rc == 0) {
                V *= pf(P.V_decay, i);
                V += fixed_to_float(a, P.inv_scale);
                if (V > pf(P.act, i)) {
                    V = pf(P.V_reset, i);
                    rc = pu(P.refr, i);
                    spike = true;
                }
            } else {
                --rc;
            }
            gp(P.V)[i] = V;
            gp(P.n_ts)[i] = rc;
        } else {
            // pre-impact (blifat_impl.h:31-65)
            uint32_t n_ts = ts_inc(gp(P.n_ts)[i]);
            float dyn = gp(P.dyn)[i] * pf(P.thr_decay, i);
            float trace = gp(P.trace)[i] * pf(P.trace_decay, i);
            float g = gp(P.g)[i] * pf(P.g_decay, i);
            float V = gp(P.V)[i] * pf(P.V_decay, i);
            uint32_t bphase = gp(P.bphase)[i];
            if (bphase == 1) {
                if (KIND != POP_RESOURCE) bphase = 0;  // resource neurons keep the phase (blifat_impl.h:62)
                V += pf(P.reflexive, i);
            }
            const float V_pre = V;
            gp(P.Vpre)[i] = V_pre;
            // impacts (blifat_impl.h:68-107): exact sums of this step's impacts
            int64_t tbp = P.track_tbp ? gp(P.tbp)[i] : INT64_MAX;
            if (P.acc_g) {
                int64_t ag = gp(P.acc_g)[ai];
                gp(P.acc_g)[ai] = 0;
                if (ag) g += fixed_to_float(ag, P.inv_scale);
            }
            float dop_now = 0.f;
            int64_t dop_fix = 0;
            if (P.acc_d) {
                dop_fix = gp(P.acc_d)[ai];
                gp(P.acc_d)[ai] = 0;
                dop_now = fixed_to_float(dop_fix, P.inv_scale);
            }
            uint32_t forced = 0;
            if (P.acc_f) {
                forced = gp(P.acc_f)[ai];
                gp(P.acc_f)[ai] = 0;
            }
            if (P.acc_b) {
                uint64_t key = gp(P.acc_b)[ai];
                gp(P.acc_b)[ai] = 0;
                if (key) tbp = blocking_value(net, key);
            }
            V = (V_pre == 0.f) ? fixed_to_float(a, P.inv_scale) : V_pre + fixed_to_float(a, P.inv_scale);
            bool exact = (V_pre == 0.f);
            // post-impact (blifat_impl.h:119-168)
            if (tbp <= 0) {
                V = V_pre;
                exact = false;
                a = 0;
            }
            if (tbp < 0) {
                ++tbp;
                if (tbp == 0) tbp = INT64_MAX;
            } else if (tbp > 0) {
                tbp -= 1;
            }
            if (g < 1.f) {
                if (g != 0.f) exact = false;
                V -= (V - pf(P.E_inh, i)) * g;
            } else {
                V = pf(P.E_inh, i);
                exact = false;
            }
            // threshold (blifat_impl.h:113-115): the static part alone when the dynamic threshold is 0 (exact
            // tests), else static + dynamic (+ trained additional threshold) in float-float
            bool above;
            if (KIND != POP_RESOURCE && dyn == 0.f) {
                above = exact ? a >= pi64(P.thr_fix, i) : V >= pf(P.thr_up, i);
            } else {
                float ts, tl, add = 0.f, add_lo = 0.f;
                if (KIND == POP_RESOURCE) split_double(gp(P.res.add_thr)[i], add, add_lo);
                threshold_ff(pf(P.thr_hi, i), pf(P.thr_lo, i), dyn, add, add_lo, ts, tl);
                above = exact ? fixed_ge_ff(a, ts, tl, P.scale) : float_ge_ff(V, ts, tl);
            }
            spike = (n_ts > pu(P.refr, i)) && above;
            if (spike) {
                dyn += pf(P.thr_inc, i);
                trace += pf(P.trace_inc, i);
                V = pf(P.V_reset, i);
                bphase = pu(P.bperiod, i);
                n_ts = 0;
            }
            const float vmin = pf(P.V_min, i);
            if (V < vmin) V = vmin;
            gp(P.V)[i] = V;
            gp(P.n_ts)[i] = n_ts;
            gp(P.dyn)[i] = dyn;
            gp(P.trace)[i] = trace;
            gp(P.g)[i] = g;
            gp(P.bphase)[i] = bphase;
            if (P.track_tbp) gp(P.tbp)[i] = tbp;
            if (KIND == POP_RESOURCE) {
                cand = resource_candidate(P, i, t, spike, dop_now, dop_fix, forced);
            } else if (P.acc_d) {
                gp(P.dop)[i] += dop_now;  // plain BLIFAT keeps summing dopamine (never reset for it)
            }
        }
    }
    return spike;
}

template <uint32_t KIND, uint32_t NT = kBlock, bool FUSED = false, bool COHERENT = false, bool COPY = false>
__device__ __forceinline__ uint32_t population_step(const NetDev *__restrict__ net, uint32_t p, uint32_t local_block,
                                                    uint64_t t, Ctl *ctl, uint32_t *seg_count, int64_t a_in = 0,
                                                    const uint32_t *vp_pre = nullptr, const PopDev *P_copy = nullptr) {
    const PopDev &P = entry_of<COPY>(net->pops, p, P_copy);
    const uint32_t i = local_block * kBlock + threadIdx.x;
    const bool active = threadIdx.x < (uint32_t)kBlock && i < P.n;
    bool cand;
    const bool spike = population_update<KIND, FUSED>(net, P, i, active, t, a_in, cand);
    // spike compaction and bitmaps (whole wave participates)
    const uint32_t rows = block_segment_write<NT, COHERENT>(
        net, t, spike, i, P.spike_list + (size_t)(t & 1) * net->list_pstride + (size_t)local_block * kSegStride,
        P.seg_base + local_block, vp_pre);
    const uint64_t mask = __ballot(spike);
    if (threadIdx.x < (uint32_t)kBlock) wave_bitmap(mask, local_block * kBlock + (threadIdx.x & ~63u), P, t);
    if (KIND == POP_RESOURCE) block_append<NT>(cand, i, &ctl->cand_count[p], P.res.cand);
    return rows;
}

// Does channel c of source S fire at step t (counter-based Bernoulli draw against the channel's threshold
// for the current pattern)? thr_pre: the threshold if the caller holds it (single-pattern sources), else
// UINT32_MAX and it is loaded.
// The same with the step's phase (t mod steps_per_pattern) and pattern ((t / steps_per_pattern) mod
// n_patterns) given by a caller that keeps them from step to step.
__device__ __forceinline__ bool source_spike_at(const SourceDev &S, uint32_t c, bool active, uint64_t t,
                                                uint32_t phase, uint32_t pattern, uint32_t thr_pre) {
    if (!active || c >= S.n_ch || phase < S.lo || phase > S.hi) return false;
    const uint32_t thr = thr_pre != 0xFFFFFFFFu ? thr_pre : gp(S.thr)[(size_t)pattern * S.n_ch + c];
    return bernoulli_hash(S.seed, t, S.hash_offset + c) < thr;
}

__device__ __forceinline__ bool source_spike(const SourceDev &S, uint32_t c, bool active, uint64_t t,
                                             uint32_t thr_pre) {
    // one pattern of one step (a plain Bernoulli channel group) needs no 64-bit division
    const uint64_t q = S.steps_per_pattern == 1 ? t : t / S.steps_per_pattern;
    const uint32_t phase = (uint32_t)(t - q * S.steps_per_pattern);
    const uint32_t pattern = S.n_patterns == 1 ? 0u : (uint32_t)(q % S.n_patterns);
    return source_spike_at(S, c, active, t, phase, pattern, thr_pre);
}

// thr_pre as in source_spike; vp_pre as in block_segment_write.
template <uint32_t NT = kBlock, bool COHERENT = false, bool COPY = false>
__device__ __forceinline__ uint32_t source_step(const NetDev *__restrict__ net, uint32_t s, uint32_t local_block,
                                                uint64_t t, uint32_t *seg_count, uint32_t thr_pre = 0xFFFFFFFFu,
                                                const uint32_t *vp_pre = nullptr, const SourceDev *S_copy = nullptr) {
    const SourceDev &S = entry_of<COPY>(net->srcs, s, S_copy);
    const uint32_t c = local_block * kBlock + threadIdx.x;
    const bool spike = source_spike(S, c, threadIdx.x < (uint32_t)kBlock, t, thr_pre);
    return block_segment_write<NT, COHERENT>(
        net, t, spike, S.first + c, S.list + (size_t)(t & 1) * net->list_pstride + (size_t)local_block * kSegStride,
        S.seg_base + local_block, vp_pre);
}

__global__ __launch_bounds__(kBlock) void k_phase1(const NetDev *__restrict__ net, uint64_t t) {
    t = resolve_step(net, t);
    Ctl *ctl = &net->ctl[t & 1];
    uint32_t *seg_count = net->seg_count + (size_t)(t & 1) * net->n_seg;
    if (blockIdx.x == 0) {
        KNP_GLOBAL uint32_t *nxt = gp((uint32_t *)&net->ctl[(t + 1) & 1]);
        for (uint32_t k = threadIdx.x; k < sizeof(Ctl) / 4; k += blockDim.x) nxt[k] = 0;
        // next step's counts: host-pushed segments must start empty (every other segment is rewritten)
        KNP_GLOBAL uint32_t *nc = gp(net->seg_count) + (size_t)((t + 1) & 1) * net->n_seg;
        KNP_GLOBAL uint32_t *nv = gp(net->vcount) + (size_t)((t + 1) & 1) * net->n_vseg;
        for (uint32_t k = threadIdx.x; k < net->n_projs; k += blockDim.x) {
            const uint32_t ps = net->projs[k].push_seg;
            nc[ps] = 0;
            gp(net->projs[k].in_list)[(size_t)((t + 1) & 1) * net->list_pstride] = 0;  // segment count word
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t vp = gp(net->seg_vpos)[4 * ps + j];
                if (vp != 0xFFFFFFFFu) nv[vp] = 0;
            }
        }
    }
    const uint32_t b = blockIdx.x;
    uint32_t task = 0;
    for (uint32_t k = 0; k < net->n_tasks; ++k)
        if (b >= net->tasks[k].block_begin) task = k;
    const TaskDev &T = net->tasks[task];
    const uint32_t lb = b - T.block_begin;
    if (T.kind == 1) {
        source_step(net, T.index, lb, t, seg_count);
        return;
    }
    switch (net->pops[T.index].kind) {
        case POP_LEAN: population_step<POP_LEAN>(net, T.index, lb, t, ctl, seg_count); break;
        case POP_FULL: population_step<POP_FULL>(net, T.index, lb, t, ctl, seg_count); break;
        default: population_step<POP_RESOURCE>(net, T.index, lb, t, ctl, seg_count); break;
    }
}

// ---- scatter: spike rows -> impacts in the delay rings (delta_impl.h:54-114) -------------------------
// Every workgroup first scans the spike counts of all virtual segments into LDS (one parallel load per
// thread), so the spiking rows of the step are numbered 0..total-1 without any global synchronization.
// One workgroup then expands one row at a time: each thread loads up to 4 synapses before issuing its
// atomics, so a 1,000-synapse row costs one memory round trip.
// BLOCKING keys order impacts as the reference delivers them (the "last" impact wins, blifat_impl.h:86-88):
// projection load order, then creation step (later = shorter delay), then message, then presynaptic index,
// then position in the row.  key = 1 + [blocking projection:6][64-delay:6][message rank:5][pre:29][pos:17].
// Exact accumulation: a 64-bit integer atomic per impact (non-returning). Splitting it into two 32-bit
// atomics on the same 8 bytes measured slower: device-wide atomics are executed past the L2, one transaction
// each, so the count of atomics, not their width, sets the cost.
__device__ __forceinline__ void atomic_add_i64(KNP_GLOBAL int64_t *p, long long v) {
    atomicAdd((unsigned long long *)p, (unsigned long long)v);
}

// Where one row's impacts go, copied out of the tables once per row (global-memory pointers: no flat atomics, and
// no reload of the tables after every atomic).
struct ImpactDst {
    KNP_GLOBAL int64_t *acc_v, *acc_g, *acc_d;
    KNP_GLOBAL uint32_t *acc_f;
    KNP_GLOBAL uint64_t *acc_b;
    uint32_t D, n, slot0;
    float scale;
    bool forcing;
    uint64_t key_hi;  // BLOCKING key bits of the projection and message rank
};

__device__ __forceinline__ ImpactDst impact_dst(const PopDev &P, const ProjDev &Q, uint32_t rank, uint64_t t) {
    ImpactDst x;
    x.acc_v = gp(P.acc_v);
    x.acc_g = gp(P.acc_g);
    x.acc_d = gp(P.acc_d);
    x.acc_f = gp(P.acc_f);
    x.acc_b = gp(P.acc_b);
    x.D = P.D;
    x.n = P.n;
    x.slot0 = (uint32_t)(t % P.D);
    x.scale = P.scale;
    x.forcing = Q.forcing && P.acc_f;
    x.key_hi = ((uint64_t)Q.blocking_id << 57) | ((uint64_t)rank << 46);
    return x;
}

__device__ __forceinline__ void impact(const ImpactDst &X, uint32_t pre, uint64_t pos, uint32_t pt, uint32_t d,
                                       float w) {
    const uint32_t post = pt & 0x1FFFFFFFu;
    const uint32_t type = pt >> 29;
    uint32_t slot = X.slot0 + d;
    if (slot >= X.D) slot -= X.D;
    const size_t ai = (size_t)slot * X.n + post;
    const long long v = __float2ll_rn(w * X.scale);
    switch (type) {
        case 0:  // EXCITATORY
            atomic_add_i64(X.acc_v + ai, v);
            if (X.forcing) atomicOr((uint32_t *)(X.acc_f + ai), 1u);
            break;
        case 1:  // INHIBITORY_CURRENT
            atomic_add_i64(X.acc_v + ai, -v);
            break;
        case 2:  // INHIBITORY_CONDUCTANCE
            atomic_add_i64(X.acc_g + ai, v);
            break;
        case 3:  // DOPAMINE
            atomic_add_i64(X.acc_d + ai, v);
            break;
        default: {  // BLOCKING
            const uint64_t key = 1ull + (X.key_hi | ((uint64_t)(64u - d) << 51) | ((uint64_t)pre << 17) | pos);
            atomicMax((unsigned long long *)(X.acc_b + ai), (unsigned long long)key);
            break;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(const NetDev *__restrict__ net, uint64_t t) {
    t = resolve_step(net, t);
    extern __shared__ uint32_t pfx[];  // [n_vseg + 1]
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t nv = net->n_vseg_atomic;
    const KNP_GLOBAL uint32_t *vc = gp(net->vcount) + (size_t)(t & 1) * net->n_vseg;
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    for (uint32_t v = threadIdx.x; v < nv; v += kBlock) pfx[v] = vc[v];
    const KNP_GLOBAL uint32_t *vkey = gp(net->vkey);
    if (vkey) {
        // the multi-threaded backend's semantics: each projection expands only its first message of the step,
        // the nonempty one of smallest key (every workgroup finds the same)
        __shared__ uint32_t first_key[kMaxProjs];
        for (uint32_t q = threadIdx.x; q < net->n_projs; q += kBlock) first_key[q] = 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < nv; v += kBlock)
            if (pfx[v]) atomicMin(&first_key[net->rows[net->vrs[v]].proj], vkey[v]);
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < nv; v += kBlock)
            if (vkey[v] != first_key[net->rows[net->vrs[v]].proj]) pfx[v] = 0;
    }
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nv; base += kBlock) {
        const uint32_t v = base + threadIdx.x;
        const uint32_t c = v < nv ? pfx[v] : 0u;
        uint32_t x = c;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t woff = 0, btot = 0;
        for (uint32_t k = 0; k < kBlock / 64; ++k) {
            if (k < w) woff += wsum[k];
            btot += wsum[k];
        }
        if (v < nv) pfx[v] = carry + woff + x - c;
        carry += btot;
        __syncthreads();
    }
    if (threadIdx.x == 0) pfx[nv] = carry;
    __syncthreads();
    const uint32_t total = pfx[nv];
    unsigned long long my_ev = 0, my_rows = 0;
    for (uint32_t row = blockIdx.x; row < total; row += gridDim.x) {
        uint32_t lo = 0, hi = nv;  // largest v with pfx[v] <= row
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pfx[mid] <= row) lo = mid;
            else hi = mid;
        }
        const uint32_t seg = gp(net->vseg)[lo];
        const KNP_GLOBAL RowSource *R = gp(net->rows) + gp(net->vrs)[lo];
        uint32_t pre = gp(net->seg_list[seg])[(size_t)(t & 1) * net->list_pstride + 1 + row - pfx[lo]];
        // pushed entries under the multi-threaded semantics: the index's count in the message, which multiplies
        // the weight in float (delta_impl.h:144)
        float mult = 1.f;
        if (vkey && vkey[lo] == kMsgKeyPush) {
            mult = (float)(pre >> kMsgCountShift);
            pre &= kMsgIndexMask;
        }
        const ProjDev &Q = net->projs[R->proj];
        if (pre >= Q.n_pre) continue;
        const uint64_t b = gp(Q.row_ptr)[pre], e = gp(Q.row_ptr)[pre + 1];
        my_ev += e - b;
        my_rows += 1;
        const ImpactDst X = impact_dst(net->pops[Q.post_pop], Q, R->rank, t);
        const KNP_GLOBAL uint32_t *post_type = gp(Q.post_type);
        const KNP_GLOBAL uint8_t *delay = gp(Q.delay);
        const KNP_GLOBAL float *w_csr = gp(Q.w_csr), *W = gp((const float *)Q.W);
        const KNP_GLOBAL uint32_t *csr_syn = gp(Q.csr_syn);
        const bool delta = Q.kind == PROJ_DELTA;
        if (Q.kind == PROJ_RESOURCE && threadIdx.x == 0) gp(Q.row_spike)[pre] = t;  // init_synapse of the row
        for (uint64_t base = b; base < e; base += 4 * kBlock) {
            uint32_t pt[4], d[4];
            float wv[4];
            bool ok[4];
            for (int j = 0; j < 4; ++j) {
                const uint64_t k = base + threadIdx.x + (uint64_t)j * kBlock;
                ok[j] = k < e;
                if (ok[j]) {
                    pt[j] = post_type[k];
                    d[j] = delay[k];
                    wv[j] = delta ? w_csr[k] : W[csr_syn[k]];
                }
            }
            for (int j = 0; j < 4; ++j)
                if (ok[j]) impact(X, pre, base + threadIdx.x + (uint64_t)j * kBlock - b, pt[j], d[j], wv[j] * mult);
        }
    }
    if (threadIdx.x == 0) {
        gp(net->part)[2 * blockIdx.x] += my_ev;
        gp(net->part)[2 * blockIdx.x + 1] += my_rows;
    }
}

// ---- tile-owner scatter: exact sums per 256-neuron tile in LDS, no global atomics ----------------------
// Device-wide atomics are executed past the L2 at one memory transaction each, which bounds the atomic
// scatter at roughly 10^10 impacts/s. A tiled population instead gives each of its 256-neuron tiles to one
// workgroup: the workgroup reads the part of every spiking row that lands in its tile (rows are sorted by
// postsynaptic neuron and carry a tile directory), adds the impacts into a [D][256] int64 LDS ring with LDS
// atomics, and finally adds the non-zero sums into the population's ring in HBM; it is the only writer of
// those cells, so plain loads and stores suffice. The sums are the same exact integers as the atomic scatter's.
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *wsum, uint32_t *total) {
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    uint32_t inc = x;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if ((int)lane >= off) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (uint32_t k = 0; k < NT / 64; ++k) {
        const uint32_t v = wsum[k];
        if (k < w) woff += v;
        tot += v;
    }
    __syncthreads();  // wsum may be reused by the next scan
    *total = tot;
    return woff + inc - x;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Per-projection pointers the tile scatter needs, staged in LDS.
struct TileProj {
    const uint64_t *syn8;
    const uint64_t *tile_dir;
    const RowMeta *row_meta;
    const uint64_t *row_dir;
    const uint64_t *row_ptr;
    uint32_t stride, n_pre;
};

// A thread's row part whose records are added later: [b, b + len) of projection q.
struct TilePart {
    uint32_t b, len, q;
};

// Dynamic LDS of the tile kernels: the ring sums [tile_max_d][256] (int64), then for the tile's segments
// their spike lists [tile_max_list] (pointers, parity 0), row-count prefix [tile_max_list + 1] and
// projection [tile_max_list].
__host__ __device__ inline size_t tile_lds_bytes(uint32_t max_d, uint32_t max_list) {
    return (size_t)max_d * kBlock * 8 + (size_t)max_list * 8 + ((size_t)max_list + 1) * 4 + max_list;
}

// State of one tile workgroup: its tile, its population's fields (in registers: stores through the ring
// would otherwise force reloads of the table) and its LDS tables.
struct TileCtx {
    TileDev td;
    uint32_t D, n_pop, tile_lo, kind;
    int64_t *acc_v;
    unsigned long long *ring;
    const uint32_t **vlist;
    uint32_t *vpfx;  // prefix of the segments' counts this step
    uint8_t *vproj;
    uint32_t nv;
};

// Shared-memory scratch of the tile kernels. One pass of the gather takes kGatherRows rows (one per thread);
// the segment entries stashed by the count stage (kStash per segment) fill the row tables of the first pass.
constexpr uint32_t kGatherRows = kTileThreads, kStash = 3;
constexpr uint32_t kRowLoad = 0x80;  // rq flag: the row's entry is loaded in stage 2 (rval = its segment)
struct TileShared {
    uint32_t rval[kGatherRows];  // stage 2: presynaptic index (or segment) of each row; 3b: CSR position
    uint32_t rpfx[kGatherRows];  // 3b: exclusive prefix of the long parts' lengths
    uint8_t rq[kGatherRows];     // projection of each row (| kRowLoad)
    uint32_t wsum[kTileThreads / 64];
    unsigned long long cnt[4];
    TileProj tp[kMaxProjs];
};

// The tile of workgroup b. Workgroups are dispatched to the 8 XCDs in turn (workgroup b on XCD b mod 8), so
// the tiles are dealt out in 8 contiguous ranges, one per XCD: neighbouring tiles read neighbouring
// directory entries and row parts (rows are sorted by postsynaptic neuron), which then share the XCD's L2.
__device__ __forceinline__ uint32_t tile_of_block(uint32_t b, uint32_t n_tiles) {
    const uint32_t q = n_tiles / 8, r = n_tiles % 8, x = b % 8;
    return x * q + min(x, r) + b / 8;
}

// Loads the tile's description, the projection pointers and the list of segments that can reach the tile.
__device__ __forceinline__ TileCtx tile_setup(const NetDev *__restrict__ net, unsigned long long *dyn,
                                              TileShared &sh) {
    TileCtx c;
    {
        const KNP_GLOBAL TileDev *tdp = gp(net->tiles) + tile_of_block(blockIdx.x, gridDim.x);
        c.td.pop = tdp->pop;
        c.td.tile = tdp->tile;
        c.td.vbeg = tdp->vbeg;
        c.td.vend = tdp->vend;
    }
    const PopDev &P = net->pops[c.td.pop];
    c.D = P.D;
    c.n_pop = P.n;
    c.kind = P.kind;
    c.acc_v = P.acc_v;
    c.tile_lo = c.td.tile * kBlock;
    const uint32_t L = net->tile_max_list;
    c.ring = dyn;
    c.vlist = reinterpret_cast<const uint32_t **>(dyn + (size_t)net->tile_max_d * kBlock);
    c.vpfx = reinterpret_cast<uint32_t *>(c.vlist + L);
    c.vproj = reinterpret_cast<uint8_t *>(c.vpfx + L + 1);
    c.nv = c.td.vend - c.td.vbeg;
    const uint32_t tid = threadIdx.x;
    if (tid < 4) sh.cnt[tid] = 0;
    if (tid < net->n_projs) {
        const ProjDev &Q = net->projs[tid];
        sh.tp[tid] = TileProj{Q.syn8, Q.tile_dir, Q.row_meta, Q.row_dir, Q.row_ptr, Q.tile_stride, Q.n_pre};
    }
    const TileSeg *ts = net->tile_seg + c.td.vbeg;
    for (uint32_t v = tid; v < c.nv; v += kTileThreads) {
        const KNP_GLOBAL TileSeg *e = gp(ts) + v;
        c.vlist[v] = e->list;
        c.vproj[v] = (uint8_t)e->proj;
    }
    return c;
}

// Exact sum of one packed synapse record of a step whose ring slot is slot0 into the LDS ring (false: a
// neuron of another tile). The record holds the signed fixed-point impact as m * 2^s.
__device__ __forceinline__ bool add_record(const TileCtx &c, uint32_t slot0, uint64_t rec) {
    const uint32_t j = (uint32_t)(rec >> 41) - c.tile_lo;
    if (j >= (uint32_t)kBlock) return false;  // a short row read whole: neuron of another tile
    uint32_t slot = slot0 + ((uint32_t)(rec >> 32) & 0xFFu);
    if (slot >= c.D) slot -= c.D;
    const int32_t m = (int32_t)((uint32_t)rec << 7) >> 7;
    const unsigned long long v = (unsigned long long)(long long)m << (((uint32_t)rec >> 25) & 63u);
    atomicAdd(&c.ring[slot * kBlock + j], v);
    return true;
}

// The thread adds the records [b, b + n) of syn8, 8 loads in flight at a time. PAIRS: 16 bytes (an aligned
// pair of records) per load, for long parts (fewer memory requests); the array holds one record past its
// end, so the pair of its last record is readable. Else 8 bytes per load (fewer instructions for the short
// parts of sparse networks).
template <bool PAIRS>
__device__ __forceinline__ void thread_add_records(const TileCtx &c, uint32_t slot0, const KNP_GLOBAL uint64_t *syn8,
                                                   uint32_t b, uint32_t n, uint32_t &my_ev) {
    if constexpr (!PAIRS) {
        const KNP_GLOBAL uint64_t *src = syn8 + b;
        for (uint32_t k0 = 0; k0 < n; k0 += 8) {
            uint64_t rec[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u)
                if (k0 + u < n) rec[u] = src[k0 + u];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u)
                if (k0 + u < n && add_record(c, slot0, rec[u])) ++my_ev;
        }
        return;
    }
    const uint32_t end = b + n;
    for (uint32_t p0 = n ? (b & ~1u) : end; p0 < end; p0 += 8) {
        v4u pr[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
            if (p0 + 2 * u < end) pr[u] = *reinterpret_cast<const KNP_GLOBAL v4u *>(syn8 + p0 + 2 * u);
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                const uint32_t k = p0 + 2 * u + h;
                if (k >= b && k < end) {
                    const uint64_t rec = h ? ((uint64_t)pr[u].w << 32 | pr[u].z) : ((uint64_t)pr[u].y << 32 | pr[u].x);
                    if (add_record(c, slot0, rec)) ++my_ev;
                }
            }
    }
}

// Stage 3: every thread's row part (records [b, b + len) of projection q) into the ring. Short parts (the
// common case) by their own thread; long parts spread over the workgroup, one thread per record, through
// the row tables. Ends with a workgroup barrier (after which the row tables are free again). STEPWISE (the
// one-launch-per-step kernel): records loaded in 16-byte pairs, parts up to kTileShortPartStepwise short.
template <bool STEPWISE>
__device__ __forceinline__ void tile_add_parts(const TileCtx &c, TileShared &sh, uint32_t slot0, uint32_t b,
                                               uint32_t len, uint32_t q, uint32_t &my_ev) {
    constexpr uint32_t kShort = STEPWISE ? kTileShortPartStepwise : kTileShortPart;
    const uint32_t tid = threadIdx.x;
    if (len <= kShort) thread_add_records<STEPWISE>(c, slot0, gp(sh.tp[q].syn8), b, len, my_ev);
    const uint32_t llen = len > kShort ? len : 0u;
    if (__syncthreads_or(llen != 0)) {
        uint32_t n_syn;
        const uint32_t ex = block_excl_scan<kTileThreads>(llen, sh.wsum, &n_syn);
        sh.rval[tid] = b;
        sh.rpfx[tid] = ex;
        sh.rq[tid] = (uint8_t)q;
        __syncthreads();
        for (uint32_t s = tid; s < n_syn; s += kTileThreads) {
            uint32_t lo = 0, hi = kTileThreads;  // largest r with rpfx[r] <= s: a row whose part holds s
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sh.rpfx[mid] <= s) lo = mid;
                else hi = mid;
            }
            if (add_record(c, slot0, gp(sh.tp[sh.rq[lo]].syn8)[sh.rval[lo] + (s - sh.rpfx[lo])])) ++my_ev;
        }
        __syncthreads();
    }
}

// Where row pre's part in the tile lies: records [b, b + len) of the projection, the first `urgent` of
// them with delay 1. No part: len = 0. One 8-byte directory entry per part (two dependent loads with a
// ranged directory).
__device__ __forceinline__ void part_of(const TileProj &Q, uint32_t pre, uint32_t tile, uint32_t &b, uint32_t &len,
                                        uint32_t &urgent) {
    uint64_t e = 0;
    uint32_t base = 0;
    if (Q.tile_dir) {
        e = gp(Q.tile_dir)[(size_t)pre * Q.stride + tile];
    } else if (Q.row_meta) {
        const KNP_GLOBAL RowMeta *m = gp(Q.row_meta) + pre;
        const uint32_t m_begin = m->begin, m_first = m->first_tile, m_n = m->n_tiles, m_dir = m->dir;
        const uint32_t k = tile - m_first;  // wraps above n_tiles for tiles before the row
        if (k < m_n) e = gp(Q.row_dir)[m_dir + k];
        base = m_begin;
    } else {
        const uint64_t r0 = gp(Q.row_ptr)[pre], r1 = gp(Q.row_ptr)[pre + 1];
        b = (uint32_t)r0;
        len = urgent = (uint32_t)(r1 - r0);
        return;
    }
    b = base + (uint32_t)e;
    len = (uint32_t)(e >> 32) & ((1u << kPartLenBits) - 1u);
    const uint32_t u = (uint32_t)(e >> 56);
    urgent = u == kUrgentAll ? len : u;
}

// Stages 1-3 of the tile scatter of step t: numbers the spiking rows that may reach the tile (prefix over
// the counts of its segments), finds every row's part in the tile, and adds the part's impacts into the LDS
// ring with LDS atomics. Ends with a workgroup barrier. Adds the impacts (ev) and row parts (parts).
// COHERENT (persistent launch): the population and source segments of step t are written by other
// workgroups during the launch, and each is read once its stamp says step t is complete; channel segments
// are read only when `channels` (the launch's first step: the host pushes channel messages before the
// launch). Returns false if a wait ended without the stamp (the launch is being aborted).
// SPLIT: of the first pass's row parts only the delay-1 records are added (the ring slot of step t + 1 is
// then complete); the rest of each thread's part is left in `part`, for tile_add_parts after phase 1.
template <bool COHERENT, bool SPLIT = false>
__device__ __forceinline__ bool tile_gather(const NetDev *__restrict__ net, const TileCtx &c, TileShared &sh,
                                            uint64_t t, uint32_t &my_ev, uint32_t &my_parts, bool channels = true,
                                            TilePart *part = nullptr, uint32_t slot_t = 0xFFFFFFFFu) {
    if (SPLIT) *part = TilePart{0u, 0u, 0u};
    const uint32_t tid = threadIdx.x;
    const size_t par_off = (size_t)(t & 1) * net->list_pstride;
    // the pool as a buffer: one 16-byte load gives a segment's count and first kStash entries (device-scope
    // in a persistent launch)
    const __amdgpu_buffer_rsrc_t pool = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t *>(net->list_pool), 0, (int)(2u * net->list_pstride * 4u), 0x00020000);
    constexpr int kAux = COHERENT ? 16 : 0;
    const uint32_t stamp = list_stamp(t);
    bool ok = true;
    uint32_t total = 0;
    // 1. counts and stashed entries of every segment; the rows of the first pass into the row tables (the
    // row's presynaptic index and projection, or kRowLoad and its segment when the entry is not stashed)
    for (uint32_t base = 0; base < c.nv; base += kTileThreads) {
        const uint32_t v = base + tid;
        uint32_t cv = 0, w[kStash];
        const bool chan = v < c.nv && (c.vproj[v] & kTileSegChannel) != 0;
        const bool stamped = COHERENT && !chan;
        if (v < c.nv && (!COHERENT || !chan || channels)) {
            const uint32_t off = (uint32_t)((const char *)(c.vlist[v] + par_off) - (const char *)net->list_pool);
            v4u h = __builtin_amdgcn_raw_buffer_load_b128(pool, off, 0, kAux);
            if (stamped) {
                uint32_t spins = 0;
                while ((h.x >> kStampShift) != stamp) {  // the writer has not published step t yet
                    __builtin_amdgcn_s_sleep(16);  // polls are spaced out: frequent ones slow the stores
                    if ((++spins & 63u) == 0 &&
                        (spins > net->spin_limit || ld_u32<true>(net->grid_bar + kBarAbort) != 0)) {
                        st_u32<true>(net->grid_bar + kBarAbort, 1u);
                        ok = false;
                        break;
                    }
                    h = __builtin_amdgcn_raw_buffer_load_b128(pool, off, 0, kAux);
                }
            }
            cv = chan ? h.x : h.x & kCountMask;
            w[0] = h.y, w[1] = h.z, w[2] = h.w;
        }
        if (!ok) cv = 0;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kTileThreads>(cv, sh.wsum, &tot);
        if (v < c.nv) {
            const uint32_t r0 = total + ex, q = c.vproj[v] & (kTileSegChannel - 1u);
            c.vpfx[v] = r0;
            const uint32_t n = r0 < kGatherRows ? min(cv, kGatherRows - r0) : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kStash; ++k)
                if (k < n) {
                    sh.rval[r0 + k] = w[k];
                    sh.rq[r0 + k] = (uint8_t)q;
                }
            for (uint32_t k = kStash; k < n; ++k) {
                sh.rval[r0 + k] = v;
                sh.rq[r0 + k] = (uint8_t)(q | kRowLoad);
            }
        }
        total += tot;
    }
    if (COHERENT) {
        if (!__syncthreads_and(ok)) return false;
    } else {
        __syncthreads();
    }
    const uint32_t slot0 = slot_t != 0xFFFFFFFFu ? slot_t : (uint32_t)(t % c.D);  // ring slot of step t
    for (uint32_t rb = 0; rb < total; rb += kGatherRows) {
        // 2. one thread per row: its presynaptic index and projection, and where its part for this tile lies
        const uint32_t row = rb + tid;
        uint32_t b = 0, len = 0, q = 0, pre = 0xFFFFFFFFu, urgent = 0;
        if (row < total) {
            uint32_t lo = 0xFFFFFFFFu;  // the row's segment when its entry is loaded here
            if (row < kGatherRows) {
                const uint32_t rq = sh.rq[row];
                q = rq & (kTileSegChannel - 1u);
                if (rq & kRowLoad) lo = sh.rval[row];
                else pre = sh.rval[row];
            } else {
                uint32_t hi = c.nv;  // largest v with vpfx[v] <= row
                lo = 0;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (c.vpfx[mid] <= row) lo = mid;
                    else hi = mid;
                }
                q = c.vproj[lo] & (kTileSegChannel - 1u);
            }
            if (lo != 0xFFFFFFFFu) pre = ld_u32<COHERENT>(c.vlist[lo] + par_off + 1 + (row - c.vpfx[lo]));
            if (pre < sh.tp[q].n_pre) part_of(sh.tp[q], pre, c.td.tile, b, len, urgent);
        }
        my_parts += len ? 1u : 0u;
        if (SPLIT && rb == 0) {
            // the part's delay-1 records now, the rest after phase 1 of t + 1
            thread_add_records<false>(c, slot0, gp(sh.tp[q].syn8), b, urgent, my_ev);
            *part = TilePart{b + urgent, len - urgent, q};
            __syncthreads();  // the sums are complete before phase 1 reads them
            continue;
        }
        tile_add_parts<!COHERENT>(c, sh, slot0, b, len, q, my_ev);
    }
    return true;
}

// The channel counts of step parity t1 cleared by one thread (k_small_run).
__device__ __forceinline__ void clear_push_counts_one(const NetDev *__restrict__ net, uint64_t t1) {
    uint32_t *nc = net->seg_count + (size_t)(t1 & 1) * net->n_seg;
    uint32_t *nvc = net->vcount + (size_t)(t1 & 1) * net->n_vseg;
    for (uint32_t k = 0; k < net->n_projs; ++k) {
        const uint32_t ps = net->projs[k].push_seg;
        nc[ps] = 0u;
        net->projs[k].in_list[(size_t)(t1 & 1) * net->list_pstride] = 0u;
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t vp = gp(net->seg_vpos)[4 * ps + j];
            if (vp != 0xFFFFFFFFu) nvc[vp] = 0u;
        }
    }
}

// Channel segments of step t + 1 start empty (the host writes the counts of pushed messages before that
// step); run by one workgroup.
template <bool COHERENT>
__device__ __forceinline__ void clear_push_counts(const NetDev *__restrict__ net, uint64_t t1) {
    uint32_t *nc = net->seg_count + (size_t)(t1 & 1) * net->n_seg;
    uint32_t *nvc = net->vcount + (size_t)(t1 & 1) * net->n_vseg;
    for (uint32_t k = threadIdx.x; k < net->n_projs; k += kTileThreads) {
        const uint32_t ps = net->projs[k].push_seg;
        st_u32<COHERENT>(&nc[ps], 0u);
        st_u32<COHERENT>(net->projs[k].in_list + (size_t)(t1 & 1) * net->list_pstride, 0u);  // segment count word
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t vp = gp(net->seg_vpos)[4 * ps + j];
            if (vp != 0xFFFFFFFFu) st_u32<COHERENT>(&nvc[vp], 0u);
        }
    }
}

// Phase 1 of step t1 for the tile's neurons (a = exact sum of the impacts arriving at t1 for this thread's
// neuron) and the Bernoulli sources of t1 whose 256-channel blocks fall to this workgroup (block b is run by
// tile b mod n_tiles).
template <bool COHERENT>
__device__ __forceinline__ uint32_t tile_phase1(const NetDev *__restrict__ net, const TileCtx &c, uint64_t t1,
                                                int64_t a) {
    Ctl *ctl = &net->ctl[t1 & 1];
    uint32_t *seg_count = net->seg_count + (size_t)(t1 & 1) * net->n_seg;
    uint32_t rows;
    if (c.kind == POP_LEAN)
        rows = population_step<POP_LEAN, kTileThreads, true, COHERENT>(net, c.td.pop, c.td.tile, t1, ctl, seg_count, a);
    else
        rows = population_step<POP_FULL, kTileThreads, true, COHERENT>(net, c.td.pop, c.td.tile, t1, ctl, seg_count, a);
    uint32_t b0 = 0;
    for (uint32_t s = 0; s < net->n_srcs; ++s) {
        const uint32_t nb = max(1u, (net->srcs[s].n_ch + kBlock - 1) / kBlock);
        for (uint32_t b = blockIdx.x; b < b0 + nb; b += net->n_tiles)
            if (b >= b0) rows += source_step<kTileThreads, COHERENT>(net, s, b - b0, t1, nullptr);
        b0 += nb;  // the next source's blocks continue the round robin
    }
    return rows;  // rows published for step t1 (thread 0)
}

// What phase 1 of a persistent workgroup writes to, fixed for the launch, kept in LDS so that a step's
// phase 1 loads nothing but its own state: the virtual positions of its population segment and of its
// source blocks (block b of the sources is run by tile b mod n_tiles), copies of their table entries (LDS
// loads cannot alias the step's global stores) and every thread's channel threshold for single-pattern
// sources. The host uses the persistent launch only when no tile runs more than kCachedSrcBlocks blocks.
constexpr uint32_t kCachedSrcBlocks = 8;
struct PhaseCache {
    uint32_t pop_vp[4];
    uint32_t n_src;
    uint32_t src[kCachedSrcBlocks], blk[kCachedSrcBlocks];
    uint32_t vp[kCachedSrcBlocks][4];
    PopDev pop;
    SourceDev srcd[kCachedSrcBlocks];
    uint32_t thr[kCachedSrcBlocks][kBlock];  // UINT32_MAX: load the threshold (several patterns)
    // the table fields the publish reads (LDS reads instead of dependent scalar loads from the table)
    uint32_t *seg_count, *vcount;
    uint32_t n_seg, n_vseg, hold_block;
    uint64_t hold_until;
};

// Word-wise copy of a table entry into LDS by the workgroup (the caller synchronizes).
template <typename T>
__device__ __forceinline__ void lds_copy(T *dst, const T *src) {
    static_assert(sizeof(T) % 4 == 0, "word copy");
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    const KNP_GLOBAL uint32_t *g = gp(reinterpret_cast<const uint32_t *>(src));
    for (uint32_t k = threadIdx.x; k < sizeof(T) / 4; k += blockDim.x) d[k] = g[k];
}

__device__ __forceinline__ void phase_cache(const NetDev *__restrict__ net, const TileCtx &c, PhaseCache &pc) {
    const uint32_t pseg = net->pops[c.td.pop].seg_base + c.td.tile;
    if (threadIdx.x < 4) pc.pop_vp[threadIdx.x] = gp(net->seg_vpos)[4 * pseg + threadIdx.x];
    lds_copy(&pc.pop, &net->pops[c.td.pop]);
    uint32_t n = 0, b0 = 0;
    for (uint32_t s = 0; s < net->n_srcs; ++s) {
        const SourceDev &S = net->srcs[s];
        const uint32_t nb = max(1u, (S.n_ch + kBlock - 1) / kBlock);
        for (uint32_t b = blockIdx.x; b < b0 + nb; b += net->n_tiles) {
            if (b < b0) continue;
            const uint32_t j = n++;
            if (j >= kCachedSrcBlocks) continue;  // not reached: the host checks the bound
            if (threadIdx.x == 0) {
                pc.src[j] = s;
                pc.blk[j] = b - b0;
            }
            if (threadIdx.x < 4) pc.vp[j][threadIdx.x] = gp(net->seg_vpos)[4 * (S.seg_base + b - b0) + threadIdx.x];
            lds_copy(&pc.srcd[j], &net->srcs[s]);
            const uint32_t ch = (b - b0) * kBlock + threadIdx.x;
            if (threadIdx.x < (uint32_t)kBlock)
                pc.thr[j][threadIdx.x] = (S.n_patterns == 1 && ch < S.n_ch) ? gp(S.thr)[ch] : 0xFFFFFFFFu;
        }
        b0 += nb;
    }
    if (threadIdx.x == 0) {
        pc.n_src = n < kCachedSrcBlocks ? n : kCachedSrcBlocks;
        pc.seg_count = net->seg_count;
        pc.vcount = net->vcount;
        pc.n_seg = net->n_seg;
        pc.n_vseg = net->n_vseg;
        pc.hold_block = net->hold_block;
        pc.hold_until = net->hold_until;
    }
    __syncthreads();
}

// Population kind of every tile of a persistent launch, or kMixedKinds (the kind is read per tile; a larger
// kernel).
constexpr uint32_t kMixedKinds = 0xFF;

// Compaction of the two 256-thread halves of a 512-thread workgroup into two segments at once: threads
// 0-255 into seg_a, threads 256-511 into seg_b (null: that half has no candidates), each in ascending order,
// with one pair of barriers. Returns the rows published (threads 0 and 256).
template <bool COHERENT>
__device__ __forceinline__ uint32_t dual_segment_write(const PhaseCache &pc, uint64_t t, bool pred,
                                                       uint32_t value, uint32_t *seg_a, uint32_t sa,
                                                       const uint32_t *vp_a, uint32_t *seg_b, uint32_t sb,
                                                       const uint32_t *vp_b) {
    __shared__ uint32_t wcnt[kTileThreads / 64];
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    const bool upper = w >= 4;
    const uint32_t w0 = upper ? 4u : 0u;
    uint32_t off = 0, total = 0;
    for (uint32_t k = w0; k < w0 + 4; ++k) {
        if (k < w) off += wcnt[k];
        total += wcnt[k];
    }
    uint32_t *seg = upper ? seg_b : seg_a;
    if (pred) st_u32<COHERENT>(&seg[1 + off + __popcll(mask & ((1ull << lane) - 1ull))], value);
    if (COHERENT) {
        __builtin_amdgcn_s_waitcnt(0);  // the entries have reached memory before the stamped counts
        __syncthreads();
    }
    uint32_t rows = 0;
    const uint32_t h = threadIdx.x & (kBlock - 1);
    if (h < 4 && seg != nullptr) {
        const uint32_t sg = upper ? sb : sa;
        const uint32_t *vp = upper ? vp_b : vp_a;
        const size_t par = (size_t)(t & 1);
        if (h == 0) {
            if (!COHERENT || blockIdx.x != pc.hold_block || t >= pc.hold_until)  // (a test may have one workgroup withhold)
                st_u32<COHERENT>(&seg[0], COHERENT ? total | list_stamp(t) << kStampShift : total);
            st_u32<COHERENT>(&pc.seg_count[par * pc.n_seg + sg], total);
            for (uint32_t k = 0; k < 4; ++k) rows += vp[k] != 0xFFFFFFFFu ? total : 0u;
        }
        if (vp[h] != 0xFFFFFFFFu) st_u32<COHERENT>(&pc.vcount[par * pc.n_vseg + vp[h]], total);
    }
    // wcnt is next written at the following call, after the caller's barriers in between
    return rows;
}

// Phase 1 of step t1 for a persistent workgroup with both halves busy: the tile's neurons on threads 0-255
// and its first source block on threads 256-511, compacted together; further source blocks follow.
// KIND POP_LEAN: the potential and last spike step of the thread's neuron stay in V_keep / ls_keep.
template <uint32_t KIND>
__device__ __forceinline__ uint32_t tile_phase1_dual(const NetDev *__restrict__ net, const TileCtx &c,
                                                     const PhaseCache &pc, uint64_t t1, int64_t a, float &V_keep,
                                                     uint32_t &ls_keep) {
    const uint32_t tid = threadIdx.x;
    const size_t par_off = (size_t)(t1 & 1) * net->list_pstride;
    const uint32_t n_src = pc.n_src;
    bool pred = false;
    uint32_t value = 0;
    if (tid < (uint32_t)kBlock) {
        const uint32_t i = c.tile_lo + tid;
        bool cand;
        if (KIND == POP_LEAN)
            pred = population_update<POP_LEAN, true, true>(net, pc.pop, i, i < c.n_pop, t1, a, cand, &V_keep, &ls_keep);
        else if (KIND == kMixedKinds && c.kind == POP_LEAN)
            pred = population_update<POP_LEAN, true>(net, pc.pop, i, i < c.n_pop, t1, a, cand);
        else
            pred = population_update<POP_FULL, true>(net, pc.pop, i, i < c.n_pop, t1, a, cand);
        value = i;
    } else if (n_src > 0) {
        const uint32_t ch = pc.blk[0] * kBlock + (tid - kBlock);
        pred = source_spike(pc.srcd[0], ch, true, t1, pc.thr[0][tid - kBlock]);
        value = pc.srcd[0].first + ch;
    }
    uint32_t *seg_a = pc.pop.spike_list + par_off + (size_t)c.td.tile * kSegStride;
    uint32_t *seg_b = n_src > 0 ? pc.srcd[0].list + par_off + (size_t)pc.blk[0] * kSegStride : nullptr;
    const uint32_t sb = n_src > 0 ? pc.srcd[0].seg_base + pc.blk[0] : 0u;
    uint32_t rows = dual_segment_write<true>(pc, t1, pred, value, seg_a, pc.pop.seg_base + c.td.tile, pc.pop_vp,
                                             seg_b, sb, pc.vp[0]);
    const uint64_t mask = __ballot(pred);
    if (tid < (uint32_t)kBlock) wave_bitmap(mask, c.tile_lo + (tid & ~63u), pc.pop, t1);
    for (uint32_t j = 1; j < n_src; ++j) {
        const uint32_t thr = tid < (uint32_t)kBlock ? pc.thr[j][tid] : 0xFFFFFFFFu;
        rows += source_step<kTileThreads, true, true>(net, pc.src[j], pc.blk[j], t1, nullptr, thr, pc.vp[j], &pc.srcd[j]);
    }
    return rows;
}

// Counter totals of the workgroup: impacts, row parts, ring cells (block 0 also adds the rows of the step).
__device__ __forceinline__ void tile_count(TileShared &sh, uint32_t ev, uint32_t parts, uint32_t cells) {
    ev = wave_sum_u32(ev);
    parts = wave_sum_u32(parts);
    cells = wave_sum_u32(cells);
    if (lane_id() == 0) {
        atomicAdd(&sh.cnt[0], (unsigned long long)ev);
        atomicAdd(&sh.cnt[1], (unsigned long long)parts);
        atomicAdd(&sh.cnt[2], (unsigned long long)cells);
    }
}

// Rows expanded at step t by the tiled path (sum of the counts of every tiled segment), added by block 0.
template <bool COHERENT>
__device__ __forceinline__ void tile_count_rows(const NetDev *__restrict__ net, TileShared &sh, uint64_t t) {
    const uint32_t *vc = net->vcount + (size_t)(t & 1) * net->n_vseg;
    uint32_t r = 0;
    for (uint32_t v = net->n_vseg_atomic + threadIdx.x; v < net->n_vseg; v += kTileThreads) r += ld_u32<COHERENT>(&vc[v]);
    r = wave_sum_u32(r);
    if (lane_id() == 0) atomicAdd(&sh.cnt[3], (unsigned long long)r);
}

__device__ __forceinline__ void tile_flush_counts(const NetDev *__restrict__ net, TileShared &sh) {
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long *tpart = net->tile_part + 4 * (size_t)blockIdx.x;
        tpart[0] += sh.cnt[0];
        tpart[1] += sh.cnt[3];
        tpart[2] += sh.cnt[1];
        tpart[3] += sh.cnt[2];
    }
}

// One step's tile scatter (the ring in HBM, this step's sums in LDS). PHASE1: the same launch also runs
// phase 1 of step t + 1 for the tile's neurons (their ring slot for t + 1 is complete once this step's
// impacts are summed) and the sources of t + 1; the host uses it when every population is tiled, saving
// one launch per step. WIDE: at most 80 VGPRs (6 waves per SIMD), so three 512-thread workgroups share a CU
// instead of two, at the cost of a few spilled registers; the host picks it when the tiles do not all fit
// on the device at once (many-instance networks: 129 -> 126 us a step; with one round of tiles it is
// slower).
// SHALLOW: every tiled population's ring is at most kTileShallowD deep (half the registers for the ring cells'
// old values, which otherwise spill in the 6-wave build).
constexpr uint32_t kTileShallowD = 16;
template <bool PHASE1, bool WIDE, bool SHALLOW>
__global__ __launch_bounds__(kTileThreads, WIDE ? 6 : 1) void k_tile_scatter(const NetDev *__restrict__ net,
                                                                              uint64_t t) {
    t = resolve_step(net, t);
    extern __shared__ unsigned long long tacc[];
    __shared__ TileShared sh;
    if (blockIdx.x == 0) clear_push_counts<false>(net, t + 1);
    const TileCtx c = tile_setup(net, tacc, sh);
    const uint32_t tid = threadIdx.x;
    const uint32_t s1 = PHASE1 ? (uint32_t)((t + 1) % c.D) : 0xFFFFFFFFu;  // slot consumed by phase 1 of t + 1
    // The ring cells of the tile are written by no one else during this kernel: their values are loaded now,
    // overlapping the rest, so that adding this step's sums at the end needs no further round trip.
    constexpr uint32_t kCells = (SHALLOW ? kTileShallowD : kTileMaxD) * kBlock / kTileThreads;
    int64_t old[kCells];
#pragma unroll
    for (uint32_t u = 0; u < kCells; ++u) {
        const uint32_t cc = tid + u * kTileThreads;
        const uint32_t i = c.tile_lo + (cc % kBlock);
        old[u] = 0;
        if (cc < c.D * kBlock && i < c.n_pop && cc / kBlock != s1) old[u] = gp(c.acc_v)[(size_t)(cc / kBlock) * c.n_pop + i];
    }
    int64_t a1 = 0;  // earlier steps' sum in slot s1 for this thread's neuron (phase 1 of t + 1)
    if (PHASE1 && tid < (uint32_t)kBlock && c.tile_lo + tid < c.n_pop)
        a1 = gp(c.acc_v)[(size_t)s1 * c.n_pop + c.tile_lo + tid];
    for (uint32_t k = tid; k < c.D * kBlock; k += kTileThreads) c.ring[k] = 0;
    __syncthreads();
    uint32_t my_ev = 0, my_parts = 0, my_cells = 0;
    tile_gather<false>(net, c, sh, t, my_ev, my_parts);
    // 4. add the tile's sums into the ring (old values loaded at the start)
#pragma unroll
    for (uint32_t u = 0; u < kCells; ++u) {
        const uint32_t cc = tid + u * kTileThreads;
        const uint32_t i = c.tile_lo + (cc % kBlock);
        if (cc < c.D * kBlock && i < c.n_pop && cc / kBlock != s1) {
            const unsigned long long x = c.ring[cc];
            if (x != 0) {
                gp(c.acc_v)[(size_t)(cc / kBlock) * c.n_pop + i] = old[u] + (long long)x;
                ++my_cells;
            }
        }
    }
    tile_count(sh, my_ev, my_parts, my_cells);
    if (blockIdx.x == 0) tile_count_rows<false>(net, sh, t);
    tile_flush_counts(net, sh);
    if (PHASE1) {
        // 5. phase 1 of step t + 1: slot s1 = earlier steps' sums in HBM + this step's sums in LDS
        const uint32_t i = c.tile_lo + tid;
        int64_t a = 0;
        if (tid < (uint32_t)kBlock && i < c.n_pop) {
            a = a1 + (int64_t)c.ring[s1 * kBlock + tid];
            gp(c.acc_v)[(size_t)s1 * c.n_pop + i] = 0;
        }
        tile_phase1<false>(net, c, t + 1, a);
    }
}

// Persistent run of steps t0 .. t0 + n - 1 when every population is tiled and all tiles fit on the device at
// once (cooperative launch): the ring lives in LDS for the whole launch (loaded from HBM at the start,
// stored back at the end). There is no grid barrier: the scatter of step t waits for the stamp of step t on
// each segment it reads (written through to memory by phase 1 of t in the segment's workgroup). A
// workgroup overwrites the segments of step t (same parity) at phase 1 of t + 2, after its own scatter of
// t + 1 has seen every segment of t + 1, whose writers had finished reading step t before publishing them.
template <uint32_t KIND>
__global__ __launch_bounds__(kTileThreads) void k_tile_run(const NetDev *__restrict__ net, uint64_t t0, uint32_t n) {
    extern __shared__ unsigned long long tacc[];
    __shared__ TileShared sh;
    // an earlier launch of the same call gave up a wait: the state is undefined, nothing more runs
    if (ld_u32<true>(net->grid_bar + kBarAbort) != 0) return;
    const TileCtx c = tile_setup(net, tacc, sh);
    const uint32_t tid = threadIdx.x;
    // the ring of the tile from HBM
    for (uint32_t cc = tid; cc < c.D * kBlock; cc += kTileThreads) {
        const uint32_t i = c.tile_lo + (cc % kBlock);
        c.ring[cc] = i < c.n_pop ? (unsigned long long)gp(c.acc_v)[(size_t)(cc / kBlock) * c.n_pop + i] : 0ull;
    }
    __syncthreads();
    __shared__ PhaseCache pc;
    phase_cache(net, c, pc);
    // default BLIFAT: the thread's neuron's potential and last spike step, kept in registers for the launch
    const bool keeper = KIND == POP_LEAN && tid < (uint32_t)kBlock && c.tile_lo + tid < c.n_pop;
    float V_keep = 0.f;
    uint32_t ls_keep = 0;
    if (keeper) {
        V_keep = gp(pc.pop.V)[c.tile_lo + tid];
        ls_keep = gp(pc.pop.last_spike)[c.tile_lo + tid];
    }
    uint32_t my_ev = 0, my_parts = 0;
    unsigned long long my_rows = 0;  // thread 0: rows published by this workgroup's phase 1
    if (blockIdx.x == 0) {
        // Channel messages exist only for t0 (pushed by the host before the launch): their rows are counted
        // here, and the channel counts of t0 + 1 (which the launch never reads) are cleared now; those of t0
        // are cleared by k_clear_push_counts after the launch, when no workgroup can still be reading them.
        if (threadIdx.x == 0)
            for (uint32_t q = 0; q < net->n_projs; ++q) {
                const uint32_t vp = gp(net->seg_vpos)[4 * net->projs[q].push_seg];
                if (vp != 0xFFFFFFFFu) my_rows += gp(net->vcount)[(size_t)(t0 & 1) * net->n_vseg + vp];
            }
        clear_push_counts<true>(net, t0 + 1);
    }
    // phase 1 of t0
    {
        const uint32_t s0 = (uint32_t)(t0 % c.D);
        int64_t a = 0;
        if (tid < (uint32_t)kBlock) {
            a = (int64_t)c.ring[s0 * kBlock + tid];
            c.ring[s0 * kBlock + tid] = 0;
        }
        my_rows += tile_phase1_dual<KIND>(net, c, pc, t0, a, V_keep, ls_keep);
    }
    uint32_t slot_t = (uint32_t)(t0 % c.D);  // ring slot of step t, advanced without a division per step
    for (uint32_t k = 0; k < n; ++k) {
        const uint64_t t = t0 + k;
        const uint32_t s1 = slot_t + 1 == c.D ? 0u : slot_t + 1;  // slot of t + 1
        // the impacts of step t that phase 1 of t + 1 needs (delay 1), then phase 1, then the others of step
        // t (delays 2 and up: later slots), while the other workgroups finish phase 1
        TilePart part;
        if (!tile_gather<true, true>(net, c, sh, t, my_ev, my_parts, k == 0, &part, slot_t)) break;
        if (k + 1 < n) {
            int64_t a = 0;
            if (tid < (uint32_t)kBlock) {
                a = (int64_t)c.ring[s1 * kBlock + tid];
                c.ring[s1 * kBlock + tid] = 0;
            }
            my_rows += tile_phase1_dual<KIND>(net, c, pc, t + 1, a, V_keep, ls_keep);
        }
        tile_add_parts<false>(c, sh, slot_t, part.b, part.len, part.q, my_ev);
        slot_t = s1;
    }
    // the ring and the kept neuron state back to HBM
    if (keeper) {
        gp(pc.pop.V)[c.tile_lo + tid] = V_keep;
        gp(pc.pop.last_spike)[c.tile_lo + tid] = ls_keep;
    }
    __syncthreads();
    uint32_t my_cells = 0;
    for (uint32_t cc = tid; cc < c.D * kBlock; cc += kTileThreads) {
        const uint32_t i = c.tile_lo + (cc % kBlock);
        if (i < c.n_pop) gp(c.acc_v)[(size_t)(cc / kBlock) * c.n_pop + i] = (int64_t)c.ring[cc];
    }
    tile_count(sh, my_ev, my_parts, my_cells);
    if (my_rows) atomicAdd(&sh.cnt[3], my_rows);
    tile_flush_counts(net, sh);
}

// Whole steps t0 .. t0 + n - 1 of a small network in one workgroup (SmallRun). Per step: phase 1 of every
// neuron (its ring slot in LDS) and every source channel; the spiking rows of every projection into an LDS
// list; then the rows' records (row_ptr, then the packed syn8 records in CSR order) added with LDS atomics
// into the rings. Channel messages exist only for t0 (pushed before the launch). The sums are exact integers,
// so the order of the additions does not matter and the results equal the tile kernels'.
__device__ __forceinline__ void small_append(uint32_t *rows, uint32_t *n_rows, bool pred, uint32_t value) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0ull) return;
    const uint32_t lane = lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(n_rows, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) rows[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

struct SmallProj {  // what the record pass needs of a projection, staged in LDS
    uint32_t row_base;     // the projection's row offsets in the launch's LDS offset table from here
    const uint64_t *syn8;
    uint32_t n_pre, ring_off, D, n_post, slot0;  // slot0: the step's ring slot of its postsynaptic population
};

// Rows (projection << 24 | presynaptic index) rows[0 .. count) of step t into the rings, a wave per row (the
// waves take rows in turn): the row's records are contiguous, so its lanes load 64 of them at a time, two
// such loads in flight, and add them with LDS atomics. Adds the impacts to ev. rows: LDS, or a channel list
// in global memory (then q is q_fixed). No barrier inside: the caller synchronizes.
template <bool GLOBAL_LIST>
__device__ __forceinline__ void small_expand(const SmallProj *sp, const uint32_t *rowp, const uint32_t *list,
                                             uint32_t count, uint32_t q_fixed, unsigned long long *ring, uint32_t &ev) {
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    constexpr uint32_t W = kSmallThreads / 64;
    for (uint32_t row = w; row < count; row += W) {
        const uint32_t e = GLOBAL_LIST ? gp(list)[row] : list[row];
        const uint32_t q = GLOBAL_LIST ? q_fixed : e >> 24;
        const uint32_t r = GLOBAL_LIST ? e : (e & 0xFFFFFFu);
        const SmallProj &P = sp[q];
        if (r >= P.n_pre) continue;
        const uint32_t b = rowp[P.row_base + r], end = rowp[P.row_base + r + 1];
        const KNP_GLOBAL uint64_t *syn8 = gp(P.syn8);
        for (uint32_t k0 = b + lane; k0 < end; k0 += 128) {
            const uint32_t k1 = k0 + 64;
            const uint64_t rec0 = syn8[k0];
            const uint64_t rec1 = k1 < end ? syn8[k1] : 0ull;
#pragma unroll
            for (uint32_t u = 0; u < 2; ++u) {
                if (u == 1 && k1 >= end) break;
                const uint64_t rec = u ? rec1 : rec0;
                const uint32_t post = (uint32_t)(rec >> 41);
                uint32_t slot = P.slot0 + ((uint32_t)(rec >> 32) & 0xFFu);
                if (slot >= P.D) slot -= P.D;
                const int32_t m = (int32_t)((uint32_t)rec << 7) >> 7;
                const unsigned long long v = (unsigned long long)(long long)m << (((uint32_t)rec >> 25) & 63u);
                atomicAdd(&ring[P.ring_off + slot * P.n_post + post], v);
                ++ev;
            }
        }
    }
}

template <uint32_t KIND>
__global__ __launch_bounds__(kSmallThreads) void k_small_run(const NetDev *__restrict__ net, SmallRun L, uint64_t t0,
                                                             uint32_t n) {
    extern __shared__ unsigned long long ring[];
    uint32_t *const rowp = reinterpret_cast<uint32_t *>(ring + L.ring_cells);
    uint32_t *const thr_lds = rowp + L.row_words;
    __shared__ uint32_t rows[kSmallRows];
    __shared__ uint32_t n_rows_buf[2];  // the rows of step k are counted in n_rows_buf[k & 1]
    __shared__ PopDev pops[kSmallMaxPops];
    __shared__ SmallProj sp[kMaxProjs];
    const uint32_t tid = threadIdx.x;
    for (uint32_t p = 0; p < L.n_pops; ++p) lds_copy(&pops[p], &net->pops[p]);
    if (tid < net->n_projs) {
        const ProjDev &Q = net->projs[tid];
        const PopDev &Pp = net->pops[Q.post_pop];
        sp[tid] = SmallProj{L.row_base[tid], Q.syn8, Q.n_pre, L.ring_off[Q.post_pop], Pp.D, Pp.n, 0u};
    }
    // the row offsets and single-pattern thresholds into LDS
    for (uint32_t q = 0; q < net->n_projs; ++q) {
        const ProjDev &Q = net->projs[q];
        for (uint32_t r = tid; r <= Q.n_pre; r += kSmallThreads) rowp[L.row_base[q] + r] = (uint32_t)gp(Q.row_ptr)[r];
    }
    for (uint32_t s = 0; s < net->n_srcs; ++s) {
        if (L.thr_base[s] == 0xFFFFFFFFu) continue;
        const SourceDev &S = net->srcs[s];
        for (uint32_t c = tid; c < S.n_ch; c += kSmallThreads) thr_lds[L.thr_base[s] + c] = gp(S.thr)[c];
    }
    if (tid == 0) n_rows_buf[0] = n_rows_buf[1] = 0;
    __syncthreads();
    // the rings from HBM
    for (uint32_t p = 0; p < L.n_pops; ++p) {
        const PopDev &P = pops[p];
        for (uint32_t c = tid; c < P.D * P.n; c += kSmallThreads) ring[L.ring_off[p] + c] = (unsigned long long)gp(P.acc_v)[c];
    }
    // this thread's neuron: its population (the last one starting at or before the thread) and index
    uint32_t mp = 0;
    for (uint32_t p = 1; p < L.n_pops; ++p)
        if (tid >= L.pop_off[p]) mp = p;
    const PopDev &MP = pops[mp];
    const uint32_t mi = tid - L.pop_off[mp];
    const bool mine = mi < MP.n;
    // the thread's ring and its population's projections, read from the launch's arguments once
    const uint32_t my_ring = L.ring_off[mp];
    const uint64_t my_projs = L.proj_of_pop[mp];
    float V_keep = 0.f;
    uint32_t ls_keep = 0;
    if (KIND == POP_LEAN && mine) {
        V_keep = gp(MP.V)[mi];
        ls_keep = gp(MP.last_spike)[mi];
    }
    __shared__ SourceDev srcs[kSmallMaxSrcs];
    const uint32_t n_srcs = net->n_srcs;
    for (uint32_t s = 0; s < n_srcs; ++s) lds_copy(&srcs[s], &net->srcs[s]);
    __syncthreads();
    // the step's ring slots and source phases, kept from step to step (no 64-bit division per step)
    uint32_t my_slot = (uint32_t)(t0 % MP.D);
    uint32_t proj_slot = tid < net->n_projs ? (uint32_t)(t0 % sp[tid].D) : 0u;
    uint32_t phase[kSmallMaxSrcs], pattern[kSmallMaxSrcs];
#pragma unroll
    for (uint32_t s = 0; s < kSmallMaxSrcs; ++s)
        if (s < n_srcs) {
            phase[s] = (uint32_t)(t0 % srcs[s].steps_per_pattern);
            pattern[s] = (uint32_t)((t0 / srcs[s].steps_per_pattern) % srcs[s].n_patterns);
        }
    uint32_t ev = 0, rows_done = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint64_t t = t0 + k;
        uint32_t *n_rows = &n_rows_buf[k & 1];
        if (tid < net->n_projs) sp[tid].slot0 = proj_slot;  // read after the step's first barrier
        // phase 1 (blifat_impl.h:31-168): the neuron's exact sum of the impacts arriving at t
        bool spike = false;
        {
            int64_t a = 0;
            if (mine) {
                unsigned long long &cell = ring[my_ring + my_slot * MP.n + mi];
                a = (int64_t)cell;
                cell = 0ull;
            }
            bool cand;
            if (KIND == POP_LEAN)
                spike = population_update<POP_LEAN, true, true>(net, MP, mi, mine, t, a, cand, &V_keep, &ls_keep);
            else
                spike = population_update<POP_FULL, true>(net, MP, mi, mine, t, a, cand);
            const uint64_t mask = __ballot(spike);
            if (mi - lane_id() < MP.n) wave_bitmap(mask, mi - lane_id(), MP, t);  // a wave holding neurons
            uint64_t qs = my_projs;
            while (qs) {  // uniform over the wave: every lane of it belongs to population mp
                const uint32_t q = (uint32_t)__ffsll((unsigned long long)qs) - 1u;
                qs &= qs - 1ull;
                small_append(rows, n_rows, spike, q << 24 | mi);
            }
        }
        // sources (their channels' spikes are rows of their projections)
#pragma unroll
        for (uint32_t s = 0; s < kSmallMaxSrcs; ++s) {
            if (s >= n_srcs) break;
            const SourceDev &S = srcs[s];
            for (uint32_t c0 = 0; c0 < S.n_ch; c0 += kSmallThreads) {
                const uint32_t c = c0 + tid;
                const uint32_t thr = (L.thr_base[s] != 0xFFFFFFFFu && c < S.n_ch) ? thr_lds[L.thr_base[s] + c] : 0xFFFFFFFFu;
                const bool sp_ = source_spike_at(S, c, true, t, phase[s], pattern[s], thr);
                small_append(rows, n_rows, sp_, S.proj << 24 | (S.first + c));
            }
        }
        __syncthreads();
        const uint32_t count = *n_rows;
        small_expand<false>(sp, rowp, rows, count, 0, ring, ev);
        rows_done += count;
        if (k == 0) {
            // channel messages pushed for t0, then their counts cleared for the steps after the launch
            for (uint32_t q = 0; q < net->n_projs; ++q) {
                const uint32_t *lst = net->projs[q].in_list + (size_t)(t0 & 1) * net->list_pstride;
                const uint32_t cnt = gp(lst)[0];
                if (cnt == 0) continue;
                small_expand<true>(sp, rowp, lst + 1, cnt, q, ring, ev);
                rows_done += cnt;
            }
            __syncthreads();
            if (tid == 0) {
                clear_push_counts_one(net, t0);
                clear_push_counts_one(net, t0 + 1);
            }
        }
        // the next step's slots and source phases
        if (++my_slot == MP.D) my_slot = 0;
        if (tid < net->n_projs && ++proj_slot == sp[tid].D) proj_slot = 0;
#pragma unroll
        for (uint32_t s = 0; s < kSmallMaxSrcs; ++s)
            if (s < n_srcs && ++phase[s] == srcs[s].steps_per_pattern) {
                phase[s] = 0;
                if (++pattern[s] == srcs[s].n_patterns) pattern[s] = 0;
            }
        // the next step's row count starts at 0; its last readers (step k - 1) are past two barriers
        if (tid == 0) n_rows_buf[(k + 1) & 1] = 0;
        __syncthreads();  // the step's sums are in the rings before phase 1 of the next step reads them
    }
    // the rings and the kept neuron state back to HBM; the counters into the first tile's slots
    for (uint32_t p = 0; p < L.n_pops; ++p) {
        const PopDev &P = pops[p];
        for (uint32_t c = tid; c < P.D * P.n; c += kSmallThreads) gp(P.acc_v)[c] = (int64_t)ring[L.ring_off[p] + c];
    }
    if (KIND == POP_LEAN && mine) {
        gp(MP.V)[mi] = V_keep;
        gp(MP.last_spike)[mi] = ls_keep;
    }
    ev = wave_sum_u32(ev);
    if (lane_id() == 0 && ev) atomicAdd(&net->tile_part[0], (unsigned long long)ev);
    if (tid == 0) net->tile_part[1] += rows_done;
}

// ---- synaptic-resource STDP (blifat_stdp.h:35-172, training/stdp.h:40-186) ---------------------------
// float -> bf16, rounded to nearest even (NaN stays a NaN).
__host__ __device__ inline uint16_t bf16_rne(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// The low BF16 part of a weight split in two: bf16(w - float(hi)) (w - float(hi) is exact in FP32).
__host__ __device__ inline uint16_t bf16_lo(float f, uint16_t hi) {
    return bf16_rne(f - __builtin_bit_cast(float, (uint32_t)hi << 16));
}


__device__ __forceinline__ float recalc_weight(float r, float w_min, float w_max) {
    // recalculate_synapse_weights (training/stdp.h:40-53)
    const float syn_w = r > 0.f ? r : 0.f;
    const float weight_diff = w_max - w_min;
    return w_min + weight_diff * syn_w / (weight_diff + syn_w);
}

__device__ __forceinline__ float stability_factor(float stability) {
    // std::min(static_cast<float>(std::pow(2, -stability)), 1.F): the power in double, rounded once to float
    // (stability_ > 0: below 1; otherwise 1, or NaN for a NaN stability, as std::min returns it)
    if (!(stability > 0.f)) return stability == stability ? 1.f : stability;
    const float v = (float)exp2(-(double)stability);
    return 1.f < v ? 1.f : v;
}

__device__ __forceinline__ bool in_interval(uint64_t b, uint64_t e, uint64_t p) {
    if (b > e) {
        uint64_t tmp = b;
        b = e;
        e = tmp;
    }
    return b <= p && p <= e;
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// The reference updates free_synaptic_resource_ once per synapse in synapse order; the float result
// depends on that order, so the wave replays it lane by lane: free += d_u of lane l; if lane l contributes:
// free += delta. du_lane is each lane's d_u, read across lanes with readlane (no memory round trip); every
// lane runs the same additions in the same order, so all lanes end with the same value. Adding a zero d_u
// changes nothing unless free is -0 (-0 + +0 = +0), and a non-negative-zero free never becomes -0 in the
// replay, so lanes with a zero d_u (not in `nonzero`) that do not contribute are skipped unless free starts
// as -0.
__device__ __forceinline__ float serial_free(float freer, uint64_t active, uint64_t nonzero, uint64_t contrib,
                                             float du_lane, float delta, bool use_du) {
    const bool neg_zero = freer == 0.f && signbit(freer);
    uint64_t m = use_du ? (neg_zero ? active : (nonzero | contrib)) : contrib;
    while (m) {
        const int l = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        if (use_du) freer += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(du_lane), l));
        if ((contrib >> l) & 1ull) freer += delta;
    }
    return freer;
}

// The per-neuron state of a training candidate (SynapticResourceSTDPNeuron's fields, blifat_stdp.h).
struct TrainNeuron {
    uint32_t status, isi_max;
    uint64_t last_step, lss, fis;
    float stab, freer;
    int64_t dop_fix;
    double dop;
    bool forced, spiked;
};

// Load candidate n's state; a spiking candidate's ISI status moves on (process_spiking_neurons_impl,
// blifat_stdp.h:64-103: status machine, last spike step, stability change at the start of an ISI period).
__device__ __forceinline__ TrainNeuron train_begin(const PopDev &P, uint32_t n, uint64_t t) {
    const ResDev &R = P.res;
    TrainNeuron s;
    s.spiked = (gp(P.bitmap)[n >> 5] >> (n & 31)) & 1u;
    // dopamine_value_ (a double sum of float impacts in the reference) from the exact accumulator sum
    s.dop_fix = gp(R.dop_fix)[n];
    s.dop = (double)s.dop_fix * (double)P.inv_scale;
    s.forced = gp(R.forced_now)[n] != 0;
    s.status = gp(R.isi_status)[n];
    s.last_step = gp(R.last_step)[n];
    s.lss = gp(R.last_spike_step)[n];
    s.fis = gp(R.first_isi_spike)[n];
    s.stab = gp(R.stability)[n];
    s.freer = gp(R.free_res)[n];
    s.isi_max = pu(R.isi_max, n);
    if (s.spiked) {
        s.lss = t;
        if (s.forced) {
            s.status = 0;  // is_forced
        } else {
            switch (s.status) {
                case 3:  // not_in_period
                case 0:  // is_forced
                    s.status = 1;
                    s.fis = t;
                    break;
                case 1:  // period_started
                    if (s.last_step - t < (uint64_t)s.isi_max) s.status = 2;
                    break;
                case 2:  // period_continued
                    if (s.last_step - t >= (uint64_t)s.isi_max || s.dop_fix != 0) {
                        s.status = 1;
                        s.fis = t;
                    }
                    break;
                default: break;
            }
            s.last_step = t;
        }
        if (s.status == 1) s.stab -= pf(R.change_at_isi, n);
    }
    return s;
}

// Does the dopamine pass run (do_dopamine_plasticity_impl's condition), and its float resource_change
// (a double product, rounded once).
__device__ __forceinline__ bool dopamine_runs(const TrainNeuron &s) {
    return s.dop > 0.0 || (s.dop < 0.0 && s.status != 0);
}

// stability_ after the dopamine pass, in the reference's double promotions (blifat_stdp.h:135-150), rounded once.
__device__ __forceinline__ float dopamine_stability(const ResDev &R, uint32_t n, const TrainNeuron &s, uint64_t t) {
    float stab = s.stab;
    if (s.forced || s.dop < 0.0) {
        stab = (float)((double)stab - s.dop * (double)pf(R.change_param, n));
        stab = stab < 0.f ? 0.f : stab;  // std::max(stability_, 0.0F)
    } else {
        const double difference = (double)(t - s.fis - (uint64_t)s.isi_max);  // u64 arithmetic, as fabs sees it
        double f = 2.0 - difference / (double)s.isi_max;
        f = f < -1.0 ? -1.0 : f;  // std::max(..., -1.0)
        stab = (float)((double)stab + (double)pf(R.change_param, n) * s.dop * f);
    }
    return stab;
}

// Does renormalize_resource run after the passes (outside an ISI period, |free resource| at the threshold)?
__device__ __forceinline__ bool renorm_runs(const ResDev &R, uint32_t n, const TrainNeuron &s, uint64_t t) {
    const bool in_isi = (t - s.last_step <= (uint64_t)s.isi_max) && s.status != 0;
    return !in_isi && fabsf(s.freer) >= pf(R.res_thr, n);
}

// additional_threshold_ = (the sum of the positive weights, in double) * synapse_sum_threshold_coefficient_
// (blifat_stdp.h:53-67): the exact sum times the coefficient, rounded once in double as the reference rounds it.
__device__ __forceinline__ void store_add_thr(const PopDev &P, uint32_t n, long long sum) {
    gp(P.res.add_thr)[n] = (double)sum * (double)P.inv_scale * (double)pf(P.res.coeff, n);
}

__device__ __forceinline__ void train_end(const ResDev &R, uint32_t n, const TrainNeuron &s) {
    gp(R.isi_status)[n] = (uint8_t)s.status;
    gp(R.last_step)[n] = s.last_step;
    gp(R.last_spike_step)[n] = s.lss;
    gp(R.first_isi_spike)[n] = s.fis;
    gp(R.stability)[n] = s.stab;
    gp(R.free_res)[n] = s.freer;
}

// One working projection's column of training candidate n as the passes read it: the table's pointers as
// global-memory pointers and the rule's shared values, copied out of the table once per column (otherwise every
// store into the synapse arrays makes the compiler reload them from the table, one dependent round trip each),
// the column's CSC range and its place in the projection's BF16 copy.
struct TrainCol {
    KNP_GLOBAL float *W, *r;
    KNP_GLOBAL uint8_t *flags;
    const KNP_GLOBAL float *w_min, *w_max, *d_u;
    const KNP_GLOBAL uint32_t *period, *col_syn, *col_pre;
    const KNP_GLOBAL uint64_t *lss, *row_spike;
    const KNP_GLOBAL uint8_t *delay_syn;
    KNP_GLOBAL uint16_t *dn_w, *dn_w_lo;  // the BF16 copy and its low parts (null: none)
    uint64_t cb, ce;                      // the column's CSC positions
    int64_t pre_off;                      // presynaptic index - CSC position (columns of consecutive indexes)
    uint64_t dn_col;                      // copy position of the column's weight from the instance's first input
    uint32_t dn_stride;                   // npad * 8: from one octet of inputs to the next
    uint32_t dn_pre0;                     // the instance's first presynaptic index
    uint32_t uni;                         // kUni* bits
    bool ident, pre0, by_row;
    float u_wmin, u_wmax, u_du;
    uint32_t u_period, u_delay;

    // synapses stored by postsynaptic neuron, consecutive presynaptic indexes, every rule parameter shared and
    // the last spike steps given by the rows (the layout of all-to-all projections): a synapse's own loads are
    // its weight, flags and resource only
    __device__ __forceinline__ bool fast() const { return ident && pre0 && by_row && (uni & kUniAll) == kUniAll; }
    __device__ __forceinline__ uint32_t syn(uint64_t k) const { return ident ? (uint32_t)k : col_syn[k]; }
    __device__ __forceinline__ uint32_t pre(uint64_t k) const {
        return pre0 ? (uint32_t)((int64_t)k + pre_off) : col_pre[k];
    }
    // a rule parameter: the projection's one value when every synapse carries the same one (no per-synapse
    // load), else the synapse's
    __device__ __forceinline__ float wmin(uint32_t s) const { return (uni & kUniWmin) ? u_wmin : w_min[s]; }
    __device__ __forceinline__ float wmax(uint32_t s) const { return (uni & kUniWmax) ? u_wmax : w_max[s]; }
    // a new weight of synapse s (presynaptic index pre) into W and into the BF16 copy
    __device__ __forceinline__ void store_weight(uint32_t s, uint32_t pre, float w) const {
        W[s] = w;
        if (dn_w) {
            const uint32_t i = pre - dn_pre0;
            const uint64_t pos = dn_col + (uint64_t)(i >> 3) * dn_stride + (i & 7);
            const uint16_t hi = bf16_rne(w);
            dn_w[pos] = hi;
            if (dn_w_lo) dn_w_lo[pos] = bf16_lo(w, hi);
        }
    }
};

// A value every lane of the wave holds alike, as a scalar (loop bounds and offsets then live in scalar registers).
__device__ __forceinline__ uint32_t wave_uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t wave_uniform(uint64_t x) {
    return (uint64_t)wave_uniform((uint32_t)x) | (uint64_t)wave_uniform((uint32_t)(x >> 32)) << 32;
}

// n: the same neuron in every lane of the wave.
__device__ __forceinline__ TrainCol train_col(const ProjDev &Q, uint32_t n) {
    TrainCol c;
    c.W = gp(Q.W);
    c.r = gp(Q.r);
    c.flags = gp(Q.flags);
    c.w_min = gp((const float *)Q.w_min);
    c.w_max = gp((const float *)Q.w_max);
    c.d_u = gp((const float *)Q.d_u);
    c.period = gp((const uint32_t *)Q.period);
    c.col_syn = gp(Q.col_syn);
    c.col_pre = gp(Q.col_pre);
    c.lss = gp((const uint64_t *)Q.last_spike_step);
    c.row_spike = gp((const uint64_t *)Q.row_spike);
    c.delay_syn = gp(Q.delay_syn);
    c.cb = wave_uniform(gp(Q.col_ptr)[n]);
    c.ce = wave_uniform(gp(Q.col_ptr)[n + 1]);
    c.pre0 = Q.col_pre0 != nullptr;
    c.pre_off = c.pre0 ? (int64_t)wave_uniform(gp(Q.col_pre0)[n]) - (int64_t)c.cb : 0;
    c.uni = Q.rule_uniform;
    c.ident = Q.csc_identity != 0;
    c.by_row = Q.lss_by_row != 0;
    c.u_wmin = Q.u_wmin;
    c.u_wmax = Q.u_wmax;
    c.u_du = Q.u_du;
    c.u_period = Q.u_period;
    c.u_delay = Q.u_delay;
    c.dn_w = nullptr;
    c.dn_w_lo = nullptr;
    c.dn_col = 0;
    c.dn_stride = 0;
    c.dn_pre0 = 0;
    if (Q.dense && c.cb < c.ce) {  // a neuron with synapses in a dense projection lies in one of its instances
        const KNP_GLOBAL DenseInst *I = gp(Q.dn_inst_tab) + gp(Q.dn_post_inst)[n];
        c.dn_w = gp(Q.dn_w);
        c.dn_w_lo = gp(Q.dn_w_lo);
        c.dn_col = wave_uniform(I->base + (uint64_t)(n - I->post_begin) * 8);
        c.dn_stride = wave_uniform(I->npad * 8);
        c.dn_pre0 = wave_uniform(I->pre_begin);
    }
    return c;
}

// Spiking pass of synapse k (CSC position) of the candidate (blifat_stdp.h:105-131): had_contribution /
// had_hebbian_update flags, resource -= d_u (and += d_h on a new contribution), weight recalculated. Adds the
// old weight's fixed-point value to sum when positive; returns d_u in du (0 when the neuron is forced) and
// whether the synapse contributed.
struct SpikingSyn {  // what the spiking pass reads of one synapse
    uint32_t sy, pre;
    float w, r, du, wmin, wmax;
    uint64_t per, spk;
    uint8_t fl;
};

// The presynaptic index of a column of consecutive indexes is k + pre_off, with no load in front of row_spike.
__device__ __forceinline__ SpikingSyn spiking_load(const TrainCol &C, uint64_t k, uint32_t status) {
    SpikingSyn x;
    x.sy = C.syn(k);
    x.w = C.W[x.sy];
    x.fl = C.flags[x.sy];
    x.per = (C.uni & kUniPeriod) ? C.u_period : C.period[x.sy];
    x.pre = C.pre(k);
    const uint64_t rs = C.row_spike[x.pre];
    const uint32_t delay = (C.uni & kUniDelay) ? C.u_delay : C.delay_syn[x.sy];
    x.spk = (rs != UINT64_MAX ? rs : C.lss[x.sy]) + delay - 1;
    x.r = C.r[x.sy];
    x.du = status != 0 ? ((C.uni & kUniDu) ? C.u_du : C.d_u[x.sy]) : 0.f;
    x.wmin = C.wmin(x.sy);
    x.wmax = C.wmax(x.sy);
    return x;
}

// What the spiking pass writes back for one synapse (stores of unchanged values are left out; the BF16 copy of
// an unchanged weight is current too).
struct SpikingOut {
    float rr, wn;
    uint8_t nf;
    bool st_r, st_w, st_f;
};

__device__ __forceinline__ bool spiking_compute(const SpikingSyn &x, uint64_t t, uint32_t status, float d_h,
                                                float scale, long long &sum, SpikingOut &o) {
    const float w = x.w, r = x.r, du = x.du;
    const uint8_t fl = x.fl;
    if (w > 0.f) sum += __float2ll_rn(w * scale);
    bool hc = fl & 2, hh = fl & 1;
    const bool had_spike = in_interval(t - x.per, t, x.spk);
    if (status != 2 || had_spike) hc = had_spike;
    if (status != 2) hh = false;
    float rr = r;
    bool contributing = false;
    if (status != 0) {
        rr -= du;
        if (hc && !hh) {
            rr += d_h;
            hh = true;
            contributing = true;
        }
    }
    o.rr = rr;
    o.st_r = __float_as_uint(rr) != __float_as_uint(r);
    o.wn = recalc_weight(rr, x.wmin, x.wmax);
    o.st_w = __float_as_uint(o.wn) != __float_as_uint(w);
    o.nf = (uint8_t)((hc ? 2 : 0) | (hh ? 1 : 0));
    o.st_f = o.nf != fl;
    return contributing;
}

__device__ __forceinline__ void spiking_store(const TrainCol &C, uint32_t sy, uint32_t pre, const SpikingOut &o) {
    if (o.st_r) C.r[sy] = o.rr;
    if (o.st_w) C.store_weight(sy, pre, o.wn);
    if (o.st_f) C.flags[sy] = o.nf;
}

__device__ __forceinline__ bool spiking_update(const TrainCol &C, const SpikingSyn &x, uint64_t t, uint32_t status,
                                               float d_h, float scale, long long &sum) {
    SpikingOut o;
    const bool contributing = spiking_compute(x, t, status, d_h, scale, sum, o);
    spiking_store(C, x.sy, x.pre, o);
    return contributing;
}

// Dopamine pass of synapse k (blifat_stdp.h:133-156): resource += change where the synapse contributed and the
// neuron spiked recently; the weight is recalculated in every case. Returns whether the resource changed.
__device__ __forceinline__ bool dopamine_synapse(const TrainCol &C, uint64_t k, bool time_ok, float rc) {
    const uint32_t sy = C.syn(k);
    const float r = C.r[sy];
    const float w = C.W[sy];
    const uint8_t fl = C.flags[sy];
    const float wmin = C.wmin(sy), wmax = C.wmax(sy);
    float rr = r;
    bool contributing = false;
    if (time_ok && (fl & 2)) {
        rr += rc;
        if (__float_as_uint(rr) != __float_as_uint(r)) C.r[sy] = rr;
        contributing = true;
    }
    const float wn = recalc_weight(rr, wmin, wmax);
    if (__float_as_uint(wn) != __float_as_uint(w)) C.store_weight(sy, C.pre(k), wn);
    return contributing;
}

// renormalize_resource for synapse k (blifat_stdp.h:174-190): resource += add, weight recalculated.
// r0 and w: the synapse's resource and weight as loaded.
__device__ __forceinline__ void renorm_update(const TrainCol &C, uint64_t k, float r0, float w, float add) {
    const uint32_t s = C.syn(k);
    const float r = r0 + add;
    if (__float_as_uint(r) != __float_as_uint(r0)) C.r[s] = r;
    const float wn = recalc_weight(r, C.wmin(s), C.wmax(s));
    if (__float_as_uint(wn) != __float_as_uint(w)) C.store_weight(s, C.pre(k), wn);
}

__device__ __forceinline__ void renorm_synapse(const TrainCol &C, uint64_t k, float add) {
    const uint32_t s = C.syn(k);
    renorm_update(C, k, C.r[s], C.W[s], add);
}

// The free-resource replay of one chunk of the spiking pass (serial_free; lane 0 is active in every chunk, and
// one d_u over the chunk takes the closed form). Every lane of the wave calls it.
__device__ __forceinline__ float replay_chunk(float freer, bool act, float du, bool contributing, float d_h) {
    const uint64_t am = __ballot(act), nz = __ballot(act && du != 0.f), cm = __ballot(contributing);
    const float du0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(du)));
    if (du0 != 0.f && __ballot(act && __float_as_int(du) != __float_as_int(du0)) == 0)
        return replay_uniform(freer, am, cm, du0, -d_h);
    return serial_free(freer, am, nz, cm, du, -d_h, true);
}

// What the spiking pass loads of U chunks of a fast column (TrainCol::fast).
template <uint32_t U>
struct FastChunks {
    float w[U], r[U];
    uint64_t rs[U];  // the row's last spike step (the synapse's last_spike_step_)
    uint8_t fl[U];
};

template <uint32_t U>
__device__ __forceinline__ void fast_load(const TrainCol &C, uint64_t base, uint32_t lane, FastChunks<U> &g) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
        uint64_t k = base + 64 * u + lane;
        k = k < C.ce ? k : C.ce - 1;  // lanes past the column's end load its last synapse and store nothing
        g.w[u] = C.W[k];
        g.fl[u] = C.flags[k];
        g.r[u] = C.r[k];
        g.rs[u] = C.row_spike[(uint32_t)((int64_t)k + C.pre_off)];
    }
}

// The spiking pass over a fast column, U chunks at a time, software-pipelined: the next chunks' loads are issued,
// the current chunks are updated in registers and their free-resource sums replayed in synapse order, then the
// wave waits for the next chunks' loads and only after that issues the current chunks' stores. (Stores of
// unchanged values are skipped, so the number of stores in flight is unknown to the compiler; a load issued
// before them could only be waited for together with them.)
template <uint32_t U>
__device__ __forceinline__ void spiking_column_fast(const TrainCol &C, TrainNeuron &s, uint64_t t, float d_h,
                                                    float scale, long long &sum, uint32_t lane) {
    constexpr uint64_t kStep = 64 * U;
    FastChunks<U> cur, nxt;
    fast_load<U>(C, C.cb, lane, cur);
    nxt = cur;
    for (uint64_t base = C.cb; base < C.ce; base += kStep) {
        const bool more = base + kStep < C.ce;
        if (more) fast_load<U>(C, base + kStep, lane, nxt);
        SpikingOut o[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (base + 64 * u >= C.ce) break;
            const uint64_t k = base + 64 * u + lane;
            const bool act = k < C.ce;
            SpikingSyn x;
            x.w = act ? cur.w[u] : 0.f;  // (a lane past the end adds nothing to the weight sum)
            x.r = cur.r[u];
            x.fl = cur.fl[u];
            x.per = C.u_period;
            x.spk = cur.rs[u] + C.u_delay - 1;
            x.du = s.status != 0 ? C.u_du : 0.f;
            x.wmin = C.u_wmin;
            x.wmax = C.u_wmax;
            const bool contributing = spiking_compute(x, t, s.status, d_h, scale, sum, o[u]) && act;
            if (!act) o[u].st_r = o[u].st_w = o[u].st_f = false;
            if (s.status != 0) s.freer = replay_chunk(s.freer, act, act ? x.du : 0.f, contributing, d_h);
        }
        // s_waitcnt vmcnt(0): every vector memory access issued so far (the next chunks' loads) complete
        __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (base + 64 * u >= C.ce) break;
            const uint64_t k = base + 64 * u + lane;
            spiking_store(C, (uint32_t)k, (uint32_t)((int64_t)k + C.pre_off), o[u]);
        }
        cur = nxt;
    }
}

// What the dopamine and renormalization passes load of U chunks of a fast column.
template <uint32_t U>
struct FastWRF {
    float w[U], r[U];
    uint8_t fl[U];
};

template <uint32_t U>
__device__ __forceinline__ void wrf_load(const TrainCol &C, uint64_t base, uint32_t lane, FastWRF<U> &g) {
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
        uint64_t k = base + 64 * u + lane;
        k = k < C.ce ? k : C.ce - 1;  // lanes past the column's end load its last synapse and store nothing
        g.w[u] = C.W[k];
        g.r[u] = C.r[k];
        g.fl[u] = C.flags[k];
    }
}

// A pass over a fast column that changes resources and weights (dopamine, renormalization), pipelined as the
// spiking pass: update(act, w, r, fl, &r_new) gives the synapse's new resource; the weight is its recalculation.
// Every lane calls update for every chunk, in synapse order.
template <uint32_t U, typename F>
__device__ __forceinline__ void resource_column_fast(const TrainCol &C, uint32_t lane, F update) {
    constexpr uint64_t kStep = 64 * U;
    FastWRF<U> cur, nxt;
    wrf_load<U>(C, C.cb, lane, cur);
    nxt = cur;
    for (uint64_t base = C.cb; base < C.ce; base += kStep) {
        if (base + kStep < C.ce) wrf_load<U>(C, base + kStep, lane, nxt);
        float rn[U], wn[U];
        bool st_r[U], st_w[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (base + 64 * u >= C.ce) break;
            const bool act = base + 64 * u + lane < C.ce;
            rn[u] = update(act, cur.w[u], cur.r[u], cur.fl[u]);
            wn[u] = recalc_weight(rn[u], C.u_wmin, C.u_wmax);
            st_r[u] = act && __float_as_uint(rn[u]) != __float_as_uint(cur.r[u]);
            st_w[u] = act && __float_as_uint(wn[u]) != __float_as_uint(cur.w[u]);
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // the next chunks' loads complete before this group's stores
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (base + 64 * u >= C.ce) break;
            const uint64_t k = base + 64 * u + lane;
            if (st_r[u]) C.r[k] = rn[u];
            if (st_w[u]) C.store_weight((uint32_t)k, (uint32_t)((int64_t)k + C.pre_off), wn[u]);
        }
        cur = nxt;
    }
}

#ifndef KNP_TRAIN_FAST_UNROLL
#define KNP_TRAIN_FAST_UNROLL 2
#endif
// Chunks of a fast column loaded together (the pipeline holds twice as many).
constexpr uint32_t kTrainFastUnroll = KNP_TRAIN_FAST_UNROLL;

#ifndef KNP_TRAIN_UNROLL
#define KNP_TRAIN_UNROLL 2
#endif
// Chunks whose loads the wave form issues together (latency: a batch's candidates fill the device with waves
// that each walk one neuron's chunks).
constexpr uint32_t kTrainUnroll = KNP_TRAIN_UNROLL;

// One training candidate on one wave: the synapses of each working projection (load order, then CSC order)
// 64 at a time, the order-dependent free-resource sum replayed after every chunk.
__device__ __forceinline__ void train_candidate_wave(const NetDev *__restrict__ net, const PopDev &P, uint32_t n,
                                                     uint64_t t, uint32_t lane) {
    const ResDev &R = P.res;
    TrainNeuron s = train_begin(P, n, t);
    if (s.spiked) {
        const float d_h = pf(R.d_h, n) * stability_factor(s.stab);
        long long sum = 0;
        for (uint32_t j = 0; j < R.n_work; ++j) {
            const TrainCol C = train_col(net->projs[R.work[j]], n);
            if (C.fast()) {
                spiking_column_fast<kTrainFastUnroll>(C, s, t, d_h, P.scale, sum, lane);
                continue;
            }
            // kTrainUnroll chunks at a time: all their loads are issued before the first chunk's stores (the
            // chunks hold different synapses), then each chunk is updated and its free-resource sum replayed
            // in order
            for (uint64_t base = C.cb; base < C.ce; base += 64 * kTrainUnroll) {
                SpikingSyn x[kTrainUnroll];
                bool act[kTrainUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kTrainUnroll; ++u) {
                    act[u] = base + 64 * u + lane < C.ce;
                    if (act[u]) x[u] = spiking_load(C, base + 64 * u + lane, s.status);
                }
#pragma unroll
                for (uint32_t u = 0; u < kTrainUnroll; ++u) {
                    if (base + 64 * u >= C.ce) break;
                    const bool contributing = act[u] && spiking_update(C, x[u], t, s.status, d_h, P.scale, sum);
                    if (s.status != 0)
                        s.freer = replay_chunk(s.freer, act[u], act[u] ? x[u].du : 0.f, contributing, d_h);
                }
            }
        }
        sum = wave_sum_i64(sum);
        if (lane == 0) store_add_thr(P, n, sum);
    }
    bool passed = s.spiked;  // a pass over every working synapse recalculated the weights from the resources
    if (dopamine_runs(s)) {
        const bool time_ok = (t - s.lss) <= (uint64_t)pu(R.dpt, n);
        const float rc = (float)(s.dop * (double)stability_factor(s.stab));
        // without time_ok no resource changes; the pass then only recalculates weights, which changes none
        // once a pass has made every weight its recalculation
        const bool needed = time_ok || !gp(R.w_current)[n];
        passed |= needed;
        for (uint32_t j = 0; j < R.n_work && needed; ++j) {
            const TrainCol C = train_col(net->projs[R.work[j]], n);
            if (C.fast()) {
                uint64_t contrib = 0;
                resource_column_fast<kTrainFastUnroll>(C, lane, [&](bool act, float, float r, uint8_t fl) {
                    const bool c = time_ok && (fl & 2);
                    contrib += (uint64_t)__popcll(__ballot(act && c));
                    return c ? r + rc : r;
                });
                // free -= change once per contributing synapse: the same float operation each time
                s.freer = repeat_add(s.freer, -rc, contrib);
                continue;
            }
            for (uint64_t base = C.cb; base < C.ce; base += 64) {
                const uint64_t k = base + lane;
                const uint64_t cm = __ballot(k < C.ce && dopamine_synapse(C, k, time_ok, rc));
                s.freer = repeat_add(s.freer, -rc, (uint64_t)__popcll(cm));
            }
        }
        s.stab = dopamine_stability(R, n, s, t);
    }
    if (renorm_runs(R, n, s, t)) {
        uint64_t k_total = 0;
        for (uint32_t j = 0; j < R.n_work; ++j) {
            const KNP_GLOBAL uint64_t *cp = gp(net->projs[R.work[j]].col_ptr);
            k_total += cp[n + 1] - cp[n];
        }
        const float add = s.freer / (float)(k_total + (uint64_t)pu(R.drain, n));
        for (uint32_t j = 0; j < R.n_work; ++j) {
            const TrainCol C = train_col(net->projs[R.work[j]], n);
            if (C.fast()) {
                resource_column_fast<kTrainFastUnroll>(C, lane,
                                                       [&](bool, float, float r, uint8_t) { return r + add; });
                continue;
            }
            for (uint64_t k = C.cb + lane; k < C.ce; k += 64 * kTrainUnroll) {
                float r0[kTrainUnroll], w0[kTrainUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kTrainUnroll; ++u)
                    if (k + 64 * u < C.ce) {
                        const uint32_t sy = C.syn(k + 64 * u);
                        r0[u] = C.r[sy];
                        w0[u] = C.W[sy];
                    }
#pragma unroll
                for (uint32_t u = 0; u < kTrainUnroll; ++u)
                    if (k + 64 * u < C.ce) renorm_update(C, k + 64 * u, r0[u], w0[u], add);
            }
        }
        s.freer = 0.f;
        passed = true;
    }
    if (lane == 0) {
        train_end(R, n, s);
        if (passed) gp(R.w_current)[n] = 1;
    }
}

// Synaptic-resource STDP of one population's training candidates (train_population_impl, blifat_stdp.h:158-172).
// Every per-synapse quantity depends only on the neuron and the synapse, and candidates own disjoint synapses;
// only the free-resource sum of a candidate depends on the order of its float additions.
//
// k_train_wave: a wave per candidate, which replays its own sum between chunks, so no wave waits for another
// (populations large enough that a step's candidates fill the device: batches of instances).
__global__ __launch_bounds__(kBlock) void k_train_wave(const NetDev *__restrict__ net, uint32_t p, uint64_t t) {
    t = resolve_step(net, t);
    const PopDev &P = net->pops[p];
    const uint32_t count = net->ctl[t & 1].cand_count[p];
    const uint32_t lane = lane_id();
    constexpr uint32_t W = kBlock / 64;
    // the candidate is the wave's (scalar loads of its state and its columns' table entries)
    const uint32_t c0 = __builtin_amdgcn_readfirstlane(blockIdx.x * W + threadIdx.x / 64);
    for (uint32_t c = c0; c < count; c += gridDim.x * W)
        train_candidate_wave(net, P, wave_uniform(gp(P.res.cand)[c]), t, lane);
}

// k_train: a workgroup per candidate (smaller populations, whose few candidates a step cannot fill the device
// with): the candidate's synapses (working projections in load order, then CSC order) are cut into 64-synapse
// chunks, the sixteen waves update sixteen consecutive chunks at a time, and wave 0 then replays the
// free-resource sum chunk by chunk in synapse order from the d_u values and masks the waves left in LDS.
constexpr uint32_t kTrainBlock = 1024;  // 16 waves: a 784-input neuron's 13 chunks in one round

__global__ __launch_bounds__(kTrainBlock) void k_train(const NetDev *__restrict__ net, uint32_t p, uint64_t t) {
    t = resolve_step(net, t);
    const PopDev &P = net->pops[p];
    const ResDev &R = P.res;
    const uint32_t count = net->ctl[t & 1].cand_count[p];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / 64;
    constexpr uint32_t W = kTrainBlock / 64;
    __shared__ float s_du[W][64];
    __shared__ unsigned long long s_mask[W][3];  // active, nonzero d_u, contributing
    __shared__ float s_uni[W];                   // the chunk's d_u when all its synapses carry it (else 0)
    __shared__ long long s_sum[W];
    __shared__ uint32_t s_cnt[W];
    __shared__ uint64_t s_cb[kMaxWorking], s_ce[kMaxWorking];
    __shared__ uint32_t s_nch[kMaxWorking];
    __shared__ float s_free;
    // chunk g of the candidate: its working projection, CSC position of its first synapse and the column's end
    auto chunk = [&](uint32_t g, uint32_t &j, uint64_t &base, uint64_t &end) -> bool {
        for (j = 0; j < R.n_work; ++j) {
            if (g < s_nch[j]) {
                base = s_cb[j] + 64ull * g;
                end = s_ce[j];
                return true;
            }
            g -= s_nch[j];
        }
        return false;
    };
    for (uint32_t c = blockIdx.x; c < count; c += gridDim.x) {
        const uint32_t n = wave_uniform(gp(R.cand)[c]);
        __syncthreads();  // the previous candidate's LDS reads are done
        if (threadIdx.x < R.n_work) {
            const ProjDev &Q = net->projs[R.work[threadIdx.x]];
            s_cb[threadIdx.x] = gp(Q.col_ptr)[n];
            s_ce[threadIdx.x] = gp(Q.col_ptr)[n + 1];
            s_nch[threadIdx.x] = (uint32_t)((s_ce[threadIdx.x] - s_cb[threadIdx.x] + 63) / 64);
        }
        __syncthreads();
        uint32_t G = 0;
        for (uint32_t j = 0; j < R.n_work; ++j) G += s_nch[j];
        TrainNeuron s = train_begin(P, n, t);
        if (s.spiked) {
            const float d_h = pf(R.d_h, n) * stability_factor(s.stab);
            long long sum = 0;
            for (uint32_t r0 = 0; r0 < G; r0 += W) {
                uint64_t am = 0, nz = 0, cm = 0;
                float du = 0.f;
                uint32_t j;
                uint64_t base, end;
                if (r0 + wv < G && chunk(r0 + wv, j, base, end)) {
                    const TrainCol C = train_col(net->projs[R.work[j]], n);
                    const uint64_t k = base + lane;
                    const bool act = k < end;
                    bool contributing = false;
                    if (act) {
                        const SpikingSyn x = spiking_load(C, k, s.status);
                        du = x.du;
                        contributing = spiking_update(C, x, t, s.status, d_h, P.scale, sum);
                    }
                    am = __ballot(act);
                    nz = __ballot(act && du != 0.f);
                    cm = __ballot(contributing);
                }
                // lane 0 is active in every chunk; one d_u over the chunk takes the closed form
                const float du0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(du)));
                const bool uni = du0 != 0.f && __ballot(am != 0 && ((am >> lane) & 1) &&
                                                        __float_as_int(du) != __float_as_int(du0)) == 0;
                s_du[wv][lane] = du;
                if (lane == 0) {
                    s_mask[wv][0] = am;
                    s_mask[wv][1] = nz;
                    s_mask[wv][2] = cm;
                    s_uni[wv] = uni ? du0 : 0.f;
                }
                __syncthreads();
                if (wv == 0 && s.status != 0)
                    for (uint32_t w = 0; w < W && r0 + w < G; ++w) {
                        const float u = s_uni[w];
                        if (u != 0.f)
                            s.freer = replay_uniform(s.freer, s_mask[w][0], s_mask[w][2], u, -d_h);
                        else
                            s.freer = serial_free(s.freer, s_mask[w][0], s_mask[w][1], s_mask[w][2], s_du[w][lane],
                                                  -d_h, true);
                    }
                __syncthreads();
            }
            sum = wave_sum_i64(sum);
            if (lane == 0) s_sum[wv] = sum;
            if (threadIdx.x == 0) s_free = s.freer;
            __syncthreads();
            s.freer = s_free;  // the replay's result (wave 0) for every thread
            if (threadIdx.x == 0) {
                long long total = 0;
                for (uint32_t w = 0; w < W; ++w) total += s_sum[w];
                store_add_thr(P, n, total);
            }
        }
        bool passed = s.spiked;  // a pass over every working synapse recalculated the weights
        if (dopamine_runs(s)) {
            const bool time_ok = (t - s.lss) <= (uint64_t)pu(R.dpt, n);
            const float rc = (float)(s.dop * (double)stability_factor(s.stab));
            // without time_ok the pass only recalculates weights, which changes none once they are current
            const bool needed = time_ok || !gp(R.w_current)[n];
            passed |= needed;
            uint32_t contrib = 0;
            for (uint32_t g = wv; g < G && needed; g += W) {
                uint32_t j;
                uint64_t base, end;
                if (!chunk(g, j, base, end)) break;
                const TrainCol C = train_col(net->projs[R.work[j]], n);
                const uint64_t k = base + lane;
                contrib += (uint32_t)__popcll(__ballot(k < end && dopamine_synapse(C, k, time_ok, rc)));
            }
            if (lane == 0) s_cnt[wv] = contrib;
            __syncthreads();
            // free -= change once per contributing synapse: the same float operation each time, in any order
            uint32_t total = 0;
            for (uint32_t w = 0; w < W; ++w) total += s_cnt[w];
            if (threadIdx.x == 0) {
                s.freer = repeat_add(s.freer, -rc, total);
                s_free = s.freer;
            }
            __syncthreads();
            s.freer = s_free;
            s.stab = dopamine_stability(R, n, s, t);
        }
        if (renorm_runs(R, n, s, t)) {
            uint64_t k_total = 0;
            for (uint32_t j = 0; j < R.n_work; ++j) k_total += s_ce[j] - s_cb[j];
            const float add = s.freer / (float)(k_total + (uint64_t)pu(R.drain, n));
            for (uint32_t g = wv; g < G; g += W) {
                uint32_t j;
                uint64_t base, end;
                if (!chunk(g, j, base, end)) break;
                const TrainCol C = train_col(net->projs[R.work[j]], n);
                const uint64_t k = base + lane;
                if (k < end) renorm_synapse(C, k, add);
            }
            s.freer = 0.f;
            passed = true;
        }
        if (threadIdx.x == 0) {
            train_end(R, n, s);
            if (passed) gp(R.w_current)[n] = 1;
        }
    }
}

// ---- additive STDP (additive_stdp.h:93-196) ----------------------------------------------------------
// init_projection: spikes of the projection's stdp_populations_ append the step to the spike histories of
// every synapse whose POSTsynaptic index equals the spiking neuron's index (additive_stdp.h:149-160).
__global__ __launch_bounds__(kBlock) void k_additive_append(const NetDev *__restrict__ net, uint32_t q, uint64_t t) {
    t = resolve_step(net, t);
    const ProjDev &Q = net->projs[q];
    const uint32_t *cnt = net->seg_count + (size_t)(t & 1) * net->n_seg;
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (kBlock / 64);
    const uint32_t wave = blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
    const uint32_t n_post = net->pops[Q.post_pop].n;
    for (uint32_t m = 0; m < Q.n_stdp; ++m) {
        const PopDev &SP = net->pops[Q.stdp_pop[m]];
        const uint32_t chunks = (SP.n + kBlock - 1) / kBlock;
        // one wave per spike: walk (chunk, index) pairs
        for (uint32_t item = wave; item < chunks * kBlock; item += waves) {
            const uint32_t chunk = item / kBlock, r = item % kBlock;
            if (r >= cnt[SP.seg_base + chunk]) continue;
            const uint32_t nidx = SP.spike_list[(size_t)(t & 1) * net->list_pstride + (size_t)chunk * kSegStride + 1 + r];
            if (nidx >= n_post) continue;
            const uint64_t cb = Q.col_ptr[nidx], ce = Q.col_ptr[nidx + 1];
            for (uint64_t k = cb + lane; k < ce; k += 64) {
                const uint32_t s = Q.col_syn[k];
                const uint32_t lim = Q.lim[s];
                uint32_t old = atomicAdd(&Q.n_post_t[s], 1u);
                if (old < lim) Q.post_times[(size_t)s * Q.hist + old] = t;
                else atomicSub(&Q.n_post_t[s], 1u);
                if (Q.stdp_mode[m] == KNP_STDP_AND_SPIKE) {
                    old = atomicAdd(&Q.n_pre_t[s], 1u);
                    if (old < lim) Q.pre_times[(size_t)s * Q.hist + old] = t;
                    else atomicSub(&Q.n_pre_t[s], 1u);
                }
            }
        }
    }
}

// modify_weights: synapses whose two histories reached tau_plus_ + tau_minus_ get
// w += sum_pre sum_post stdp_w(t_post - t_pre) and their histories cleared (additive_stdp.h:173-196).
__global__ __launch_bounds__(kBlock) void k_additive_modify(const NetDev *__restrict__ net, uint32_t q) {
    const ProjDev &Q = net->projs[q];
    for (uint64_t s = blockIdx.x * (uint64_t)kBlock + threadIdx.x; s < Q.n_syn; s += (uint64_t)gridDim.x * kBlock) {
        const float tp = Q.tau_plus[s], tm = Q.tau_minus[s];
        const float period = tp + tm;
        const uint32_t np = Q.n_pre_t[s], nq = Q.n_post_t[s];
        if ((float)np >= period && (float)nq >= period) {
            float w_j = 0.f;
            for (uint32_t a = 0; a < np; ++a) {
                const uint64_t tf = Q.pre_times[s * Q.hist + a];
                for (uint32_t b2 = 0; b2 < nq; ++b2) {
                    const float diff = (float)(Q.post_times[s * Q.hist + b2] - tf);
                    w_j += diff > 0.f ? expf(-diff / tp) : expf(diff / tm);
                }
            }
            Q.W[s] += w_j;
            Q.n_pre_t[s] = 0;
            Q.n_post_t[s] = 0;
        }
    }
}


// ---- dense input projections on MFMA ---------------------------------------------------------------
// The binary spike matrix of a time block (T steps x n_in inputs per instance) is exact in FP8 (1.0 = 0x38)
// and BF16 (1.0 = 0x3F80); the products with the weights are exact and FP32 accumulation of a few hundred
// of them is exact for weights on a 2^-k grid, so impacts equal the reference's sum of the same weights.
typedef float v16f __attribute__((ext_vector_type(16)));
typedef short v8s __attribute__((ext_vector_type(8)));

// Spike bits of steps t0 .. t0+T-1: byte b of (step, instance) holds inputs 8b .. 8b+7 (bytes from 2 ksteps
// up to the row pitch are 0). Counts the dense synaptic events and rows of the block (one atomic pair per
// wave).
__global__ __launch_bounds__(kBlock) void k_dense_spikes(const NetDev *__restrict__ net, uint32_t d, uint64_t t0) {
    t0 = resolve_step(net, t0);
    const DenseDev &Dn = net->dense[d];
    const uint32_t pitch = Dn.pitch, used = 2u * Dn.ksteps;
    const uint64_t total = (uint64_t)Dn.T * Dn.n_inst * pitch;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    unsigned long long ev = 0, rows = 0;
    for (uint64_t idx = (uint64_t)blockIdx.x * kBlock + threadIdx.x; idx < total; idx += stride) {
        const uint32_t byte = (uint32_t)(idx % pitch);
        const uint64_t rest = idx / pitch;
        const uint32_t inst = (uint32_t)(rest % Dn.n_inst);
        const uint64_t t = t0 + rest / Dn.n_inst;
        uint32_t v = 0;
        if (byte < used)
            for (uint32_t b = 0; b < 8; ++b) {
                const uint32_t i = byte * 8 + b;
                if (i < Dn.n_in && bernoulli_hash(Dn.seed, t, Dn.hash_offset + inst * Dn.n_in + i) < Dn.threshold)
                    v |= 1u << b;
            }
        gp(Dn.bits)[idx] = (uint8_t)v;
        const uint32_t c = __popc(v);
        rows += c;
        ev += (unsigned long long)c * gp(Dn.inst_post_count)[inst];
    }
    for (int off = 32; off > 0; off >>= 1) {
        ev += __shfl_xor(ev, off);
        rows += __shfl_xor(rows, off);
    }
    if (lane_id() == 0 && rows) {
        atomicAdd((unsigned long long *)gp(net->events), ev);
        atomicAdd((unsigned long long *)gp(net->events) + 1, rows);
    }
}

// A operands from spike bits. Four bits x (x < 16) to one byte each: x * 0x204081 places bit k of x at bit
// 8k (the shifted copies x << 7k do not overlap, so the product has no carries), the mask keeps those bits.
__device__ __forceinline__ uint32_t spread4_bytes(uint32_t x) { return __umul24(x, 0x204081u) & 0x01010101u; }

// 8 spike bits (bit e = input e) to 8 fp8 e4m3 values: 1.0 (0x38) where the bit is set, else 0.
__device__ __forceinline__ long spread_fp8_ones(uint32_t byte) {
    const uint32_t lo = spread4_bytes(byte & 0xFu) * 0x38u, hi = spread4_bytes((byte >> 4) & 0xFu) * 0x38u;
    return (long)((uint64_t)hi << 32 | lo);
}

// 8 spike bits to 8 bf16 values: 1.0 (0x3F80) where the bit is set, else 0. Two bits x (x < 4) to the two
// halves of a word: x * 0x8001 places bit 1 at bit 16, the mask keeps bits 0 and 16.
__device__ __forceinline__ v8s spread_bf16_ones(uint32_t byte) {
    typedef unsigned int v4w __attribute__((ext_vector_type(4)));
    v4w w;
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (__umul24((byte >> (2 * k)) & 3u, 0x8001u) & 0x00010001u) * 0x3F80u;
    return __builtin_bit_cast(v8s, w);
}

// One wave per (32-column tile, group of kDenseGroup 32-step row blocks): A = the spike bits of 32 steps (row
// m = step, K = inputs), B = the tile's weight fragments, read once per group and shared by its row blocks.
// Each lane loads the spike bits of its step rows 16 bytes (8 K-steps) at a time. Output element (row m,
// column n) of a 32x32 accumulator sits in lane n + 32 * ((m / 4) % 2), register 4 * (m / 8) + m % 4.
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_dense_mfma(const NetDev *__restrict__ net, uint32_t d, uint64_t t0) {
    t0 = resolve_step(net, t0);
    const DenseDev &Dn = net->dense[d];
    const PopDev &P = net->pops[Dn.post_pop];
    constexpr uint32_t FB = FMT == 0 ? 8 : 16;  // fragment bytes per lane per K-step
    constexpr uint32_t G = kDenseGroup;
    const uint32_t lane = lane_id(), h = lane >> 5, col = lane & 31;
    const uint32_t n_rb = Dn.T / 32, n_grp = (n_rb + G - 1) / G;
    const uint64_t units = (uint64_t)Dn.n_tiles * n_grp;
    const uint64_t ring = 2ull * Dn.T;
    const uint64_t waves = (uint64_t)gridDim.x * (kBlock / 64);
    for (uint64_t u = (uint64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64; u < units; u += waves) {
        const uint32_t tile = (uint32_t)(u / n_grp), rb0 = (uint32_t)(u % n_grp) * G;
        const uint32_t nb = min(G, n_rb - rb0);
        const uint32_t inst = gp(Dn.tile_inst)[tile], j0 = gp(Dn.tile_j0)[tile], ncol = gp(Dn.tile_ncol)[tile];
        const KNP_GLOBAL uint8_t *fr = gp(Dn.frag) + (size_t)tile * Dn.ksteps * 64 * FB + (size_t)lane * FB;
        const KNP_GLOBAL uint8_t *brow[G];
#pragma unroll
        for (uint32_t g = 0; g < G; ++g)
            brow[g] = gp(Dn.bits) + ((size_t)((rb0 + min(g, nb - 1)) * 32 + col) * Dn.n_inst + inst) * Dn.pitch;
        v16f acc[G];
#pragma unroll
        for (uint32_t g = 0; g < G; ++g)
            for (int v = 0; v < 16; ++v) acc[g][v] = 0.f;
        // Chunks of 8 K-steps, double-buffered: the loads of chunk c + 1 are in flight while chunk c is
        // multiplied. Loads are unconditional (beyond the last chunk they repeat it, and K-steps beyond the
        // last repeat its fragment), so the wait before a chunk's products covers only that chunk's loads.
        typedef typename std::conditional<FMT == 0, long, v8s>::type Frag;
        const uint32_t nck = (Dn.ksteps + 7) / 8;
        struct Chunk {
            v4u sb[G];
            Frag b[8];
        };
        auto load = [&](uint32_t ck, Chunk &x) {
            ck = min(ck, nck - 1);
#pragma unroll
            for (uint32_t g = 0; g < G; ++g) x.sb[g] = *(const KNP_GLOBAL v4u *)(brow[g] + 16 * ck);
#pragma unroll
            for (uint32_t e = 0; e < 8; ++e)
                x.b[e] = *(const KNP_GLOBAL Frag *)(fr + (size_t)min(8 * ck + e, Dn.ksteps - 1) * 64 * FB);
        };
        auto multiply = [&](uint32_t ck, const Chunk &x) {
            const uint32_t ne = min(8u, Dn.ksteps - 8 * ck);
#pragma unroll
            for (uint32_t e = 0; e < 8; ++e) {
                if (e >= ne) break;
#pragma unroll
                for (uint32_t g = 0; g < G; ++g) {
                    const uint32_t byte = (x.sb[g][e >> 1] >> (16u * (e & 1u) + 8u * h)) & 0xFFu;
                    if constexpr (FMT == 0)
                        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(spread_fp8_ones(byte), x.b[e], acc[g], 0, 0, 0);
                    else
                        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(spread_bf16_ones(byte), x.b[e], acc[g], 0, 0, 0);
                }
            }
        };
        Chunk c0, c1;
        load(0, c0);
        for (uint32_t ck = 0; ck < nck; ck += 2) {
            load(ck + 1, c1);
            multiply(ck, c0);
            load(ck + 2, c0);
            if (ck + 1 < nck) multiply(ck + 1, c1);
        }
        if (col < ncol) {
            const float sc = gp(Dn.row_scale)[j0 + col];
#pragma unroll
            for (uint32_t g = 0; g < G; ++g) {
                if (g >= nb) break;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const uint32_t m = 8u * (v / 4) + 4u * h + (v % 4);
                    const uint64_t arrival = t0 + (uint64_t)(rb0 + g) * 32 + m + 1;
                    gp(Dn.ff)[(size_t)(arrival % ring) * P.n + j0 + col] = acc[g][v] * sc;
                }
            }
        }
    }
}


// ---- projections expanded on MFMA (ProjDev::dense) ---------------------------------------------------
// A projection whose synapses form dense blocks (instances) is expanded each step as a product of the
// instance's spike counts (K = presynaptic indexes) and a BF16 copy of its weights (N = postsynaptic
// neurons) on the matrix cores, with FP32 accumulation; training keeps the FP32 weights and refreshes the copy.

// Spike counts of step t for a set of MFMA-expanded projections, one launch for all of them: every spiking
// row of a projection's virtual segments adds 1 to its index's byte in the octet's 64-bit count word (a
// non-returning atomic; a channel message repeating an index adds more), synaptic-resource projections note
// the step each row spiked (init_synapse, synaptic_resource_stdp.h:36-62), and the next step's words (read by
// the previous step's product) are cleared. set.cnt: each projection's count buffer ([2][dn_oct][8] bytes:
// its own, or a region of the library's device scratch for knp_mi355x_step_noalloc). The workgroup's impact and row totals go to its own
// counter slot (no shared addresses: every workgroup would queue on them).
__global__ __launch_bounds__(kBlock) void k_dense_counts(const NetDev *__restrict__ net, DenseCountSet set, uint64_t t) {
    t = resolve_step(net, t);
    const uint32_t *vc = net->vcount + (size_t)(t & 1) * net->n_vseg;
    const size_t par = (size_t)(t & 1) * net->list_pstride;
    unsigned long long ev = 0, rows = 0;
    for (uint32_t m = 0; m < set.n; ++m) {
        const ProjDev &Q = net->projs[set.proj[m]];
        const size_t words = (size_t)Q.dn_oct * 2;
        uint32_t *cur = set.cnt[m] + (size_t)(t & 1) * words;
        uint32_t *nxt = set.cnt[m] + (size_t)((t + 1) & 1) * words;
        for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < words; k += (size_t)gridDim.x * kBlock)
            gp(nxt)[k] = 0;
        for (uint32_t v = Q.dn_v0 + blockIdx.x; v < Q.dn_v1; v += gridDim.x) {
            const uint32_t cnt = gp(vc)[v];
            const uint32_t *list = net->seg_list[gp(net->vseg)[v]] + par + 1;
            for (uint32_t e = threadIdx.x; e < cnt; e += kBlock) {
                const uint32_t pre = gp(list)[e];
                if (pre >= Q.n_pre) continue;
                const uint64_t info = gp(Q.dn_pre_info)[pre];
                if (info == UINT64_MAX) continue;
                atomicAdd((unsigned long long *)&cur[(size_t)(uint32_t)info * 2], 1ull << (8u * ((info >> 32) & 7u)));
                if (Q.kind == PROJ_RESOURCE) gp(Q.row_spike)[pre] = t;
                ev += info >> 35;
                rows += 1;
            }
        }
    }
    __shared__ unsigned long long red[2][kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        ev += __shfl_xor(ev, off);
        rows += __shfl_xor(rows, off);
    }
    if (lane_id() == 0) {
        red[0][threadIdx.x / 64] = ev;
        red[1][threadIdx.x / 64] = rows;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned long long x = 0;
        for (int w = 0; w < kBlock / 64; ++w) x += red[threadIdx.x][w];
        if (x) gp(net->projs[set.proj[0]].dn_part)[4 * blockIdx.x + threadIdx.x] += x;
    }
}

// The channel counts of step parity t1 cleared (one workgroup, after a persistent launch whose first step read
// them).
__global__ __launch_bounds__(kTileThreads) void k_clear_push_counts(const NetDev *__restrict__ net, uint64_t t1) {
    clear_push_counts<false>(net, t1);
}

// The device's current step: set by the host before graph replays, advanced at the end of every captured step.
__global__ void k_set_step(const NetDev *__restrict__ net, uint64_t t) {
    if (threadIdx.x == 0) *gp(net->step_ptr) = t;
}
__global__ void k_step_advance(const NetDev *__restrict__ net) {
    if (threadIdx.x == 0) *gp(net->step_ptr) += 1;
}

// p[0 .. n) = v (a kernel, not a memset node, so that captured sequences replay it correctly).
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (uint64_t)gridDim.x * kBlock) gp(p)[k] = v;
}

// 8 spike counts (one byte each, below 256) to 8 bf16 values (exact).
__device__ __forceinline__ v8s counts_to_bf16(uint64_t c) {
    typedef unsigned int v4w __attribute__((ext_vector_type(4)));
    v4w w;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t a = (uint32_t)(c >> (16 * k)) & 0xFFu, b = (uint32_t)(c >> (16 * k + 8)) & 0xFFu;
        const uint32_t fa = __builtin_bit_cast(uint32_t, (float)a) >> 16, fb = __builtin_bit_cast(uint32_t, (float)b) >> 16;
        w[k] = fa | fb << 16;
    }
    return __builtin_bit_cast(v8s, w);
}

// The sum over one projection of a wave's 32-neuron tile: C (32 x 32) += A (16 K: spike counts, the same in
// every row) x B (16 K x 32 neurons: BF16 weights) over the instance's octets, two octets per K-step (lanes
// 0-31 the first, 32-63 the second), four K-steps per round with their loads issued first. SKIP: only octets
// holding a spike. Every column of C is the neuron's sum: the lane's column of row 0 is returned (hi + lo for
// two-part weights). mine0: this lane's count word of the first 64 octets, loaded by the caller.
template <bool SKIP, bool SPLIT>
__device__ __forceinline__ float dense_tile_sum(const ProjDev &Q, const DenseTile &T, const KNP_GLOBAL uint64_t *cnt,
                                                uint64_t mine0, uint32_t lane, uint32_t &read_oct) {
    const uint32_t h = lane >> 5, col = lane & 31, n_oct = T.n_oct;
    const KNP_GLOBAL uint16_t *wb = gp(Q.dn_w) + T.base + (size_t)(T.j0 + col) * 8;
    const KNP_GLOBAL uint16_t *wl = SPLIT ? gp(Q.dn_w_lo) + T.base + (size_t)(T.j0 + col) * 8 : nullptr;
    const size_t ostride = (size_t)T.npad * 8;
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    v16f acc, acc_lo;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = acc_lo[v] = 0.f;
    const v8s zero = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t c0 = 0; c0 < n_oct; c0 += 64) {
        const uint32_t o = c0 + lane;
        const uint64_t mine = c0 == 0 ? mine0 : (o < n_oct ? cnt[o] : 0ull);
        uint64_t mask = SKIP ? __ballot(mine != 0ull) : __ballot(o < n_oct);
        read_oct += (uint32_t)__popcll(mask);
        while (mask) {
            uint32_t oct[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                uint32_t a = kNone, b = kNone;
                if (mask) {
                    a = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
                    mask &= mask - 1ull;
                }
                if (mask) {
                    b = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
                    mask &= mask - 1ull;
                }
                oct[g] = h ? b : a;
            }
            v8s frag[4], frag_lo[4];
            uint64_t cv[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const uint32_t my = oct[g];
                cv[g] = __shfl(mine, (int)(my & 63u));
                if (my == kNone) cv[g] = 0ull;
                frag[g] = my != kNone ? *(const KNP_GLOBAL v8s *)(wb + (size_t)(c0 + my) * ostride) : zero;
                if (SPLIT) frag_lo[g] = my != kNone ? *(const KNP_GLOBAL v8s *)(wl + (size_t)(c0 + my) * ostride) : zero;
            }
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const v8s a = counts_to_bf16(cv[g]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, frag[g], acc, 0, 0, 0);
                if (SPLIT) acc_lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, frag_lo[g], acc_lo, 0, 0, 0);
            }
        }
    }
    return SPLIT ? acc[0] + acc_lo[0] : acc[0];
}

// dense_tile_sum for an instance of at most kDenseStagedOct octets whose count words the wave has staged in LDS
// (cw): the octets holding a spike are taken eight to a round across the whole instance (no round is cut short
// at a 64-octet boundary, and no count word is loaded in the loop).
constexpr uint32_t kDenseStagedOct = 128;

template <bool SKIP, bool SPLIT>
__device__ __forceinline__ float dense_tile_sum_staged(const ProjDev &Q, const DenseTile &T, const uint64_t *cw,
                                                       uint8_t *list, uint32_t lane, uint32_t &read_oct) {
    const uint32_t h = lane >> 5, col = lane & 31, n_oct = T.n_oct;
    // the instance's weights from one base (the instance holds fewer than 2^32 of them): 32-bit lane offsets
    const KNP_GLOBAL uint16_t *wb = gp(Q.dn_w) + T.base;
    const KNP_GLOBAL uint16_t *wl = SPLIT ? gp(Q.dn_w_lo) + T.base : nullptr;
    const uint32_t off0 = (T.j0 + col) * 8, ostride = T.npad * 8;
    // the octets to multiply, in ascending order, listed in LDS (each lane places its own)
    const bool s0 = SKIP ? cw[lane] != 0ull : lane < n_oct;
    const bool s1 = SKIP ? cw[64 + lane] != 0ull : 64 + lane < n_oct;
    const uint64_t m0 = __ballot(s0), m1 = __ballot(s1), below = (1ull << lane) - 1ull;
    const uint32_t n0 = (uint32_t)__popcll(m0), total = n0 + (uint32_t)__popcll(m1);
    if (s0) list[__popcll(m0 & below)] = (uint8_t)lane;
    if (s1) list[n0 + __popcll(m1 & below)] = (uint8_t)(64 + lane);
    __builtin_amdgcn_wave_barrier();
    read_oct += total;
    v16f acc, acc_lo;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = acc_lo[v] = 0.f;
    // rounds of eight octets: lanes 0-31 take the even places of the round, lanes 32-63 the odd ones; a lane
    // without an octet loads octet 0 and multiplies it by zero counts (no branch around the loads)
    for (uint32_t r = 0; r < total; r += 8) {
        v8s frag[4], frag_lo[4];
        uint64_t cv[4];
        uint32_t my[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint32_t idx = r + 2 * g + h;
            my[g] = list[idx < total ? idx : 0u];  // (total > 0 inside the loop)
            if (idx >= total) my[g] = 0u;
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint32_t off = off0 + my[g] * ostride;
            frag[g] = *(const KNP_GLOBAL v8s *)(wb + off);
            if (SPLIT) frag_lo[g] = *(const KNP_GLOBAL v8s *)(wl + off);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) cv[g] = r + 2 * g + h < total ? cw[my[g]] : 0ull;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const v8s a = counts_to_bf16(cv[g]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, frag[g], acc, 0, 0, 0);
            if (SPLIT) acc_lo = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, frag_lo[g], acc_lo, 0, 0, 0);
        }
    }
    return SPLIT ? acc[0] + acc_lo[0] : acc[0];
}

// One wave per 32-neuron tile for a group of MFMA-expanded projections into the same neurons with the same
// delay (their instances cover the same postsynaptic ranges): every tile entry and first count word of the
// group is loaded at once, then each projection's sum is taken in turn, converted to fixed point on its own
// (as it would be alone) and added to the ring cell, which the lane owns (no other launch runs meanwhile) and
// loaded at the start, so that the total is stored without a further round trip.
// STAGED: every instance of the group has at most kDenseStagedOct octets (count words staged in LDS).
template <bool STAGED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_dense_expand(const NetDev *__restrict__ net, DenseGroup grp, uint64_t t) {
    t = resolve_step(net, t);
    const ProjDev &Q0 = net->projs[grp.proj[0]];
    const uint32_t tile = wave_uniform(blockIdx.x * (kBlock / 64) + threadIdx.x / 64);
    if (tile >= Q0.dn_tiles) return;
    const PopDev &P = net->pops[Q0.post_pop];
    const uint32_t lane = lane_id(), h = lane >> 5, col = lane & 31;
    // the count words of the wave's instances (STAGED)
    __shared__ uint64_t s_cw[STAGED ? kBlock / 64 : 1][kDenseGroupMax][kDenseStagedOct];
    __shared__ uint8_t s_list[STAGED ? kBlock / 64 : 1][kDenseGroupMax][kDenseStagedOct];
    uint64_t *cw_wave = &s_cw[STAGED ? threadIdx.x / 64 : 0][0][0];
    uint8_t *list_wave = &s_list[STAGED ? threadIdx.x / 64 : 0][0][0];
    DenseTile T[kDenseGroupMax];
    uint64_t mine0[kDenseGroupMax];
#pragma unroll
    for (uint32_t m = 0; m < kDenseGroupMax; ++m)
        if (m < grp.n) {
            // one tile entry in two 16-byte loads (cbase, n_oct, npad, post_begin | post_count, j0, base)
            const KNP_GLOBAL v4u *e = (const KNP_GLOBAL v4u *)gp(net->projs[grp.proj[m]].dn_tile_info) + 2 * tile;
            const v4u a = e[0], b = e[1];
            T[m] = DenseTile{a[0], a[1], a[2], a[3], b[0], b[1], (uint64_t)b[2] | (uint64_t)b[3] << 32};
        }
    if (STAGED) {
        // every count word of the group loaded at once (lanes past an instance's octets load its last one),
        // then staged in LDS
        uint64_t w[kDenseGroupMax][2];
#pragma unroll
        for (uint32_t m = 0; m < kDenseGroupMax; ++m)
            if (m < grp.n) {
                const ProjDev &Q = net->projs[grp.proj[m]];
                const KNP_GLOBAL uint64_t *cnt =
                    (const KNP_GLOBAL uint64_t *)gp(grp.cnt[m]) + (size_t)(t & 1) * Q.dn_oct + T[m].cbase;
                const uint32_t last = T[m].n_oct - 1;  // an instance has at least one octet
                w[m][0] = cnt[lane < last ? lane : last];
                w[m][1] = cnt[64 + lane < last ? 64 + lane : last];
            }
#pragma unroll
        for (uint32_t m = 0; m < kDenseGroupMax; ++m)
            if (m < grp.n) {
                cw_wave[m * kDenseStagedOct + lane] = lane < T[m].n_oct ? w[m][0] : 0ull;
                cw_wave[m * kDenseStagedOct + 64 + lane] = 64 + lane < T[m].n_oct ? w[m][1] : 0ull;
            }
        __builtin_amdgcn_wave_barrier();  // the wave's count words are in LDS before any lane reads another's
    } else {
#pragma unroll
        for (uint32_t m = 0; m < kDenseGroupMax; ++m)
            if (m < grp.n) {
                const ProjDev &Q = net->projs[grp.proj[m]];
                const KNP_GLOBAL uint64_t *cnt =
                    (const KNP_GLOBAL uint64_t *)gp(grp.cnt[m]) + (size_t)(t & 1) * Q.dn_oct + T[m].cbase;
                mine0[m] = lane < T[m].n_oct ? cnt[lane] : 0ull;
            }
    }
    const uint32_t j0 = T[0].j0;
    const bool writer = h == 0 && j0 + col < T[0].post_count;
    // the ring cell (lanes that write none load the tile's first one)
    const size_t ai = (size_t)((t + Q0.dn_delay) % P.D) * P.n + (T[0].post_begin + j0 + (writer ? col : 0));
    int64_t total = gp(P.acc_v)[ai];
#pragma unroll
    for (uint32_t m = 0; m < kDenseGroupMax; ++m) {
        if (m >= grp.n) break;
        const ProjDev &Q = net->projs[grp.proj[m]];
        const KNP_GLOBAL uint64_t *cnt =
            (const KNP_GLOBAL uint64_t *)gp(grp.cnt[m]) + (size_t)(t & 1) * Q.dn_oct + T[m].cbase;
        const bool skip = (Q.dense - 1) & kDenseSkipEmpty, split = Q.dn_w_lo != nullptr;
        uint32_t read_oct = 0;  // octets whose weights this wave reads
        float sum;
        const uint64_t *cw = cw_wave + m * kDenseStagedOct;
        uint8_t *list = list_wave + m * kDenseStagedOct;
        if (STAGED) {
            if (skip && split) sum = dense_tile_sum_staged<true, true>(Q, T[m], cw, list, lane, read_oct);
            else if (skip) sum = dense_tile_sum_staged<true, false>(Q, T[m], cw, list, lane, read_oct);
            else if (split) sum = dense_tile_sum_staged<false, true>(Q, T[m], cw, list, lane, read_oct);
            else sum = dense_tile_sum_staged<false, false>(Q, T[m], cw, list, lane, read_oct);
        } else if (skip && split) sum = dense_tile_sum<true, true>(Q, T[m], cnt, mine0[m], lane, read_oct);
        else if (skip) sum = dense_tile_sum<true, false>(Q, T[m], cnt, mine0[m], lane, read_oct);
        else if (split) sum = dense_tile_sum<false, true>(Q, T[m], cnt, mine0[m], lane, read_oct);
        else sum = dense_tile_sum<false, false>(Q, T[m], cnt, mine0[m], lane, read_oct);
        const long long v = __float2ll_rn(sum * P.scale);
        total += Q.dn_sign < 0.f ? -v : v;
        // weight bytes read (32 neurons x 8 inputs x 2 bytes per octet, twice for two-part weights): the
        // traffic of the skipping schedule, counted in the slots the counting workgroups use
        if (skip && lane == 0 && read_oct)
            atomicAdd((unsigned long long *)gp(Q.dn_part) + 4 * (blockIdx.x % kDenseCountBlocks) + 2,
                      (unsigned long long)read_oct * (split ? 1024ull : 512ull));
    }
    if (writer) gp(P.acc_v)[ai] = total;
}

}  // namespace knp_mi355x
