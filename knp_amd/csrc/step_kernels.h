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

#include "device_tables.h"

namespace knp_mi355x {

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

__device__ __forceinline__ float pf(const ParamF &p, uint32_t i) { return p.a ? p.a[i] : p.v; }
__device__ __forceinline__ uint32_t pu(const ParamU &p, uint32_t i) { return p.a ? p.a[i] : p.v; }

// n_time_steps_since_last_firing_ kept in 32 bits: 0xFFFFFFFF stands for SIZE_MAX (wraps to 0 on the next
// increment, blifat_impl.h:34 with the default of blifat.h:53); counts saturate at 0xFFFFFFFE.
__device__ __forceinline__ uint32_t ts_inc(uint32_t x) {
    return x == 0xFFFFFFFFu ? 0u : (x >= 0xFFFFFFFDu ? 0xFFFFFFFEu : x + 1u);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Exact "sum >= threshold" when the potential before impacts is 0: the accumulator holds the exact sum
// on a 2^-frac grid, so comparing against the smallest grid point >= thr decides exactly.
__device__ __forceinline__ bool fixed_ge(int64_t a, float thr, float scale, float v_float) {
    float ts = thr * scale;
    if (fabsf(ts) < 9.0e18f) return a >= (long long)ceilf(ts);
    return v_float >= thr;
}

__device__ __forceinline__ float fixed_to_float(int64_t a, float inv_scale) {
    return __ll2float_rn((long long)a) * inv_scale;
}

// Wave-level compaction: lanes with `pred` append `value` to list[*counter ...] with one atomic per wave.
__device__ __forceinline__ void wave_append(bool pred, uint32_t value, uint32_t *counter, uint32_t *list) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return;
    uint32_t lane = lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = value;
}

// Ordered compaction of a workgroup's 256 candidates into its own segment: ascending order, no global
// atomics, one count store. NT = threads of the workgroup (threads from 256 on carry no candidate).
template <uint32_t NT = kBlock>
__device__ __forceinline__ void block_segment_write(const NetDev *__restrict__ net, uint64_t t, bool pred,
                                                    uint32_t value, uint32_t *segment, uint32_t seg) {
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
    if (pred) segment[off + __popcll(mask & ((1ull << lane) - 1ull))] = value;
    if (threadIdx.x < 4) {
        const size_t par = (size_t)(t & 1);
        if (threadIdx.x == 0) net->seg_count[par * net->n_seg + seg] = total;
        const uint32_t vp = net->seg_vpos[4 * seg + threadIdx.x];
        if (vp != 0xFFFFFFFFu) net->vcount[par * net->n_vseg + vp] = total;
    }
    __syncthreads();
}

// Write the wave's 64-neuron spike mask into the bitmap words (and the record), bit i = neuron i.
__device__ __forceinline__ void wave_bitmap(uint64_t mask, uint32_t first_neuron, const PopDev &P, uint64_t t) {
    uint32_t lane = lane_id();
    uint32_t w0 = first_neuron >> 5;
    if (lane < 2 && w0 + lane < P.words) {
        uint32_t word = lane == 0 ? (uint32_t)mask : (uint32_t)(mask >> 32);
        P.bitmap[w0 + lane] = word;
        if (P.rec && t >= P.rec_t0 && t - P.rec_t0 < P.rec_cap) P.rec[(t - P.rec_t0) * P.words + w0 + lane] = word;
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
template <uint32_t KIND, uint32_t NT = kBlock, bool FUSED = false>
__device__ __forceinline__ void population_step(const NetDev *__restrict__ net, uint32_t p, uint32_t local_block,
                                                uint64_t t, Ctl *ctl, uint32_t *seg_count, int64_t a_in = 0) {
    const PopDev &P = net->pops[p];
    const uint32_t i = local_block * kBlock + threadIdx.x;
    const bool active = threadIdx.x < (uint32_t)kBlock && i < P.n;
    const uint32_t slot = (uint32_t)(t % P.D);
    const size_t ai = (size_t)slot * P.n + i;
    bool spike = false;
    bool cand = false;
    if (active) {
        int64_t a = a_in;
        if (!FUSED && P.acc_v) {
            a = P.acc_v[ai];
            P.acc_v[ai] = 0;
        }
        if (P.ff) a += __float2ll_rn(P.ff[(size_t)(t % (2ull * P.ff_T)) * P.n + i] * P.scale);
        if (KIND == POP_LEAN) {
            // default BLIFAT (blifat.h defaults): V_pre = 0, no decays, no refractory period, no blocking.
            const bool can = !(P.lean_ts_max && t == P.t_first);  // step 0 rule (wrap of SIZE_MAX)
            const float thr = pf(P.act, i) + pf(P.add_thr, i);
            float V = fixed_to_float(a, P.inv_scale);
            spike = can && fixed_ge(a, thr, P.scale, V);
            if (spike) {
                V = pf(P.V_reset, i);
                P.last_spike[i] = (uint32_t)t;
            }
            const float vmin = pf(P.V_min, i);
            if (V < vmin) V = vmin;
            P.V[i] = V;
        } else {
            // pre-impact (blifat_impl.h:31-65)
            uint32_t n_ts = ts_inc(P.n_ts[i]);
            float dyn = P.dyn[i] * pf(P.thr_decay, i);
            float trace = P.trace[i] * pf(P.trace_decay, i);
            float g = P.g[i] * pf(P.g_decay, i);
            float V = P.V[i] * pf(P.V_decay, i);
            uint32_t bphase = P.bphase[i];
            if (bphase == 1) {
                if (KIND != POP_RESOURCE) bphase = 0;  // resource neurons keep the phase (blifat_impl.h:62)
                V += pf(P.reflexive, i);
            }
            const float V_pre = V;
            P.Vpre[i] = V_pre;
            // impacts (blifat_impl.h:68-107): exact sums of this step's impacts
            int64_t tbp = P.track_tbp ? P.tbp[i] : INT64_MAX;
            if (P.acc_g) {
                int64_t ag = P.acc_g[ai];
                P.acc_g[ai] = 0;
                if (ag) g += fixed_to_float(ag, P.inv_scale);
            }
            float dop_now = 0.f;
            if (P.acc_d) {
                int64_t ad = P.acc_d[ai];
                P.acc_d[ai] = 0;
                dop_now = fixed_to_float(ad, P.inv_scale);
            }
            uint32_t forced = 0;
            if (P.acc_f) {
                forced = P.acc_f[ai];
                P.acc_f[ai] = 0;
            }
            if (P.acc_b) {
                uint64_t key = P.acc_b[ai];
                P.acc_b[ai] = 0;
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
            const float add = (KIND == POP_RESOURCE) ? P.res.add_thr[i] : pf(P.add_thr, i);
            const float thr = pf(P.act, i) + dyn + add;
            const bool above = exact ? fixed_ge(a, thr, P.scale, V) : (V >= thr);
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
            P.V[i] = V;
            P.n_ts[i] = n_ts;
            P.dyn[i] = dyn;
            P.trace[i] = trace;
            P.g[i] = g;
            P.bphase[i] = bphase;
            if (P.track_tbp) P.tbp[i] = tbp;
            if (KIND == POP_RESOURCE) {
                const ResDev &R = P.res;
                R.dop_now[i] = dop_now;  // dopamine_value_ is reset every step for resource neurons
                R.forced_now[i] = (uint8_t)(forced != 0);
                // training candidate: spiked, has dopamine, or due for resource renormalization
                // (the renormalization test only depends on state training leaves unchanged for others)
                bool renorm = false;
                const float fr = R.free_res[i];
                if (fabsf(fr) >= pf(R.res_thr, i)) {
                    const uint8_t st = R.isi_status[i];
                    const bool in_isi = (t - R.last_step[i] <= (uint64_t)pu(R.isi_max, i)) && st != 0;
                    renorm = !in_isi;
                }
                cand = spike || dop_now != 0.f || renorm;
            } else if (P.acc_d) {
                P.dop[i] += dop_now;  // plain BLIFAT keeps summing dopamine (never reset for it)
            }
        }
    }
    // spike compaction and bitmaps (whole wave participates)
    block_segment_write<NT>(net, t, spike, i, P.spike_list + (size_t)(t & 1) * net->list_pstride + (size_t)local_block * kBlock,
                            P.seg_base + local_block);
    const uint64_t mask = __ballot(spike);
    if (threadIdx.x < (uint32_t)kBlock) wave_bitmap(mask, local_block * kBlock + (threadIdx.x & ~63u), P, t);
    if (KIND == POP_RESOURCE) wave_append(cand, i, &ctl->cand_count[p], P.res.cand);
}

template <uint32_t NT = kBlock>
__device__ __forceinline__ void source_step(const NetDev *__restrict__ net, uint32_t s, uint32_t local_block,
                                            uint64_t t, uint32_t *seg_count) {
    const SourceDev &S = net->srcs[s];
    const uint32_t c = local_block * kBlock + threadIdx.x;
    const uint32_t phase = (uint32_t)(t % S.steps_per_pattern);
    bool spike = false;
    if (threadIdx.x < (uint32_t)kBlock && c < S.n_ch && phase >= S.lo && phase <= S.hi) {
        const uint32_t pattern = (uint32_t)((t / S.steps_per_pattern) % S.n_patterns);
        spike = bernoulli_hash(S.seed, t, S.hash_offset + c) < S.thr[(size_t)pattern * S.n_ch + c];
    }
    block_segment_write<NT>(net, t, spike, S.first + c, S.list + (size_t)(t & 1) * net->list_pstride + (size_t)local_block * kBlock,
                            S.seg_base + local_block);
}

__global__ __launch_bounds__(kBlock) void k_phase1(const NetDev *__restrict__ net, uint64_t t) {
    Ctl *ctl = &net->ctl[t & 1];
    uint32_t *seg_count = net->seg_count + (size_t)(t & 1) * net->n_seg;
    if (blockIdx.x == 0) {
        uint32_t *nxt = (uint32_t *)&net->ctl[(t + 1) & 1];
        for (uint32_t k = threadIdx.x; k < sizeof(Ctl) / 4; k += blockDim.x) nxt[k] = 0;
        // next step's counts: host-pushed segments must start empty (every other segment is rewritten)
        uint32_t *nc = net->seg_count + (size_t)((t + 1) & 1) * net->n_seg;
        uint32_t *nv = net->vcount + (size_t)((t + 1) & 1) * net->n_vseg;
        for (uint32_t k = threadIdx.x; k < net->n_projs; k += blockDim.x) {
            const uint32_t ps = net->projs[k].push_seg;
            nc[ps] = 0;
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t vp = net->seg_vpos[4 * ps + j];
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
__device__ __forceinline__ void atomic_add_i64(int64_t *p, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}

__device__ __forceinline__ void impact(const PopDev &P, const ProjDev &Q, const RowSource &R, uint32_t slot0,
                                       uint32_t pre, uint64_t pos, uint32_t pt, uint32_t d, float w) {
    const uint32_t post = pt & 0x1FFFFFFFu;
    const uint32_t type = pt >> 29;
    uint32_t slot = slot0 + d;
    if (slot >= P.D) slot -= P.D;
    const size_t ai = (size_t)slot * P.n + post;
    const long long v = __float2ll_rn(w * P.scale);
    switch (type) {
        case 0:  // EXCITATORY
            atomic_add_i64(&P.acc_v[ai], v);
            if (Q.forcing && P.acc_f) atomicOr(&P.acc_f[ai], 1u);
            break;
        case 1:  // INHIBITORY_CURRENT
            atomic_add_i64(&P.acc_v[ai], -v);
            break;
        case 2:  // INHIBITORY_CONDUCTANCE
            atomic_add_i64(&P.acc_g[ai], v);
            break;
        case 3:  // DOPAMINE
            atomic_add_i64(&P.acc_d[ai], v);
            break;
        default: {  // BLOCKING
            const uint64_t key = 1ull + (((uint64_t)Q.blocking_id << 57) | ((uint64_t)(64u - d) << 51) |
                                         ((uint64_t)R.rank << 46) | ((uint64_t)pre << 17) | pos);
            atomicMax((unsigned long long *)&P.acc_b[ai], (unsigned long long)key);
            break;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(const NetDev *__restrict__ net, uint64_t t) {
    extern __shared__ uint32_t pfx[];  // [n_vseg + 1]
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t nv = net->n_vseg_atomic;
    const uint32_t *vc = net->vcount + (size_t)(t & 1) * net->n_vseg;
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    for (uint32_t v = threadIdx.x; v < nv; v += kBlock) pfx[v] = vc[v];
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
        const uint32_t seg = net->vseg[lo];
        const RowSource &R = net->rows[net->vrs[lo]];
        const uint32_t pre = net->seg_list[seg][(size_t)(t & 1) * net->list_pstride + row - pfx[lo]];
        const ProjDev &Q = net->projs[R.proj];
        if (pre >= Q.n_pre) continue;
        const uint64_t b = Q.row_ptr[pre], e = Q.row_ptr[pre + 1];
        my_ev += e - b;
        my_rows += 1;
        const PopDev &P = net->pops[Q.post_pop];
        const uint32_t slot0 = (uint32_t)(t % P.D);
        for (uint64_t base = b; base < e; base += 4 * kBlock) {
            uint32_t pt[4], d[4];
            float wv[4];
            bool ok[4];
            for (int j = 0; j < 4; ++j) {
                const uint64_t k = base + threadIdx.x + (uint64_t)j * kBlock;
                ok[j] = k < e;
                if (ok[j]) {
                    pt[j] = Q.post_type[k];
                    d[j] = Q.delay[k];
                    if (Q.kind == PROJ_DELTA) {
                        wv[j] = Q.w_csr[k];
                    } else {
                        const uint32_t s = Q.csr_syn[k];
                        wv[j] = Q.W[s];
                        if (Q.kind == PROJ_RESOURCE) Q.last_spike_step[s] = t;  // init_synapse
                    }
                }
            }
            for (int j = 0; j < 4; ++j)
                if (ok[j]) impact(P, Q, R, slot0, pre, base + threadIdx.x + (uint64_t)j * kBlock - b, pt[j], d[j], wv[j]);
        }
    }
    if (threadIdx.x == 0) {
        net->part[2 * blockIdx.x] += my_ev;
        net->part[2 * blockIdx.x + 1] += my_rows;
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
    const uint32_t *tile_dir;
    const RowMeta *row_meta;
    const uint32_t *row_dir;
    const uint64_t *row_ptr;
    uint32_t stride, n_pre;
};

// Dynamic LDS of k_tile_scatter: the ring sums [tile_max_d][256] (int64), then for the tile's segments their
// spike lists [tile_max_list] (pointers), row-count prefix [tile_max_list + 1] and projection [tile_max_list].
__host__ __device__ inline size_t tile_lds_bytes(uint32_t max_d, uint32_t max_list) {
    return (size_t)max_d * kBlock * 8 + (size_t)max_list * 8 + ((size_t)max_list + 1) * 4 + max_list;
}

// PHASE1: the same launch also runs phase 1 of step t + 1 for the tile's neurons (their ring slot for t + 1
// is complete once this step's impacts are summed) and, in the workgroups after the tiles, the Bernoulli
// sources of step t + 1; the host uses it when every population is tiled, saving one launch per step.
template <bool PHASE1>
__global__ __launch_bounds__(kTileThreads) void k_tile_scatter(const NetDev *__restrict__ net, uint64_t t) {
    extern __shared__ unsigned long long tacc[];
    __shared__ uint32_t rbeg[kTileThreads];  // CSR position of each row's part in the tile
    __shared__ uint32_t rpfx[kTileThreads];  // exclusive prefix of the parts' lengths
    __shared__ uint8_t rproj[kTileThreads];  // projection of each row
    __shared__ uint32_t wsum[kTileThreads / 64];
    __shared__ unsigned long long cnt[3];
    __shared__ TileProj tp[kMaxProjs];
    if (blockIdx.x >= net->n_tiles) {  // Bernoulli sources of step t + 1 (PHASE1 grids only)
        uint32_t b = blockIdx.x - net->n_tiles;
        for (uint32_t s = 0; s < net->n_srcs; ++s) {
            const uint32_t nb = max(1u, (net->srcs[s].n_ch + kBlock - 1) / kBlock);
            if (b < nb) {
                source_step<kTileThreads>(net, s, b, t + 1, nullptr);
                break;
            }
            b -= nb;
        }
        return;
    }
    if (blockIdx.x == 0) {
        // channel segments of step t + 1 start empty (the host writes the counts of pushed messages later)
        uint32_t *nc = net->seg_count + (size_t)((t + 1) & 1) * net->n_seg;
        uint32_t *nvc = net->vcount + (size_t)((t + 1) & 1) * net->n_vseg;
        for (uint32_t k = threadIdx.x; k < net->n_projs; k += kTileThreads) {
            const uint32_t ps = net->projs[k].push_seg;
            nc[ps] = 0;
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t vp = net->seg_vpos[4 * ps + j];
                if (vp != 0xFFFFFFFFu) nvc[vp] = 0;
            }
        }
    }
    const TileDev td = net->tiles[blockIdx.x];
    const PopDev &P = net->pops[td.pop];
    // population fields in registers: stores through the ring would otherwise force reloads of the table
    const uint32_t D = P.D, n_pop = P.n;
    int64_t *const acc_v = P.acc_v;
    const float scale = P.scale;
    const uint32_t tid = threadIdx.x;
    const uint32_t L = net->tile_max_list;
    unsigned long long *ring = tacc;
    const uint32_t **vlist = reinterpret_cast<const uint32_t **>(tacc + (size_t)net->tile_max_d * kBlock);
    uint32_t *vpfx = reinterpret_cast<uint32_t *>(vlist + L);
    uint8_t *vproj = reinterpret_cast<uint8_t *>(vpfx + L + 1);
    for (uint32_t k = tid; k < D * kBlock; k += kTileThreads) ring[k] = 0;
    if (tid < 3) cnt[tid] = 0;
    if (tid < net->n_projs) {
        const ProjDev &Q = net->projs[tid];
        tp[tid] = TileProj{Q.syn8, Q.tile_dir, Q.row_meta, Q.row_dir, Q.row_ptr, Q.tile_stride, Q.n_pre};
    }
    // 1. the spiking rows that may reach the tile, numbered through a prefix over its segments' counts
    const uint32_t nv = td.vend - td.vbeg;
    const TileSeg *ts = net->tile_seg + td.vbeg;
    const uint32_t *vc = net->vcount + (size_t)(t & 1) * net->n_vseg;
    uint32_t total = 0;
    for (uint32_t base = 0; base < nv; base += kTileThreads) {
        const uint32_t v = base + tid;
        uint32_t c = 0;
        if (v < nv) {
            const TileSeg e = ts[v];
            vlist[v] = e.list + (size_t)(t & 1) * net->list_pstride;
            vproj[v] = (uint8_t)e.proj;
            c = vc[e.vpos];
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan<kTileThreads>(c, wsum, &tot);
        if (v < nv) vpfx[v] = total + ex;
        total += tot;
    }
    __syncthreads();
    const uint32_t tile_lo = td.tile * kBlock;
    const uint32_t slot0 = (uint32_t)(t % D);
    uint32_t my_ev = 0, my_parts = 0;
    // exact sum of one packed synapse record into the LDS ring (false: a neuron of another tile)
    auto add_record = [&](uint64_t rec) -> bool {
        const uint32_t j = (uint32_t)(rec >> 41) - tile_lo;
        if (j >= (uint32_t)kBlock) return false;  // a short row read whole: neuron of another tile
        uint32_t slot = slot0 + ((uint32_t)(rec >> 32) & 0xFFu);
        if (slot >= D) slot -= D;
        long long v = __float2ll_rn(__uint_as_float((uint32_t)rec) * scale);
        if ((rec >> 40) & 1u) v = -v;  // INHIBITORY_CURRENT, the only other type a tiled population receives
        atomicAdd(&ring[slot * kBlock + j], (unsigned long long)v);
        return true;
    };
    for (uint32_t rb = 0; rb < total; rb += kTileThreads) {
        // 2. one thread per row: where the row's part for this tile lies
        const uint32_t row = rb + tid;
        uint32_t b = 0, len = 0, q = 0;
        if (row < total) {
            uint32_t lo = 0, hi = nv;  // largest v with vpfx[v] <= row
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (vpfx[mid] <= row) lo = mid;
                else hi = mid;
            }
            q = vproj[lo];
            const uint32_t pre = vlist[lo][row - vpfx[lo]];
            const TileProj &Q = tp[q];
            if (pre < Q.n_pre) {
                if (Q.tile_dir) {
                    const uint32_t *d = Q.tile_dir + (size_t)pre * Q.stride + td.tile;
                    const uint32_t d0 = d[0], d1 = d[1];
                    b = d0;
                    len = d1 - d0;
                } else if (Q.row_meta) {
                    const RowMeta m = Q.row_meta[pre];
                    const uint32_t k = td.tile - m.first_tile;  // wraps above n_tiles for tiles before the row
                    if (k < m.n_tiles) {
                        const uint32_t o0 = Q.row_dir[m.dir + k], o1 = Q.row_dir[m.dir + k + 1];
                        b = m.begin + o0;
                        len = o1 - o0;
                    }
                } else {
                    const uint64_t r0 = Q.row_ptr[pre], r1 = Q.row_ptr[pre + 1];
                    b = (uint32_t)r0;
                    len = (uint32_t)(r1 - r0);
                }
            }
        }
        my_parts += len ? 1u : 0u;
        // 3a. short parts (the common case): the row's thread loads its records, 8 at a time, then adds them
        if (len <= kTileShortPart) {
            const uint64_t *src = tp[q].syn8 + b;
            for (uint32_t k0 = 0; k0 < len; k0 += 8) {
                uint64_t rec[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u)
                    if (k0 + u < len) rec[u] = src[k0 + u];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u)
                    if (k0 + u < len && add_record(rec[u])) ++my_ev;
            }
        }
        // 3b. long parts: spread over the workgroup, one thread per synapse
        const uint32_t llen = len > kTileShortPart ? len : 0u;
        if (__syncthreads_or(llen != 0)) {
            uint32_t n_syn;
            const uint32_t ex = block_excl_scan<kTileThreads>(llen, wsum, &n_syn);
            rbeg[tid] = b;
            rpfx[tid] = ex;
            rproj[tid] = (uint8_t)q;
            __syncthreads();
            const uint32_t nrows = min((uint32_t)kTileThreads, total - rb);
            for (uint32_t s = tid; s < n_syn; s += kTileThreads) {
                uint32_t lo = 0, hi = nrows;  // largest r with rpfx[r] <= s: a row whose part holds s
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (rpfx[mid] <= s) lo = mid;
                    else hi = mid;
                }
                if (add_record(tp[rproj[lo]].syn8[rbeg[lo] + (s - rpfx[lo])])) ++my_ev;
            }
        }
        __syncthreads();
    }
    // 4. add the tile's sums into the ring; no other thread of the device writes these cells in this kernel.
    // All loads first, then all stores, so the round trips overlap.
    constexpr uint32_t kCells = kTileMaxD * kBlock / kTileThreads;
    unsigned long long x[kCells];
    int64_t old[kCells];
    int64_t *cell[kCells];
    uint32_t my_cells = 0;
    const uint32_t s1 = PHASE1 ? (uint32_t)((t + 1) % D) : 0xFFFFFFFFu;  // consumed below, not written back
#pragma unroll
    for (uint32_t u = 0; u < kCells; ++u) {
        const uint32_t c = tid + u * kTileThreads;
        const uint32_t i = tile_lo + (c % kBlock);
        x[u] = 0;
        cell[u] = nullptr;
        if (c < D * kBlock && i < n_pop && c / kBlock != s1) {
            x[u] = ring[c];
            cell[u] = acc_v + (size_t)(c / kBlock) * n_pop + i;
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < kCells; ++u) old[u] = x[u] ? *cell[u] : 0;
#pragma unroll
    for (uint32_t u = 0; u < kCells; ++u)
        if (x[u]) {
            *cell[u] = old[u] + (long long)x[u];
            ++my_cells;
        }
    my_ev = wave_sum_u32(my_ev);
    my_parts = wave_sum_u32(my_parts);
    my_cells = wave_sum_u32(my_cells);
    if (lane_id() == 0) {
        atomicAdd(&cnt[0], (unsigned long long)my_ev);
        atomicAdd(&cnt[1], (unsigned long long)my_parts);
        atomicAdd(&cnt[2], (unsigned long long)my_cells);
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long *tpart = net->tile_part + 4 * (size_t)blockIdx.x;
        tpart[0] += cnt[0];
        tpart[2] += cnt[1];
        tpart[3] += cnt[2];
    }
    // rows expanded this step by the tiled path: block 0 sums the counts of every tiled segment
    if (blockIdx.x == 0) {
        uint32_t r = 0;
        for (uint32_t v = net->n_vseg_atomic + tid; v < net->n_vseg; v += kTileThreads) r += vc[v];
        r = wave_sum_u32(r);
        if (lane_id() == 0) wsum[tid / 64] = r;
        __syncthreads();
        if (tid == 0) {
            unsigned long long rows = 0;
            for (uint32_t k = 0; k < kTileThreads / 64; ++k) rows += wsum[k];
            net->tile_part[1] += rows;
        }
    }
    if (PHASE1) {
        // 5. phase 1 of step t + 1: slot s1 = earlier steps' sums in HBM + this step's sums in LDS
        const uint64_t t1 = t + 1;
        const uint32_t i = tile_lo + tid;
        int64_t a = 0;
        if (tid < (uint32_t)kBlock && i < n_pop) {
            int64_t *c1 = acc_v + (size_t)s1 * n_pop + i;
            a = *c1 + (int64_t)ring[s1 * kBlock + tid];
            *c1 = 0;
        }
        Ctl *ctl = &net->ctl[t1 & 1];
        uint32_t *seg_count = net->seg_count + (size_t)(t1 & 1) * net->n_seg;
        if (P.kind == POP_LEAN)
            population_step<POP_LEAN, kTileThreads, true>(net, td.pop, td.tile, t1, ctl, seg_count, a);
        else
            population_step<POP_FULL, kTileThreads, true>(net, td.pop, td.tile, t1, ctl, seg_count, a);
    }
}

// ---- synaptic-resource STDP (blifat_stdp.h:35-172, training/stdp.h:40-186) ---------------------------
__device__ __forceinline__ float recalc_weight(float r, float w_min, float w_max) {
    // recalculate_synapse_weights (training/stdp.h:40-53)
    const float syn_w = r > 0.f ? r : 0.f;
    const float weight_diff = w_max - w_min;
    return w_min + weight_diff * syn_w / (weight_diff + syn_w);
}

__device__ __forceinline__ float stability_factor(float stability) {
    // std::min(static_cast<float>(std::pow(2, -stability)), 1.F)
    const float v = exp2f(-stability);
    return v < 1.f ? v : 1.f;
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
// depends on that order, so one lane replays it: free += d_u[l]; if contributing: free += delta.
__device__ __forceinline__ float serial_free(float freer, uint64_t active, uint64_t contrib, const float *du_lds,
                                             float delta, bool use_du) {
    if (lane_id() == 0) {
        uint64_t m = use_du ? active : contrib;
        while (m) {
            const int l = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            if (use_du) freer += du_lds[l];
            if ((contrib >> l) & 1ull) freer += delta;
        }
    }
    return __shfl(freer, 0);
}

__global__ __launch_bounds__(kBlock) void k_train(const NetDev *__restrict__ net, uint32_t p, uint64_t t) {
    __shared__ float du_lds[kBlock / 64][64];
    const PopDev &P = net->pops[p];
    const ResDev &R = P.res;
    const Ctl *ctl = &net->ctl[t & 1];
    const uint32_t count = ctl->cand_count[p];
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x / 64;
    float *du_w = du_lds[wv];
    const uint32_t waves = gridDim.x * (kBlock / 64);
    for (uint32_t c = blockIdx.x * (kBlock / 64) + wv; c < count; c += waves) {
        const uint32_t n = R.cand[c];
        const bool spiked = (P.bitmap[n >> 5] >> (n & 31)) & 1u;
        const float dop = R.dop_now[n];
        const bool forced = R.forced_now[n] != 0;
        uint32_t status = R.isi_status[n];
        uint64_t last_step = R.last_step[n], lss = R.last_spike_step[n], fis = R.first_isi_spike[n];
        float stab = R.stability[n];
        float freer = R.free_res[n];
        const uint32_t isi_max = pu(R.isi_max, n);
        // ---- process_spiking_neurons_impl
        if (spiked) {
            lss = t;
            if (forced) {
                status = 0;  // is_forced
            } else {
                switch (status) {
                    case 3:  // not_in_period
                    case 0:  // is_forced
                        status = 1;
                        fis = t;
                        break;
                    case 1:  // period_started
                        if (last_step - t < (uint64_t)isi_max) status = 2;
                        break;
                    case 2:  // period_continued
                        if (last_step - t >= (uint64_t)isi_max || dop != 0.f) {
                            status = 1;
                            fis = t;
                        }
                        break;
                    default: break;
                }
                last_step = t;
            }
            if (status == 1) stab -= pf(R.change_at_isi, n);
            const float d_h = pf(R.d_h, n) * stability_factor(stab);
            long long sum = 0;
            for (uint32_t j = 0; j < R.n_work; ++j) {
                const ProjDev &Q = net->projs[R.work[j]];
                const uint64_t cb = Q.col_ptr[n], ce = Q.col_ptr[n + 1];
                for (uint64_t base = cb; base < ce; base += 64) {
                    const uint64_t k = base + lane;
                    const bool act = k < ce;
                    bool contributing = false;
                    float du = 0.f;
                    if (act) {
                        const uint32_t s = Q.col_syn[k];
                        const float w = Q.W[s];
                        if (w > 0.f) sum += __float2ll_rn(w * P.scale);
                        uint8_t fl = Q.flags[s];
                        bool hc = fl & 2, hh = fl & 1;
                        const uint64_t per = Q.period[s];
                        const bool had_spike = in_interval(t - per, t, Q.last_spike_step[s] + Q.delay_syn[s] - 1);
                        if (status != 2 || had_spike) hc = had_spike;
                        if (status != 2) hh = false;
                        float r = Q.r[s];
                        if (status != 0) {
                            du = Q.d_u[s];
                            r -= du;
                            if (hc && !hh) {
                                r += d_h;
                                hh = true;
                                contributing = true;
                            }
                            Q.r[s] = r;
                        }
                        Q.W[s] = recalc_weight(r, Q.w_min[s], Q.w_max[s]);
                        Q.flags[s] = (uint8_t)((hc ? 2 : 0) | (hh ? 1 : 0));
                    }
                    if (status != 0) {
                        du_w[lane] = du;
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                        const uint64_t am = __ballot(act), cm = __ballot(contributing);
                        freer = serial_free(freer, am, cm, du_w, -d_h, true);
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
            sum = wave_sum_i64(sum);
            if (lane == 0) R.add_thr[n] = fixed_to_float(sum, P.inv_scale) * pf(R.coeff, n);
        }
        // ---- do_dopamine_plasticity_impl
        if (dop > 0.f || (dop < 0.f && status != 0)) {
            const bool time_ok = (t - lss) <= (uint64_t)pu(R.dpt, n);
            const float rc = dop * stability_factor(stab);
            for (uint32_t j = 0; j < R.n_work; ++j) {
                const ProjDev &Q = net->projs[R.work[j]];
                const uint64_t cb = Q.col_ptr[n], ce = Q.col_ptr[n + 1];
                for (uint64_t base = cb; base < ce; base += 64) {
                    const uint64_t k = base + lane;
                    bool contributing = false;
                    if (k < ce) {
                        const uint32_t s = Q.col_syn[k];
                        float r = Q.r[s];
                        if (time_ok && (Q.flags[s] & 2)) {
                            r += rc;
                            Q.r[s] = r;
                            contributing = true;
                        }
                        Q.W[s] = recalc_weight(r, Q.w_min[s], Q.w_max[s]);
                    }
                    const uint64_t cm = __ballot(contributing);
                    freer = serial_free(freer, cm, cm, du_w, -rc, false);
                }
            }
            if (forced || dop < 0.f) {
                stab = fmaf(-dop, pf(R.change_param, n), stab);
                stab = stab > 0.f ? stab : 0.f;
            } else {
                const float difference = (float)(t - fis - (uint64_t)isi_max);
                float f = 2.f - difference / (float)isi_max;
                f = f > -1.f ? f : -1.f;
                stab = stab + pf(R.change_param, n) * dop * f;
            }
        }
        // ---- renormalize_resource
        const bool in_isi = (t - last_step <= (uint64_t)isi_max) && status != 0;
        if (!in_isi && fabsf(freer) >= pf(R.res_thr, n)) {
            uint64_t k_total = 0;
            for (uint32_t j = 0; j < R.n_work; ++j) {
                const ProjDev &Q = net->projs[R.work[j]];
                k_total += Q.col_ptr[n + 1] - Q.col_ptr[n];
            }
            const float add = freer / (float)(k_total + (uint64_t)pu(R.drain, n));
            for (uint32_t j = 0; j < R.n_work; ++j) {
                const ProjDev &Q = net->projs[R.work[j]];
                const uint64_t cb = Q.col_ptr[n], ce = Q.col_ptr[n + 1];
                for (uint64_t k = cb + lane; k < ce; k += 64) {
                    const uint32_t s = Q.col_syn[k];
                    const float r = Q.r[s] + add;
                    Q.r[s] = r;
                    Q.W[s] = recalc_weight(r, Q.w_min[s], Q.w_max[s]);
                }
            }
            freer = 0.f;
        }
        if (lane == 0) {
            R.isi_status[n] = (uint8_t)status;
            R.last_step[n] = last_step;
            R.last_spike_step[n] = lss;
            R.first_isi_spike[n] = fis;
            R.stability[n] = stab;
            R.free_res[n] = freer;
        }
    }
}

// ---- additive STDP (additive_stdp.h:93-196) ----------------------------------------------------------
// init_projection: spikes of the projection's stdp_populations_ append the step to the spike histories of
// every synapse whose POSTsynaptic index equals the spiking neuron's index (additive_stdp.h:149-160).
__global__ __launch_bounds__(kBlock) void k_additive_append(const NetDev *__restrict__ net, uint32_t q, uint64_t t) {
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
            const uint32_t nidx = SP.spike_list[(size_t)(t & 1) * net->list_pstride + (size_t)chunk * kBlock + r];
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

// Spike bits of steps t0 .. t0+T-1: byte b of (step, instance) holds inputs 8b .. 8b+7.
__global__ __launch_bounds__(kBlock) void k_dense_spikes(const NetDev *__restrict__ net, uint32_t d, uint64_t t0) {
    const DenseDev &Dn = net->dense[d];
    const uint32_t bytes_per = 2u * Dn.ksteps;
    const uint64_t total = (uint64_t)Dn.T * Dn.n_inst * bytes_per;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < total; base += stride) {
        const uint64_t idx = base + threadIdx.x;
        uint32_t v = 0;
        unsigned long long ev = 0, rows = 0;
        if (idx < total) {
            const uint32_t byte = (uint32_t)(idx % bytes_per);
            const uint64_t rest = idx / bytes_per;
            const uint32_t inst = (uint32_t)(rest % Dn.n_inst);
            const uint64_t t = t0 + rest / Dn.n_inst;
            for (uint32_t b = 0; b < 8; ++b) {
                const uint32_t i = byte * 8 + b;
                if (i < Dn.n_in && bernoulli_hash(Dn.seed, t, Dn.hash_offset + inst * Dn.n_in + i) < Dn.threshold)
                    v |= 1u << b;
            }
            Dn.bits[idx] = (uint8_t)v;
            rows = __popc(v);
            ev = rows * Dn.inst_post_count[inst];
        }
        for (int off = 32; off > 0; off >>= 1) {
            ev += __shfl_xor(ev, off);
            rows += __shfl_xor(rows, off);
        }
        if (lane_id() == 0 && rows) {
            atomicAdd(&net->events[0], ev);
            atomicAdd(&net->events[1], rows);
        }
    }
}

__device__ __forceinline__ long spread_fp8_ones(uint32_t byte) {
    uint64_t x = byte;
    x = (x | (x << 28)) & 0x0000000F0000000Full;
    x = (x | (x << 14)) & 0x0003000300030003ull;
    x = (x | (x << 7)) & 0x0101010101010101ull;
    return (long)(x * 0x38ull);  // fp8 e4m3 1.0 in every byte whose bit is set
}

__device__ __forceinline__ v8s spread_bf16_ones(uint32_t byte) {
    v8s r;
    for (int e = 0; e < 8; ++e) r[e] = ((byte >> e) & 1u) ? (short)0x3F80 : (short)0;
    return r;
}

// One wave per 32-column tile; it walks the block's 32-step row tiles. Output element (row m, column n) of
// the 32x32 accumulator sits in lane n + 32 * ((m / 4) % 2), register 4 * (m / 8) + m % 4.
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_dense_mfma(const NetDev *__restrict__ net, uint32_t d, uint64_t t0) {
    const DenseDev &Dn = net->dense[d];
    const PopDev &P = net->pops[Dn.post_pop];
    constexpr uint32_t FB = FMT == 0 ? 8 : 16;  // fragment bytes per lane per K-step
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (kBlock / 64);
    const uint32_t bytes_per = 2u * Dn.ksteps;
    const uint64_t ring = 2ull * Dn.T;
    for (uint32_t tile = blockIdx.x * (kBlock / 64) + threadIdx.x / 64; tile < Dn.n_tiles; tile += waves) {
        const uint32_t inst = Dn.tile_inst[tile];
        const uint32_t j0 = Dn.tile_j0[tile];
        const uint32_t ncol = Dn.tile_ncol[tile];
        const uint8_t *fr = Dn.frag + (size_t)tile * Dn.ksteps * 64 * FB + (size_t)lane * FB;
        const uint32_t col = lane & 31;
        const float sc = col < ncol ? Dn.row_scale[j0 + col] : 0.f;
        for (uint32_t tt = 0; tt < Dn.T / 32; ++tt) {
            v16f acc;
            for (int v = 0; v < 16; ++v) acc[v] = 0.f;
            const uint32_t row = tt * 32 + (lane & 31);
            const uint8_t *brow = Dn.bits + ((size_t)row * Dn.n_inst + inst) * bytes_per + (lane >> 5);
            for (uint32_t kk = 0; kk < Dn.ksteps; ++kk) {
                const uint32_t sb = brow[2 * kk];
                if (FMT == 0) {
                    const long a = spread_fp8_ones(sb);
                    const long b = *(const long *)(fr + (size_t)kk * 64 * FB);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a, b, acc, 0, 0, 0);
                } else {
                    const v8s a = spread_bf16_ones(sb);
                    const v8s b = *(const v8s *)(fr + (size_t)kk * 64 * FB);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
                }
            }
            if (col < ncol) {
                for (int v = 0; v < 16; ++v) {
                    const uint32_t m = 8u * (v / 4) + 4u * (lane >> 5) + (v % 4);
                    const uint64_t arrival = t0 + tt * 32 + m + 1;
                    Dn.ff[(size_t)(arrival % ring) * P.n + j0 + col] = acc[v] * sc;
                }
            }
        }
    }
}

}  // namespace knp_mi355x
