// knp_mi355x.hip - host runtime and C ABI of the MI355X backend (see include/knp_mi355x.h).
//
// The runtime owns one HIP stream per network handle. Loading copies the caller's reference-layout
// (AoS) populations and projections, _init (knp_mi355x_init) converts them into device structure-of-arrays
// state, CSR/CSC synapse tables and per-population delay rings, and every step then runs on the device
// with no host synchronization (reference step: knp/backends/cpu/cpu-single-threaded-backend/impl/
// backend.cpp:89-141).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <fstream>
#include <limits>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/knp_mi355x.h"
#include "device_tables.h"
#include "step_kernels.h"

using namespace knp_mi355x;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return fail(KNP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct HostPop {
    bool resource = false;
    uint32_t model = MODEL_BLIFAT;  // LIF: state in nl, AltAI-LIF: in na (nb stays empty for both)
    uint32_t n = 0;
    std::vector<knp_blifat_neuron> nb;           // initial state as loaded
    std::vector<knp_resource_blifat_neuron> nr;  // resource populations
    std::vector<knp_lif_neuron> nl;              // LIF populations
    std::vector<knp_altai_lif_neuron> na;        // AltAI-LIF populations
    uint32_t kind = POP_FULL;
    uint32_t D = 1;
    bool has_cond = false, has_dop = false, has_block = false, has_forcing = false;
    int32_t frac_bits = 48;
    uint64_t inexact = 0;  // loaded synapse weights into it that are not multiples of 2^-frac_bits
    bool tiled = false;  // impacts summed by the tile-owner scatter
    uint64_t t_first = 0;
    bool ran = false;
    uint64_t rec_cap = 0, rec_t0 = 0;
};

struct HostProj {
    uint32_t kind = PROJ_DELTA;
    uint32_t pre_pop = KNP_NO_POPULATION, n_pre = 0, post_pop = 0;
    uint64_t n_syn = 0;
    std::vector<uint32_t> pre, post, delay;
    std::vector<float> w;
    std::vector<uint8_t> type;
    std::vector<knp_resource_rule> rules;
    std::vector<float> tau_p, tau_m;
    std::vector<uint32_t> stdp_pops;
    std::vector<uint8_t> stdp_modes;
    bool locked = true;
    uint32_t max_delay = 1;
    bool has_blocking = false;
    std::vector<std::vector<uint32_t>> pending;  // channel messages for the next step
    uint32_t gen_channels = 0;                   // channels of device sources feeding it
    int dense_format = -1;                       // knp_mi355x_set_projection_dense (-1: scatter)
    uint32_t dense_flags = 0;
    std::vector<uint64_t> dense_sig;  // MFMA expansion: per tile, first neuron << 32 | neurons (grouping)
};

// The dense blocks (instances) of a projection expanded on MFMA.
struct DenseBlocks {
    uint32_t n_oct = 0, delay = 1;  // octets over all instances
    float sign = 1.f;
    std::vector<DenseInst> inst;
    std::vector<uint32_t> pre_inst, post_inst;  // instance of every presynaptic / postsynaptic index
};

struct HostDense {
    uint32_t post_pop = 0, n_inst = 0, n_in = 0, format = 0, threshold = 0, hash_offset = 0, T = 0;
    uint64_t seed = 0;
    std::vector<uint32_t> post_begin, post_count;
    std::vector<uint8_t> weights;  // row-major [sum post_count][n_in], 1 or 2 bytes per value
    std::vector<float> row_scale;  // per row of `weights`
};

struct HostSource {
    uint32_t proj, first, n_ch, n_patterns, spp, lo, hi, hash_offset;
    uint64_t seed;
    std::vector<uint32_t> thr;
};

}  // namespace

struct knp_net {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<HostPop> pops;
    std::vector<HostProj> projs;
    std::vector<HostSource> srcs;
    std::vector<HostDense> dense;
    bool initialized = false;
    uint64_t step = 0;
    NetDev table{};
    NetDev *dev_table = nullptr;
    std::vector<void *> allocations;
    std::vector<size_t> alloc_bytes;  // size of each allocation (state snapshots of the step tuner)
    uint64_t allocated = 0, cap = 0;
    uint32_t scatter_grid = 1;        // workgroups of the atomic scatter (the step tuner may lower it)
    uint32_t scatter_grid_alloc = 1;  // ... its largest value: the per-workgroup counters allocated
    int scatter_mode = KNP_SCATTER_AUTO;
    int semantics = KNP_SEMANTICS_SINGLE_THREADED;  // knp_mi355x_set_message_semantics
    bool fused = false;     // every population tiled: phase 1 of step t + 1 runs inside the tile kernel of t
    bool persistent = false;  // fused, and every tile fits on the device at once (cooperative launch)
    bool small = false;       // ... and the whole network fits one workgroup: k_small_run (small_run, small_lds)
    SmallRun small_run{};
    uint32_t small_lds = 0;
    bool wide_tiles = false;  // tiles outnumber the workgroups resident at once: the 6-waves tile kernel
    // what init found possible (the step tuner may run a network persistent-eligible stepwise, etc.): the
    // persistent launch of every tile, the one-workgroup launch
    bool can_persistent = false, can_small = false;
    // the MFMA expansions' spike counts were last staged in the device scratch (knp_mi355x_step_noalloc): the
    // per-projection buffers must be cleared before the next knp_mi355x_step uses them
    bool counts_in_scratch = false;
    // captured step graph (knp_mi355x_set_step_graph): the instantiated graph, its steps, whether its MFMA
    // expansions stage in the device scratch, replays so far, and the steps per graph without dense projections
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_steps = 0;
    bool graph_scratch = false;
    uint64_t graph_replays = 0;
    uint32_t graph_default = 32;
    uint32_t test_hold_block = 0xFFFFFFFFu, test_spin_limit = kSpinLimit;  // knp_mi355x_set_wait_test
    uint64_t test_hold_until = UINT64_MAX;
    uint32_t train_wg_limit = 16384;  // knp_mi355x_set_train_workgroup_limit
    bool tiles_shallow = false;       // every tiled population's ring is at most kTileShallowD deep
    uint32_t run_kind = POP_LEAN;  // population kind of the persistent launch
    uint32_t tile_lds = 0;  // dynamic LDS bytes of the tile scatter (largest tiled ring)
    std::vector<uint32_t *> seg_list_host;
    std::vector<uint32_t> push_vpos;  // virtual position of each projection's push segment
    std::vector<uint32_t> staging;
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timed[KNP_KERNEL_COUNT];
    double time_ms[KNP_KERNEL_COUNT] = {};
    uint64_t launches[KNP_KERNEL_COUNT] = {};
};

namespace {

int dev_alloc(knp_net *net, void **out, size_t bytes, bool zero = true) {
    *out = nullptr;
    if (bytes == 0) bytes = 4;
    if (net->cap && net->allocated + bytes > net->cap)
        return fail(KNP_ERR_OUT_OF_MEMORY, "device memory cap of " + std::to_string(net->cap) + " bytes exceeded");
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(KNP_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    net->allocations.push_back(*out);
    net->alloc_bytes.push_back(bytes);
    net->allocated += bytes;
    if (zero) HIP_TRY(hipMemsetAsync(*out, 0, bytes, net->stream));
    return KNP_OK;
}

template <typename T>
int upload(knp_net *net, T **out, const T *src, size_t count) {
    int rc = dev_alloc(net, (void **)out, count * sizeof(T), false);
    if (rc) return rc;
    if (count) HIP_TRY(hipMemcpyAsync(*out, src, count * sizeof(T), hipMemcpyHostToDevice, net->stream));
    // the source is host memory owned by the caller's frame: finish before it goes away
    HIP_TRY(hipStreamSynchronize(net->stream));
    return KNP_OK;
}

// A parameter field: one value if uniform over the population, else a device array.
template <typename Get>
int make_param_f(knp_net *net, uint32_t n, Get get, ParamF *out) {
    out->a = nullptr;
    out->v = n ? get(0) : 0.f;
    bool uniform = true;
    for (uint32_t i = 1; i < n && uniform; ++i) {
        float a = get(i), b = out->v;
        uniform = (std::memcmp(&a, &b, sizeof(float)) == 0);
    }
    if (uniform) return KNP_OK;
    std::vector<float> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = get(i);
    float *d = nullptr;
    int rc = upload(net, &d, v.data(), n);
    out->a = d;
    return rc;
}

template <typename Get>
int make_param_u(knp_net *net, uint32_t n, Get get, ParamU *out) {
    out->a = nullptr;
    out->v = n ? get(0) : 0u;
    bool uniform = true;
    for (uint32_t i = 1; i < n && uniform; ++i) uniform = get(i) == out->v;
    if (uniform) return KNP_OK;
    std::vector<uint32_t> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = get(i);
    uint32_t *d = nullptr;
    int rc = upload(net, &d, v.data(), n);
    out->a = d;
    return rc;
}

int make_param_i64(knp_net *net, const std::vector<int64_t> &v, ParamI64 *out) {
    out->a = nullptr;
    out->v = v.empty() ? 0 : v[0];
    bool uniform = true;
    for (size_t i = 1; i < v.size() && uniform; ++i) uniform = v[i] == out->v;
    if (uniform) return KNP_OK;
    int64_t *d = nullptr;
    int rc = upload(net, &d, v.data(), v.size());
    out->a = d;
    return rc;
}

// The static part of a BLIFAT neuron's spike threshold, activation_threshold_ + additional_threshold_, in the
// forms the device compares with (PopDev::thr_fix / thr_up / thr_hi / thr_lo). T is the reference's double sum.
struct ThresholdForms {
    int64_t fix;  // ceil(T * 2^frac_bits), kept within +-2^61 (accumulator sums stay below 2^60)
    float up, hi, lo;
};
ThresholdForms threshold_forms(double T, int32_t frac_bits) {
    ThresholdForms f;
    const double lim = std::ldexp(1.0, 61);
    const double x = std::ceil(std::ldexp(T, frac_bits));
    f.fix = x >= lim ? (int64_t(1) << 61) : x <= -lim ? -(int64_t(1) << 61) : (int64_t)x;
    f.up = (float)T;
    if ((double)f.up < T) f.up = std::nextafter(f.up, std::numeric_limits<float>::infinity());
    f.hi = (float)T;
    f.lo = (float)(T - (double)f.hi);
    return f;
}

uint32_t ts_to_device(uint64_t v) {
    if (v == std::numeric_limits<uint64_t>::max()) return 0xFFFFFFFFu;
    if (v >= 0xFFFFFFFEull) return 0xFFFFFFFEu;
    return (uint32_t)v;
}

uint64_t ts_from_device(uint32_t v) {
    if (v == 0xFFFFFFFFu) return std::numeric_limits<uint64_t>::max();
    return v;
}

// activation_threshold_ + additional_threshold_ (a double in the reference) as the smallest float not below
// it: for a float potential V, V >= T exactly when V >= that float (altai_impl.h:116).
float altai_threshold(const knp_altai_lif_neuron &x) {
    const double T = (double)x.activation_threshold + x.additional_threshold;
    float f = (float)T;
    if ((double)f < T) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

// The resource fields (everything after `base`) of one resource neuron struct into another.
template <class A, class B>
void copy_resource_fields(const A &s, B &d) {
    d.synapse_sum_threshold_coefficient = s.synapse_sum_threshold_coefficient;
    d.dopamine_plasticity_time = s.dopamine_plasticity_time;
    d.free_synaptic_resource = s.free_synaptic_resource;
    d.synaptic_resource_threshold = s.synaptic_resource_threshold;
    d.resource_drain_coefficient = s.resource_drain_coefficient;
    d.stability = s.stability;
    d.stability_change_parameter = s.stability_change_parameter;
    d.stability_change_at_isi = s.stability_change_at_isi;
    d.isi_max = s.isi_max;
    d.d_h = s.d_h;
    d.isi_status = s.isi_status;
    d.last_step = s.last_step;
    d.last_spike_step = s.last_spike_step;
    d.first_isi_spike = s.first_isi_spike;
    d.is_being_forced = s.is_being_forced;
}

bool lean_eligible(const HostPop &P) {
    if (P.resource || P.model != MODEL_BLIFAT || P.has_cond || P.has_dop || P.has_block) return false;
    bool any_max = false, any_other = false;
    for (const auto &n : P.nb) {
        if (n.n_time_steps_since_last_firing == std::numeric_limits<uint64_t>::max()) any_max = true;
        else any_other = true;
        if (n.dynamic_threshold != 0.0 || n.threshold_increment != 0.0) return false;
        if (!std::isfinite(n.threshold_decay) || !std::isfinite(n.postsynaptic_trace_decay)) return false;
        if (n.postsynaptic_trace != 0.0 || n.postsynaptic_trace_increment != 0.0) return false;
        if (n.inhibitory_conductance != 0.0 || !std::isfinite(n.inhibitory_conductance_decay)) return false;
        if (n.potential_decay != 0.0 || !std::isfinite(n.potential)) return false;
        if (n.bursting_phase != 0 || n.bursting_period != 0) return false;
        if (n.absolute_refractory_period != 0) return false;
        if (n.total_blocking_period < (int64_t(1) << 62)) return false;
    }
    return !(any_max && any_other);
}

// Fraction bits of a population's exact accumulators: the largest sum of |weight| a neuron can receive in
// one step must stay below 2^(62 - frac_bits) with a factor-4 margin; at most 52 bits.
int32_t choose_frac_bits(double bound) {
    int e = 0;
    std::frexp(bound + 1.0, &e);  // bound + 1 < 2^e
    int f = 61 - e - 2;
    return std::min(52, f);
}

// Words of a channel (host-pushed) segment: the count, up to 4 entries per presynaptic index (each index at
// most 4 times per step, checked when pushed), rounded to 64 bytes.
uint64_t channel_seg_words(uint32_t n_pre) {
    return (1ull + 4ull * std::max(n_pre, 1u) + 15) / 16 * 16;
}

// The persistent kernel for the population kind of every tile (or kMixedKinds).
const void *tile_run_kernel(uint32_t kind) {
    if (kind == POP_LEAN) return reinterpret_cast<const void *>(&k_tile_run<POP_LEAN>);
    if (kind == POP_FULL) return reinterpret_cast<const void *>(&k_tile_run<POP_FULL>);
    return reinterpret_cast<const void *>(&k_tile_run<kMixedKinds>);
}

int upload_working_sets(knp_net *net);
int upload_table(knp_net *net);
void dev_free(knp_net *net, void *p);

// Value of a dense weight as a float (fp8 e4m3 decoded exactly, bf16 widened).
double dense_value(const HostDense &D, uint64_t row, uint32_t i) {
    const size_t idx = (size_t)row * D.n_in + i;
    if (D.format == KNP_DENSE_BF16) {
        uint32_t bits = (uint32_t)(D.weights[2 * idx] | (D.weights[2 * idx + 1] << 8)) << 16;
        float f;
        std::memcpy(&f, &bits, 4);
        return f;
    }
    const uint8_t b = D.weights[idx];
    const int e = (b >> 3) & 15, m = b & 7;
    if ((b & 0x7F) == 0x7F) return std::numeric_limits<double>::quiet_NaN();
    const double mag = e == 0 ? std::ldexp(m / 8.0, -6) : std::ldexp(1.0 + m / 8.0, e - 7);
    return (b & 0x80) ? -mag : mag;
}

uint32_t dense_ksteps(const HostDense &D) { return (D.n_in + 15) / 16; }
uint32_t dense_pitch(const HostDense &D) { return (2 * dense_ksteps(D) + 15) / 16 * 16; }

int alloc_dense_buffers(knp_net *net, uint32_t d, uint32_t T) {
    DenseDev &Dn = net->table.dense[d];
    const HostDense &Dh = net->dense[d];
    const uint32_t n = net->pops[Dh.post_pop].n;
    dev_free(net, Dn.bits);
    dev_free(net, Dn.ff);
    Dn.bits = nullptr;
    Dn.ff = nullptr;
    Dn.T = T;
    int rc;
    if ((rc = dev_alloc(net, (void **)&Dn.bits, (size_t)T * Dh.n_inst * dense_pitch(Dh)))) return rc;
    if ((rc = dev_alloc(net, (void **)&Dn.ff, (size_t)2 * T * n * sizeof(float)))) return rc;
    PopDev &P = net->table.pops[Dh.post_pop];
    P.ff = Dn.ff;
    P.ff_T = T;
    return KNP_OK;
}

void launch_dense(knp_net *net, uint32_t d, uint64_t t0) {
    const DenseDev &Dn = net->table.dense[d];
    const uint64_t bytes = (uint64_t)Dn.T * Dn.n_inst * Dn.pitch;
    const uint32_t g1 = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(8192, (bytes + kBlock - 1) / kBlock));
    k_dense_spikes<<<g1, kBlock, 0, net->stream>>>(net->dev_table, d, t0);
    // one wave per (tile, group of kDenseGroup row blocks), four waves per workgroup
    const uint64_t units = (uint64_t)Dn.n_tiles * ((Dn.T / 32 + kDenseGroup - 1) / kDenseGroup);
    const uint32_t g2 = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(65536, (units + 3) / 4));
    if (Dn.format == 0) k_dense_mfma<0><<<g2, kBlock, 0, net->stream>>>(net->dev_table, d, t0);
    else k_dense_mfma<1><<<g2, kBlock, 0, net->stream>>>(net->dev_table, d, t0);
}

std::string tune_cache_path() {
    if (const char *p = std::getenv("KNP_MI355X_TUNE_CACHE")) return p;
    if (const char *h = std::getenv("HOME")) return std::string(h) + "/.knp_mi355x_tune.txt";
    return "";
}

// Picks the time block of a dense projection by timing the candidates on the loaded weights (cost per step
// = block time / block length) and caches the choice per shape.
int autotune_dense(knp_net *net, uint32_t d, uint32_t *best_T) {
    const HostDense &Dh = net->dense[d];
    DenseDev &Dn = net->table.dense[d];
    char key[160];
    std::snprintf(key, sizeof key, "dense fmt=%u n_in=%u tiles=%u inst=%u", Dh.format, Dh.n_in, Dn.n_tiles, Dh.n_inst);
    const std::string path = tune_cache_path();
    if (!path.empty()) {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            const size_t sp = line.rfind(' ');
            if (sp != std::string::npos && line.substr(0, sp) == key) {
                *best_T = (uint32_t)std::strtoul(line.c_str() + sp + 1, nullptr, 10);
                if (*best_T >= 32 && *best_T % 32 == 0) return KNP_OK;
            }
        }
    }
    const uint32_t n = net->pops[Dh.post_pop].n;
    double best = 1e30;
    *best_T = 64;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    for (uint32_t T : {32u, 64u, 128u, 256u, 512u}) {
        if ((uint64_t)2 * T * n * 4 > (uint64_t)8 << 30) break;  // keep the impact ring under 8 GiB
        int rc = alloc_dense_buffers(net, d, T);
        if (rc) return rc;
        if ((rc = upload_table(net))) return rc;
        launch_dense(net, d, 0);  // warm
        HIP_TRY(hipEventRecord(a, net->stream));
        for (int r = 0; r < 3; ++r) launch_dense(net, d, (uint64_t)r * T);
        HIP_TRY(hipEventRecord(b, net->stream));
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        const double per_step = ms / (3.0 * T);
        if (per_step < best) {
            best = per_step;
            *best_T = T;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIP_TRY(hipMemsetAsync(&net->table.events[0], 0, 16, net->stream));  // tuning spikes are not events
    if (!path.empty()) {
        std::ofstream out(path, std::ios::app);
        if (out) out << key << " " << *best_T << "\n";
    }
    return KNP_OK;
}

int init_dense(knp_net *net, uint32_t d) {
    const HostDense &Dh = net->dense[d];
    DenseDev &Dn = net->table.dense[d];
    const HostPop &P = net->pops[Dh.post_pop];
    if (P.resource) return fail(KNP_ERR_UNSUPPORTED, "dense input projections target BLIFAT populations only");
    for (uint32_t e = 0; e < d; ++e)
        if (net->dense[e].post_pop == Dh.post_pop)
            return fail(KNP_ERR_UNSUPPORTED, "one dense input projection per population");
    const uint32_t ks = dense_ksteps(Dh);
    const uint32_t FB = Dh.format == KNP_DENSE_FP8_E4M3 ? 8 : 16;
    const uint32_t vb = Dh.format == KNP_DENSE_FP8_E4M3 ? 1 : 2;  // bytes per value
    std::vector<uint32_t> tinst, tj0, tncol;
    std::vector<uint64_t> trow;  // first weight row of each tile
    uint64_t row = 0;
    for (uint32_t k = 0; k < Dh.n_inst; ++k) {
        for (uint32_t c = 0; c < Dh.post_count[k]; c += 32) {
            tinst.push_back(k);
            tj0.push_back(Dh.post_begin[k] + c);
            tncol.push_back(std::min(32u, Dh.post_count[k] - c));
            trow.push_back(row + c);
        }
        row += Dh.post_count[k];
    }
    const uint32_t n_tiles = (uint32_t)tinst.size();
    std::vector<uint8_t> frag((size_t)n_tiles * ks * 64 * FB, 0);
    for (uint32_t tl = 0; tl < n_tiles; ++tl)
        for (uint32_t kk = 0; kk < ks; ++kk)
            for (uint32_t lane = 0; lane < 64; ++lane) {
                const uint32_t col = lane & 31;
                if (col >= tncol[tl]) continue;
                uint8_t *dst = &frag[(((size_t)tl * ks + kk) * 64 + lane) * FB];
                const uint64_t r = trow[tl] + col;
                for (uint32_t e = 0; e < 8; ++e) {
                    const uint32_t i = 16 * kk + 8 * (lane >> 5) + e;
                    if (i >= Dh.n_in) continue;
                    std::memcpy(dst + e * vb, &Dh.weights[((size_t)r * Dh.n_in + i) * vb], vb);
                }
            }
    std::vector<float> scale(P.n, 1.f);
    row = 0;
    for (uint32_t k = 0; k < Dh.n_inst; ++k)
        for (uint32_t j = 0; j < Dh.post_count[k]; ++j, ++row) scale[Dh.post_begin[k] + j] = Dh.row_scale[row];
    int rc;
    uint8_t *fr;
    uint32_t *a1, *a2, *a3, *a4;
    float *sc;
    if ((rc = upload(net, &fr, frag.data(), frag.size()))) return rc;
    if ((rc = upload(net, &a1, tinst.data(), n_tiles))) return rc;
    if ((rc = upload(net, &a2, tj0.data(), n_tiles))) return rc;
    if ((rc = upload(net, &a3, tncol.data(), n_tiles))) return rc;
    if ((rc = upload(net, &a4, Dh.post_count.data(), Dh.n_inst))) return rc;
    if ((rc = upload(net, &sc, scale.data(), scale.size()))) return rc;
    Dn.format = Dh.format;
    Dn.post_pop = Dh.post_pop;
    Dn.n_in = Dh.n_in;
    Dn.ksteps = ks;
    Dn.n_inst = Dh.n_inst;
    Dn.n_tiles = n_tiles;
    Dn.pitch = dense_pitch(Dh);
    Dn.threshold = Dh.threshold;
    Dn.hash_offset = Dh.hash_offset;
    Dn.seed = Dh.seed;
    Dn.frag = fr;
    Dn.tile_inst = a1;
    Dn.tile_j0 = a2;
    Dn.tile_ncol = a3;
    Dn.inst_post_count = a4;
    Dn.row_scale = sc;
    uint32_t T = Dh.T;
    if (T == 0) {
        if ((rc = autotune_dense(net, d, &T))) return rc;
    }
    net->dense[d].T = T;
    if ((rc = alloc_dense_buffers(net, d, T))) return rc;
    return KNP_OK;
}

// Checks that the synapses of a projection marked for the MFMA expansion form dense blocks and finds them:
// one delay and one output type (EXCITATORY or INHIBITORY_CURRENT; a plain EXCITATORY projection onto a
// synaptic-resource population would also carry forcing, which the expansion does not); every postsynaptic
// neuron with synapses takes each presynaptic index of one range [lo, hi] exactly once; consecutive neurons
// with the same range form an instance; instance ranges do not overlap.
int dense_blocks(const knp_net *net, const HostProj &H, DenseBlocks &B) {
    const uint32_t n_post = net->pops[H.post_pop].n;
    if (H.kind == PROJ_ADDITIVE) return fail(KNP_ERR_UNSUPPORTED, "additive STDP projections are not expanded on MFMA");
    if (H.dense_format != KNP_DENSE_BF16) return fail(KNP_ERR_UNSUPPORTED, "the MFMA expansion takes BF16 weights");
    if (H.n_syn == 0) return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion of a projection without synapses");
    const uint8_t ty = H.type[0];
    const uint32_t d = H.delay[0];
    for (uint64_t s = 0; s < H.n_syn; ++s)
        if (H.type[s] != ty || H.delay[s] != d)
            return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: all synapses need one delay and one output type");
    if (ty != KNP_EXCITATORY && ty != KNP_INHIBITORY_CURRENT)
        return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: EXCITATORY or INHIBITORY_CURRENT synapses only");
    if (H.kind == PROJ_DELTA && ty == KNP_EXCITATORY && net->pops[H.post_pop].resource)
        return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: a forcing projection onto a synaptic-resource population");
    std::vector<uint32_t> lo(n_post, 0xFFFFFFFFu), hi(n_post, 0), cnt(n_post, 0);
    std::vector<uint64_t> first(n_post + 1, 0);  // bit offsets of every neuron's range in `seen`
    for (uint64_t s = 0; s < H.n_syn; ++s) {
        const uint32_t j = H.post[s], i = H.pre[s];
        lo[j] = std::min(lo[j], i);
        hi[j] = std::max(hi[j], i);
        cnt[j]++;
    }
    for (uint32_t j = 0; j < n_post; ++j) {
        const uint64_t w = cnt[j] ? (uint64_t)hi[j] - lo[j] + 1 : 0;
        if (w != cnt[j]) return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: every neuron needs a full range of presynaptic indexes");
        first[j + 1] = first[j] + w;
    }
    std::vector<uint64_t> seen((first[n_post] + 63) / 64, 0);
    for (uint64_t s = 0; s < H.n_syn; ++s) {
        const uint64_t b = first[H.post[s]] + (H.pre[s] - lo[H.post[s]]);
        if (seen[b >> 6] >> (b & 63) & 1) return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: a repeated synapse");
        seen[b >> 6] |= 1ull << (b & 63);
    }
    B = DenseBlocks{};
    B.delay = d;
    B.sign = ty == KNP_INHIBITORY_CURRENT ? -1.f : 1.f;
    B.post_inst.assign(n_post, 0xFFFFFFFFu);
    B.pre_inst.assign(std::max(H.n_pre, 1u), 0xFFFFFFFFu);
    uint64_t base = 0;
    for (uint32_t j = 0; j < n_post; ++j) {
        if (!cnt[j]) continue;
        const bool same = !B.inst.empty() && B.inst.back().pre_begin == lo[j] && B.inst.back().n_in == cnt[j] &&
                          B.inst.back().post_begin + B.inst.back().post_count == j;
        if (!same) {
            if (!B.inst.empty()) {
                base += (uint64_t)B.inst.back().n_oct * B.inst.back().npad * 8;
                B.n_oct += B.inst.back().n_oct;
            }
            DenseInst I{};
            I.pre_begin = lo[j];
            I.post_begin = j;
            I.n_in = cnt[j];
            I.n_oct = (cnt[j] + 7) / 8;
            I.cbase = B.n_oct;
            I.base = base;
            B.inst.push_back(I);
        }
        DenseInst &I = B.inst.back();
        I.post_count++;
        I.npad = (I.post_count + 31) / 32 * 32;
        B.post_inst[j] = (uint32_t)B.inst.size() - 1;
    }
    B.n_oct += B.inst.back().n_oct;
    if (B.inst.size() >= (1u << 16)) return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: 65536 or more instances");
    if (B.n_oct >= (1u << 30)) return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: too many presynaptic indexes");
    for (uint32_t k = 0; k < B.inst.size(); ++k) {
        if (B.inst[k].post_count > (1u << 21)) return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: instance too large");
        for (uint32_t i = 0; i < B.inst[k].n_in; ++i) {
            uint32_t &x = B.pre_inst[B.inst[k].pre_begin + i];
            if (x != 0xFFFFFFFFu) return fail(KNP_ERR_UNSUPPORTED, "MFMA expansion: instances share presynaptic indexes");
            x = k;
        }
    }
    return KNP_OK;
}

// The multi-threaded backend's semantics cover that backend's types only
// (cpu-multi-threaded-backend/include/knp/backends/cpu-multi-threaded/backend.h:80-88), and no BLOCKING
// synapses: that backend delivers one projection's impacts in the completion order of its threads' parts.
int check_multi_threaded(const knp_net *net) {
    for (const auto &P : net->pops)
        if (P.model == MODEL_ALTAI) return fail(KNP_ERR_UNSUPPORTED, "Type is not supported.");
    for (const auto &Q : net->projs) {
        if (Q.kind == PROJ_ADDITIVE) return fail(KNP_ERR_UNSUPPORTED, "Type is not supported.");
        if (Q.has_blocking)
            return fail(KNP_ERR_UNSUPPORTED, "BLOCKING synapses under the multi-threaded message semantics");
        if (Q.dense_format >= 0)
            return fail(KNP_ERR_UNSUPPORTED, "MFMA expansions under the multi-threaded message semantics");
        if (Q.gen_channels && !Q.pending.empty())
            return fail(KNP_ERR_UNSUPPORTED, "under the multi-threaded message semantics a projection fed by device "
                                             "sources takes no pushed messages");
        if (Q.n_pre > kMsgIndexMask + 1u)
            return fail(KNP_ERR_UNSUPPORTED, "presynaptic ranges above 2^24 under the multi-threaded message semantics");
    }
    return KNP_OK;
}

// In a persistent launch a workgroup waits, at each step, for the stamps of the segments its tile reads, and
// overwrites its own segments of step t at step t + 2. That is safe when every workgroup whose segment another
// workgroup reads in turn reads a segment of that reader: it then cannot publish step t + 2 before the reader
// has published step t + 1, which the reader does only after reading step t. Segments and their writers:
// population segment -> the workgroup of its tile (tile_of_block), source block j -> workgroup j mod tiles;
// channel segments are written by the host before the launch. Networks where the relation is not symmetric
// run stepwise.
bool waits_cover_readers(const knp_net *net, const std::vector<std::vector<uint32_t>> &tile_reads,
                         const std::unordered_map<uint32_t, uint32_t> &seg_tile) {
    const NetDev &T = net->table;
    const uint32_t n = T.n_tiles;
    std::vector<uint32_t> wg_of_tile(n);
    for (uint32_t b = 0; b < n; ++b) {
        const uint32_t q = n / 8, r = n % 8, x = b % 8;  // tile_of_block
        wg_of_tile[x * q + std::min(x, r) + b / 8] = b;
    }
    std::unordered_map<uint32_t, uint32_t> writer;  // segment -> workgroup
    for (const auto &kv : seg_tile) writer[kv.first] = wg_of_tile[kv.second];
    uint32_t j = 0;
    for (uint32_t s2 = 0; s2 < T.n_srcs; ++s2) {
        const uint32_t nb = std::max(1u, (net->srcs[s2].n_ch + kBlock - 1) / kBlock);
        for (uint32_t c = 0; c < nb; ++c, ++j) writer[T.srcs[s2].seg_base + c] = j % n;
    }
    std::vector<std::vector<uint32_t>> reads(n);  // workgroup -> workgroups whose segments it reads
    for (uint32_t i = 0; i < n; ++i) {
        auto &rd = reads[wg_of_tile[i]];
        for (uint32_t sg : tile_reads[i]) {
            auto it = writer.find(sg);
            if (it != writer.end()) rd.push_back(it->second);
        }
        std::sort(rd.begin(), rd.end());
        rd.erase(std::unique(rd.begin(), rd.end()), rd.end());
    }
    for (uint32_t y = 0; y < n; ++y)
        for (uint32_t x : reads[y])
            if (x != y && !std::binary_search(reads[x].begin(), reads[x].end(), y)) return false;
    return true;
}

int init_network(knp_net *net) {
    HIP_TRY(hipSetDevice(net->device));
    const uint32_t n_pops = (uint32_t)net->pops.size(), n_projs = (uint32_t)net->projs.size();
    const bool mt = net->semantics == KNP_SEMANTICS_MULTI_THREADED;
    if (mt) {
        int rc = check_multi_threaded(net);
        if (rc) return rc;
    }
    if (n_pops > (uint32_t)kMaxPops) return fail(KNP_ERR_UNSUPPORTED, "too many populations");
    if (n_projs > (uint32_t)kMaxProjs) return fail(KNP_ERR_UNSUPPORTED, "too many projections");
    if (net->srcs.size() > (size_t)kMaxSources) return fail(KNP_ERR_UNSUPPORTED, "too many sources");
    NetDev &T = net->table;
    std::memset(&T, 0, sizeof(T));
    T.n_pops = n_pops;
    T.n_projs = n_projs;
    T.n_srcs = (uint32_t)net->srcs.size();

    // ---- incoming-projection facts per population
    std::vector<std::vector<double>> bound(n_pops);
    for (uint32_t p = 0; p < n_pops; ++p) bound[p].assign(net->pops[p].n, 0.0);
    uint32_t n_blocking = 0;
    for (uint32_t q = 0; q < n_projs; ++q) {
        HostProj &Q = net->projs[q];
        HostPop &P = net->pops[Q.post_pop];
        P.D = std::max(P.D, Q.max_delay);
        for (uint64_t s = 0; s < Q.n_syn; ++s) {
            const uint8_t ty = Q.type[s];
            if (ty == KNP_INHIBITORY_CONDUCTANCE) P.has_cond = true;
            if (ty == KNP_DOPAMINE) P.has_dop = true;
            if (ty == KNP_BLOCKING) P.has_block = true;
            if (ty == KNP_EXCITATORY && Q.kind == PROJ_DELTA && P.resource) P.has_forcing = true;
            double mag = std::fabs((double)Q.w[s]);
            if (Q.kind == PROJ_RESOURCE)
                mag = std::max({mag, std::fabs((double)Q.rules[s].w_min), std::fabs((double)Q.rules[s].w_max)});
            if (Q.kind == PROJ_ADDITIVE) mag = 2.0 * mag + 1.0;
            bound[Q.post_pop][Q.post[s]] += mag;
        }
        if (Q.has_blocking) {
            if (Q.kind != PROJ_DELTA)
                return fail(KNP_ERR_UNSUPPORTED, "BLOCKING synapses are supported on Delta projections only");
            if (Q.max_delay > 64) return fail(KNP_ERR_UNSUPPORTED, "BLOCKING projection with delay > 64");
            if (n_blocking >= 64) return fail(KNP_ERR_UNSUPPORTED, "more than 64 projections with BLOCKING synapses");
            T.blk_proj[n_blocking] = q;
            T.projs[q].blocking_id = n_blocking++;
        }
    }
    if (n_pops > 31 && n_blocking) return fail(KNP_ERR_UNSUPPORTED, "BLOCKING ranks support up to 31 populations");
    // projections expanded on MFMA: their dense blocks
    std::vector<DenseBlocks> dense_info(n_projs);
    for (uint32_t q = 0; q < n_projs; ++q)
        if (net->projs[q].dense_format >= 0) {
            int rc = dense_blocks(net, net->projs[q], dense_info[q]);
            if (rc) return rc;
        }
    for (const auto &Dh : net->dense) {
        uint64_t row = 0;
        for (uint32_t k = 0; k < Dh.n_inst; ++k)
            for (uint32_t j = 0; j < Dh.post_count[k]; ++j, ++row) {
                double sum = 0;
                for (uint32_t i = 0; i < Dh.n_in; ++i) sum += std::fabs(dense_value(Dh, row, i));
                bound[Dh.post_pop][Dh.post_begin[k] + j] += sum * std::fabs((double)Dh.row_scale[row]);
            }
    }

    // ---- spike lists: one pool holding every segment (population, source and channel lists) twice, by step
    // parity, so a step's lists can be written while the previous step's are still being read
    uint32_t *pool_cursor = nullptr;
    {
        uint64_t half = 0;
        for (uint32_t p = 0; p < n_pops; ++p) half += (uint64_t)std::max(1u, (net->pops[p].n + kBlock - 1) / kBlock) * kSegStride;
        for (const auto &Sh : net->srcs) half += (uint64_t)std::max(1u, (Sh.n_ch + kBlock - 1) / kBlock) * kSegStride;
        for (const auto &Qh : net->projs) half += channel_seg_words(Qh.n_pre);
        // byte offsets into the pool (both parities) are 32-bit in the tile kernels' buffer loads
        if (2 * half * 4 >= (uint64_t(1) << 32)) return fail(KNP_ERR_UNSUPPORTED, "spike lists above 2^29 entries");
        int rc;
        if ((rc = dev_alloc(net, (void **)&pool_cursor, (size_t)2 * half * 4))) return rc;
        T.list_pstride = (uint32_t)half;
        T.list_pool = pool_cursor;
    }

    // ---- populations
    for (uint32_t p = 0; p < n_pops; ++p) {
        HostPop &H = net->pops[p];
        PopDev &P = T.pops[p];
        const uint32_t n = H.n;
        double maxb = 0;
        for (double b : bound[p]) maxb = std::max(maxb, b);
        H.frac_bits = choose_frac_bits(maxb);
        if (H.frac_bits < 16)
            return fail(KNP_ERR_RANGE, "population " + std::to_string(p) +
                                           ": incoming |weight| sums too large for the exact accumulator");
        // weights off the accumulator's grid (magnitudes below 2^(23 - frac_bits) with low mantissa bits) are
        // rounded when they become impacts: counted, so that a caller can see when the sums are not exact
        H.inexact = 0;
        for (const auto &Q : net->projs) {
            if (Q.post_pop != p) continue;
            for (uint64_t s2 = 0; s2 < Q.n_syn; ++s2) {
                const double v = std::ldexp((double)Q.w[s2], H.frac_bits);
                if (v != std::nearbyint(v)) ++H.inexact;
            }
        }
        H.kind = H.resource ? POP_RESOURCE : (lean_eligible(H) ? POP_LEAN : POP_FULL);
        P.n = n;
        P.kind = H.kind;
        P.D = H.D;
        P.words = (n + 31) / 32;
        P.frac_bits = H.frac_bits;
        P.scale = std::ldexp(1.0f, H.frac_bits);
        P.inv_scale = std::ldexp(1.0f, -H.frac_bits);
        int rc;
        const size_t ring = (size_t)H.D * n;
        if ((rc = dev_alloc(net, (void **)&P.acc_v, ring * 8))) return rc;
        if (H.has_cond && (rc = dev_alloc(net, (void **)&P.acc_g, ring * 8))) return rc;
        if (H.has_dop && (rc = dev_alloc(net, (void **)&P.acc_d, ring * 8))) return rc;
        if (H.has_forcing && (rc = dev_alloc(net, (void **)&P.acc_f, ring * 4))) return rc;
        if (H.has_block && (rc = dev_alloc(net, (void **)&P.acc_b, ring * 8))) return rc;
        P.model = H.model;
        if (H.model == MODEL_ALTAI) {
            // AltAILIF (altai_lif.h:128-354) on the FULL state arrays (see PopModel)
            const auto &na = H.na;
            std::vector<float> f(n);
            std::vector<int64_t> at(n);
            for (uint32_t i = 0; i < n; ++i) f[i] = na[i].potential;
            if ((rc = upload(net, &P.V, f.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) f[i] = na[i].pre_impact_potential;
            if ((rc = upload(net, &P.Vpre, f.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) f[i] = (float)na[i].dopamine_value;
            if ((rc = upload(net, &P.dop, f.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) at[i] = na[i].activity_time;
            if ((rc = upload(net, &P.tbp, at.data(), n))) return rc;
            // resource neurons add their trained additional threshold on the device (res.add_thr)
            auto act = [&](uint32_t i) { return H.resource ? (float)na[i].activation_threshold : altai_threshold(na[i]); };
            if ((rc = make_param_f(net, n, act, &P.act))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)na[i].potential_reset_value; }, &P.V_reset))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)na[i].potential_leak; }, &P.V_decay))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return -(float)na[i].negative_activation_threshold; }, &P.V_min))) return rc;
            auto flags = [&](uint32_t i) {
                const auto &x = na[i];
                return (x.is_diff ? kAltDiff : 0u) | (x.is_reset ? kAltReset : 0u) | (x.leak_rev ? kAltLeakRev : 0u) |
                       (x.saturate ? kAltSaturate : 0u) | (x.do_not_save ? kAltDoNotSave : 0u);
            };
            if ((rc = make_param_u(net, n, flags, &P.alt_flags))) return rc;
        } else if (H.model == MODEL_LIF) {
            // LIFNeuron (lif.h:72-117) on the FULL state arrays: V = potential_, n_ts = refract_counter_,
            // V_decay = leak_coefficient_, refr = refract_period_
            const auto &nl = H.nl;
            std::vector<float> Vh(n);
            std::vector<uint32_t> rcnt(n);
            for (uint32_t i = 0; i < n; ++i) {
                Vh[i] = nl[i].potential;
                rcnt[i] = nl[i].refract_counter;
            }
            if ((rc = upload(net, &P.V, Vh.data(), n))) return rc;
            if ((rc = upload(net, &P.n_ts, rcnt.data(), n))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nl[i].activation_threshold; }, &P.act))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nl[i].potential_reset_value; }, &P.V_reset))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nl[i].leak_coefficient; }, &P.V_decay))) return rc;
            if ((rc = make_param_u(net, n, [&](uint32_t i) { return nl[i].refract_period; }, &P.refr))) return rc;
        } else {
            const auto &nb = H.nb;
            std::vector<float> Vh(n);
            for (uint32_t i = 0; i < n; ++i) Vh[i] = (float)nb[i].potential;
            if ((rc = upload(net, &P.V, Vh.data(), n))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].activation_threshold; }, &P.act))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].potential_reset_value; }, &P.V_reset))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].min_potential; }, &P.V_min))) return rc;
            {
                // static threshold: activation + additional threshold (resource neurons train the additional one)
                std::vector<ThresholdForms> tf(n);
                std::vector<int64_t> fix(n);
                for (uint32_t i = 0; i < n; ++i) {
                    const double T = nb[i].activation_threshold + (H.kind == POP_RESOURCE ? 0.0 : nb[i].additional_threshold);
                    tf[i] = threshold_forms(T, H.frac_bits);
                    fix[i] = tf[i].fix;
                }
                if ((rc = make_param_i64(net, fix, &P.thr_fix))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return tf[i].up; }, &P.thr_up))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return tf[i].hi; }, &P.thr_hi))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return tf[i].lo; }, &P.thr_lo))) return rc;
            }
            if (H.kind == POP_LEAN) {
                P.lean_ts_max = nb.empty() ? 0 : (nb[0].n_time_steps_since_last_firing == std::numeric_limits<uint64_t>::max());
                std::vector<uint32_t> ls(n, 0xFFFFFFFFu);
                if ((rc = upload(net, &P.last_spike, ls.data(), n))) return rc;
            } else {
                std::vector<float> f(n);
                std::vector<uint32_t> u(n);
                auto upf = [&](float **dst, auto get) {
                    for (uint32_t i = 0; i < n; ++i) f[i] = get(i);
                    return upload(net, dst, f.data(), n);
                };
                auto upu = [&](uint32_t **dst, auto get) {
                    for (uint32_t i = 0; i < n; ++i) u[i] = get(i);
                    return upload(net, dst, u.data(), n);
                };
                if ((rc = upf(&P.Vpre, [&](uint32_t i) { return (float)nb[i].pre_impact_potential; }))) return rc;
                if ((rc = upf(&P.dyn, [&](uint32_t i) { return (float)nb[i].dynamic_threshold; }))) return rc;
                if ((rc = upf(&P.trace, [&](uint32_t i) { return (float)nb[i].postsynaptic_trace; }))) return rc;
                if ((rc = upf(&P.g, [&](uint32_t i) { return (float)nb[i].inhibitory_conductance; }))) return rc;
                if ((rc = upf(&P.dop, [&](uint32_t i) { return (float)nb[i].dopamine_value; }))) return rc;
                if ((rc = upu(&P.n_ts, [&](uint32_t i) { return ts_to_device(nb[i].n_time_steps_since_last_firing); }))) return rc;
                if ((rc = upu(&P.bphase, [&](uint32_t i) { return nb[i].bursting_phase; }))) return rc;
                bool track = H.has_block;
                for (const auto &x : nb) track |= x.total_blocking_period < (int64_t(1) << 62);
                P.track_tbp = track ? 1 : 0;
                if (track) {
                    std::vector<int64_t> tb(n);
                    for (uint32_t i = 0; i < n; ++i) tb[i] = nb[i].total_blocking_period;
                    if ((rc = upload(net, &P.tbp, tb.data(), n))) return rc;
                }
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].threshold_decay; }, &P.thr_decay))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].threshold_increment; }, &P.thr_inc))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].postsynaptic_trace_decay; }, &P.trace_decay))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].postsynaptic_trace_increment; }, &P.trace_inc))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].inhibitory_conductance_decay; }, &P.g_decay))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].potential_decay; }, &P.V_decay))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].reflexive_weight; }, &P.reflexive))) return rc;
                if ((rc = make_param_f(net, n, [&](uint32_t i) { return (float)nb[i].reversal_inhibitory_potential; }, &P.E_inh))) return rc;
                if ((rc = make_param_u(net, n, [&](uint32_t i) { return nb[i].bursting_period; }, &P.bperiod))) return rc;
                if ((rc = make_param_u(net, n, [&](uint32_t i) { return nb[i].absolute_refractory_period; }, &P.refr))) return rc;
            }
        }
        if (H.kind == POP_RESOURCE) {
            ResDev &R = P.res;
            const auto &nr = H.nr;
            std::vector<float> f(n);
            std::vector<uint64_t> u64(n);
            std::vector<uint8_t> u8(n);
            std::vector<double> add(n);
            for (uint32_t i = 0; i < n; ++i)
                add[i] = H.model == MODEL_ALTAI ? H.na[i].additional_threshold : H.nb[i].additional_threshold;
            if ((rc = upload(net, &R.add_thr, add.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) f[i] = nr[i].free_synaptic_resource;
            if ((rc = upload(net, &R.free_res, f.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) f[i] = nr[i].stability;
            if ((rc = upload(net, &R.stability, f.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) u8[i] = (uint8_t)nr[i].isi_status;
            if ((rc = upload(net, &R.isi_status, u8.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) u64[i] = nr[i].last_step;
            if ((rc = upload(net, &R.last_step, u64.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) u64[i] = nr[i].last_spike_step;
            if ((rc = upload(net, &R.last_spike_step, u64.data(), n))) return rc;
            for (uint32_t i = 0; i < n; ++i) u64[i] = nr[i].first_isi_spike;
            if ((rc = upload(net, &R.first_isi_spike, u64.data(), n))) return rc;
            R.dop_now = P.dop;
            if ((rc = dev_alloc(net, (void **)&R.dop_fix, (size_t)n * 8))) return rc;
            if ((rc = dev_alloc(net, (void **)&R.forced_now, n))) return rc;
            if ((rc = dev_alloc(net, (void **)&R.cand, (size_t)n * 4))) return rc;
            if ((rc = dev_alloc(net, (void **)&R.w_current, n))) return rc;  // zero: not known current
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nr[i].synapse_sum_threshold_coefficient; }, &R.coeff))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nr[i].synaptic_resource_threshold; }, &R.res_thr))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nr[i].stability_change_parameter; }, &R.change_param))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nr[i].stability_change_at_isi; }, &R.change_at_isi))) return rc;
            if ((rc = make_param_f(net, n, [&](uint32_t i) { return nr[i].d_h; }, &R.d_h))) return rc;
            if ((rc = make_param_u(net, n, [&](uint32_t i) { return nr[i].dopamine_plasticity_time; }, &R.dpt))) return rc;
            if ((rc = make_param_u(net, n, [&](uint32_t i) { return nr[i].resource_drain_coefficient; }, &R.drain))) return rc;
            if ((rc = make_param_u(net, n, [&](uint32_t i) { return nr[i].isi_max; }, &R.isi_max))) return rc;
        }
        const uint32_t chunks = std::max(1u, (n + kBlock - 1) / kBlock);
        P.spike_list = pool_cursor;
        pool_cursor += (size_t)chunks * kSegStride;
        if ((rc = dev_alloc(net, (void **)&P.bitmap, (size_t)std::max(P.words, 1u) * 4))) return rc;
        if (H.rec_cap) {
            if ((rc = dev_alloc(net, (void **)&P.rec, (size_t)H.rec_cap * P.words * 4))) return rc;
            P.rec_cap = H.rec_cap;
            P.rec_t0 = net->step;
            H.rec_t0 = net->step;
        }
        P.t_first = net->step;
        H.t_first = net->step;
    }

    // ---- tile-owner scatter: which populations are tiled, which projections get tile directories
    // A population is tiled when every impact it receives is a plain Delta EXCITATORY / INHIBITORY_CURRENT
    // impact (one int64 ring), its ring fits in LDS and its indexes fit the packed synapse record (23 bits).
    // Rows longer than 4 synapses on average get a tile directory: dense ([n_pre][tiles + 1]) unless rows span
    // few of many tiles (ranged, per row); shorter rows are read whole by every tile they may reach and
    // filtered.
    enum { DIR_NONE = 0, DIR_RANGED = 1, DIR_DENSE = 2 };
    std::vector<uint8_t> proj_dir(n_projs, DIR_NONE);
    std::vector<std::vector<uint32_t>> row_lo(n_projs), row_hi(n_projs);  // postsynaptic range of every row
    uint32_t tile_max_d = kTileMaxD;
    {
        // the LDS ring (D x 256 x 8 bytes), the segment table and the static LDS must fit one workgroup
        int lds_limit = 0;
        HIP_TRY(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, net->device));
        hipFuncAttributes fa{};
        HIP_TRY(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&k_tile_scatter<true, false, false>)));
        const int64_t room = (int64_t)lds_limit - (int64_t)fa.sharedSizeBytes - (int64_t)tile_lds_bytes(0, kTileMaxList);
        tile_max_d = (uint32_t)std::max<int64_t>(0, std::min<int64_t>(kTileMaxD, room / (kBlock * 8)));
    }
    for (uint32_t p = 0; p < n_pops; ++p) {
        HostPop &H = net->pops[p];
        H.tiled = net->scatter_mode != KNP_SCATTER_ATOMIC && !mt && !H.resource && !H.has_cond && !H.has_dop &&
                  !H.has_block && H.D <= tile_max_d && H.n > 0 && H.n < (1u << 23);
    }
    for (uint32_t q = 0; q < n_projs; ++q)
        if (net->projs[q].kind != PROJ_DELTA || net->projs[q].dense_format >= 0)
            net->pops[net->projs[q].post_pop].tiled = false;  // (MFMA expansions add into the ring in HBM)
    {
        std::vector<uint64_t> dir_entries(n_pops, 0);
        for (uint32_t q = 0; q < n_projs; ++q) {
            const HostProj &H = net->projs[q];
            if (!net->pops[H.post_pop].tiled) continue;
            auto &lo = row_lo[q], &hi = row_hi[q];
            lo.assign(H.n_pre, 0xFFFFFFFFu);
            hi.assign(H.n_pre, 0);
            std::vector<uint32_t> row_len(H.n_pre, 0);
            for (uint64_t k = 0; k < H.n_syn; ++k) {
                const uint32_t r = H.pre[k], j = H.post[k];
                lo[r] = std::min(lo[r], j);
                hi[r] = std::max(hi[r], j);
                row_len[r]++;
            }
            // part lengths in directory entries have kPartLenBits bits
            for (uint32_t r = 0; r < H.n_pre; ++r)
                if (row_len[r] >= (1u << kPartLenBits)) net->pops[H.post_pop].tiled = false;
            if (H.n_syn <= 4ull * H.n_pre) continue;
            // directory sizes in 4-byte words: 8-byte part entries, a 16-byte meta per ranged row
            uint64_t ranged = 0;
            for (uint32_t r = 0; r < H.n_pre; ++r)
                ranged += 4 + (lo[r] <= hi[r] ? 2 * (hi[r] / kBlock - lo[r] / kBlock + 1) : 0);
            const uint64_t tiles = (net->pops[H.post_pop].n + kBlock - 1) / kBlock;
            const uint64_t dense = 2 * (uint64_t)H.n_pre * tiles;
            proj_dir[q] = (dense <= 2 * ranged) ? DIR_DENSE : DIR_RANGED;
            dir_entries[H.post_pop] += std::min(dense, ranged);
        }
        for (uint32_t p = 0; p < n_pops; ++p)
            if (dir_entries[p] > (uint64_t(1) << 28)) net->pops[p].tiled = false;  // 1 GiB of directories
        for (uint32_t q = 0; q < n_projs; ++q)
            if (!net->pops[net->projs[q].post_pop].tiled) proj_dir[q] = DIR_NONE;
    }

    // ---- projections
    for (uint32_t q = 0; q < n_projs; ++q) {
        HostProj &H = net->projs[q];
        ProjDev &Q = T.projs[q];
        const uint32_t n_post = net->pops[H.post_pop].n;
        Q.kind = H.kind;
        Q.pre_pop = H.pre_pop;
        Q.post_pop = H.post_pop;
        Q.n_pre = H.n_pre;
        Q.n_syn = H.n_syn;
        Q.forcing = (H.kind == PROJ_DELTA) ? 1u : 0u;  // stdp_dispatcher.h:54-58, synaptic_resource_stdp.h:58-62
        const uint64_t S = H.n_syn;
        if (S >= 0xFFFFFFFFull) return fail(KNP_ERR_UNSUPPORTED, "projection with 2^32 or more synapses");
        // CSR by presynaptic neuron (stable: ascending synapse index within a row)
        std::vector<uint64_t> row_ptr((size_t)H.n_pre + 1, 0);
        for (uint64_t s = 0; s < S; ++s) row_ptr[H.pre[s] + 1]++;
        for (uint32_t i = 0; i < H.n_pre; ++i) row_ptr[i + 1] += row_ptr[i];
        std::vector<uint32_t> csr_syn(S);
        {
            std::vector<uint64_t> fill(row_ptr.begin(), row_ptr.end() - 1);
            if (net->pops[H.post_pop].tiled) {
                // rows ordered by the postsynaptic tile (a tile's part of a row is contiguous), in each part the
                // delay-1 records first (a persistent launch adds them before phase 1 of the next step and the
                // others after it), then by synapse index: a stable counting sort by that key, then the stable
                // fill by presynaptic neuron
                const uint32_t n_keys = 2 * ((n_post + kBlock - 1) / kBlock);
                auto key = [&](uint64_t s) { return 2 * (H.post[s] / kBlock) + (H.delay[s] == 1 ? 0u : 1u); };
                std::vector<uint64_t> col((size_t)n_keys + 1, 0);
                for (uint64_t s = 0; s < S; ++s) col[key(s) + 1]++;
                for (uint32_t j = 0; j < n_keys; ++j) col[j + 1] += col[j];
                std::vector<uint32_t> by_key(S);
                for (uint64_t s = 0; s < S; ++s) by_key[col[key(s)]++] = (uint32_t)s;
                for (uint64_t k = 0; k < S; ++k) csr_syn[fill[H.pre[by_key[k]]]++] = by_key[k];
            } else {
                for (uint64_t s = 0; s < S; ++s) csr_syn[fill[H.pre[s]]++] = (uint32_t)s;
            }
        }
        // delay-1 records at the start of the part [b, e)
        auto urgent_count = [&](uint64_t b, uint64_t e) -> uint32_t {
            uint64_t k = b;
            while (k < e && H.delay[csr_syn[k]] == 1) ++k;
            return (uint32_t)(k - b);
        };
        if (net->pops[H.post_pop].tiled) {
            // packed records for the tile scatter: the synapse's signed fixed-point impact m * 2^s (m in the low
            // 25 bits, two's complement; s in bits 25-30) | delay << 32 | post << 41. The impact is
            // round-half-even(w * 2^F), the value the float product w * 2^F (exact) converted to int64 gives,
            // negated for INHIBITORY_CURRENT; it has at most 24 significant bits, so m * 2^s holds it exactly.
            const double scale = std::ldexp(1.0, net->pops[H.post_pop].frac_bits);
            std::vector<uint64_t> rec(S);
            for (uint64_t k = 0; k < S; ++k) {
                const uint32_t sy = csr_syn[k];
                long long v = std::llrint((double)H.w[sy] * scale);
                if (H.type[sy] == KNP_INHIBITORY_CURRENT) v = -v;
                const uint64_t mag = (uint64_t)(v < 0 ? -v : v);
                // exact: an impact of 2^24 or more is a multiple of 2^(its binade - 23)
                const uint32_t sh = mag >= (1ull << 24) ? (uint32_t)(63 - __builtin_clzll(mag)) - 23 : 0;
                v /= (1ll << sh);
                rec[k] = ((uint64_t)v & 0x1FFFFFFull) | ((uint64_t)sh << 25) | ((uint64_t)(H.delay[sy] & 0xFFu) << 32) |
                         ((uint64_t)H.post[sy] << 41);
            }
            int rc2;
            uint64_t *rd;
            rec.push_back(0);  // one record past the end: the device loads records in aligned pairs
            if ((rc2 = upload(net, &rd, rec.data(), S + 1))) return rc2;
            Q.syn8 = rd;
        }
        if (proj_dir[q] == DIR_DENSE) {
            const uint32_t tiles = (n_post + kBlock - 1) / kBlock;
            Q.tile_stride = tiles;
            std::vector<uint64_t> dir((size_t)H.n_pre * tiles);
            for (uint32_t i = 0; i < H.n_pre; ++i) {
                uint64_t k = row_ptr[i];
                const uint64_t e = row_ptr[i + 1];
                for (uint32_t x = 0; x < tiles; ++x) {
                    const uint64_t b = k;
                    while (k < e && H.post[csr_syn[k]] / kBlock <= x) ++k;
                    dir[(size_t)i * tiles + x] = dir_entry((uint32_t)b, (uint32_t)(k - b), urgent_count(b, k));
                }
            }
            int rc2;
            uint64_t *dd;
            if ((rc2 = upload(net, &dd, dir.data(), dir.size()))) return rc2;
            Q.tile_dir = dd;
        } else if (proj_dir[q] == DIR_RANGED) {
            std::vector<RowMeta> meta(H.n_pre);
            std::vector<uint64_t> dir;
            for (uint32_t i = 0; i < H.n_pre; ++i) {
                const uint64_t b = row_ptr[i], e = row_ptr[i + 1];
                RowMeta &m = meta[i];
                m.begin = (uint32_t)b;
                m.dir = (uint32_t)dir.size();
                m.first_tile = 0;
                m.n_tiles = 0;
                if (b == e) continue;
                m.first_tile = H.post[csr_syn[b]] / kBlock;
                m.n_tiles = H.post[csr_syn[e - 1]] / kBlock - m.first_tile + 1;
                uint64_t k = b;
                for (uint32_t j = 0; j < m.n_tiles; ++j) {
                    const uint64_t pb = k;
                    while (k < e && H.post[csr_syn[k]] / kBlock <= m.first_tile + j) ++k;
                    dir.push_back(dir_entry((uint32_t)(pb - b), (uint32_t)(k - pb), urgent_count(pb, k)));
                }
            }
            int rc2;
            RowMeta *md;
            if ((rc2 = upload(net, &md, meta.data(), meta.size()))) return rc2;
            Q.row_meta = md;
            uint64_t *dd;
            if ((rc2 = upload(net, &dd, dir.data(), dir.size()))) return rc2;
            Q.row_dir = dd;
        }
        if (H.has_blocking)
            for (uint32_t i = 0; i < H.n_pre; ++i)
                if (row_ptr[i + 1] - row_ptr[i] >= (1u << 17))
                    return fail(KNP_ERR_UNSUPPORTED, "BLOCKING projection row longer than 131071 synapses");
        std::vector<uint32_t> post_type(S);
        std::vector<uint8_t> dl(S);
        for (uint64_t k = 0; k < S; ++k) {
            const uint32_t s = csr_syn[k];
            post_type[k] = H.post[s] | ((uint32_t)H.type[s] << 29);
            dl[k] = (uint8_t)H.delay[s];
        }
        int rc;
        uint64_t *rp;
        if ((rc = upload(net, &rp, row_ptr.data(), row_ptr.size()))) return rc;
        Q.row_ptr = rp;
        uint32_t *pt;
        if ((rc = upload(net, &pt, post_type.data(), S))) return rc;
        Q.post_type = pt;
        uint8_t *dd;
        if ((rc = upload(net, &dd, dl.data(), S))) return rc;
        Q.delay = dd;
        uint32_t *cs;
        if ((rc = upload(net, &cs, csr_syn.data(), S))) return rc;
        Q.csr_syn = cs;
        if (H.kind == PROJ_DELTA) {
            std::vector<float> wc(S);
            for (uint64_t k = 0; k < S; ++k) wc[k] = H.w[csr_syn[k]];
            float *wd;
            if ((rc = upload(net, &wd, wc.data(), S))) return rc;
            Q.w_csr = wd;
        } else {
            if ((rc = upload(net, &Q.W, H.w.data(), S))) return rc;
            // CSC by postsynaptic neuron (stable)
            std::vector<uint64_t> col_ptr((size_t)n_post + 1, 0);
            for (uint64_t s = 0; s < S; ++s) col_ptr[H.post[s] + 1]++;
            for (uint32_t i = 0; i < n_post; ++i) col_ptr[i + 1] += col_ptr[i];
            std::vector<uint32_t> col_syn(S);
            {
                std::vector<uint64_t> fill(col_ptr.begin(), col_ptr.end() - 1);
                for (uint64_t s = 0; s < S; ++s) col_syn[fill[H.post[s]]++] = (uint32_t)s;
            }
            uint64_t *cp;
            if ((rc = upload(net, &cp, col_ptr.data(), col_ptr.size()))) return rc;
            Q.col_ptr = cp;
            uint32_t *cy;
            if ((rc = upload(net, &cy, col_syn.data(), S))) return rc;
            Q.col_syn = cy;
            bool identity = true;
            for (uint64_t k = 0; k < S && identity; ++k) identity = col_syn[k] == (uint32_t)k;
            Q.csc_identity = identity ? 1u : 0u;
            std::vector<uint8_t> ds(S);
            for (uint64_t s = 0; s < S; ++s) ds[s] = (uint8_t)H.delay[s];
            uint8_t *dsd;
            if ((rc = upload(net, &dsd, ds.data(), S))) return rc;
            Q.delay_syn = dsd;
            std::vector<uint32_t> cpre(S);
            for (uint64_t k = 0; k < S; ++k) cpre[k] = H.pre[col_syn[k]];
            uint32_t *cpd;
            if ((rc = upload(net, &cpd, cpre.data(), S))) return rc;
            Q.col_pre = cpd;
            // columns of consecutive presynaptic indexes (all-to-all blocks): the index follows from the position
            bool consecutive = true;
            for (uint32_t j = 0; j < n_post && consecutive; ++j)
                for (uint64_t k = col_ptr[j] + 1; k < col_ptr[j + 1] && consecutive; ++k)
                    consecutive = cpre[k] == cpre[k - 1] + 1;
            if (consecutive) {
                std::vector<uint32_t> c0(std::max(1u, n_post), 0);
                for (uint32_t j = 0; j < n_post; ++j) c0[j] = col_ptr[j] < col_ptr[j + 1] ? cpre[col_ptr[j]] : 0;
                uint32_t *c0d;
                if ((rc = upload(net, &c0d, c0.data(), c0.size()))) return rc;
                Q.col_pre0 = c0d;
            }
        }
        if (H.kind == PROJ_RESOURCE) {
            std::vector<float> f(S);
            std::vector<uint32_t> u(S);
            std::vector<uint64_t> u64(S);
            std::vector<uint8_t> fl(S);
            for (uint64_t s = 0; s < S; ++s) f[s] = H.rules[s].synaptic_resource;
            if ((rc = upload(net, &Q.r, f.data(), S))) return rc;
            for (uint64_t s = 0; s < S; ++s) f[s] = H.rules[s].w_min;
            if ((rc = upload(net, &Q.w_min, f.data(), S))) return rc;
            for (uint64_t s = 0; s < S; ++s) f[s] = H.rules[s].w_max;
            if ((rc = upload(net, &Q.w_max, f.data(), S))) return rc;
            for (uint64_t s = 0; s < S; ++s) f[s] = H.rules[s].d_u;
            if ((rc = upload(net, &Q.d_u, f.data(), S))) return rc;
            for (uint64_t s = 0; s < S; ++s) u[s] = H.rules[s].dopamine_plasticity_period;
            if ((rc = upload(net, &Q.period, u.data(), S))) return rc;
            // parameters every synapse carries alike are read from the projection entry instead
            auto same = [&](auto get) {
                for (uint64_t s = 1; s < S; ++s)
                    if (!(get(s) == get(0))) return false;
                return S > 0;
            };
            auto fbits = [](float x) {
                uint32_t b;
                std::memcpy(&b, &x, 4);
                return b;
            };
            Q.rule_uniform = 0;
            if (same([&](uint64_t s) { return fbits(H.rules[s].w_min); })) Q.rule_uniform |= kUniWmin;
            if (same([&](uint64_t s) { return fbits(H.rules[s].w_max); })) Q.rule_uniform |= kUniWmax;
            if (same([&](uint64_t s) { return fbits(H.rules[s].d_u); })) Q.rule_uniform |= kUniDu;
            if (same([&](uint64_t s) { return H.rules[s].dopamine_plasticity_period; })) Q.rule_uniform |= kUniPeriod;
            if (same([&](uint64_t s) { return (uint8_t)H.delay[s]; })) Q.rule_uniform |= kUniDelay;  // as delay_syn
            if (S > 0) {
                Q.u_wmin = H.rules[0].w_min;
                Q.u_wmax = H.rules[0].w_max;
                Q.u_du = H.rules[0].d_u;
                Q.u_period = H.rules[0].dopamine_plasticity_period;
                Q.u_delay = (uint8_t)H.delay[0];
            }
            for (uint64_t s = 0; s < S; ++s) u64[s] = H.rules[s].last_spike_step;
            if ((rc = upload(net, &Q.last_spike_step, u64.data(), S))) return rc;
            // each row's last_spike_step until it spikes: the value its synapses share, else UINT64_MAX
            std::vector<uint64_t> row0((size_t)std::max(1u, H.n_pre), UINT64_MAX);
            std::vector<uint8_t> seen(row0.size(), 0), mixed(row0.size(), 0);
            for (uint64_t s = 0; s < S; ++s) {
                const uint32_t r = H.pre[s];
                if (!seen[r]) {
                    seen[r] = 1;
                    row0[r] = u64[s];
                } else if (row0[r] != u64[s]) {
                    mixed[r] = 1;
                }
            }
            Q.lss_by_row = 1;
            for (size_t r = 0; r < row0.size(); ++r)
                if (mixed[r]) {
                    row0[r] = UINT64_MAX;
                    Q.lss_by_row = 0;
                }
            if ((rc = upload(net, &Q.row_spike, row0.data(), row0.size()))) return rc;
            for (uint64_t s = 0; s < S; ++s)
                fl[s] = (uint8_t)((H.rules[s].had_hebbian_update ? 1 : 0) | (H.rules[s].has_contributed ? 2 : 0));
            if ((rc = upload(net, &Q.flags, fl.data(), S))) return rc;
        }
        if (H.kind == PROJ_ADDITIVE) {
            if ((rc = upload(net, &Q.tau_plus, H.tau_p.data(), S))) return rc;
            if ((rc = upload(net, &Q.tau_minus, H.tau_m.data(), S))) return rc;
            std::vector<uint32_t> lim(S);
            uint32_t hist = 1;
            for (uint64_t s = 0; s < S; ++s) {
                // push while size() < tau_minus_ + tau_plus_ (float comparison, additive_stdp.h:108)
                const float sum = H.tau_m[s] + H.tau_p[s];
                uint32_t l = 0;
                while ((float)l < sum && l < 256) ++l;
                lim[s] = l;
                hist = std::max(hist, l);
            }
            if ((rc = upload(net, &Q.lim, lim.data(), S))) return rc;
            Q.hist = hist;
            if ((rc = dev_alloc(net, (void **)&Q.pre_times, (size_t)S * hist * 8))) return rc;
            if ((rc = dev_alloc(net, (void **)&Q.post_times, (size_t)S * hist * 8))) return rc;
            if ((rc = dev_alloc(net, (void **)&Q.n_pre_t, (size_t)S * 4))) return rc;
            if ((rc = dev_alloc(net, (void **)&Q.n_post_t, (size_t)S * 4))) return rc;
            if (H.stdp_pops.size() > 4) return fail(KNP_ERR_UNSUPPORTED, "more than 4 stdp populations");
            Q.n_stdp = (uint32_t)H.stdp_pops.size();
            for (uint32_t m = 0; m < Q.n_stdp; ++m) {
                Q.stdp_pop[m] = H.stdp_pops[m];
                Q.stdp_mode[m] = H.stdp_modes[m];
            }
        }
        if (H.dense_format >= 0) {
            // MFMA expansion: instance table, index maps, tiles, the BF16 weight copy and the spike counts
            const DenseBlocks &B = dense_info[q];
            std::vector<uint32_t> tiles;
            for (uint32_t k = 0; k < B.inst.size(); ++k)
                for (uint32_t j = 0; j < B.inst[k].post_count; j += 32) tiles.push_back(k << 16 | j / 32);
            const DenseInst &L = B.inst.back();
            const bool split = H.dense_flags & KNP_DENSE_SPLIT;
            std::vector<uint16_t> wb(L.base + (uint64_t)L.n_oct * L.npad * 8, 0), wlo(split ? wb.size() : 0, 0);
            for (uint64_t s2 = 0; s2 < S; ++s2) {
                const DenseInst &I = B.inst[B.post_inst[H.post[s2]]];
                const uint32_t i = H.pre[s2] - I.pre_begin, j = H.post[s2] - I.post_begin;
                const uint64_t pos = I.base + ((uint64_t)(i >> 3) * I.npad + j) * 8 + (i & 7);
                wb[pos] = bf16_rne(H.w[s2]);
                if (split) wlo[pos] = bf16_lo(H.w[s2], wb[pos]);
            }
            if (split && (rc = upload(net, &Q.dn_w_lo, wlo.data(), wlo.size()))) return rc;
            DenseInst *it;
            uint32_t *pi, *po, *tl;
            uint16_t *wd;
            if ((rc = upload(net, &it, B.inst.data(), B.inst.size()))) return rc;
            if ((rc = upload(net, &pi, B.pre_inst.data(), B.pre_inst.size()))) return rc;
            if ((rc = upload(net, &po, B.post_inst.data(), B.post_inst.size()))) return rc;
            if ((rc = upload(net, &tl, tiles.data(), tiles.size()))) return rc;
            // one-load tables of the kernels: per tile its instance's fields, per input its count byte
            std::vector<DenseTile> tinfo(tiles.size());
            for (size_t x = 0; x < tiles.size(); ++x) {
                const DenseInst &I = B.inst[tiles[x] >> 16];
                tinfo[x] = DenseTile{I.cbase, I.n_oct, I.npad, I.post_begin, I.post_count, (tiles[x] & 0xFFFFu) * 32, I.base};
            }
            std::vector<uint64_t> pinfo(B.pre_inst.size(), UINT64_MAX);
            for (size_t i = 0; i < B.pre_inst.size(); ++i) {
                if (B.pre_inst[i] == 0xFFFFFFFFu) continue;
                const DenseInst &I = B.inst[B.pre_inst[i]];
                const uint32_t x = (uint32_t)i - I.pre_begin;
                pinfo[i] = (uint64_t)(I.cbase + (x >> 3)) | (uint64_t)(x & 7) << 32 | (uint64_t)I.post_count << 35;
            }
            DenseTile *ti;
            uint64_t *pf;
            if ((rc = upload(net, &ti, tinfo.data(), tinfo.size()))) return rc;
            if ((rc = upload(net, &pf, pinfo.data(), pinfo.size()))) return rc;
            Q.dn_tile_info = ti;
            Q.dn_pre_info = pf;
            H.dense_sig.resize(tinfo.size());
            for (size_t x = 0; x < tinfo.size(); ++x)
                H.dense_sig[x] = (uint64_t)(tinfo[x].post_begin + tinfo[x].j0) << 32 |
                                 std::min<uint32_t>(32u, tinfo[x].post_count - tinfo[x].j0);
            if ((rc = upload(net, &wd, wb.data(), wb.size()))) return rc;
            if ((rc = dev_alloc(net, (void **)&Q.dn_cnt, (size_t)2 * B.n_oct * 8))) return rc;
            if ((rc = dev_alloc(net, (void **)&Q.dn_part, (size_t)kDenseCountBlocks * 4 * 8))) return rc;
            Q.dense = 1 + H.dense_flags;
            Q.dn_oct = B.n_oct;
            Q.dn_max_oct = Q.dn_max_npad = 0;
            for (const DenseInst &I : B.inst) {
                Q.dn_max_oct = std::max(Q.dn_max_oct, I.n_oct);
                Q.dn_max_npad = std::max(Q.dn_max_npad, I.npad);
            }
            Q.dn_inst = (uint32_t)B.inst.size();
            Q.dn_tiles = (uint32_t)tiles.size();
            Q.dn_delay = B.delay;
            Q.dn_sign = B.sign;
            Q.dn_inst_tab = it;
            Q.dn_pre_inst = pi;
            Q.dn_post_inst = po;
            Q.dn_tile = tl;
            Q.dn_w = wd;
        }
        // channel messages of one step: every index at most 4 times (checked when pushed), so this capacity
        // is never exceeded and the list never moves (tile tables hold its address)
        Q.in_cap = 4 * std::max(H.n_pre, 1u);
        Q.in_list = pool_cursor;
        pool_cursor += channel_seg_words(H.n_pre);
    }

    // ---- sources
    for (uint32_t s = 0; s < T.n_srcs; ++s) {
        const HostSource &H = net->srcs[s];
        SourceDev &S = T.srcs[s];
        S.proj = H.proj;
        S.first = H.first;
        S.hash_offset = H.hash_offset;
        S.n_ch = H.n_ch;
        S.n_patterns = H.n_patterns;
        S.steps_per_pattern = H.spp;
        S.lo = H.lo;
        S.hi = H.hi;
        S.seed = H.seed;
        uint32_t *th;
        int rc;
        if ((rc = upload(net, &th, H.thr.data(), H.thr.size()))) return rc;
        S.thr = th;
        const uint32_t chunks = std::max(1u, (H.n_ch + kBlock - 1) / kBlock);
        S.list = pool_cursor;
        pool_cursor += (size_t)chunks * kSegStride;
    }

    // ---- spike segments and the scatter's virtual segment list
    std::vector<std::vector<uint32_t>> tile_reads;  // per tile: the population and source segments it reads
    std::unordered_map<uint32_t, uint32_t> seg_tile;  // population segment -> the tile that writes it
    bool multi_ok = false;
    {
        int rc;
        std::vector<uint32_t *> seg_list;
        std::vector<uint32_t> pop_chunks(n_pops);
        for (uint32_t p = 0; p < n_pops; ++p) {
            PopDev &P = T.pops[p];
            pop_chunks[p] = std::max(1u, (P.n + kBlock - 1) / kBlock);
            P.seg_base = (uint32_t)seg_list.size();
            for (uint32_t c = 0; c < pop_chunks[p]; ++c) seg_list.push_back(P.spike_list + (size_t)c * kSegStride);
        }
        // every (projection, segment) pair whose spikes the projection expands, with the presynaptic rows the
        // segment can hold: for every projection (load order) its channel segments, then the segments of every
        // population whose spike messages it receives
        struct VDesc {
            uint32_t seg, rs, q, row_lo, row_hi, key;  // key: message order under the multi-threaded semantics
        };
        std::vector<VDesc> vd;
        uint32_t nrs = 0;
        for (uint32_t q = 0; q < n_projs; ++q) {
            const HostProj &H = net->projs[q];
            ProjDev &Q = T.projs[q];
            // channel messages: host pushes, then every source feeding q
            Q.push_seg = (uint32_t)seg_list.size();
            seg_list.push_back(Q.in_list);
            if (nrs >= (uint32_t)kMaxRowSources) return fail(KNP_ERR_UNSUPPORTED, "too many row sources");
            T.rows[nrs] = RowSource{q, 0u};
            vd.push_back(VDesc{Q.push_seg, nrs, q, 0u, H.n_pre, kMsgKeyPush});
            for (uint32_t si = 0; si < T.n_srcs; ++si) {
                SourceDev &S = T.srcs[si];
                if (S.proj != q) continue;
                const uint32_t chunks = std::max(1u, (S.n_ch + kBlock - 1) / kBlock);
                S.seg_base = (uint32_t)seg_list.size();
                for (uint32_t c = 0; c < chunks; ++c) {
                    const uint32_t r0 = S.first + c * kBlock, r1 = S.first + std::min((c + 1) * kBlock, S.n_ch);
                    vd.push_back(VDesc{(uint32_t)seg_list.size(), nrs, q, r0, r1, kMsgKeySource + si});
                    seg_list.push_back(S.list + (size_t)c * kSegStride);
                }
            }
            ++nrs;
            // population messages in population order: the presynaptic population and, for additive STDP,
            // every stdp population whose messages also produce impacts (STDPOnly messages are emptied,
            // additive_stdp.h:162-166)
            for (uint32_t p = 0; p < n_pops; ++p) {
                bool sub = (H.pre_pop == p);
                int mode = -1;
                for (size_t m = 0; m < H.stdp_pops.size(); ++m)
                    if (H.stdp_pops[m] == p) {
                        sub = true;
                        mode = H.stdp_modes[m];
                    }
                if (!sub || mode == KNP_STDP_ONLY) continue;
                if (nrs >= (uint32_t)kMaxRowSources) return fail(KNP_ERR_UNSUPPORTED, "too many row sources");
                T.rows[nrs] = RowSource{q, 1u + p};
                for (uint32_t c = 0; c < pop_chunks[p]; ++c)
                    vd.push_back(VDesc{T.pops[p].seg_base + c, nrs, q, c * kBlock,
                                       std::min((c + 1) * kBlock, T.pops[p].n), kMsgKeyPopulation});
                ++nrs;
            }
        }
        // tiles of tiled populations and the segments that can reach each of them (through the postsynaptic
        // range of the rows a segment can hold); a population with a tile reached by more than kTileMaxList
        // segments stays on the atomic path
        std::vector<std::pair<uint32_t, uint32_t>> vd_tiles(vd.size(), {1u, 0u});  // tile range, empty = {1,0}
        for (size_t k = 0; k < vd.size(); ++k) {
            const VDesc &d = vd[k];
            const uint32_t post = net->projs[d.q].post_pop;
            if (!net->pops[post].tiled) continue;
            const auto &lo = row_lo[d.q], &hi = row_hi[d.q];
            uint32_t tlo = 0xFFFFFFFFu, thi = 0;
            for (uint32_t r = d.row_lo; r < d.row_hi && r < (uint32_t)lo.size(); ++r)
                if (lo[r] <= hi[r]) {
                    tlo = std::min(tlo, lo[r] / kBlock);
                    thi = std::max(thi, hi[r] / kBlock);
                }
            if (tlo <= thi) vd_tiles[k] = {tlo, thi};
        }
        for (uint32_t p = 0; p < n_pops; ++p) {
            if (!net->pops[p].tiled) continue;
            std::vector<uint32_t> reach((net->pops[p].n + kBlock - 1) / kBlock, 0);
            for (size_t k = 0; k < vd.size(); ++k)
                if (net->projs[vd[k].q].post_pop == p)
                    for (uint32_t x = vd_tiles[k].first; x <= vd_tiles[k].second && vd_tiles[k].first <= vd_tiles[k].second; ++x)
                        reach[x]++;
            for (uint32_t c : reach)
                if (c > (uint32_t)kTileMaxList) net->pops[p].tiled = false;
        }
        // virtual segment order: the atomic scatter's segments first, then those of the projections expanded on
        // MFMA (each projection's contiguous), then the tiled populations'
        auto vclass = [&](size_t k) {
            if (net->projs[vd[k].q].dense_format >= 0) return 1;
            return net->pops[net->projs[vd[k].q].post_pop].tiled ? 2 : 0;
        };
        std::vector<uint32_t> vseg, vrs, vkey, vidx(vd.size());
        for (int pass = 0; pass < 3; ++pass)
            for (size_t k = 0; k < vd.size(); ++k) {
                if (vclass(k) != pass) continue;
                vidx[k] = (uint32_t)vseg.size();
                vkey.push_back(vd[k].key);
                ProjDev &Qv = T.projs[vd[k].q];
                if (pass == 1) {
                    if (Qv.dn_v1 == Qv.dn_v0) Qv.dn_v0 = (uint32_t)vseg.size();
                    Qv.dn_v1 = (uint32_t)vseg.size() + 1;
                }
                vseg.push_back(vd[k].seg);
                vrs.push_back(vd[k].rs);
            }
        uint32_t n_atomic = 0;
        for (size_t k = 0; k < vd.size(); ++k)
            if (vclass(k) == 0) ++n_atomic;
        if (n_atomic > (uint32_t)kMaxVSeg) return fail(KNP_ERR_UNSUPPORTED, "too many spike segments for one step");
        T.n_vseg_atomic = n_atomic;
        std::vector<TileDev> tiles;
        std::vector<TileSeg> tile_seg;
        uint32_t maxD = 0, maxL = 0;
        for (uint32_t p = 0; p < n_pops; ++p) {
            T.pops[p].tiled = net->pops[p].tiled ? 1u : 0u;
            if (!net->pops[p].tiled) continue;
            maxD = std::max(maxD, T.pops[p].D);
            const uint32_t nt = (net->pops[p].n + kBlock - 1) / kBlock;
            std::vector<std::vector<uint32_t>> reach(nt);
            for (size_t k = 0; k < vd.size(); ++k)
                if (net->projs[vd[k].q].post_pop == p && vd_tiles[k].first <= vd_tiles[k].second)
                    for (uint32_t x = vd_tiles[k].first; x <= vd_tiles[k].second; ++x) reach[x].push_back((uint32_t)k);
            for (uint32_t x = 0; x < nt; ++x) {
                seg_tile[T.pops[p].seg_base + x] = (uint32_t)tiles.size();
                tiles.push_back(TileDev{p, x, (uint32_t)tile_seg.size(), (uint32_t)(tile_seg.size() + reach[x].size())});
                maxL = std::max(maxL, (uint32_t)reach[x].size());
                tile_reads.emplace_back();
                for (uint32_t k : reach[x]) {
                    const bool chan = vd[k].seg == T.projs[vd[k].q].push_seg;
                    tile_seg.push_back(TileSeg{seg_list[vd[k].seg], vd[k].q | (chan ? kTileSegChannel : 0u)});
                    if (!chan) tile_reads.back().push_back(vd[k].seg);
                }
            }
        }
        T.n_tiles = (uint32_t)tiles.size();
        T.tile_max_d = maxD;
        T.tile_max_list = maxL;
        net->tile_lds = (uint32_t)tile_lds_bytes(maxD, maxL);
        if (T.n_tiles) {
            TileDev *td;
            if ((rc = upload(net, &td, tiles.data(), tiles.size()))) return rc;
            T.tiles = td;
            TileSeg *tsd;
            if ((rc = upload(net, &tsd, tile_seg.data(), tile_seg.size()))) return rc;
            T.tile_seg = tsd;
        }
        if ((rc = dev_alloc(net, (void **)&T.tile_part, (size_t)4 * std::max(1u, T.n_tiles) * 8))) return rc;
        T.n_row_sources = nrs;
        T.n_seg = (uint32_t)seg_list.size();
        T.n_vseg = (uint32_t)vseg.size();
        uint32_t **sl;
        uint32_t *vs, *vr;
        if ((rc = upload(net, &sl, seg_list.data(), seg_list.size()))) return rc;
        T.seg_list = sl;
        net->seg_list_host = seg_list;
        if ((rc = upload(net, &vs, vseg.data(), vseg.size()))) return rc;
        if ((rc = upload(net, &vr, vrs.data(), vrs.size()))) return rc;
        T.vseg = vs;
        T.vrs = vr;
        T.vkey = nullptr;
        if (mt) {
            uint32_t *vk;
            if ((rc = upload(net, &vk, vkey.data(), vkey.size()))) return rc;
            T.vkey = vk;
        }
        if ((rc = dev_alloc(net, (void **)&T.seg_count, (size_t)2 * T.n_seg * 4))) return rc;
        if ((rc = dev_alloc(net, (void **)&T.vcount, (size_t)2 * std::max(1u, T.n_vseg) * 4))) return rc;
        std::vector<uint32_t> vpos((size_t)4 * T.n_seg, 0xFFFFFFFFu);
        for (uint32_t v = 0; v < T.n_vseg; ++v) {
            uint32_t *slot = &vpos[(size_t)4 * vseg[v]];
            int k = 0;
            while (k < 4 && slot[k] != 0xFFFFFFFFu) ++k;
            if (k == 4) return fail(KNP_ERR_UNSUPPORTED, "a spike source feeds more than 4 projections");
            slot[k] = v;
        }
        uint32_t *vp;
        if ((rc = upload(net, &vp, vpos.data(), vpos.size()))) return rc;
        T.seg_vpos = vp;
        net->push_vpos.assign(n_projs, 0xFFFFFFFFu);
        for (uint32_t q = 0; q < n_projs; ++q) net->push_vpos[q] = vpos[(size_t)4 * T.projs[q].push_seg];
        // segments hold at most 256 rows (pushes more, handled by the grid-stride loop)
        const uint64_t max_rows = (uint64_t)n_atomic * kBlock;
        net->scatter_grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(2048, max_rows));
        net->scatter_grid_alloc = net->scatter_grid;
        if ((rc = dev_alloc(net, (void **)&T.part, (size_t)2 * net->scatter_grid * 8))) return rc;
    }

    // ---- phase-1 tasks
    uint32_t blocks = 0, nt = 0;
    for (uint32_t p = 0; p < n_pops; ++p) {
        const uint32_t b = std::max(1u, (net->pops[p].n + kBlock - 1) / kBlock);
        T.tasks[nt++] = TaskDev{0u, p, blocks, b};
        blocks += b;
    }
    for (uint32_t s = 0; s < T.n_srcs; ++s) {
        const uint32_t b = std::max(1u, (net->srcs[s].n_ch + kBlock - 1) / kBlock);
        T.tasks[nt++] = TaskDev{1u, s, blocks, b};
        blocks += b;
    }
    T.n_tasks = nt;
    T.total_blocks = blocks;
    T.n_src_blocks = 0;
    for (uint32_t s = 0; s < T.n_srcs; ++s) T.n_src_blocks += std::max(1u, (net->srcs[s].n_ch + kBlock - 1) / kBlock);
    net->fused = T.n_tiles > 0;
    for (const auto &H : net->pops) net->fused = net->fused && H.tiled;
    // persistent launches: every population tiled, every tile on the device at once, and
    // at most kCachedSrcBlocks source blocks per tile
    net->persistent = false;
    net->run_kind = POP_LEAN;
    if (net->fused && (net->scatter_mode == KNP_SCATTER_AUTO || net->scatter_mode == KNP_SCATTER_TILES)) {
        bool one_kind = true;
        for (uint32_t p = 0; p < n_pops; ++p) one_kind = one_kind && T.pops[p].kind == T.pops[0].kind;
        net->run_kind = one_kind ? T.pops[0].kind : kMixedKinds;
        const void *fn = tile_run_kernel(net->run_kind);
        int coop = 0, cus = 0, per_cu = 0;
        HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, net->device));
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, net->device));
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kTileThreads, net->tile_lds));
        const bool src_fit = T.n_src_blocks <= (uint64_t)kCachedSrcBlocks * T.n_tiles;
        net->persistent = coop && src_fit && (uint64_t)per_cu * cus >= T.n_tiles;
        multi_ok = net->persistent && waits_cover_readers(net, tile_reads, seg_tile);
        // a network whose neurons all fit on one workgroup's threads runs its steps inside that workgroup
        net->small = false;
        if (net->persistent && net->scatter_mode == KNP_SCATTER_AUTO && one_kind &&
            (T.pops[0].kind == POP_LEAN || T.pops[0].kind == POP_FULL) &&
            n_pops <= kSmallMaxPops && net->dense.empty() && n_projs <= (uint32_t)kMaxProjs &&
            net->test_hold_block == 0xFFFFFFFFu) {
            SmallRun L{};
            L.n_pops = n_pops;
            uint64_t threads = 0, cells = 0, rows = 0;
            bool ok = true;
            for (uint32_t p = 0; p < n_pops; ++p) {
                L.pop_off[p] = (uint32_t)threads;
                L.ring_off[p] = (uint32_t)cells;
                threads += (net->pops[p].n + 63) / 64 * 64;
                cells += (uint64_t)T.pops[p].D * net->pops[p].n;
            }
            uint64_t row_words = 0, thr_words = 0;
            for (uint32_t q = 0; q < n_projs; ++q) {
                const HostProj &H = net->projs[q];
                ok = ok && H.dense_format < 0 && T.projs[q].syn8 && T.projs[q].row_ptr && H.n_pre < (1u << 24) &&
                     T.projs[q].n_syn < (uint64_t(1) << 32);
                L.row_base[q] = (uint32_t)row_words;
                row_words += (uint64_t)H.n_pre + 1;
                if (H.pre_pop != KNP_NO_POPULATION) {
                    L.proj_of_pop[H.pre_pop] |= 1ull << q;
                    rows += net->pops[H.pre_pop].n;
                }
            }
            ok = ok && T.n_srcs <= kSmallMaxSrcs;
            for (uint32_t s2 = 0; s2 < T.n_srcs && s2 < kSmallMaxSrcs; ++s2) {
                rows += net->srcs[s2].n_ch;
                L.thr_base[s2] = 0xFFFFFFFFu;
                if (T.srcs[s2].n_patterns == 1) {
                    L.thr_base[s2] = (uint32_t)thr_words;
                    thr_words += net->srcs[s2].n_ch;
                }
            }
            L.ring_cells = (uint32_t)cells;
            L.row_words = (uint32_t)row_words;
            const uint64_t lds64 = cells * 8 + row_words * 4 + thr_words * 4;
            ok = ok && lds64 <= 128 * 1024;
            const uint32_t lds = (uint32_t)lds64;
            int per_cu1 = 0;
            if (ok && threads <= kSmallThreads && cells <= kSmallRingCells && rows <= kSmallRows) {
                const void *fs = T.pops[0].kind == POP_LEAN ? reinterpret_cast<const void *>(&k_small_run<POP_LEAN>)
                                                            : reinterpret_cast<const void *>(&k_small_run<POP_FULL>);
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu1, fs, kSmallThreads, lds));
            }
            if (per_cu1 > 0) {
                net->small = true;
                net->small_run = L;
                net->small_lds = lds;
            }
        }
    }
    // the persistent launch of every tile needs waits that keep each segment's readers within a step of its
    // writer (waits_cover_readers); the one-workgroup launch runs every step in order anyway
    if (!multi_ok && !net->small) net->persistent = false;
    net->can_persistent = multi_ok;
    net->can_small = net->small;
    net->wide_tiles = false;
    net->tiles_shallow = true;
    for (const auto &H : net->pops)
        if (H.tiled && H.D > kTileShallowD) net->tiles_shallow = false;
    if (T.n_tiles) {
        int cus = 0, per_cu = 0;
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, net->device));
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(&k_tile_scatter<true, false, false>), kTileThreads, net->tile_lds));
        net->wide_tiles = (uint64_t)per_cu * cus < T.n_tiles;
    }
    {
        int rc;
        if ((rc = dev_alloc(net, (void **)&T.grid_bar, kBarWords * 4))) return rc;
        T.spin_limit = net->test_spin_limit;
        T.hold_block = net->test_hold_block;
        T.hold_until = net->test_hold_until;
    }

    // ---- control blocks
    int rc;
    if ((rc = dev_alloc(net, (void **)&T.ctl, 2 * sizeof(Ctl)))) return rc;
    if ((rc = dev_alloc(net, (void **)&T.events, 32))) return rc;
    if ((rc = dev_alloc(net, (void **)&T.step_ptr, 8))) return rc;

    // ---- dense input projections
    T.n_dense = (uint32_t)net->dense.size();
    for (uint32_t d = 0; d < T.n_dense; ++d)
        if ((rc = init_dense(net, d))) return rc;

    // launch shapes

    if ((rc = upload_working_sets(net))) return rc;
    net->initialized = true;
    return KNP_OK;
}


int upload_table(knp_net *net) {
    if (!net->dev_table) {
        int rc = dev_alloc(net, (void **)&net->dev_table, sizeof(NetDev), false);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(net->dev_table, &net->table, sizeof(NetDev), hipMemcpyHostToDevice, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    return KNP_OK;
}

// Working projections of a resource population: unlocked synaptic-resource STDP projections that target it,
// in load order (find_projection_by_type_and_postsynaptic, populations_old.h:57-83). Refreshed whenever a
// lock changes.
int upload_working_sets(knp_net *net) {
    for (uint32_t p = 0; p < net->pops.size(); ++p) {
        if (!net->pops[p].resource) continue;
        ResDev &R = net->table.pops[p].res;
        const uint32_t before = R.n_work;
        uint32_t work_before[kMaxWorking];
        std::memcpy(work_before, R.work, sizeof(work_before));
        R.n_work = 0;
        for (uint32_t q = 0; q < net->projs.size(); ++q) {
            const HostProj &H = net->projs[q];
            if (H.kind != PROJ_RESOURCE || H.post_pop != p || H.locked) continue;
            if (R.n_work >= (uint32_t)kMaxWorking)
                return fail(KNP_ERR_UNSUPPORTED, "more than 8 trainable projections into one population");
            R.work[R.n_work++] = q;
        }
        // a projection that joins the working set may hold weights that are not its resources' recalculation
        const bool changed = R.n_work != before || std::memcmp(work_before, R.work, R.n_work * 4) != 0;
        if (changed && R.w_current) {
            const uint32_t n = net->table.pops[p].n;
            HIP_TRY(hipMemsetAsync(R.w_current, 0, n, net->stream));
        }
    }
    return upload_table(net);
}

void dev_free(knp_net *net, void *p) {
    if (!p) return;
    auto it = std::find(net->allocations.begin(), net->allocations.end(), p);
    if (it != net->allocations.end()) {
        const size_t i = (size_t)(it - net->allocations.begin());
        net->allocated -= net->alloc_bytes[i];
        net->alloc_bytes.erase(net->alloc_bytes.begin() + (ptrdiff_t)i);
        net->allocations.erase(it);
    }
    (void)hipFree(p);
}

hipEvent_t take_event(knp_net *net) {
    if (net->ev_used < net->ev_pool.size()) return net->ev_pool[net->ev_used++];
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    net->ev_pool.push_back(e);
    net->ev_used++;
    return e;
}

template <typename Launch>
int timed_launch(knp_net *net, int kid, Launch launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (net->timing) {
        a = take_event(net);
        if (a) HIP_TRY(hipEventRecord(a, net->stream));
    }
    launch();
    HIP_TRY(hipGetLastError());
    if (net->timing && a) {
        b = take_event(net);
        if (b) {
            HIP_TRY(hipEventRecord(b, net->stream));
            net->timed[kid].emplace_back(a, b);
        }
    }
    return KNP_OK;
}

int collect_timing(knp_net *net) {
    HIP_TRY(hipStreamSynchronize(net->stream));
    for (int k = 0; k < KNP_KERNEL_COUNT; ++k) {
        for (auto &pr : net->timed[k]) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
            net->time_ms[k] += ms;
            net->launches[k]++;
        }
        net->timed[k].clear();
    }
    net->ev_used = 0;
    return KNP_OK;
}

int deliver_pushes(knp_net *net) {
    const uint64_t t = net->step;
    bool any = false;
    for (const auto &Q : net->projs) any |= !Q.pending.empty();
    if (!any) return KNP_OK;
    std::vector<std::vector<uint32_t>> lists(net->projs.size());
    for (uint32_t q = 0; q < net->projs.size(); ++q) {
        HostProj &H = net->projs[q];
        if (H.pending.empty()) continue;
        auto &L = lists[q];
        for (auto &m : H.pending) L.insert(L.end(), m.begin(), m.end());
        H.pending.clear();
        if (net->semantics == KNP_SEMANTICS_MULTI_THREADED) {
            // one entry per index: its count (at most 4, push_spikes checks) above the index bits
            std::sort(L.begin(), L.end());
            size_t o = 0;
            for (size_t i = 0; i < L.size();) {
                size_t j = i;
                while (j < L.size() && L[j] == L[i]) ++j;
                L[o++] = (uint32_t)(j - i) << kMsgCountShift | L[i];
                i = j;
            }
            L.resize(o);
        }
        ProjDev &Q = net->table.projs[q];
        if (L.size() > Q.in_cap) return fail(KNP_ERR_RANGE, "more channel spikes in one step than the list holds");
        const uint32_t cnt = (uint32_t)L.size();
        L.insert(L.begin(), cnt);  // segment layout: count, then the indexes
        HIP_TRY(hipMemcpyAsync(Q.in_list + (size_t)(t & 1) * net->table.list_pstride, L.data(), L.size() * 4,
                               hipMemcpyHostToDevice, net->stream));
        uint32_t *slot = net->table.seg_count + (size_t)(t & 1) * net->table.n_seg + Q.push_seg;
        HIP_TRY(hipMemcpyAsync(slot, &cnt, 4, hipMemcpyHostToDevice, net->stream));
        uint32_t *vslot = net->table.vcount + (size_t)(t & 1) * net->table.n_vseg + net->push_vpos[q];
        HIP_TRY(hipMemcpyAsync(vslot, &cnt, 4, hipMemcpyHostToDevice, net->stream));
        HIP_TRY(hipStreamSynchronize(net->stream));
    }
    return KNP_OK;
}

void fill_u32(knp_net *net, uint32_t *p, uint64_t n, uint32_t v) {
    const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(1024, (n + kBlock - 1) / kBlock));
    k_fill_u32<<<g, kBlock, 0, net->stream>>>(p, n, v);
}

// Steps per captured graph of a network's stepwise loop: the dense input projections' time block (they must all
// share it), else the chosen default; 0 = no graph.
uint32_t graph_steps(const knp_net *net) {
    if (net->graph_default == 0) return 0;
    uint32_t T = 0;
    for (const auto &Dh : net->dense) {
        if (T && Dh.T != T) return 0;
        T = Dh.T;
    }
    return T ? T : net->graph_default;
}

// Phase of step t in the dense time blocks (0 = a block starts at t; every dense projection's population
// started with the network, so they share it); 0 without dense input projections.
uint64_t graph_phase(const knp_net *net, uint64_t t, uint32_t G) {
    if (net->dense.empty()) return 0;
    return (t - net->pops[net->dense[0].post_pop].t_first) % G;
}

// Words of spike-count staging the MFMA expansions of a network use per call (both parities).
uint64_t dense_count_words(const knp_net *net) {
    uint64_t w = 0;
    for (uint32_t q = 0; q < net->projs.size(); ++q)
        if (net->table.projs[q].dense) w += (uint64_t)net->table.projs[q].dn_oct * 4;
    return w;
}

// The step loop. counts: null (the MFMA expansions stage their spike counts in their own buffers) or a device
// region of dense_count_words(net) words that holds them instead (the caller clears it first); timing: HIP
// events may be created for kernel timing.
int run_steps(knp_net *net, uint64_t n_steps, uint32_t *counts = nullptr, bool timing = true) {
    if (!net->initialized) {
        int rc = init_network(net);
        if (rc) return rc;
    }
    HIP_TRY(hipSetDevice(net->device));
    NetDev *D = net->dev_table;
    const bool timing_was = net->timing;
    net->timing = net->timing && timing;
    struct Restore {
        knp_net *n;
        bool t;
        ~Restore() { n->timing = t; }
    } restore{net, timing_was};
    if (!counts && net->counts_in_scratch) {
        // the previous call staged in the scratch: the buffers' current parity holds stale counts
        for (uint32_t q = 0; q < net->projs.size(); ++q)
            if (net->table.projs[q].dense) fill_u32(net, net->table.projs[q].dn_cnt, (uint64_t)net->table.projs[q].dn_oct * 4, 0);
        HIP_TRY(hipGetLastError());
    }
    net->counts_in_scratch = counts != nullptr;
    std::vector<uint32_t *> cnt(net->projs.size(), nullptr);
    std::vector<uint32_t> dense_projs;  // projections expanded on MFMA, in load order
    for (uint32_t q = 0; q < net->projs.size(); ++q)
        if (net->table.projs[q].dense) dense_projs.push_back(q);
    // products summed by one launch: projections into the same neurons (same tiles) with the same delay
    std::vector<std::vector<uint32_t>> dense_groups;
    for (uint32_t q : dense_projs) {
        const ProjDev &Qd = net->table.projs[q];
        bool placed = false;
        for (auto &g : dense_groups) {
            const ProjDev &G0 = net->table.projs[g[0]];
            if (g.size() < kDenseGroupMax && G0.post_pop == Qd.post_pop && G0.dn_delay == Qd.dn_delay &&
                net->projs[g[0]].dense_sig == net->projs[q].dense_sig) {
                g.push_back(q);
                placed = true;
                break;
            }
        }
        if (!placed) dense_groups.push_back({q});
    }
    {
        uint64_t off = 0;
        for (uint32_t q = 0; q < net->projs.size(); ++q) {
            const ProjDev &Qd = net->table.projs[q];
            if (!Qd.dense) continue;
            cnt[q] = counts ? counts + off : Qd.dn_cnt;
            off += (uint64_t)Qd.dn_oct * 4;
        }
    }
    // t: the step (host decisions); targ: the kernels' argument for it (t, or kStepDev + offset in a graph)
    auto launch_dense_due = [&](uint64_t t, uint64_t targ) -> int {
        for (uint32_t d = 0; d < net->dense.size(); ++d) {
            const uint64_t t_first = net->pops[net->dense[d].post_pop].t_first;
            if ((t - t_first) % net->dense[d].T != 0) continue;
            int rc = timed_launch(net, KNP_KERNEL_DENSE, [&] { launch_dense(net, d, targ); });
            if (rc) return rc;
        }
        return KNP_OK;
    };
    const bool small_run = net->small && net->test_hold_block == 0xFFFFFFFFu;
    if (net->persistent && n_steps >= 2 && (small_run || net->can_persistent)) {
        // persistent launches, split where a dense projection starts a new time block (its product must be
        // ready before phase 1 of that step). The abort flag is cleared once per call: a launch that starts
        // with it raised returns at once, and the host reads it after the last launch.
        fill_u32(net, net->table.grid_bar, kBarWords, 0);
        HIP_TRY(hipGetLastError());
        uint64_t done = 0;
        while (done < n_steps) {
            const uint64_t t = net->step;
            int rc = deliver_pushes(net);
            if (rc) return rc;
            if ((rc = launch_dense_due(t, t))) return rc;
            // 2^16 steps per launch at most: the launch's per-thread impact counters are 32-bit
            uint64_t len = std::min<uint64_t>(n_steps - done, uint64_t(1) << 16);
            for (const auto &Dh : net->dense) {
                const uint64_t t_first = net->pops[Dh.post_pop].t_first;
                const uint64_t next = t_first + ((t - t_first) / Dh.T + 1) * Dh.T;
                len = std::min<uint64_t>(len, next - t);
            }
            hipError_t le = hipSuccess;
            const uint32_t nt = net->table.n_tiles, steps = (uint32_t)len;
            uint64_t t0 = t;
            void *args[] = {(void *)&D, (void *)&t0, (void *)&steps};
            const bool small = small_run;
            if ((rc = timed_launch(net, KNP_KERNEL_TILE_RUN, [&] {
                     if (small && net->run_kind == POP_LEAN)
                         k_small_run<POP_LEAN><<<1, kSmallThreads, net->small_lds, net->stream>>>(D, net->small_run, t0, steps);
                     else if (small)
                         k_small_run<POP_FULL><<<1, kSmallThreads, net->small_lds, net->stream>>>(D, net->small_run, t0, steps);
                     else
                         le = hipLaunchCooperativeKernel(tile_run_kernel(net->run_kind), dim3(nt), dim3(kTileThreads),
                                                         args, net->tile_lds, net->stream);
                 })))
                return rc;
            if (le != hipSuccess) return fail(KNP_ERR_HIP, std::string("persistent launch: ") + hipGetErrorString(le));
            if (!small) {
                k_clear_push_counts<<<1, kTileThreads, 0, net->stream>>>(D, t0);
                HIP_TRY(hipGetLastError());
            }
            net->step += len;
            done += len;
            for (auto &P : net->pops) P.ran = true;
        }
        uint32_t flag = 0;
        HIP_TRY(hipMemcpyAsync(&flag, net->table.grid_bar + kBarAbort, 4, hipMemcpyDeviceToHost, net->stream));
        HIP_TRY(hipStreamSynchronize(net->stream));
        if (flag) return fail(KNP_ERR_HIP, "persistent launch: a wait for another workgroup's spikes timed out (state is undefined)");
        if (net->timing && net->ev_used > 8192) {
            int rc = collect_timing(net);
            if (rc) return rc;
        }
        return KNP_OK;
    }
    // The launches of one step. t: the step (host decisions); targ: the kernels' step argument (t, or kStepDev
    // in a captured graph); phase1_done: phase 1 of step t already ran (fused into the previous tile kernel);
    // p1: the tile kernel of step t also runs phase 1 of step t + 1.
    auto launch_step = [&](uint64_t t, uint64_t targ, bool phase1_done, bool p1) -> int {
        int rc;
        if (!phase1_done) {
            if ((rc = launch_dense_due(t, targ))) return rc;
            const uint32_t blocks = net->table.total_blocks;
            if ((rc = timed_launch(net, KNP_KERNEL_NEURON, [&] {
                     k_phase1<<<blocks, kBlock, 0, net->stream>>>(D, targ);
                 })))
                return rc;
        }
        for (uint32_t p = 0; p < net->pops.size(); ++p) {
            if (!net->pops[p].resource) continue;
            // a workgroup per candidate up to the limit's population size, a wave per candidate beyond it
            const uint32_t n = net->pops[p].n;
            const bool per_wave = n > net->train_wg_limit;
            if ((rc = timed_launch(net, KNP_KERNEL_TRAIN, [&] {
                     if (per_wave)
                         k_train_wave<<<std::min(1024u, (n + 3) / 4), kBlock, 0, net->stream>>>(D, p, targ);
                     else
                         k_train<<<std::max(1u, std::min(512u, n)), kTrainBlock, 0, net->stream>>>(D, p, targ);
                 })))
                return rc;
        }
        for (uint32_t q = 0; q < net->projs.size(); ++q) {
            if (net->projs[q].kind != PROJ_ADDITIVE || net->projs[q].stdp_pops.empty()) continue;
            k_additive_append<<<net->scatter_grid, kBlock, 0, net->stream>>>(D, q, targ);
            HIP_TRY(hipGetLastError());
        }
        // projections expanded on MFMA (after training: they use the step's trained weights): the spike counts
        // of all of them in one launch (per set of kDenseSetMax), then each product
        for (size_t s0 = 0; s0 < dense_projs.size(); s0 += kDenseSetMax) {
            DenseCountSet set{};
            uint64_t gc = 1;
            for (size_t m = s0; m < dense_projs.size() && set.n < kDenseSetMax; ++m) {
                const uint32_t q = dense_projs[m];
                const ProjDev &Qd = net->table.projs[q];
                const uint64_t words = (uint64_t)Qd.dn_oct * 2;
                gc = std::max<uint64_t>(gc, std::max<uint64_t>(Qd.dn_v1 - Qd.dn_v0, (words + kBlock - 1) / kBlock));
                set.proj[set.n] = q;
                set.cnt[set.n] = cnt[q];
                ++set.n;
            }
            const uint32_t g = (uint32_t)std::min<uint64_t>(kDenseCountBlocks, gc);
            if ((rc = timed_launch(net, KNP_KERNEL_DENSE_EXPAND, [&] {
                     k_dense_counts<<<g, kBlock, 0, net->stream>>>(D, set, targ);
                 })))
                return rc;
        }
        for (const auto &members : dense_groups) {
            DenseGroup grp{};
            for (uint32_t q : members) {
                grp.proj[grp.n] = q;
                grp.cnt[grp.n] = cnt[q];
                ++grp.n;
            }
            const uint32_t ge = (net->table.projs[members[0]].dn_tiles + 3) / 4;
            // every instance's count words fit the wave's LDS buffer, and its weights 32-bit offsets
            bool staged = true;
            for (uint32_t q : members) staged = staged && net->table.projs[q].dn_max_oct <= kDenseStagedOct &&
                                                net->table.projs[q].dn_max_npad <= (1u << 22);
            if ((rc = timed_launch(net, KNP_KERNEL_DENSE_EXPAND, [&] {
                     if (staged) k_dense_expand<true><<<ge, kBlock, 0, net->stream>>>(D, grp, targ);
                     else k_dense_expand<false><<<ge, kBlock, 0, net->stream>>>(D, grp, targ);
                 })))
                return rc;
        }
        const uint32_t sg = net->scatter_grid;
        if (net->table.n_vseg_atomic &&
            (rc = timed_launch(net, KNP_KERNEL_SCATTER, [&] {
                 k_scatter<<<sg, kBlock, (net->table.n_vseg_atomic + 1) * 4, net->stream>>>(D, targ);
             })))
            return rc;
        if (p1 && (rc = launch_dense_due(t + 1, targ + 1))) return rc;
        const uint32_t nt = net->table.n_tiles;
        if (nt && (rc = timed_launch(net, KNP_KERNEL_TILE_SCATTER, [&] {
                       const uint32_t lds = net->tile_lds;
                       const bool w = net->wide_tiles;
                       if (net->tiles_shallow) {
                           if (p1 && w) k_tile_scatter<true, true, true><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                           else if (p1) k_tile_scatter<true, false, true><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                           else if (w) k_tile_scatter<false, true, true><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                           else k_tile_scatter<false, false, true><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                       } else {
                           if (p1 && w) k_tile_scatter<true, true, false><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                           else if (p1) k_tile_scatter<true, false, false><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                           else if (w) k_tile_scatter<false, true, false><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                           else k_tile_scatter<false, false, false><<<nt, kTileThreads, lds, net->stream>>>(D, targ);
                       }
                   })))
            return rc;
        for (uint32_t q = 0; q < net->projs.size(); ++q) {
            if (net->projs[q].kind != PROJ_ADDITIVE) continue;
            const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (net->projs[q].n_syn + kBlock - 1) / kBlock));
            k_additive_modify<<<g, kBlock, 0, net->stream>>>(D, q);
            HIP_TRY(hipGetLastError());
        }
        return KNP_OK;
    };
    // Graph replays: G steps captured once (every step's launches with the device's step, then its advance)
    // and replayed while more than G steps remain after the call's first step, from steps whose dense time
    // blocks are in the captured phase (none without dense projections).
    const uint32_t G = net->timing ? 0 : graph_steps(net);
    const bool staged_in_scratch = counts != nullptr;
    // phase 1 of the current step already ran inside the previous step's tile kernel
    bool phase1_done = false;
    for (uint64_t it = 0; it < n_steps;) {
        const uint64_t t = net->step;
        if (G && it >= 1 && n_steps - it > G && graph_phase(net, t, G) == 0 && phase1_done == net->fused) {
            if (net->graph_exec && (net->graph_steps != G || net->graph_scratch != staged_in_scratch)) {
                (void)hipGraphExecDestroy(net->graph_exec);
                net->graph_exec = nullptr;
            }
            if (!net->graph_exec) {
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(net->stream, hipStreamCaptureModeThreadLocal));
                int rc = KNP_OK;
                for (uint32_t i = 0; i < G && !rc; ++i) {
                    rc = launch_step(t + i, kStepDev, net->fused, net->fused);
                    if (!rc) {
                        k_step_advance<<<1, 64, 0, net->stream>>>(D);
                        if (hipGetLastError() != hipSuccess) rc = fail(KNP_ERR_HIP, "graph capture: launch failed");
                    }
                }
                const hipError_t ce = hipStreamEndCapture(net->stream, &graph);
                if (rc) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc;
                }
                if (ce != hipSuccess) return fail(KNP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
                const hipError_t ie = hipGraphInstantiate(&net->graph_exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) {
                    net->graph_exec = nullptr;
                    return fail(KNP_ERR_HIP, std::string("graph instantiation: ") + hipGetErrorString(ie));
                }
                net->graph_steps = G;
                net->graph_scratch = staged_in_scratch;
            }
            k_set_step<<<1, 64, 0, net->stream>>>(D, t);
            HIP_TRY(hipGetLastError());
            while (n_steps - it > G) {
                HIP_TRY(hipGraphLaunch(net->graph_exec, net->stream));
                net->graph_replays++;
                net->step += G;
                it += G;
            }
            for (auto &P : net->pops) P.ran = true;
            phase1_done = net->fused;
            continue;
        }
        int rc = deliver_pushes(net);
        if (rc) return rc;
        // fused: the tile kernel of step t also runs phase 1 of step t + 1, except on the last step of the call
        // (the state a caller reads between calls is the state after whole steps)
        const bool p1 = net->fused && it + 1 < n_steps;
        if ((rc = launch_step(t, t, phase1_done, p1))) return rc;
        phase1_done = p1;
        net->step++;
        ++it;
        for (auto &P : net->pops) P.ran = true;
        if (net->timing && net->ev_used > 8192) {
            if ((rc = collect_timing(net))) return rc;
        }
    }
    return KNP_OK;
}

// ---- step tuner -------------------------------------------------------------------------------------
// The execution choices of multi-step calls, all with identical results: the schedule (one workgroup,
// persistent, or stepwise launches), the stepwise tile kernel's occupancy, steps per captured step graph and
// the training form.
struct StepChoice {
    int schedule;
    bool wide;
    uint32_t graph, train_limit, scatter_grid;
};

StepChoice current_choice(const knp_net *net) {
    StepChoice c{KNP_SCHEDULE_STEPWISE, net->wide_tiles, net->graph_default, net->train_wg_limit, net->scatter_grid};
    if (net->persistent) c.schedule = net->small ? KNP_SCHEDULE_SINGLE_WORKGROUP : KNP_SCHEDULE_PERSISTENT;
    return c;
}

void apply_choice(knp_net *net, const StepChoice &c) {
    net->small = net->can_small && c.schedule == KNP_SCHEDULE_SINGLE_WORKGROUP;
    net->persistent = net->small || (net->can_persistent && c.schedule == KNP_SCHEDULE_PERSISTENT);
    net->wide_tiles = c.wide;
    net->graph_default = c.graph;
    net->train_wg_limit = c.train_limit;
    net->scatter_grid = std::max(1u, std::min(c.scatter_grid, net->scatter_grid_alloc));
    if (net->graph_exec) (void)hipGraphExecDestroy(net->graph_exec);
    net->graph_exec = nullptr;
}

// A copy of every device allocation and of the host state a step call changes.
struct StateSnapshot {
    std::vector<void *> copies;
    uint64_t step = 0;
    std::vector<bool> ran;
    bool counts_in_scratch = false;
    std::vector<std::vector<std::vector<uint32_t>>> pending;
    ~StateSnapshot() {
        for (void *p : copies) (void)hipFree(p);
    }
};

int take_snapshot(knp_net *net, StateSnapshot &S) {
    if (net->cap && 2 * net->allocated > net->cap)
        return fail(KNP_ERR_OUT_OF_MEMORY, "the step tuner needs room for a copy of the network's device state");
    for (size_t i = 0; i < net->allocations.size(); ++i) {
        void *c = nullptr;
        hipError_t e = hipMalloc(&c, net->alloc_bytes[i]);
        if (e != hipSuccess) return fail(KNP_ERR_OUT_OF_MEMORY, std::string("step tuner snapshot: ") + hipGetErrorString(e));
        S.copies.push_back(c);
        HIP_TRY(hipMemcpyAsync(c, net->allocations[i], net->alloc_bytes[i], hipMemcpyDeviceToDevice, net->stream));
    }
    HIP_TRY(hipStreamSynchronize(net->stream));
    S.step = net->step;
    for (const auto &P : net->pops) S.ran.push_back(P.ran);
    S.counts_in_scratch = net->counts_in_scratch;
    for (const auto &Q : net->projs) S.pending.push_back(Q.pending);
    return KNP_OK;
}

int restore_snapshot(knp_net *net, const StateSnapshot &S) {
    for (size_t i = 0; i < S.copies.size(); ++i)
        HIP_TRY(hipMemcpyAsync(net->allocations[i], S.copies[i], net->alloc_bytes[i], hipMemcpyDeviceToDevice,
                               net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    net->step = S.step;
    for (size_t p = 0; p < net->pops.size(); ++p) net->pops[p].ran = S.ran[p];
    net->counts_in_scratch = S.counts_in_scratch;
    for (size_t q = 0; q < net->projs.size(); ++q) net->projs[q].pending = S.pending[q];
    return KNP_OK;
}

std::string tune_key(const knp_net *net) {
    uint64_t neurons = 0, syn = 0;
    uint32_t res = 0, mfma = 0;
    for (const auto &P : net->pops) {
        neurons += P.n;
        res += P.resource ? 1u : 0u;
    }
    for (const auto &Q : net->projs) {
        syn += Q.n_syn;
        mfma += Q.dense_format >= 0 ? 1u : 0u;
    }
    char key[256];
    std::snprintf(key, sizeof key, "step pops=%zu neurons=%llu projs=%zu syn=%llu tiles=%u resource=%u mfma=%u dense=%zu "
                  "persistent=%d small=%d", net->pops.size(), (unsigned long long)neurons, net->projs.size(),
                  (unsigned long long)syn, net->table.n_tiles, res, mfma, net->dense.size(), (int)net->can_persistent,
                  (int)net->can_small);
    return key;
}

// Seconds of `steps` steps under choice c from snapshot S (a few untimed steps first: graph capture and first
// launches), the state restored afterwards.
int time_choice(knp_net *net, const StateSnapshot &S, const StepChoice &c, uint32_t steps, double *sec) {
    int rc;
    if ((rc = restore_snapshot(net, S))) return rc;
    apply_choice(net, c);
    if ((rc = run_steps(net, std::max<uint32_t>(2, steps / 8), nullptr, false))) return rc;
    HIP_TRY(hipStreamSynchronize(net->stream));
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = run_steps(net, steps, nullptr, false))) return rc;
    HIP_TRY(hipStreamSynchronize(net->stream));
    *sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return KNP_OK;
}

int autotune_steps(knp_net *net, uint32_t steps) {
    const std::string key = tune_key(net), path = tune_cache_path();
    if (!path.empty()) {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            if (line.compare(0, key.size(), key) != 0 || line.size() <= key.size() || line[key.size()] != ' ') continue;
            int sched = 0, wide = 0;
            unsigned graph = 0, train = 0, grid = 0;
            if (std::sscanf(line.c_str() + key.size(), " %d %d %u %u %u", &sched, &wide, &graph, &train, &grid) == 5) {
                apply_choice(net, StepChoice{sched, wide != 0, graph, train, grid});
                return KNP_OK;
            }
        }
    }
    StateSnapshot S;
    int rc;
    if ((rc = take_snapshot(net, S))) return rc;
    const bool timing_was = net->timing;
    net->timing = false;
    StepChoice best = current_choice(net);
    double best_s = 1e30;
    auto consider = [&](const StepChoice &c) -> int {
        double sec = 0;
        int r = time_choice(net, S, c, steps, &sec);
        if (r) return r;
        if (sec < best_s) {
            best_s = sec;
            best = c;
        }
        return KNP_OK;
    };
    // 1. the schedule, with the default options
    std::vector<int> schedules{KNP_SCHEDULE_STEPWISE};
    if (net->can_persistent) schedules.push_back(KNP_SCHEDULE_PERSISTENT);
    if (net->can_small) schedules.push_back(KNP_SCHEDULE_SINGLE_WORKGROUP);
    const StepChoice base = best;
    for (int sc : schedules) {
        StepChoice c = base;
        c.schedule = sc;
        if ((rc = consider(c))) break;
    }
    // 2. the stepwise options, one at a time from the best so far
    if (!rc && best.schedule == KNP_SCHEDULE_STEPWISE) {
        if (net->table.n_tiles) {
            StepChoice c = best;
            c.wide = !c.wide;
            rc = consider(c);
        }
        const uint32_t graphs[] = {0u, 8u, 32u, 128u};
        const StepChoice g0 = best;
        for (uint32_t g : graphs)
            if (!rc && g != g0.graph && (net->dense.empty() || g <= 1u)) {
                StepChoice c = g0;
                c.graph = g;
                rc = consider(c);
            }
        bool resource = false;
        for (const auto &P : net->pops) resource = resource || (P.resource && P.n <= best.train_limit);
        if (!rc && resource) {
            StepChoice c = best;
            c.train_limit = 0;  // a wave per candidate
            rc = consider(c);
        }
        // the atomic scatter's grid: each workgroup scans every atomic segment's count, so a network whose
        // steps have few atomic rows runs faster with fewer workgroups
        if (net->table.n_vseg_atomic) {
            const StepChoice g1 = best;
            for (uint32_t div : {4u, 16u, 64u}) {
                const uint32_t g = std::max(1u, net->scatter_grid_alloc / div);
                if (rc || g == g1.scatter_grid || g < 16u) continue;
                StepChoice c = g1;
                c.scatter_grid = g;
                rc = consider(c);
            }
        }
    }
    net->timing = timing_was;
    const int rr = restore_snapshot(net, S);
    apply_choice(net, best);
    if (rc) return rc;
    if (rr) return rr;
    if (!path.empty()) {
        std::ofstream out(path, std::ios::app);
        if (out)
            out << key << " " << best.schedule << " " << (int)best.wide << " " << best.graph << " " << best.train_limit
                << " " << best.scatter_grid << "\n";
    }
    return KNP_OK;
}

int check_pop(knp_net *net, uint32_t pop) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (pop >= net->pops.size()) return fail(KNP_ERR_INVALID_ARGUMENT, "population index out of range");
    return KNP_OK;
}

int check_proj(knp_net *net, uint32_t proj) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (proj >= net->projs.size()) return fail(KNP_ERR_INVALID_ARGUMENT, "projection index out of range");
    return KNP_OK;
}

int ensure_init(knp_net *net) {
    if (!net->initialized) return init_network(net);
    return KNP_OK;
}

template <typename T>
int download(knp_net *net, T *dst, const T *src, size_t count) {
    if (count == 0) return KNP_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    return KNP_OK;
}

int add_projection_common(knp_net *net, uint32_t kind, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop,
                          const uint32_t *pre, const uint32_t *post, const float *weight, const uint32_t *delay,
                          const uint8_t *output_type, uint64_t n_syn, HostProj &H) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (net->initialized) return fail(KNP_ERR_STATE, "projections must be loaded before the first step");
    if (post_pop >= net->pops.size()) return fail(KNP_ERR_INVALID_ARGUMENT, "postsynaptic population out of range");
    if (pre_pop != KNP_NO_POPULATION) {
        if (pre_pop >= net->pops.size()) return fail(KNP_ERR_INVALID_ARGUMENT, "presynaptic population out of range");
        n_pre = net->pops[pre_pop].n;
    }
    if (n_syn && (!pre || !post || !weight || !delay)) return fail(KNP_ERR_INVALID_ARGUMENT, "null synapse array");
    const uint32_t n_post = net->pops[post_pop].n;
    if (n_post >= (1u << 29)) return fail(KNP_ERR_UNSUPPORTED, "postsynaptic population larger than 2^29");
    H.kind = kind;
    H.pre_pop = pre_pop;
    H.n_pre = n_pre;
    H.post_pop = post_pop;
    H.n_syn = n_syn;
    H.pre.assign(pre, pre + n_syn);
    H.post.assign(post, post + n_syn);
    H.w.assign(weight, weight + n_syn);
    H.delay.assign(delay, delay + n_syn);
    H.type.assign(n_syn, (uint8_t)KNP_EXCITATORY);
    if (output_type) H.type.assign(output_type, output_type + n_syn);
    for (uint64_t s = 0; s < n_syn; ++s) {
        if (H.pre[s] >= n_pre) return fail(KNP_ERR_INVALID_ARGUMENT, "presynaptic index out of range");
        if (H.post[s] >= n_post) return fail(KNP_ERR_INVALID_ARGUMENT, "postsynaptic index out of range");
        if (H.delay[s] == 0 || H.delay[s] > 255)
            return fail(KNP_ERR_UNSUPPORTED, "synaptic delay must be in 1..255 (delay 0 is lost by the reference)");
        if (H.type[s] > KNP_BLOCKING) return fail(KNP_ERR_UNSUPPORTED, "Unhandled synapse type.");
        // LIF neurons take EXCITATORY and INHIBITORY_CURRENT impacts only (lif_impl.h:44-64)
        if (net->pops[post_pop].model == MODEL_LIF && H.type[s] > KNP_INHIBITORY_CURRENT)
            return fail(KNP_ERR_UNSUPPORTED, "Unhandled synapse type.");
        // AltAI-LIF: conductance throws in the reference (altai_impl.h:80-82); BLOCKING is order-dependent
        if (net->pops[post_pop].model == MODEL_ALTAI) {
            if (H.type[s] == KNP_INHIBITORY_CONDUCTANCE) return fail(KNP_ERR_UNSUPPORTED, "Unhandled synapse type.");
            if (H.type[s] == KNP_BLOCKING)
                return fail(KNP_ERR_UNSUPPORTED, "BLOCKING synapses onto AltAI-LIF populations are not supported");
        }
        if (!std::isfinite(H.w[s])) return fail(KNP_ERR_INVALID_ARGUMENT, "non-finite weight");
        H.max_delay = std::max(H.max_delay, H.delay[s]);
        if (H.type[s] == KNP_BLOCKING) H.has_blocking = true;
    }
    return KNP_OK;
}

}  // namespace

// ---- device scratch of knp_mi355x_step_noalloc ----------------------------------------------------------
// Staging for the spike counts of the MFMA expansions (knp_mi355x_set_projection_dense) in a file-scope device
// array, so that the no-allocation step needs no buffer of its own: every GPU holds one copy until the process
// ends. In host code `scratch` names a host-side stand-in, not device memory; its device address is taken once
// per device with hipGetSymbolAddress and only that address is used.
namespace {
constexpr uint64_t kScratchBytes = 1u << 20;
__device__ __align__(256) uint32_t scratch[kScratchBytes / 4];
std::mutex g_scratch_mu;
std::vector<uint32_t *> g_scratch_ptr;  // per device, null until first use

int scratch_pointer(int device, uint32_t **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        HIP_TRY(hipSetDevice(device));  // the address is the current device's copy
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch_ptr[device];
    return KNP_OK;
}
}  // namespace

// =====================================================================================================
// C ABI
// =====================================================================================================
extern "C" {

void knp_blifat_defaults(knp_blifat_neuron *o) {
    std::memset(o, 0, sizeof(*o));
    o->n_time_steps_since_last_firing = std::numeric_limits<uint64_t>::max();
    o->activation_threshold = 1.0;
    o->reversal_inhibitory_potential = -0.3;
    o->min_potential = -1.0e9;
    o->total_blocking_period = std::numeric_limits<int64_t>::max();
}

void knp_lif_defaults(knp_lif_neuron *o) {
    std::memset(o, 0, sizeof(*o));
    o->activation_threshold = 1.f;
    o->leak_coefficient = 1.f;
}

void knp_altai_lif_defaults(knp_altai_lif_neuron *o) {
    std::memset(o, 0, sizeof(*o));
    o->is_reset = 1;
    o->leak_rev = 1;
    o->saturate = 1;
    o->activation_threshold = 1;
    o->negative_activation_threshold = 30000;
}

void knp_resource_altai_lif_defaults(knp_resource_altai_lif_neuron *o) {
    knp_resource_blifat_neuron r;
    knp_resource_blifat_defaults(&r);
    std::memset(o, 0, sizeof(*o));
    knp_altai_lif_defaults(&o->base);
    copy_resource_fields(r, *o);
}

void knp_resource_blifat_defaults(knp_resource_blifat_neuron *o) {
    std::memset(o, 0, sizeof(*o));
    knp_blifat_defaults(&o->base);
    o->dopamine_plasticity_time = 1;
    o->free_synaptic_resource = 1.f;
    o->synaptic_resource_threshold = std::numeric_limits<float>::max();
    o->isi_max = 1;
    o->d_h = 1.f;
    o->isi_status = KNP_ISI_NOT_IN_PERIOD;
}

void knp_resource_rule_defaults(knp_resource_rule *o) {
    std::memset(o, 0, sizeof(*o));
    o->w_max = 1.f;
}

uint32_t knp_bernoulli_hash(uint64_t seed, uint64_t step, uint32_t channel) {
    return knp_mi355x::bernoulli_hash(seed, step, channel);
}

const char *knp_mi355x_version(void) { return "knp-mi355x 0.1 (gfx950, wave64)"; }

const char *knp_mi355x_last_error(void) { return g_last_error.c_str(); }

int knp_mi355x_device_count(int *count) {
    if (!count) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        if (e == hipErrorNoDevice) return KNP_OK;
        return fail(KNP_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    return KNP_OK;
}

int knp_mi355x_device_name(int device, char *buf, size_t cap) {
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (buf && cap) std::snprintf(buf, cap, "%s [%s, %d CUs, %.0f GiB]", prop.name, prop.gcnArchName,
                                  prop.multiProcessorCount, (double)prop.totalGlobalMem / (1024.0 * 1024 * 1024));
    return KNP_OK;
}

int knp_mi355x_create(int device, knp_net **out) {
    if (!out) return fail(KNP_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    int count = 0;
    int rc = knp_mi355x_device_count(&count);
    if (rc) return rc;
    if (device < 0 || device >= count)
        return fail(KNP_ERR_INVALID_ARGUMENT, "device " + std::to_string(device) + " not present (" +
                                                  std::to_string(count) + " HIP devices)");
    auto *net = new knp_net();
    net->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete net;
        return fail(KNP_ERR_HIP, std::string("stream creation: ") + hipGetErrorString(e));
    }
    *out = net;
    return KNP_OK;
}

int knp_mi355x_destroy(knp_net *net) {
    if (!net) return KNP_OK;
    (void)hipSetDevice(net->device);
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    if (net->graph_exec) (void)hipGraphExecDestroy(net->graph_exec);
    for (void *p : net->allocations) (void)hipFree(p);
    for (hipEvent_t e : net->ev_pool) (void)hipEventDestroy(e);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
    return KNP_OK;
}

int knp_mi355x_set_memory_cap(knp_net *net, uint64_t bytes) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    net->cap = bytes;
    return KNP_OK;
}

int knp_mi355x_add_blifat_population(knp_net *net, const knp_blifat_neuron *neurons, uint32_t n, uint32_t *out_pop) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (net->initialized) return fail(KNP_ERR_STATE, "populations must be loaded before the first step");
    if (n && !neurons) return fail(KNP_ERR_INVALID_ARGUMENT, "null neuron array");
    if (net->pops.size() >= (size_t)kMaxPops) return fail(KNP_ERR_UNSUPPORTED, "too many populations");
    HostPop P;
    P.n = n;
    P.nb.assign(neurons, neurons + n);
    net->pops.push_back(std::move(P));
    if (out_pop) *out_pop = (uint32_t)net->pops.size() - 1;
    return KNP_OK;
}

int knp_mi355x_add_lif_population(knp_net *net, const knp_lif_neuron *neurons, uint32_t n, uint32_t *out_pop) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (net->initialized) return fail(KNP_ERR_STATE, "populations must be loaded before the first step");
    if (n && !neurons) return fail(KNP_ERR_INVALID_ARGUMENT, "null neuron array");
    if (net->pops.size() >= (size_t)kMaxPops) return fail(KNP_ERR_UNSUPPORTED, "too many populations");
    HostPop P;
    P.model = MODEL_LIF;
    P.n = n;
    P.nl.assign(neurons, neurons + n);
    net->pops.push_back(std::move(P));
    if (out_pop) *out_pop = (uint32_t)net->pops.size() - 1;
    return KNP_OK;
}

int knp_mi355x_add_altai_lif_population(knp_net *net, const knp_altai_lif_neuron *neurons, uint32_t n,
                                        uint32_t *out_pop) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (net->initialized) return fail(KNP_ERR_STATE, "populations must be loaded before the first step");
    if (n && !neurons) return fail(KNP_ERR_INVALID_ARGUMENT, "null neuron array");
    if (net->pops.size() >= (size_t)kMaxPops) return fail(KNP_ERR_UNSUPPORTED, "too many populations");
    for (uint32_t i = 0; i < n; ++i) {
        const auto &x = neurons[i];
        if (x.is_diff && (double)altai_threshold(x) != (double)x.activation_threshold + x.additional_threshold)
            return fail(KNP_ERR_UNSUPPORTED,
                        "AltAI-LIF with is_diff_: activation_threshold_ + additional_threshold_ must be a float");
    }
    HostPop P;
    P.model = MODEL_ALTAI;
    P.n = n;
    P.na.assign(neurons, neurons + n);
    net->pops.push_back(std::move(P));
    if (out_pop) *out_pop = (uint32_t)net->pops.size() - 1;
    return KNP_OK;
}

int knp_mi355x_add_resource_altai_lif_population(knp_net *net, const knp_resource_altai_lif_neuron *neurons,
                                                 uint32_t n, uint32_t *out_pop) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (net->initialized) return fail(KNP_ERR_STATE, "populations must be loaded before the first step");
    if (n && !neurons) return fail(KNP_ERR_INVALID_ARGUMENT, "null neuron array");
    if (net->pops.size() >= (size_t)kMaxPops) return fail(KNP_ERR_UNSUPPORTED, "too many populations");
    HostPop P;
    P.model = MODEL_ALTAI;
    P.resource = true;
    P.n = n;
    P.na.resize(n);
    P.nr.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        P.na[i] = neurons[i].base;
        knp_resource_blifat_defaults(&P.nr[i]);  // only the resource fields of nr are used
        copy_resource_fields(neurons[i], P.nr[i]);
    }
    net->pops.push_back(std::move(P));
    if (out_pop) *out_pop = (uint32_t)net->pops.size() - 1;
    return KNP_OK;
}

int knp_mi355x_add_resource_blifat_population(knp_net *net, const knp_resource_blifat_neuron *neurons, uint32_t n,
                                              uint32_t *out_pop) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (net->initialized) return fail(KNP_ERR_STATE, "populations must be loaded before the first step");
    if (n && !neurons) return fail(KNP_ERR_INVALID_ARGUMENT, "null neuron array");
    if (net->pops.size() >= (size_t)kMaxPops) return fail(KNP_ERR_UNSUPPORTED, "too many populations");
    HostPop P;
    P.resource = true;
    P.n = n;
    P.nr.assign(neurons, neurons + n);
    P.nb.resize(n);
    for (uint32_t i = 0; i < n; ++i) P.nb[i] = neurons[i].base;
    net->pops.push_back(std::move(P));
    if (out_pop) *out_pop = (uint32_t)net->pops.size() - 1;
    return KNP_OK;
}

int knp_mi355x_add_delta_projection(knp_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop,
                                    const uint32_t *pre, const uint32_t *post, const float *weight,
                                    const uint32_t *delay, const uint8_t *output_type, uint64_t n_syn,
                                    uint32_t *out_proj) {
    HostProj H;
    int rc = add_projection_common(net, PROJ_DELTA, pre_pop, n_pre, post_pop, pre, post, weight, delay, output_type,
                                   n_syn, H);
    if (rc) return rc;
    if (net->projs.size() >= (size_t)kMaxProjs) return fail(KNP_ERR_UNSUPPORTED, "too many projections");
    net->projs.push_back(std::move(H));
    if (out_proj) *out_proj = (uint32_t)net->projs.size() - 1;
    return KNP_OK;
}

int knp_mi355x_add_resource_stdp_projection(knp_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop,
                                            const uint32_t *pre, const uint32_t *post, const float *weight,
                                            const uint32_t *delay, const uint8_t *output_type,
                                            const knp_resource_rule *rules, uint64_t n_syn, uint32_t *out_proj) {
    HostProj H;
    int rc = add_projection_common(net, PROJ_RESOURCE, pre_pop, n_pre, post_pop, pre, post, weight, delay,
                                   output_type, n_syn, H);
    if (rc) return rc;
    if (!net->pops[post_pop].resource)
        return fail(KNP_ERR_UNSUPPORTED, "a synaptic-resource STDP projection must target a resource population");
    if (n_syn && !rules) return fail(KNP_ERR_INVALID_ARGUMENT, "null rule array");
    if (net->projs.size() >= (size_t)kMaxProjs) return fail(KNP_ERR_UNSUPPORTED, "too many projections");
    H.rules.assign(rules, rules + n_syn);
    net->projs.push_back(std::move(H));
    if (out_proj) *out_proj = (uint32_t)net->projs.size() - 1;
    return KNP_OK;
}

int knp_mi355x_add_additive_stdp_projection(knp_net *net, uint32_t pre_pop, uint32_t n_pre, uint32_t post_pop,
                                            const uint32_t *pre, const uint32_t *post, const float *weight,
                                            const uint32_t *delay, const uint8_t *output_type,
                                            const float *tau_plus, const float *tau_minus, uint64_t n_syn,
                                            const uint32_t *stdp_pops, const uint8_t *stdp_modes,
                                            uint32_t n_stdp_pops, uint32_t *out_proj) {
    HostProj H;
    int rc = add_projection_common(net, PROJ_ADDITIVE, pre_pop, n_pre, post_pop, pre, post, weight, delay,
                                   output_type, n_syn, H);
    if (rc) return rc;
    if (n_syn && (!tau_plus || !tau_minus)) return fail(KNP_ERR_INVALID_ARGUMENT, "null tau array");
    if (n_stdp_pops > 4) return fail(KNP_ERR_UNSUPPORTED, "more than 4 stdp populations");
    for (uint32_t m = 0; m < n_stdp_pops; ++m) {
        if (stdp_pops[m] >= net->pops.size()) return fail(KNP_ERR_INVALID_ARGUMENT, "stdp population out of range");
        if (stdp_modes[m] > KNP_STDP_AND_SPIKE) return fail(KNP_ERR_INVALID_ARGUMENT, "bad stdp processing type");
    }
    if (net->projs.size() >= (size_t)kMaxProjs) return fail(KNP_ERR_UNSUPPORTED, "too many projections");
    H.tau_p.assign(tau_plus, tau_plus + n_syn);
    H.tau_m.assign(tau_minus, tau_minus + n_syn);
    H.stdp_pops.assign(stdp_pops, stdp_pops + n_stdp_pops);
    H.stdp_modes.assign(stdp_modes, stdp_modes + n_stdp_pops);
    net->projs.push_back(std::move(H));
    if (out_proj) *out_proj = (uint32_t)net->projs.size() - 1;
    return KNP_OK;
}

int knp_mi355x_set_locked(knp_net *net, uint32_t proj, int locked) {
    int rc = check_proj(net, proj);
    if (rc) return rc;
    net->projs[proj].locked = locked != 0;
    if (net->initialized) return upload_working_sets(net);
    return KNP_OK;
}

int knp_mi355x_start_learning(knp_net *net) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    for (auto &Q : net->projs) Q.locked = false;
    if (net->initialized) return upload_working_sets(net);
    return KNP_OK;
}

int knp_mi355x_stop_learning(knp_net *net) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    for (auto &Q : net->projs) Q.locked = true;
    if (net->initialized) return upload_working_sets(net);
    return KNP_OK;
}

int knp_mi355x_init(knp_net *net) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    return ensure_init(net);
}

int knp_mi355x_push_spikes(knp_net *net, uint32_t proj, const uint32_t *idx, uint32_t n) {
    int rc = check_proj(net, proj);
    if (rc) return rc;
    if (n && !idx) return fail(KNP_ERR_INVALID_ARGUMENT, "null index array");
    HostProj &H = net->projs[proj];
    for (uint32_t i = 0; i < n; ++i)
        if (idx[i] >= H.n_pre) return fail(KNP_ERR_INVALID_ARGUMENT, "spike index out of the presynaptic range");
    if (net->semantics == KNP_SEMANTICS_MULTI_THREADED) {
        // the first channel message of the step would be either this or a device source's, by send order
        if (H.gen_channels)
            return fail(KNP_ERR_UNSUPPORTED, "under the multi-threaded message semantics a projection fed by device "
                                             "sources takes no pushed messages");
        // the multi-threaded backend expands the first message of a step only (an empty one included)
        if (!H.pending.empty()) return KNP_OK;
    }
    // The exact accumulators keep a factor 4 of headroom over one impact per synapse and step: an index may
    // repeat at most 4 times over the messages delivered in one step.
    std::vector<uint32_t> all(idx, idx + n);
    for (const auto &m : H.pending) all.insert(all.end(), m.begin(), m.end());
    std::sort(all.begin(), all.end());
    for (size_t i = 4; i < all.size(); ++i)
        if (all[i] == all[i - 4])
            return fail(KNP_ERR_RANGE, "a spike index repeats more than 4 times in the messages of one step");
    H.pending.emplace_back(idx, idx + n);
    return KNP_OK;
}

int knp_mi355x_add_bernoulli_source(knp_net *net, uint32_t proj, uint32_t first_index, uint32_t n_channels,
                                    const uint32_t *thresholds, uint32_t n_patterns, uint32_t steps_per_pattern,
                                    uint32_t active_lo, uint32_t active_hi, uint64_t seed, uint32_t hash_offset) {
    int rc = check_proj(net, proj);
    if (rc) return rc;
    if (net->initialized) return fail(KNP_ERR_STATE, "sources must be added before the first step");
    if (n_patterns == 0 || steps_per_pattern == 0 || !thresholds)
        return fail(KNP_ERR_INVALID_ARGUMENT, "bad source shape");
    if ((uint64_t)first_index + n_channels > net->projs[proj].n_pre)
        return fail(KNP_ERR_INVALID_ARGUMENT, "source channels exceed the presynaptic index range");
    if (net->srcs.size() >= (size_t)kMaxSources) return fail(KNP_ERR_UNSUPPORTED, "too many sources");
    HostSource S;
    S.proj = proj;
    S.first = first_index;
    S.n_ch = n_channels;
    S.n_patterns = n_patterns;
    S.spp = steps_per_pattern;
    S.lo = active_lo;
    S.hi = active_hi;
    S.seed = seed;
    S.hash_offset = hash_offset;
    S.thr.assign(thresholds, thresholds + (size_t)n_channels * n_patterns);
    net->srcs.push_back(std::move(S));
    net->projs[proj].gen_channels += n_channels;
    return KNP_OK;
}

int knp_mi355x_add_dense_input_projection(knp_net *net, uint32_t post_pop, uint32_t n_instances,
                                          const uint32_t *post_begin, const uint32_t *post_count, uint32_t n_in,
                                          int format, const void *weights, const float *row_scale,
                                          uint32_t threshold, uint64_t seed, uint32_t hash_offset,
                                          uint32_t time_block, uint32_t *out_dense) {
    int rc = check_pop(net, post_pop);
    if (rc) return rc;
    if (net->initialized) return fail(KNP_ERR_STATE, "projections must be loaded before the first step");
    if (format != KNP_DENSE_FP8_E4M3 && format != KNP_DENSE_BF16) return fail(KNP_ERR_UNSUPPORTED, "dense format");
    if (!n_instances || !post_begin || !post_count || !weights || n_in == 0)
        return fail(KNP_ERR_INVALID_ARGUMENT, "bad dense projection arguments");
    if (time_block % 32 != 0) return fail(KNP_ERR_INVALID_ARGUMENT, "time_block must be a multiple of 32");
    if (net->dense.size() >= (size_t)kMaxDense) return fail(KNP_ERR_UNSUPPORTED, "too many dense projections");
    const uint32_t n = net->pops[post_pop].n;
    HostDense D;
    D.post_pop = post_pop;
    D.n_inst = n_instances;
    D.n_in = n_in;
    D.format = (uint32_t)format;
    D.threshold = threshold;
    D.seed = seed;
    D.hash_offset = hash_offset;
    D.T = time_block;
    D.post_begin.assign(post_begin, post_begin + n_instances);
    D.post_count.assign(post_count, post_count + n_instances);
    uint64_t rows = 0;
    for (uint32_t k = 0; k < n_instances; ++k) {
        if ((uint64_t)post_begin[k] + post_count[k] > n)
            return fail(KNP_ERR_INVALID_ARGUMENT, "dense instance range outside the population");
        rows += post_count[k];
    }
    const size_t vb = format == KNP_DENSE_FP8_E4M3 ? 1 : 2;
    D.weights.assign((const uint8_t *)weights, (const uint8_t *)weights + rows * n_in * vb);
    D.row_scale.assign(rows, 1.f);
    if (row_scale) D.row_scale.assign(row_scale, row_scale + rows);
    // The matrix instructions keep about 13 binades below the largest product of each 16-input dot product;
    // rows whose nonzero weights stay within a window below the row maximum multiply exactly.
    const int window = format == KNP_DENSE_FP8_E4M3 ? 9 : 5;
    for (uint64_t r = 0; r < rows; ++r) {
        int emax = std::numeric_limits<int>::min();
        for (uint32_t i = 0; i < n_in; ++i) {
            const double v = dense_value(D, r, i);
            if (!std::isfinite(v)) return fail(KNP_ERR_INVALID_ARGUMENT, "non-finite dense weight");
            if (v != 0.0) {
                int e;
                std::frexp(v, &e);
                emax = std::max(emax, e);
            }
        }
        for (uint32_t i = 0; i < n_in; ++i) {
            const double v = dense_value(D, r, i);
            if (v == 0.0) continue;
            int e;
            std::frexp(v, &e);
            if (e < emax - window)
                return fail(KNP_ERR_UNSUPPORTED, "dense weight more than " + std::to_string(window) +
                                                     " binades below its row maximum (not exact on the matrix cores)");
        }
    }
    net->dense.push_back(std::move(D));
    if (out_dense) *out_dense = (uint32_t)net->dense.size() - 1;
    return KNP_OK;
}

int knp_mi355x_get_dense_time_block(knp_net *net, uint32_t dense, uint32_t *time_block) {
    if (!net || !time_block || dense >= net->dense.size()) return fail(KNP_ERR_INVALID_ARGUMENT, "bad dense index");
    int rc = ensure_init(net);
    if (rc) return rc;
    *time_block = net->dense[dense].T;
    return KNP_OK;
}

int knp_mi355x_step(knp_net *net, uint64_t n_steps) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    return run_steps(net, n_steps);
}

int knp_mi355x_step_noalloc(knp_net *net, uint64_t n_steps) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (!net->initialized)
        return fail(KNP_ERR_STATE, "knp_mi355x_step_noalloc needs an initialized network (knp_mi355x_init allocates)");
    const uint64_t words = dense_count_words(net);
    if (words * 4 > kScratchBytes)
        return fail(KNP_ERR_RANGE, "knp_mi355x_step_noalloc: the MFMA expansions stage " + std::to_string(words * 4) +
                                       " bytes of spike counts, above the " + std::to_string(kScratchBytes) +
                                       "-byte device scratch");
    HIP_TRY(hipSetDevice(net->device));
    uint32_t *sp = nullptr;
    int rc = scratch_pointer(net->device, &sp);
    if (rc) return rc;
    // the scratch keeps whatever the last call on this device left: the counts start at zero
    if (words) {
        fill_u32(net, sp, words, 0);
        HIP_TRY(hipGetLastError());
    }
    return run_steps(net, n_steps, sp, false);
}

int knp_mi355x_get_step(knp_net *net, uint64_t *step) {
    if (!net || !step) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    *step = net->step;
    return KNP_OK;
}

int knp_mi355x_synchronize(knp_net *net) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    HIP_TRY(hipStreamSynchronize(net->stream));
    return KNP_OK;
}

int knp_mi355x_record_spikes(knp_net *net, uint32_t pop, uint64_t capacity_steps) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    HostPop &H = net->pops[pop];
    H.rec_cap = capacity_steps;
    H.rec_t0 = net->step;
    if (!net->initialized) return KNP_OK;
    PopDev &P = net->table.pops[pop];
    HIP_TRY(hipStreamSynchronize(net->stream));
    dev_free(net, P.rec);
    P.rec = nullptr;
    if (capacity_steps) {
        if ((rc = dev_alloc(net, (void **)&P.rec, (size_t)capacity_steps * P.words * 4))) return rc;
    }
    P.rec_cap = capacity_steps;
    P.rec_t0 = net->step;
    return upload_table(net);
}

int knp_mi355x_read_spike_record(knp_net *net, uint32_t pop, uint32_t *words, uint64_t n_words, uint64_t *n_steps) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (!n_steps) return fail(KNP_ERR_INVALID_ARGUMENT, "null n_steps");
    *n_steps = 0;
    if (!net->initialized) return KNP_OK;
    PopDev &P = net->table.pops[pop];
    if (!P.rec) return fail(KNP_ERR_STATE, "population is not recorded");
    const uint64_t steps = std::min<uint64_t>(net->step - P.rec_t0, P.rec_cap);
    const uint64_t need = steps * P.words;
    if (n_words < need) return fail(KNP_ERR_RANGE, "spike record buffer too small");
    if ((rc = download(net, words, P.rec, need))) return rc;
    *n_steps = steps;
    P.rec_t0 = net->step;
    net->pops[pop].rec_t0 = net->step;
    return upload_table(net);
}

int knp_mi355x_read_last_spikes(knp_net *net, uint32_t pop, uint32_t *idx, uint32_t cap, uint32_t *n) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (!n) return fail(KNP_ERR_INVALID_ARGUMENT, "null count pointer");
    *n = 0;
    if (!net->initialized || net->step == 0) return KNP_OK;
    const PopDev &P = net->table.pops[pop];
    std::vector<uint32_t> bm(P.words);
    if ((rc = download(net, bm.data(), P.bitmap, P.words))) return rc;
    uint32_t k = 0;
    for (uint32_t w = 0; w < P.words; ++w) {
        uint32_t x = bm[w];
        while (x) {
            const uint32_t b = __builtin_ctz(x);
            x &= x - 1;
            if (idx && k < cap) idx[k] = w * 32 + b;
            ++k;
        }
    }
    *n = k;
    return KNP_OK;
}

int knp_mi355x_read_potentials(knp_net *net, uint32_t pop, float *out, uint32_t n) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (n > net->pops[pop].n || (n && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer");
    if ((rc = ensure_init(net))) return rc;
    return download(net, out, net->table.pops[pop].V, n);
}

static int export_blifat(knp_net *net, uint32_t pop, std::vector<knp_blifat_neuron> &out) {
    const HostPop &H = net->pops[pop];
    const PopDev &P = net->table.pops[pop];
    const uint32_t n = H.n;
    out = H.nb;
    int rc;
    std::vector<float> V(n);
    if ((rc = download(net, V.data(), P.V, n))) return rc;
    const uint64_t steps_run = net->step - H.t_first;
    if (H.kind == POP_LEAN) {
        std::vector<uint32_t> ls(n);
        if ((rc = download(net, ls.data(), P.last_spike, n))) return rc;
        for (uint32_t i = 0; i < n; ++i) {
            auto &o = out[i];
            o.potential = V[i];
            if (steps_run) {
                o.pre_impact_potential = 0.0;
                if (ls[i] == 0xFFFFFFFFu) o.n_time_steps_since_last_firing += steps_run;  // wraps like size_t
                else o.n_time_steps_since_last_firing = (net->step - 1) - ls[i];
                o.total_blocking_period -= (int64_t)steps_run;
                o.dynamic_threshold *= 0.0;
                o.postsynaptic_trace *= 0.0;
                o.inhibitory_conductance *= 0.0;
            }
        }
        return KNP_OK;
    }
    std::vector<float> Vp(n), dyn(n), tr(n), g(n), dop(n);
    std::vector<uint32_t> ts(n), bp(n);
    std::vector<int64_t> tbp;
    if ((rc = download(net, Vp.data(), P.Vpre, n))) return rc;
    if ((rc = download(net, dyn.data(), P.dyn, n))) return rc;
    if ((rc = download(net, tr.data(), P.trace, n))) return rc;
    if ((rc = download(net, g.data(), P.g, n))) return rc;
    if ((rc = download(net, dop.data(), P.dop, n))) return rc;
    if ((rc = download(net, ts.data(), P.n_ts, n))) return rc;
    if ((rc = download(net, bp.data(), P.bphase, n))) return rc;
    if (P.track_tbp) {
        tbp.resize(n);
        if ((rc = download(net, tbp.data(), P.tbp, n))) return rc;
    }
    for (uint32_t i = 0; i < n; ++i) {
        auto &o = out[i];
        o.potential = V[i];
        o.pre_impact_potential = Vp[i];
        o.dynamic_threshold = dyn[i];
        o.postsynaptic_trace = tr[i];
        o.inhibitory_conductance = g[i];
        o.dopamine_value = dop[i];
        o.n_time_steps_since_last_firing = ts_from_device(ts[i]);
        o.bursting_phase = bp[i];
        o.total_blocking_period = P.track_tbp ? tbp[i] : o.total_blocking_period - (int64_t)steps_run;
    }
    return KNP_OK;
}

int knp_mi355x_read_blifat_neurons(knp_net *net, uint32_t pop, knp_blifat_neuron *out, uint32_t n) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (net->pops[pop].model != MODEL_BLIFAT) return fail(KNP_ERR_INVALID_ARGUMENT, "not a BLIFAT population");
    if (n > net->pops[pop].n || (n && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer");
    if ((rc = ensure_init(net))) return rc;
    std::vector<knp_blifat_neuron> v;
    if ((rc = export_blifat(net, pop, v))) return rc;
    if (net->pops[pop].resource) {
        std::vector<double> add(net->pops[pop].n);
        if ((rc = download(net, add.data(), net->table.pops[pop].res.add_thr, add.size()))) return rc;
        for (uint32_t i = 0; i < v.size(); ++i) v[i].additional_threshold = add[i];
    }
    std::copy(v.begin(), v.begin() + n, out);
    return KNP_OK;
}

int knp_mi355x_read_lif_neurons(knp_net *net, uint32_t pop, knp_lif_neuron *out, uint32_t n) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (net->pops[pop].model != MODEL_LIF) return fail(KNP_ERR_INVALID_ARGUMENT, "not a LIF population");
    if (n > net->pops[pop].n || (n && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer");
    if ((rc = ensure_init(net))) return rc;
    const PopDev &P = net->table.pops[pop];
    std::vector<float> V(n);
    std::vector<uint32_t> rcnt(n);
    if ((rc = download(net, V.data(), P.V, n))) return rc;
    if ((rc = download(net, rcnt.data(), P.n_ts, n))) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = net->pops[pop].nl[i];
        out[i].potential = V[i];
        out[i].refract_counter = rcnt[i];
    }
    return KNP_OK;
}

int knp_mi355x_read_altai_lif_neurons(knp_net *net, uint32_t pop, knp_altai_lif_neuron *out, uint32_t n) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (net->pops[pop].model != MODEL_ALTAI) return fail(KNP_ERR_INVALID_ARGUMENT, "not an AltAI-LIF population");
    if (n > net->pops[pop].n || (n && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer");
    if ((rc = ensure_init(net))) return rc;
    const PopDev &P = net->table.pops[pop];
    std::vector<float> V(n), Vpre(n), dop(n);
    std::vector<int64_t> at(n);
    if ((rc = download(net, V.data(), P.V, n))) return rc;
    if ((rc = download(net, Vpre.data(), P.Vpre, n))) return rc;
    if ((rc = download(net, dop.data(), P.dop, n))) return rc;
    if ((rc = download(net, at.data(), P.tbp, n))) return rc;
    std::vector<double> add;
    if (net->pops[pop].resource) {
        add.resize(n);
        if ((rc = download(net, add.data(), P.res.add_thr, n))) return rc;
    }
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = net->pops[pop].na[i];
        out[i].potential = V[i];
        out[i].pre_impact_potential = Vpre[i];
        out[i].dopamine_value = dop[i];
        out[i].activity_time = at[i];
        if (!add.empty()) out[i].additional_threshold = add[i];
    }
    return KNP_OK;
}

// dopamine_value_ of a resource population's neurons once it has run: the exact sum of the last step's
// dopamine impacts (the value is reset every step, blifat_impl.h:50-65), in place of the FP32 copy.
static int exact_resource_dopamine(knp_net *net, uint32_t pop, std::vector<double> &dop) {
    dop.clear();
    if (!net->pops[pop].ran) return KNP_OK;
    const uint32_t N = net->pops[pop].n;
    std::vector<int64_t> fix(N);
    int rc = download(net, fix.data(), net->table.pops[pop].res.dop_fix, N);
    if (rc) return rc;
    dop.resize(N);
    for (uint32_t i = 0; i < N; ++i) dop[i] = std::ldexp((double)fix[i], -net->pops[pop].frac_bits);
    return KNP_OK;
}

// The resource fields of every neuron of a resource population (device state over the loaded values).
static int export_resource_fields(knp_net *net, uint32_t pop, std::vector<knp_resource_blifat_neuron> &out) {
    const uint32_t N = net->pops[pop].n;
    const ResDev &R = net->table.pops[pop].res;
    std::vector<float> fr(N), st(N);
    std::vector<uint8_t> is(N), fo(N);
    std::vector<uint64_t> ls(N), lss(N), fis(N);
    int rc;
    if ((rc = download(net, fr.data(), R.free_res, N))) return rc;
    if ((rc = download(net, st.data(), R.stability, N))) return rc;
    if ((rc = download(net, is.data(), R.isi_status, N))) return rc;
    if ((rc = download(net, fo.data(), R.forced_now, N))) return rc;
    if ((rc = download(net, ls.data(), R.last_step, N))) return rc;
    if ((rc = download(net, lss.data(), R.last_spike_step, N))) return rc;
    if ((rc = download(net, fis.data(), R.first_isi_spike, N))) return rc;
    out = net->pops[pop].nr;
    for (uint32_t i = 0; i < N; ++i) {
        out[i].free_synaptic_resource = fr[i];
        out[i].stability = st[i];
        out[i].isi_status = is[i];
        out[i].is_being_forced = fo[i];
        out[i].last_step = ls[i];
        out[i].last_spike_step = lss[i];
        out[i].first_isi_spike = fis[i];
    }
    return KNP_OK;
}

int knp_mi355x_read_resource_altai_lif_neurons(knp_net *net, uint32_t pop, knp_resource_altai_lif_neuron *out,
                                               uint32_t n) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (!net->pops[pop].resource || net->pops[pop].model != MODEL_ALTAI)
        return fail(KNP_ERR_INVALID_ARGUMENT, "not a resource AltAI-LIF population");
    if (n > net->pops[pop].n || (n && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer");
    if ((rc = ensure_init(net))) return rc;
    const uint32_t N = net->pops[pop].n;
    std::vector<knp_altai_lif_neuron> base(N);
    std::vector<knp_resource_blifat_neuron> res;
    std::vector<double> dop;
    if ((rc = knp_mi355x_read_altai_lif_neurons(net, pop, base.data(), N))) return rc;
    if ((rc = export_resource_fields(net, pop, res))) return rc;
    if ((rc = exact_resource_dopamine(net, pop, dop))) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        out[i].base = base[i];
        if (!dop.empty()) out[i].base.dopamine_value = dop[i];
        copy_resource_fields(res[i], out[i]);
    }
    return KNP_OK;
}

int knp_mi355x_read_resource_neurons(knp_net *net, uint32_t pop, knp_resource_blifat_neuron *out, uint32_t n) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (!net->pops[pop].resource || net->pops[pop].model != MODEL_BLIFAT)
        return fail(KNP_ERR_INVALID_ARGUMENT, "not a resource BLIFAT population");
    if (n > net->pops[pop].n || (n && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer");
    if ((rc = ensure_init(net))) return rc;
    const uint32_t N = net->pops[pop].n;
    std::vector<knp_blifat_neuron> base(N);
    std::vector<knp_resource_blifat_neuron> res;
    std::vector<double> dop;
    if ((rc = knp_mi355x_read_blifat_neurons(net, pop, base.data(), N))) return rc;
    if ((rc = export_resource_fields(net, pop, res))) return rc;
    if ((rc = exact_resource_dopamine(net, pop, dop))) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = res[i];
        out[i].base = base[i];
        if (!dop.empty()) out[i].base.dopamine_value = dop[i];
    }
    return KNP_OK;
}

int knp_mi355x_read_weights(knp_net *net, uint32_t proj, float *out, uint64_t n_syn) {
    int rc = check_proj(net, proj);
    if (rc) return rc;
    const HostProj &H = net->projs[proj];
    if (n_syn != H.n_syn || (n_syn && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer size");
    if ((rc = ensure_init(net))) return rc;
    const ProjDev &Q = net->table.projs[proj];
    if (H.kind != PROJ_DELTA) return download(net, out, Q.W, n_syn);
    std::vector<float> wc(n_syn);
    std::vector<uint32_t> cs(n_syn);
    if ((rc = download(net, wc.data(), Q.w_csr, n_syn))) return rc;
    if ((rc = download(net, cs.data(), Q.csr_syn, n_syn))) return rc;
    for (uint64_t k = 0; k < n_syn; ++k) out[cs[k]] = wc[k];
    return KNP_OK;
}

int knp_mi355x_read_resource_rules(knp_net *net, uint32_t proj, knp_resource_rule *out, uint64_t n_syn) {
    int rc = check_proj(net, proj);
    if (rc) return rc;
    const HostProj &H = net->projs[proj];
    if (H.kind != PROJ_RESOURCE) return fail(KNP_ERR_INVALID_ARGUMENT, "not a synaptic-resource projection");
    if (n_syn != H.n_syn || (n_syn && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer size");
    if ((rc = ensure_init(net))) return rc;
    const ProjDev &Q = net->table.projs[proj];
    std::vector<float> r(n_syn), wmin(n_syn), wmax(n_syn), du(n_syn);
    std::vector<uint32_t> per(n_syn);
    std::vector<uint64_t> lss(n_syn);
    std::vector<uint8_t> fl(n_syn);
    if ((rc = download(net, r.data(), Q.r, n_syn))) return rc;
    if ((rc = download(net, wmin.data(), Q.w_min, n_syn))) return rc;
    if ((rc = download(net, wmax.data(), Q.w_max, n_syn))) return rc;
    if ((rc = download(net, du.data(), Q.d_u, n_syn))) return rc;
    if ((rc = download(net, per.data(), Q.period, n_syn))) return rc;
    if ((rc = download(net, lss.data(), Q.last_spike_step, n_syn))) return rc;
    std::vector<uint64_t> row_spike(std::max(1u, H.n_pre));
    if ((rc = download(net, row_spike.data(), Q.row_spike, row_spike.size()))) return rc;
    for (uint64_t s = 0; s < n_syn; ++s)
        if (row_spike[H.pre[s]] != UINT64_MAX) lss[s] = row_spike[H.pre[s]];
    if ((rc = download(net, fl.data(), Q.flags, n_syn))) return rc;
    for (uint64_t s = 0; s < n_syn; ++s) {
        out[s].synaptic_resource = r[s];
        out[s].w_min = wmin[s];
        out[s].w_max = wmax[s];
        out[s].d_u = du[s];
        out[s].dopamine_plasticity_period = per[s];
        out[s].last_spike_step = lss[s];
        out[s].had_hebbian_update = fl[s] & 1;
        out[s].has_contributed = (fl[s] >> 1) & 1;
    }
    return KNP_OK;
}

static int read_counters(knp_net *net, uint64_t *events, uint64_t *rows) {
    unsigned long long c[2] = {0, 0};
    *events = *rows = 0;
    if (!net->initialized) return KNP_OK;
    int rc = download(net, c, net->table.events, 2);
    if (rc) return rc;
    std::vector<unsigned long long> part(2 * (size_t)net->scatter_grid_alloc);
    if ((rc = download(net, part.data(), net->table.part, part.size()))) return rc;
    *events = c[0];
    *rows = c[1];
    for (size_t b = 0; b < net->scatter_grid_alloc; ++b) {
        *events += part[2 * b];
        *rows += part[2 * b + 1];
    }
    for (uint32_t q = 0; q < net->projs.size(); ++q) {
        const ProjDev &Q = net->table.projs[q];
        if (!Q.dense) continue;
        std::vector<unsigned long long> dp(4 * (size_t)kDenseCountBlocks);
        if ((rc = download(net, dp.data(), Q.dn_part, dp.size()))) return rc;
        for (size_t b = 0; b < kDenseCountBlocks; ++b) {
            *events += dp[4 * b];
            *rows += dp[4 * b + 1];
        }
    }
    if (net->table.n_tiles) {
        std::vector<unsigned long long> tp(4 * (size_t)net->table.n_tiles);
        if ((rc = download(net, tp.data(), net->table.tile_part, tp.size()))) return rc;
        for (size_t b = 0; b < net->table.n_tiles; ++b) {
            *events += tp[4 * b];
            *rows += tp[4 * b + 1];
        }
    }
    return KNP_OK;
}

int knp_mi355x_get_tile_counters(knp_net *net, uint64_t *row_parts, uint64_t *ring_cells) {
    if (!net || !row_parts || !ring_cells) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    *row_parts = *ring_cells = 0;
    if (!net->initialized || !net->table.n_tiles) return KNP_OK;
    std::vector<unsigned long long> tp(4 * (size_t)net->table.n_tiles);
    int rc = download(net, tp.data(), net->table.tile_part, tp.size());
    if (rc) return rc;
    for (size_t b = 0; b < net->table.n_tiles; ++b) {
        *row_parts += tp[4 * b + 2];
        *ring_cells += tp[4 * b + 3];
    }
    return KNP_OK;
}

int knp_mi355x_set_wait_test(knp_net *net, uint32_t hold_block, uint32_t spin_limit, uint64_t hold_until) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (spin_limit < 64) return fail(KNP_ERR_INVALID_ARGUMENT, "spin limit below 64 polls");
    net->test_hold_block = hold_block;
    net->test_spin_limit = spin_limit;
    net->test_hold_until = hold_until;
    if (!net->initialized) return KNP_OK;
    net->table.hold_block = hold_block;
    net->table.spin_limit = spin_limit;
    net->table.hold_until = hold_until;
    return upload_table(net);
}

int knp_mi355x_set_train_workgroup_limit(knp_net *net, uint32_t max_neurons) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    net->train_wg_limit = max_neurons;
    if (net->graph_exec) {  // a captured step graph holds the previous choice of training kernel
        (void)hipGraphExecDestroy(net->graph_exec);
        net->graph_exec = nullptr;
    }
    return KNP_OK;
}

int knp_mi355x_set_projection_dense(knp_net *net, uint32_t proj, int format, uint32_t flags) {
    int rc = check_proj(net, proj);
    if (rc) return rc;
    if (net->initialized) return fail(KNP_ERR_STATE, "set the MFMA expansion before the first step");
    if (format != KNP_DENSE_BF16) return fail(KNP_ERR_UNSUPPORTED, "the MFMA expansion takes BF16 weights");
    if (flags & ~(uint32_t)(KNP_DENSE_SKIP_EMPTY | KNP_DENSE_SPLIT)) return fail(KNP_ERR_INVALID_ARGUMENT, "unknown flags");
    const HostProj &H = net->projs[proj];
    if (H.kind == PROJ_ADDITIVE) return fail(KNP_ERR_UNSUPPORTED, "additive STDP projections are not expanded on MFMA");
    net->projs[proj].dense_format = format;
    net->projs[proj].dense_flags = flags;
    return KNP_OK;
}

int knp_mi355x_get_dense_weight_bytes(knp_net *net, uint64_t *skip_bytes, uint64_t *full_bytes_per_step) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (skip_bytes) *skip_bytes = 0;
    if (full_bytes_per_step) *full_bytes_per_step = 0;
    if (!net->initialized) return KNP_OK;
    int rc;
    uint64_t full = 0, skip = 0;
    for (uint32_t q = 0; q < net->projs.size(); ++q) {
        const ProjDev &Q = net->table.projs[q];
        if (!Q.dense) continue;
        if ((Q.dense - 1) & kDenseSkipEmpty) {
            std::vector<unsigned long long> dp(4 * (size_t)kDenseCountBlocks);
            if ((rc = download(net, dp.data(), Q.dn_part, dp.size()))) return rc;
            for (size_t b = 0; b < kDenseCountBlocks; ++b) skip += dp[4 * b + 2];
            continue;
        }
        std::vector<DenseInst> inst(Q.dn_inst);
        if ((rc = download(net, inst.data(), Q.dn_inst_tab, inst.size()))) return rc;
        for (const auto &I : inst) full += (uint64_t)I.n_oct * I.npad * (Q.dn_w_lo ? 32 : 16);
    }
    if (skip_bytes) *skip_bytes = skip;
    if (full_bytes_per_step) *full_bytes_per_step = full;
    return KNP_OK;
}

int knp_mi355x_read_dense_weights(knp_net *net, uint32_t proj, uint16_t *out, uint16_t *out_lo, uint64_t n_syn) {
    int rc = check_proj(net, proj);
    if (rc) return rc;
    const HostProj &H = net->projs[proj];
    if (n_syn != H.n_syn || (n_syn && !out)) return fail(KNP_ERR_INVALID_ARGUMENT, "bad output buffer size");
    if (H.dense_format < 0) return fail(KNP_ERR_INVALID_ARGUMENT, "projection not expanded on MFMA");
    if ((rc = ensure_init(net))) return rc;
    const ProjDev &Q = net->table.projs[proj];
    std::vector<DenseInst> inst(Q.dn_inst);
    std::vector<uint32_t> post_inst(net->pops[H.post_pop].n);
    if ((rc = download(net, inst.data(), Q.dn_inst_tab, inst.size()))) return rc;
    if ((rc = download(net, post_inst.data(), Q.dn_post_inst, post_inst.size()))) return rc;
    const DenseInst &L = inst.back();
    std::vector<uint16_t> wb(L.base + (uint64_t)L.n_oct * L.npad * 8), wlo;
    if ((rc = download(net, wb.data(), Q.dn_w, wb.size()))) return rc;
    if (out_lo && Q.dn_w_lo) {
        wlo.resize(wb.size());
        if ((rc = download(net, wlo.data(), Q.dn_w_lo, wlo.size()))) return rc;
    }
    for (uint64_t s = 0; s < n_syn; ++s) {
        const DenseInst &I = inst[post_inst[H.post[s]]];
        const uint32_t i = H.pre[s] - I.pre_begin, j = H.post[s] - I.post_begin;
        const uint64_t pos = I.base + ((uint64_t)(i >> 3) * I.npad + j) * 8 + (i & 7);
        out[s] = wb[pos];
        if (out_lo) out_lo[s] = wlo.empty() ? 0 : wlo[pos];
    }
    return KNP_OK;
}

int knp_mi355x_set_step_graph(knp_net *net, uint32_t steps) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (steps > 4096) return fail(KNP_ERR_INVALID_ARGUMENT, "at most 4096 steps per graph");
    net->graph_default = steps;
    return KNP_OK;
}

int knp_mi355x_get_step_graph(knp_net *net, uint32_t *steps, uint64_t *replays) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (steps) *steps = net->graph_exec ? net->graph_steps : 0;
    if (replays) *replays = net->graph_replays;
    return KNP_OK;
}

int knp_mi355x_set_scatter_mode(knp_net *net, int mode) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (mode != KNP_SCATTER_AUTO && mode != KNP_SCATTER_ATOMIC && mode != KNP_SCATTER_STEPWISE && mode != KNP_SCATTER_TILES)
        return fail(KNP_ERR_INVALID_ARGUMENT, "unknown scatter mode");
    if (net->initialized) return fail(KNP_ERR_STATE, "the scatter mode must be set before the first step");
    net->scatter_mode = mode;
    return KNP_OK;
}

int knp_mi355x_set_message_semantics(knp_net *net, int semantics) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (semantics != KNP_SEMANTICS_SINGLE_THREADED && semantics != KNP_SEMANTICS_MULTI_THREADED)
        return fail(KNP_ERR_INVALID_ARGUMENT, "unknown message semantics");
    if (net->initialized) return fail(KNP_ERR_STATE, "the message semantics must be set before the first step");
    net->semantics = semantics;
    return KNP_OK;
}

int knp_mi355x_get_tiled(knp_net *net, uint32_t pop, int *tiled) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (!tiled) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    if ((rc = ensure_init(net))) return rc;
    *tiled = net->pops[pop].tiled ? 1 : 0;
    return KNP_OK;
}

int knp_mi355x_get_schedule(knp_net *net, int *schedule) {
    if (!net || !schedule) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    int rc;
    if ((rc = ensure_init(net))) return rc;
    if (net->persistent && net->small && net->test_hold_block == 0xFFFFFFFFu) *schedule = KNP_SCHEDULE_SINGLE_WORKGROUP;
    else if (net->persistent && net->can_persistent) *schedule = KNP_SCHEDULE_PERSISTENT;
    else if (net->table.n_tiles || net->table.n_vseg_atomic == 0) *schedule = KNP_SCHEDULE_STEPWISE;
    else *schedule = KNP_SCHEDULE_ATOMIC;
    return KNP_OK;
}

int knp_mi355x_autotune(knp_net *net, uint32_t trial_steps) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    if (trial_steps < 16 || trial_steps > 65536) return fail(KNP_ERR_INVALID_ARGUMENT, "trial_steps must be 16..65536");
    int rc;
    if ((rc = ensure_init(net))) return rc;
    HIP_TRY(hipSetDevice(net->device));
    return autotune_steps(net, trial_steps);
}

int knp_mi355x_get_tuning(knp_net *net, int *schedule, int *wide_tiles, uint32_t *graph_steps, uint32_t *train_limit,
                          uint32_t *scatter_grid) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    int rc;
    if ((rc = ensure_init(net))) return rc;
    const StepChoice c = current_choice(net);
    if (schedule) *schedule = c.schedule;
    if (wide_tiles) *wide_tiles = c.wide ? 1 : 0;
    if (graph_steps) *graph_steps = c.graph;
    if (train_limit) *train_limit = c.train_limit;
    if (scatter_grid) *scatter_grid = c.scatter_grid;
    return KNP_OK;
}

int knp_mi355x_get_event_count(knp_net *net, uint64_t *events) {
    if (!net || !events) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    uint64_t rows;
    return read_counters(net, events, &rows);
}

int knp_mi355x_get_counters(knp_net *net, uint64_t *events, uint64_t *rows) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    uint64_t e, r;
    int rc = read_counters(net, &e, &r);
    if (events) *events = e;
    if (rows) *rows = r;
    return rc;
}

int knp_mi355x_get_inexact_weights(knp_net *net, uint32_t pop, uint64_t *count) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (!count) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    if ((rc = ensure_init(net))) return rc;
    *count = net->pops[pop].inexact;
    return KNP_OK;
}

int knp_mi355x_get_accumulator_bits(knp_net *net, uint32_t pop, int32_t *frac_bits) {
    int rc = check_pop(net, pop);
    if (rc) return rc;
    if (!frac_bits) return fail(KNP_ERR_INVALID_ARGUMENT, "null pointer");
    if ((rc = ensure_init(net))) return rc;
    *frac_bits = net->pops[pop].frac_bits;
    return KNP_OK;
}

int knp_mi355x_enable_timing(knp_net *net, int enable) {
    if (!net) return fail(KNP_ERR_INVALID_ARGUMENT, "null handle");
    int rc = collect_timing(net);
    if (rc) return rc;
    net->timing = enable != 0;
    for (int k = 0; k < KNP_KERNEL_COUNT; ++k) {
        net->time_ms[k] = 0;
        net->launches[k] = 0;
    }
    return KNP_OK;
}

int knp_mi355x_kernel_time(knp_net *net, int kernel_id, double *avg_ms, uint64_t *launches) {
    if (!net || kernel_id < 0 || kernel_id >= KNP_KERNEL_COUNT) return fail(KNP_ERR_INVALID_ARGUMENT, "bad kernel id");
    int rc = collect_timing(net);
    if (rc) return rc;
    if (launches) *launches = net->launches[kernel_id];
    if (avg_ms) *avg_ms = net->launches[kernel_id] ? net->time_ms[kernel_id] / net->launches[kernel_id] : 0.0;
    return KNP_OK;
}

}  // extern "C"

// =====================================================================================================
// Network file (save / load of populations and projections with their current state)
// =====================================================================================================
namespace {

constexpr char kNetMagic[8] = {'K', 'N', 'P', 'N', 'E', 'T', '0', '1'};
enum NetFileModel : uint32_t { NF_BLIFAT = 0, NF_RES_BLIFAT = 1, NF_LIF = 2, NF_ALTAI = 3, NF_RES_ALTAI = 4 };

struct FileCloser {
    FILE *f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

bool put(FILE *f, const void *p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
template <class T>
bool put_v(FILE *f, const T &v) { return put(f, &v, sizeof v); }
template <class T>
bool put_vec(FILE *f, const std::vector<T> &v) { return put(f, v.data(), v.size() * sizeof(T)); }
bool get(FILE *f, void *p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }
template <class T>
bool get_v(FILE *f, T &v) { return get(f, &v, sizeof v); }
template <class T>
bool get_vec(FILE *f, std::vector<T> &v, uint64_t n) {
    v.resize(n);
    return get(f, v.data(), n * sizeof(T));
}

// One population: model tag, count, struct size, then the C ABI structs with the current state.
template <class T>
int save_pop_structs(knp_net *net, FILE *f, uint32_t pop, uint32_t tag,
                     int (*read)(knp_net *, uint32_t, T *, uint32_t)) {
    const uint32_t n = net->pops[pop].n, size = sizeof(T);
    std::vector<T> v(n);
    int rc = read(net, pop, v.data(), n);
    if (rc) return rc;
    if (!put_v(f, tag) || !put_v(f, n) || !put_v(f, size) || !put_vec(f, v))
        return fail(KNP_ERR_INVALID_ARGUMENT, "network file: write failed");
    return KNP_OK;
}

template <class T>
int load_pop_structs(knp_net *net, FILE *f, uint32_t n, uint32_t size,
                     int (*add)(knp_net *, const T *, uint32_t, uint32_t *)) {
    if (size != sizeof(T)) return fail(KNP_ERR_INVALID_ARGUMENT, "network file: neuron struct size mismatch");
    std::vector<T> v;
    if (!get_vec(f, v, n)) return fail(KNP_ERR_INVALID_ARGUMENT, "network file: truncated population");
    return add(net, v.data(), n, nullptr);
}

}  // namespace

extern "C" {

int knp_mi355x_save_network(knp_net *net, const char *path) {
    if (!net || !path) return fail(KNP_ERR_INVALID_ARGUMENT, "null argument");
    int rc;
    if ((rc = ensure_init(net))) return rc;
    FileCloser fc{std::fopen(path, "wb")};
    FILE *f = fc.f;
    if (!f) return fail(KNP_ERR_INVALID_ARGUMENT, std::string("network file: cannot open ") + path);
    const uint32_t version = 1, n_pops = (uint32_t)net->pops.size(), n_projs = (uint32_t)net->projs.size();
    if (!put(f, kNetMagic, 8) || !put_v(f, version) || !put_v(f, n_pops) || !put_v(f, n_projs))
        return fail(KNP_ERR_INVALID_ARGUMENT, "network file: write failed");
    for (uint32_t p = 0; p < n_pops; ++p) {
        const HostPop &H = net->pops[p];
        if (H.model == MODEL_LIF) rc = save_pop_structs(net, f, p, NF_LIF, knp_mi355x_read_lif_neurons);
        else if (H.model == MODEL_ALTAI && H.resource)
            rc = save_pop_structs(net, f, p, NF_RES_ALTAI, knp_mi355x_read_resource_altai_lif_neurons);
        else if (H.model == MODEL_ALTAI) rc = save_pop_structs(net, f, p, NF_ALTAI, knp_mi355x_read_altai_lif_neurons);
        else if (H.resource) rc = save_pop_structs(net, f, p, NF_RES_BLIFAT, knp_mi355x_read_resource_neurons);
        else rc = save_pop_structs(net, f, p, NF_BLIFAT, knp_mi355x_read_blifat_neurons);
        if (rc) return rc;
    }
    for (uint32_t q = 0; q < n_projs; ++q) {
        const HostProj &H = net->projs[q];
        std::vector<float> w(H.n_syn);
        if ((rc = knp_mi355x_read_weights(net, q, w.data(), H.n_syn))) return rc;
        const uint32_t locked = H.locked ? 1 : 0;
        bool ok = put_v(f, H.kind) && put_v(f, H.pre_pop) && put_v(f, H.n_pre) && put_v(f, H.post_pop) &&
                  put_v(f, H.n_syn) && put_v(f, locked) && put_vec(f, H.pre) && put_vec(f, H.post) && put_vec(f, w) &&
                  put_vec(f, H.delay) && put_vec(f, H.type);
        if (ok && H.kind == PROJ_RESOURCE) {
            std::vector<knp_resource_rule> rules(H.n_syn);
            if ((rc = knp_mi355x_read_resource_rules(net, q, rules.data(), H.n_syn))) return rc;
            const uint32_t size = sizeof(knp_resource_rule);
            ok = put_v(f, size) && put_vec(f, rules);
        }
        if (ok && H.kind == PROJ_ADDITIVE) {
            const uint32_t n_stdp = (uint32_t)H.stdp_pops.size();
            ok = put_vec(f, H.tau_p) && put_vec(f, H.tau_m) && put_v(f, n_stdp) && put_vec(f, H.stdp_pops) &&
                 put_vec(f, H.stdp_modes);
        }
        if (!ok) return fail(KNP_ERR_INVALID_ARGUMENT, "network file: write failed");
    }
    if (std::fflush(f) != 0) return fail(KNP_ERR_INVALID_ARGUMENT, "network file: write failed");
    return KNP_OK;
}

int knp_mi355x_load_network(knp_net *net, const char *path) {
    if (!net || !path) return fail(KNP_ERR_INVALID_ARGUMENT, "null argument");
    if (net->initialized || !net->pops.empty() || !net->projs.empty())
        return fail(KNP_ERR_STATE, "network file: load into an empty handle");
    FileCloser fc{std::fopen(path, "rb")};
    FILE *f = fc.f;
    if (!f) return fail(KNP_ERR_INVALID_ARGUMENT, std::string("network file: cannot open ") + path);
    char magic[8];
    uint32_t version = 0, n_pops = 0, n_projs = 0;
    if (!get(f, magic, 8) || std::memcmp(magic, kNetMagic, 8) != 0 || !get_v(f, version) || version != 1 ||
        !get_v(f, n_pops) || !get_v(f, n_projs))
        return fail(KNP_ERR_INVALID_ARGUMENT, "network file: not a version-1 network file");
    int rc;
    for (uint32_t p = 0; p < n_pops; ++p) {
        uint32_t tag = 0, n = 0, size = 0;
        if (!get_v(f, tag) || !get_v(f, n) || !get_v(f, size))
            return fail(KNP_ERR_INVALID_ARGUMENT, "network file: truncated population");
        switch (tag) {
            case NF_BLIFAT: rc = load_pop_structs(net, f, n, size, knp_mi355x_add_blifat_population); break;
            case NF_RES_BLIFAT: rc = load_pop_structs(net, f, n, size, knp_mi355x_add_resource_blifat_population); break;
            case NF_LIF: rc = load_pop_structs(net, f, n, size, knp_mi355x_add_lif_population); break;
            case NF_ALTAI: rc = load_pop_structs(net, f, n, size, knp_mi355x_add_altai_lif_population); break;
            case NF_RES_ALTAI:
                rc = load_pop_structs(net, f, n, size, knp_mi355x_add_resource_altai_lif_population);
                break;
            default: return fail(KNP_ERR_INVALID_ARGUMENT, "network file: unknown neuron model");
        }
        if (rc) return rc;
    }
    for (uint32_t q = 0; q < n_projs; ++q) {
        uint32_t kind = 0, pre_pop = 0, n_pre = 0, post_pop = 0, locked = 0;
        uint64_t n = 0;
        if (!get_v(f, kind) || !get_v(f, pre_pop) || !get_v(f, n_pre) || !get_v(f, post_pop) || !get_v(f, n) ||
            !get_v(f, locked) || n > (uint64_t(1) << 40))
            return fail(KNP_ERR_INVALID_ARGUMENT, "network file: truncated projection");
        std::vector<uint32_t> pre, post, delay;
        std::vector<float> w;
        std::vector<uint8_t> type;
        if (!get_vec(f, pre, n) || !get_vec(f, post, n) || !get_vec(f, w, n) || !get_vec(f, delay, n) ||
            !get_vec(f, type, n))
            return fail(KNP_ERR_INVALID_ARGUMENT, "network file: truncated projection");
        if (kind == PROJ_DELTA) {
            rc = knp_mi355x_add_delta_projection(net, pre_pop, n_pre, post_pop, pre.data(), post.data(), w.data(),
                                                 delay.data(), type.data(), n, nullptr);
        } else if (kind == PROJ_RESOURCE) {
            uint32_t size = 0;
            std::vector<knp_resource_rule> rules;
            if (!get_v(f, size) || size != sizeof(knp_resource_rule) || !get_vec(f, rules, n))
                return fail(KNP_ERR_INVALID_ARGUMENT, "network file: bad synaptic-resource rules");
            rc = knp_mi355x_add_resource_stdp_projection(net, pre_pop, n_pre, post_pop, pre.data(), post.data(),
                                                         w.data(), delay.data(), type.data(), rules.data(), n, nullptr);
        } else if (kind == PROJ_ADDITIVE) {
            std::vector<float> tp, tm;
            std::vector<uint32_t> sp;
            std::vector<uint8_t> sm;
            uint32_t n_stdp = 0;
            if (!get_vec(f, tp, n) || !get_vec(f, tm, n) || !get_v(f, n_stdp) || n_stdp > 4 || !get_vec(f, sp, n_stdp) ||
                !get_vec(f, sm, n_stdp))
                return fail(KNP_ERR_INVALID_ARGUMENT, "network file: bad additive STDP parameters");
            rc = knp_mi355x_add_additive_stdp_projection(net, pre_pop, n_pre, post_pop, pre.data(), post.data(),
                                                         w.data(), delay.data(), type.data(), tp.data(), tm.data(), n,
                                                         sp.data(), sm.data(), n_stdp, nullptr);
        } else {
            return fail(KNP_ERR_INVALID_ARGUMENT, "network file: unknown projection kind");
        }
        if (rc) return rc;
        if ((rc = knp_mi355x_set_locked(net, q, (int)locked))) return rc;
    }
    return KNP_OK;
}

}  // extern "C"
