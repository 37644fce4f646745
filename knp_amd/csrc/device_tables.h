// device_tables.h - POD tables describing a loaded network on the device.
//
// A network is a set of populations (BLIFAT or synaptic-resource BLIFAT) and projections (Delta,
// synaptic-resource STDP Delta, additive STDP Delta), in load order, exactly as a KNP backend holds them
// (knp/backends/cpu/cpu-single-threaded-backend/include/.../backend.h:107-122). The host builds these
// tables once at init and re-uploads them only when learning locks or observation settings change.
#pragma once
#include <stdint.h>

namespace knp_mi355x {

constexpr int kMaxPops = 32;
constexpr int kMaxProjs = 64;
constexpr int kMaxSources = 32;
constexpr int kMaxRowSources = 128;
constexpr int kMaxVSeg = 16384;   // virtual segments the scatter scans in LDS
constexpr int kMaxTasks = kMaxPops + kMaxSources;
constexpr int kMaxWorking = 8;
// ProjDev::rule_uniform bits: w_min, w_max, d_u, dopamine_plasticity_period, delay the same on every synapse
constexpr uint32_t kUniWmin = 1, kUniWmax = 2, kUniDu = 4, kUniPeriod = 8, kUniDelay = 16;
constexpr uint32_t kUniAll = kUniWmin | kUniWmax | kUniDu | kUniPeriod | kUniDelay;
constexpr int kMaxDense = 4;
constexpr int kBlock = 256;  // 4 wave64 per workgroup
// A spike segment: word 0 = count, words 1.. = ascending spike indexes. Population and source segments hold
// 256 entries; the stride keeps every segment 64-byte aligned, so the count and the first entries share one
// line and can be read with one 16-byte load.
constexpr uint32_t kSegStride = 272;
constexpr int kTileThreads = 512;   // workgroup of the tile-owner scatter (8 wave64)
constexpr int kTileMaxD = 32;       // ring depth limit of tile-owned populations (LDS: D x 256 x 8 bytes)
constexpr int kTileMaxList = 2048;  // virtual segments that may reach one tile
// Row parts up to this length are expanded by the row's own thread (longer ones by the whole workgroup):
// in the persistent kernel, and in the one-launch-per-step kernel (where more workgroups share each CU).
constexpr uint32_t kTileShortPart = 16, kTileShortPartStepwise = 32;

enum PopKind : uint32_t { POP_LEAN = 0, POP_FULL = 1, POP_RESOURCE = 2 };
// Neuron model of a FULL population. LIF: V, n_ts = refract counter, act, V_reset, V_decay = leak, refr.
// AltAI-LIF: V, Vpre, tbp = activity time, dop, act = activation + additional threshold, V_reset,
// V_decay = potential leak, V_min = -negative threshold, alt_flags.
enum PopModel : uint32_t { MODEL_BLIFAT = 0, MODEL_LIF = 1, MODEL_ALTAI = 2 };
constexpr uint32_t kAltDiff = 1, kAltReset = 2, kAltLeakRev = 4, kAltSaturate = 8, kAltDoNotSave = 16;
enum ProjKind : uint32_t { PROJ_DELTA = 0, PROJ_RESOURCE = 1, PROJ_ADDITIVE = 2 };

// A parameter that is either one value for the whole population or one value per neuron.
struct ParamF {
    float v;
    const float *a;
};
struct ParamU {
    uint32_t v;
    const uint32_t *a;
};
struct ParamI64 {
    int64_t v;
    const int64_t *a;
};

// Synaptic-resource neuron state and parameters (neuron_parameters<SynapticResourceSTDPNeuron<...>>).
struct ResDev {
    double *add_thr;     // additional_threshold_ (written by training)
    float *free_res;     // free_synaptic_resource_
    float *stability;    // stability_
    uint8_t *isi_status; // ISIPeriodType
    uint64_t *last_step, *last_spike_step, *first_isi_spike;
    float *dop_now;      // dopamine_value_ of this step (written by the population kernel)
    int64_t *dop_fix;    // the same as the exact accumulator sum (2^-frac_bits units), read by training
    uint8_t *forced_now; // is_being_forced_ of this step
    ParamF coeff, res_thr, change_param, change_at_isi, d_h;
    ParamU dpt, drain, isi_max;
    uint32_t *cand;      // training candidates of this step (neuron indexes)
    uint8_t *w_current;  // 1: every working synapse's weight is the recalculation of its resource (0 at load
                         // and whenever the working projections change)
    uint32_t n_work;     // working projections: unlocked resource-STDP projections targeting the population
    uint32_t work[kMaxWorking];
};

struct PopDev {
    uint32_t n, kind, D, words;  // D = ring depth (max delay of incoming projections), words = ceil(n/32)
    int32_t frac_bits;            // fixed-point fraction bits of the exact accumulators
    float scale, inv_scale;       // 2^frac_bits, 2^-frac_bits
    int64_t *acc_v;               // [D][n] EXCITATORY - INHIBITORY_CURRENT
    int64_t *acc_g;               // [D][n] INHIBITORY_CONDUCTANCE (null if no such projection)
    int64_t *acc_d;               // [D][n] DOPAMINE (null if none)
    uint32_t *acc_f;              // [D][n] forcing flags (resource populations only)
    uint64_t *acc_b;              // [D][n] BLOCKING "last impact" keys (null if none)
    // state (FULL / RESOURCE)
    float *V, *Vpre, *dyn, *trace, *g, *dop;
    uint32_t *n_ts, *bphase;
    int64_t *tbp;
    // state (LEAN): step of the last spike, UINT32_MAX = none since load
    uint32_t *last_spike;
    ParamF act, thr_decay, thr_inc, trace_decay, trace_inc, g_decay, V_decay, reflexive, E_inh, V_reset, V_min;
    // BLIFAT spike threshold (blifat_impl.h:113-115): its static part T = activation_threshold_ +
    // additional_threshold_ (synaptic-resource neurons: activation_threshold_ alone, their additional threshold
    // is trained), summed in double as the reference sums it, in three forms: thr_fix = ceil(T * 2^frac_bits)
    // (an exact accumulator sum a reaches T iff a >= thr_fix), thr_up = the smallest float >= T (a float
    // potential V reaches T iff V >= thr_up), and thr_hi + thr_lo = T to 48 bits (thresholds with a dynamic
    // part, compared in float-float arithmetic).
    ParamI64 thr_fix;
    ParamF thr_up, thr_hi, thr_lo;
    ParamU bperiod, refr;
    ParamU alt_flags;      // AltAI-LIF: kAlt* bits of is_diff_, is_reset_, leak_rev_, saturate_, do_not_save_
    uint32_t lean_ts_max;  // LEAN: every neuron started with n_time_steps_since_last_firing = SIZE_MAX
    uint32_t seg_base;     // first spike segment (one per 256 neurons)
    const float *ff;       // dense input impacts [2 ff_T][n] by arrival step (null if none)
    uint32_t ff_T;
    uint32_t track_tbp;    // FULL: total_blocking_period_ is streamed
    uint64_t t_first;      // first step this population runs
    // outputs
    uint32_t *spike_list;  // spikes of the current step: segment c holds neurons 256c.. (ascending), see seg_count
    uint32_t *bitmap;      // spikes of the current step, bit i = neuron i
    uint32_t *rec;         // recorded bitmaps [rec_cap][words]
    uint64_t rec_t0, rec_cap;
    uint32_t tiled;        // impacts into this population are added by the tile-owner scatter
    uint32_t model;        // FULL: MODEL_BLIFAT, MODEL_LIF or MODEL_ALTAI (see population_update)
    ResDev res;
};

// Spike lists are kept in segments: every phase-1 workgroup writes the spikes of its 256 neurons (or
// channels) in ascending order into its own fixed segment and stores the segment's count, so compaction needs
// no global atomics. Host-pushed channel messages of a projection form one more segment. The scatter walks a
// "virtual" segment list: for every projection (load order), its channel segments, then the segments of every
// population whose spike messages it receives, with the message rank of each.
// Message order of a projection's virtual segments under the multi-threaded backend's semantics: the
// population's message, then host pushes, then device source si (kMsgKeySource + si); the smallest key with
// spikes is the step's only message. Pushed entries then carry the index's count above kMsgCountShift.
constexpr uint32_t kMsgKeyPopulation = 0, kMsgKeyPush = 1, kMsgKeySource = 2;
constexpr uint32_t kMsgCountShift = 24, kMsgIndexMask = (1u << kMsgCountShift) - 1u;

struct RowSource {
    uint32_t proj;
    uint32_t rank;      // 0 = channel messages, 1 + population index for population messages
};

struct alignas(16) RowMeta {
    uint32_t begin;      // first CSR position of the row
    uint32_t first_tile; // tile of the row's first postsynaptic neuron
    uint32_t n_tiles;    // tiles spanned (first .. last postsynaptic neuron)
    uint32_t dir;        // offset of the row's directory in row_dir
};

struct ProjDev {
    uint32_t kind, pre_pop, post_pop, n_pre;
    uint64_t n_syn;
    const uint64_t *row_ptr;     // [n_pre + 1], CSR by presynaptic neuron, ascending synapse index
    const uint32_t *post_type;   // [n_syn] CSR order: post | (output_type << 29)
    const uint8_t *delay;        // [n_syn] CSR order
    const float *w_csr;          // [n_syn] CSR order (plain Delta)
    const uint32_t *csr_syn;     // [n_syn] CSR position -> synapse index (plastic projections)
    float *W;                    // [n_syn] weights in synapse-index order (plastic projections)
    // synaptic-resource rule state, synapse-index order
    float *r, *w_min, *w_max, *d_u;
    uint32_t *period;
    // last_spike_step_ (init_synapse, synaptic_resource_stdp.h:36-62, gives every synapse of a spiking row
    // the step): the row's last spike step once the row has spiked since load, else the loaded value
    uint64_t *last_spike_step;   // [n_syn] loaded values, synapse-index order
    uint64_t *row_spike;         // [n_pre] last step each presynaptic row spiked; before its first spike since
                                 // load, the loaded last_spike_step that all the row's synapses carry, or
                                 // UINT64_MAX where they differ (each synapse's own loaded value then counts)
    uint32_t lss_by_row;         // no row's synapses differ: row_spike alone gives every last_spike_step_
    const uint32_t *col_pre;     // [n_syn] CSC order: presynaptic neuron of col_syn[k]
    const uint32_t *col_pre0;    // [n_post] first presynaptic index of each column when every column holds
                                 // consecutive indexes (col_pre[k] = col_pre0[post] + k - col_ptr[post]), else null
    uint8_t *flags;              // bit0 had_hebbian_update_, bit1 has_contributed_
    const uint8_t *delay_syn;    // [n_syn] delay in synapse-index order
    const uint64_t *col_ptr;     // [n_post + 1] CSC by postsynaptic neuron
    const uint32_t *col_syn;     // [n_syn] synapse indexes, ascending within a column
    uint32_t csc_identity;       // col_syn[k] == k (synapses stored by postsynaptic neuron): not loaded
    // synaptic-resource rule parameters carried by every synapse alike (kUni* bits): the one value
    uint32_t rule_uniform;
    float u_wmin, u_wmax, u_du;
    uint32_t u_period, u_delay;
    // additive STDP
    float *tau_plus, *tau_minus;
    uint32_t *lim;               // spike-history cap per synapse
    uint64_t *pre_times, *post_times;  // [n_syn][hist]
    uint32_t *n_pre_t, *n_post_t;
    uint32_t hist;
    // host-pushed channel messages of the current step (one segment)
    uint32_t *in_list;
    uint32_t in_cap;
    uint32_t push_seg;
    uint32_t forcing;            // messages of this projection are forcing (plain Delta)
    uint32_t blocking_id;        // rank of this projection among projections with BLOCKING synapses
    uint32_t n_stdp;
    uint32_t stdp_pop[4];
    uint32_t stdp_mode[4];
    // tile-owner scatter (plain Delta into a tiled population): rows are ordered by postsynaptic tile, so a
    // tile's part of a row is contiguous, with the part's delay-1 records first; the owner of a tile finds
    // the part through a directory of part entries (dir_entry). Dense directory: [n_pre][tile_stride].
    // Ranged directory (rows spanning few of many tiles): row_meta + row_dir. Neither: rows are short and
    // every tile reads them whole and filters.
    const uint64_t *syn8;        // [n_syn] CSR order: impact m * 2^s (m: bits 0-24 signed, s: 25-30) | delay << 32 | post << 41
    const uint64_t *tile_dir;    // dense directory (null if none): CSR positions of the parts
    uint32_t tile_stride;        // tiles of the postsynaptic population
    const struct RowMeta *row_meta;  // [n_pre] ranged directory (null if none)
    const uint64_t *row_dir;     // per row: one entry per tile spanned, positions from the row start
    // Dense expansion on MFMA (knp_mi355x_set_projection_dense): the synapses form dense blocks (instances,
    // DenseInst), every pair of an instance's presynaptic and postsynaptic ranges once, one delay and one
    // output type. Each step
    // the spikes of every presynaptic index are counted (k_dense_counts: 0 or 1, more when a channel message
    // repeats an index) and the impacts, (spike counts) x (a BF16 copy of the weights), are summed on MFMA
    // with FP32 accumulation (k_dense_expand) into the ring slot of their arrival step.
    uint32_t dense;              // 0: expanded by the scatter; else 1 + kDense* flags
    uint32_t dn_oct;             // octets (groups of 8 presynaptic indexes) over all instances
    uint32_t dn_inst, dn_tiles;  // instances, 32-neuron tiles over all instances
    uint32_t dn_max_oct;         // octets of the largest instance
    uint32_t dn_max_npad;        // the largest instance's padded neuron count
    uint32_t dn_delay;           // the projection's delay
    float dn_sign;               // +1 EXCITATORY, -1 INHIBITORY_CURRENT
    uint32_t dn_v0, dn_v1;       // the projection's virtual segments [dn_v0, dn_v1)
    const struct DenseInst *dn_inst_tab;  // [dn_inst]
    const uint32_t *dn_pre_inst;  // [n_pre] instance of each presynaptic index (UINT32_MAX: none)
    const uint32_t *dn_post_inst; // [n_post] instance of each postsynaptic neuron (UINT32_MAX: none)
    const uint32_t *dn_tile;      // [dn_tiles] instance << 16 | first neuron of the tile / 32
    const struct DenseTile *dn_tile_info;  // [dn_tiles] what the expansion of a tile reads, in one entry
    const uint64_t *dn_pre_info;  // [n_pre] count word (octet) | byte << 32 | instance post_count << 35
                                  // (UINT64_MAX: the index is in no instance)
    uint16_t *dn_w;              // BF16 weights; instance k from base: [octet][neuron (npad)][8 inputs]
    uint16_t *dn_w_lo;           // kDenseSplit: BF16 of w - dn_w (same layout), else null
    unsigned long long *dn_part; // [kDenseCountBlocks][4] impacts, rows, weight bytes (skip) per counts workgroup
    uint32_t *dn_cnt;            // [2][dn_oct][8] spikes of each presynaptic index (one byte each) in the
                                 // step, by step parity; instance k from octet cbase
};

// An instance (dense block) of a projection expanded on MFMA: presynaptic indexes pre_begin .. + n_in - 1
// (n_oct octets) to postsynaptic neurons post_begin .. + post_count - 1; npad = post_count rounded up to 32.
struct alignas(8) DenseInst {
    uint32_t pre_begin, post_begin, post_count, npad, n_in, n_oct, cbase, pad_;
    uint64_t base;  // first BF16 weight of the instance in dn_w
};

// A 32-neuron tile of an instance: the instance's octets (cbase, n_oct), weights (base, npad), neurons
// (post_begin, post_count) and the tile's first neuron j0 within them.
struct alignas(16) DenseTile {
    uint32_t cbase, n_oct, npad, post_begin, post_count, j0;
    uint64_t base;
};

// The MFMA-expanded projections whose spike counts one launch of k_dense_counts takes (their count
// buffers: the projection's own, or a region of the device scratch).
constexpr uint32_t kDenseSetMax = 8;
struct DenseCountSet {
    uint32_t n;
    uint32_t proj[kDenseSetMax];
    uint32_t *cnt[kDenseSetMax];
};

// Small networks (every population tiled, all neurons on the threads of one workgroup): k_small_run runs
// whole steps in one workgroup with the rings and the step's spiking rows in LDS, so a step needs no
// device-scope store or poll. Population p's neurons sit on threads [pop_off[p], pop_off[p] + n) (pop_off a
// multiple of 64, so each wave holds one population's neurons); its ring at LDS cells [ring_off[p], + D * n).
constexpr uint32_t kSmallThreads = 1024, kSmallMaxPops = 4, kSmallRows = 4096;
constexpr uint32_t kSmallRingCells = 12 * 1024;  // 96 KiB of 8-byte ring cells
// Dynamic LDS: the rings (8-byte cells), then every projection's row offsets (row_ptr as 32-bit words from
// row_base[q], n_pre + 1 each), then the thresholds of single-pattern sources (from thr_base[s]).
constexpr uint32_t kSmallMaxSrcs = 8;
struct SmallRun {
    uint32_t n_pops;
    uint32_t pop_off[kSmallMaxPops];
    uint32_t ring_off[kSmallMaxPops];
    uint64_t proj_of_pop[kSmallMaxPops];  // bit q: projection q's presynaptic population is p
    uint32_t ring_cells;                  // words of the rings (8 bytes each)
    uint32_t row_words;                   // words of the row offsets (4 bytes each)
    uint32_t row_base[kMaxProjs];
    uint32_t thr_base[kSmallMaxSrcs];     // UINT32_MAX: several patterns, thresholds read per step
};

// MFMA-expanded projections that one launch of k_dense_expand takes together: they reach the same neurons
// through instances with the same postsynaptic ranges, with the same delay (one wave per tile sums them all).
constexpr uint32_t kDenseGroupMax = 2;
struct DenseGroup {
    uint32_t n;
    uint32_t proj[kDenseGroupMax];
    const uint32_t *cnt[kDenseGroupMax];
};
constexpr uint32_t kDenseCountBlocks = 1024;  // workgroups of k_dense_counts at most
constexpr uint32_t kDenseSkipEmpty = 1;  // flag: multiply only the octets of presynaptic indexes that spiked
constexpr uint32_t kDenseSplit = 2;      // flag: each weight as two BF16 values (hi + lo), two products

// A part entry of a tile directory: start | length << 32 | delay-1 records << 56 (255: all of the part's
// records count as delay 1, at 255 or more of them). Lengths stay below 2^24.
constexpr uint32_t kPartLenBits = 24, kUrgentAll = 255;
__host__ __device__ inline uint64_t dir_entry(uint32_t start, uint32_t len, uint32_t urgent) {
    return (uint64_t)start | (uint64_t)len << 32 | (uint64_t)(urgent < kUrgentAll ? urgent : kUrgentAll) << 56;
}

struct SourceDev {
    uint32_t proj, first, n_ch, n_patterns, steps_per_pattern, lo, hi, hash_offset;
    uint64_t seed;
    const uint32_t *thr;
    uint32_t *list;        // spiking channels (presynaptic indexes), segment c = channels 256c..
    uint32_t seg_base;
};

// Phase-1 task: a population update or a Bernoulli source, covering `blocks` workgroups.
struct TaskDev {
    uint32_t kind;  // 0 population, 1 source
    uint32_t index;
    uint32_t block_begin, blocks;
};

// Dense input projection from per-instance Bernoulli channel blocks, evaluated on MFMA one time block at a time.
// Weights are stored in MFMA fragment order: for tile c (32 postsynaptic neurons of one instance) and K-step
// kk (16 inputs), 64 lanes x (8 fp8 | 8 bf16) values, lane l holding rows j0 + (l % 32), inputs
// 16 kk + 8 (l / 32) + 0..7.
// Row blocks (32 steps each) of a dense projection's time block that one wave multiplies with the same
// weight fragments.
constexpr uint32_t kDenseGroup = 4;
struct DenseDev {
    uint32_t format;       // 0 fp8 e4m3, 1 bf16
    uint32_t post_pop;
    uint32_t n_in, ksteps;  // inputs, ceil(n_in / 16)
    uint32_t n_inst, n_tiles, T;  // instances, 32-column tiles, steps per time block
    uint32_t pitch;        // bytes per (step, instance) row of spike bits: 2 * ksteps rounded up to 16
    uint32_t threshold, hash_offset;
    uint64_t seed;
    const uint8_t *frag;   // [n_tiles][ksteps][64][8 or 16 bytes]
    const uint32_t *tile_inst;  // [n_tiles] instance of the tile
    const uint32_t *tile_j0;    // [n_tiles] first postsynaptic neuron (population index)
    const uint32_t *tile_ncol;  // [n_tiles] valid columns (<= 32)
    const float *row_scale;     // [population size] (1 outside instances)
    const uint32_t *inst_post_count;  // [n_inst]
    uint8_t *bits;         // [T][n_inst][pitch] spike bits of the current block (bit b of byte = input 8 byte + b)
    float *ff;             // [2T][population size] impacts by arrival step
};

// One 256-neuron tile of a tiled population: the workgroup that owns it reads every spiking row that can
// reach the tile (the virtual segments tile_v[vbeg..vend)), accumulates the tile's impacts in LDS and adds
// them to the ring without global atomics.
struct TileDev {
    uint32_t pop, tile, vbeg, vend;
};

// A spike segment that may reach a tile: its list and its projection (with kTileSegChannel set for a
// channel segment, whose messages the host pushes).
constexpr uint32_t kTileSegChannel = 0x80;
static_assert(kMaxProjs <= (int)kTileSegChannel / 2, "projection indexes stay below the channel and row flags");
struct TileSeg {
    const uint32_t *list;
    uint32_t proj;
};

struct Ctl {
    uint32_t cand_count[kMaxPops];
};

struct NetDev {
    uint32_t n_pops, n_projs, n_srcs, n_tasks, n_row_sources, total_blocks;
    PopDev pops[kMaxPops];
    ProjDev projs[kMaxProjs];
    SourceDev srcs[kMaxSources];
    TaskDev tasks[kMaxTasks];
    RowSource rows[kMaxRowSources];
    Ctl *ctl;                  // [2], indexed by step parity
    uint32_t n_seg, n_vseg;
    uint32_t *seg_count;       // [2][n_seg] spikes per segment, indexed by step parity
    uint32_t *const *seg_list; // [n_seg] first element of every segment (step parity 0)
    uint32_t list_pstride;     // offset of the parity-1 copy of every list
    const uint32_t *list_pool; // base of the spike-list pool (both parities)
    const uint32_t *vseg;      // [n_vseg] segment of every virtual segment
    const uint32_t *vrs;       // [n_vseg] row source of every virtual segment
    const uint32_t *vkey;      // [n_vseg] message key (multi-threaded semantics only, else null)
    const uint32_t *seg_vpos;  // [n_seg][4] virtual positions of every segment (UINT32_MAX = none)
    uint32_t *vcount;          // [2][n_vseg] spikes per virtual segment, indexed by step parity
    unsigned long long *part;  // [2 * scatter grid] per-workgroup impact and row counters
    uint32_t n_vseg_atomic;    // virtual segments [0, n_vseg_atomic) go through the atomic scatter
    uint32_t n_tiles;          // tiles of tiled populations (one tile-scatter workgroup each)
    const TileDev *tiles;      // [n_tiles]
    const TileSeg *tile_seg;   // segments reaching each tile (ranges in TileDev)
    uint32_t tile_max_list;    // longest such range
    uint32_t tile_max_d;       // deepest ring of a tiled population
    uint32_t n_src_blocks;     // workgroups of the Bernoulli sources (256 channels each)
    uint32_t *grid_bar;        // abort flag of persistent launches (word kBarAbort)
    uint32_t spin_limit;       // polls before a persistent launch's wait gives up (kSpinLimit)
    uint32_t hold_block;       // testing: this persistent workgroup withholds its spike counts (none: ~0u)
    uint64_t hold_until;       // testing: ... at steps before this one
    unsigned long long *tile_part;  // [4 * n_tiles] impacts, rows, row segments, ring cells updated
    uint32_t blk_proj[64];     // projection index of each projection with BLOCKING synapses
    uint32_t n_dense;
    DenseDev dense[kMaxDense];
    unsigned long long *events;  // [0] impacts, [1] rows expanded by the dense path (scatter: part)
    unsigned long long *step_ptr;  // the current step on the device (kernels launched from a captured graph read it)
};

}  // namespace knp_mi355x
