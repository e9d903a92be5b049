// naer2_hip.hip -- the fused time-driven step of a spiking network on MI355X (gfx950), and its C-ABI.
//
// One kernel launch per step of dt; one 64-lane wavefront (one workgroup) per target group of up to 64
// neurons of one population.  A workgroup owns the ring cells of its own targets, so the whole step --
// intake and delivery of the previous step's spikes, the ring read, the LIF update and the threshold --
// needs no grid-wide synchronisation beyond the kernel boundary between steps.
//
// Reference behaviour followed (MatteoSchmider/naer2):
//   pull intake     naer2/transpiler/templates/fused.py:1626-1965 (_pull): a target group reads the
//                   spikes of its sources and finds each spike's segment of links into the group
//                   (fused.py:252-267, _segments_of); here through a dense (source, group) offset table.
//   delivery        fused.py:2308-2604 (_deliver) and _walk fused.py:939-991: the links of a slot are
//                   walked in record order and the target lane runs the receive hook; here the walk
//                   accumulates into per-(delay, target) partial sums in LDS, added to the ring once.
//   ring            naer2/ring.py:132-245: arrival slots k % n_slots per target group; here fp32 input
//                   accumulators ring[n % d_slots][neuron] (tools/brunel_reference.py:29, buf[dsteps, N]).
//   update + send   tools/brunel_reference.py:34-48 (delta LIF) and test/fused_reference.py:86-107,
//                   test/networks/PD14/neuron.py:37-60 (exp-synapse LIF, exact propagator).
//   emission        fused.py:2611-2644 (_emit_spikes): a group appends its spiking source slots to the
//                   step's list with one atomic (as the reference's flat list, fused.py:2637) and writes its
//                   ballot word to the step's bitmap; the consumer sorts the list (<= 64 spikes) or scans
//                   the bitmap (bursts), so the delivery order is the source order either way.
//
// Summation order (identical in the CPU oracle): for target i and arrival step n, for each emission step s
// ascending, p = +0.0f, p += w over the links into i of sources spiking at s with s + d = n in ascending
// canonical link index; then ring[n % D][i] = ring[n % D][i] + p.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdarg>
#include <new>

#include "naer2_hip.h"
#include "twofloat.h"
#include "philox.h"

namespace {

constexpr int kWave = 64;
constexpr int kCapS = 256;   // spiking sources staged in LDS per pass (single-plan kernel)
constexpr int kCapL = 512;   // links staged in LDS per pass (LDS per group: 9.2 KB + 256 B per ring slot)
constexpr int kRowU = 8;     // row-intake 16-byte cells in flight per lane and batch
// The batched kernel stages less per pass and keeps fewer row cells in flight: 96 VGPRs and < 8 KB of LDS
// per group, so that 20 groups are resident per CU instead of 16 (the batch is throughput-bound).
constexpr int kBatchCapS = 128;
constexpr int kBatchCapL = 384;
constexpr int kBatchRowU = 4;
constexpr int kListSlots = 4;
constexpr int kMaxBatch = 256;  // plans in one batched launch
constexpr int kDenseWords = 1536;  // row intake: spike bitmaps of up to 1535 groups (98,240 neurons) in LDS, plus
                                   // the sentinel word (source slots 64 n_groups .. +63: never spiking)
// Single-plan launches deal consecutive target groups to one XCD by default (n2h_set_tuning): neighbouring
// groups read neighbouring entries of a spiking source's offset row and neighbouring link segments, so
// their cache lines are fetched once into one L2 instead of once into each of the 8.
constexpr int kDefaultXcds = 8;
constexpr int kTraceEvents = 8;  // entry, list read, sources ranked, segments, staged, walked, updated, exit

// In-kernel phase stamps (100 MHz wall clock), lane 0 of each group, when tracing is on: kept in LDS and
// written out at the end, so that the stamps add no global stores to the memory counter the kernel waits on.
#define N2H_STAMP(ev)                                                                    \
    do {                                                                                 \
        if (a.trace && lane == 0) st.stamps[(ev)] = wall_clock64();                      \
    } while (0)  // spike lists of steps s-1 (read), s (written), s+1 (cleared), s-2 (idle)

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define N2H_HIP(call)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(N2H_EHIP, "%s: %s", #call, hipGetErrorString(e_));   \
    } while (0)

struct StepArgs {
    int32_t n_neurons;
    int32_t n_groups;
    int32_t d_slots;
    uint32_t instance;
    int64_t bit_slots;
    int64_t list_stride;     // int32 words per spike-list slot: 1 + 64 * n_groups
    uint32_t key0, key1;
    const int32_t* group_first;
    const int32_t* group_info;
    const int64_t* seg_off;  // [64 * n_groups * n_groups + 1], indexed by (source slot) * n_groups + group
    const uint2* links;
    const n2h_pop_params* pops;
    float* v;
    float* isyn;
    int32_t* ref;
    float* ring;
    unsigned long long* bits;
    int32_t* spikes;         // [4][list_stride]: count, then source slots in any order
    const int64_t* step_ctr; // graph launches: the window's first step; eager launches: null
    int64_t step_imm;        // eager launches: the first step of the launch sequence
    int32_t phase;           // (first step of the sequence) mod 4: the list slots without a load
    int32_t slot_bits;       // bits of a source slot: ceil(log2(64 * n_groups))
    unsigned long long* trace;  // null, or [n_groups][kTraceEvents] wall-clock stamps of the last step
    const uint4* rows;          // row intake (null: off): each target's in-links in contract order
    const int64_t* row_base;    // [n_groups]: first 16-byte cell of group g's block (cell q of lane l: + 64 q + l)
    const int32_t* row_depth;   // [n_groups]: cells per lane
    const float* row_wtab;      // compact rows: the weight table [256]; slot16: {w_lo, w_hi, split slot bits}
    int32_t row_format;         // N2H_ROWS_WIDE / N2H_ROWS_COMPACT / N2H_ROWS_SLOT16
    int32_t row_uniform;        // every link's delay cell (delay - 1) when all delays are equal, else -1
    int32_t dense_min;          // row intake for steps following at least this many spikes
    int32_t acc_base;           // the delay cell held at acc[0]: row_uniform for uniform delays (one cell), else 0
    float* vtrace;              // null, or [vtrace_steps][n_neurons]: the potential before the threshold minus the
    int64_t vtrace_step0;       //   threshold, of steps vtrace_step0 .. + vtrace_steps - 1 (+inf on refractory steps)
    int32_t vtrace_steps;
    float* vlo;                 // LIF_EXP, null or [n_neurons] each: the remainders of V and I that fp32 rounding
    float* ilo;                 //   dropped (tracked exactly per operation; V and I themselves are unchanged)
    int32_t* log_rec;           // delivery log (the logging kernel variant only): [log_cap][5] records, null: off
    int32_t* log_count;         //   records appended so far (grows past log_cap; nothing written there)
    const int16_t* log_conn;    //   connection index of each link, parallel to links (null: -1)
    int64_t log_cap;
    int32_t* status;            // null, or [n_groups][4] status rows, written by every step
    const float* edge_attr;     // edge plans (the edge kernel variant): each link's second attribute (gain or
                                //   factor), parallel to links; link word x bits 8-9 the edge kind, bit 10 the port
    float* ring_v;              //   [d_slots][n]: the inputs of the potential port (null: no potential-port links)
    int32_t v_slots;            //   the potential port's LDS cells: its links' largest delay
    uint2* links_rw;            //   the link table, written by plastic links (same address as links)
    int32_t multirate;          // some population updates every stride > 1 steps (its parameter block's stride)
    int32_t model;              // N2H_MODEL_*: the network's neuron model (mixed batches branch on it per tile)
    int32_t xcd;                // single-plan launches: the XCD mapping (xcd_tile) of blocks to target groups
    int32_t xcd_run;
};

// The tile that block b runs under an XCD mapping (k, run) of n_tiles tiles (-1: a padding block).  The
// hardware deals a launch's blocks to its k XCDs in turn (block b to XCD b mod k).  k <= 1: block b runs
// tile b.  run <= 0: each XCD runs one contiguous chunk of ceil(n_tiles / k) tiles.  run > 0: runs of `run`
// consecutive tiles are dealt to the k XCDs in turn, so that each XCD's tiles are runs of neighbours and
// the XCDs' shares interleave over the whole tile order.
__host__ __device__ inline int64_t xcd_blocks(int64_t n_tiles, int k, int run) {
    if (k <= 1) return n_tiles;
    if (run <= 0) return static_cast<int64_t>(k) * ((n_tiles + k - 1) / k);
    const int64_t span = static_cast<int64_t>(k) * run;
    return span * ((n_tiles + span - 1) / span);
}
__device__ __forceinline__ int xcd_tile(int b, int n_tiles, int k, int run) {
    if (k <= 1) return b;
    int t;
    if (run <= 0) {
        t = (b % k) * ((n_tiles + k - 1) / k) + b / k;
    } else {
        const int i = b / k;
        t = ((i / run) * k + b % k) * run + i % run;
    }
    return t < n_tiles ? t : -1;
}

// The target group of this block in a single-plan launch (-1: a padding block of the XCD mapping).
__device__ __forceinline__ int plan_group(const StepArgs& a) {
    return xcd_tile(static_cast<int>(blockIdx.x), a.n_groups, a.xcd, a.xcd_run);
}

// Inclusive prefix sum over the 64 lanes of the wavefront, in DPP lane moves (no LDS round trip): a
// Hillis-Steele scan inside each row of 16 lanes (row_shr 1, 2, 4, 8; lanes shifted in from outside the row
// add 0), then row 0's total into rows 1 and 3 (row_bcast:15) and row 1's running total into rows 2 and 3
// (row_bcast:31).
__device__ __forceinline__ int wave_incl_scan(int x, int /*lane*/) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return x;
}

// Inclusive max-scan over the 64 lanes in DPP lane moves, for values whose running maximum from lane 0 is
// never below 0: lanes shifted in from outside a row contribute 0, which then changes no result.
__device__ __forceinline__ int wave_incl_max_nonneg(int x) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15, rows 1 and 3
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31, rows 2 and 3
    return x;
}

// The value of lane - 1 (wave_shr:1), `first` in lane 0.
__device__ __forceinline__ int wave_shift_up1(int x, int first) {
    return __builtin_amdgcn_update_dpp(first, x, 0x138, 0xf, 0xf, false);
}

// The value of lane 63 (the total of an inclusive scan), as a scalar.
__device__ __forceinline__ int last_lane(int x) { return __builtin_amdgcn_readlane(x, kWave - 1); }

// Rank of each valid lane's key among the valid lanes' keys (keys distinct, < 2^bits), from one ballot
// per key bit, most significant first: no cross-lane data movement.
__device__ __forceinline__ int wave_rank(int key, bool valid, int bits) {
    unsigned long long agree = __ballot(valid);  // valid lanes equal to this key on the bits seen so far
    int rank = 0;
    for (int p = bits - 1; p >= 0; --p) {
        const bool one = (key >> p) & 1;
        const unsigned long long b = __ballot(valid && one);
        if (one) rank += __popcll(agree & ~b);
        agree &= one ? b : ~b;
    }
    return rank;
}

// Shared-memory staging of one workgroup.  BM: 64-bit words of the row intake's spike bitmap (the
// groups' words, then the sentinel word); kernel variants size it for the groups they serve.
// CAPS / CAPL: spiking sources and links staged per pass; RU: row cells in flight per lane and batch.
template <int BM, int CAPS = kCapS, int CAPL = kCapL, int RU = kRowU, bool LOG = false, bool EDGE = false,
          bool PORTS = false>
struct Stage {
    static constexpr int kS = CAPS;
    static constexpr int kL = CAPL;
    static constexpr int kRU = RU;
    static constexpr bool kLog = LOG;
    static constexpr bool kEdge = EDGE;
    static constexpr bool kPorts = PORTS;         // edge plans with potential-port links
    static constexpr bool kCarry = LOG || EDGE;  // staged links carry their table index through the sort
    union {
        struct {                     // segment intake
            int32_t sp[CAPS];            // spiking source slots of this pass, ascending
            int64_t seg_lo[CAPS];        // first link of each source's segment into this group
            int32_t seg_start[CAPS + 1];  // exclusive scan of the segment lengths
            int16_t owner[CAPL];         // source (index into sp) of each staged link
            uint2 sorted[CAPL];          // the staged links regrouped by target, each target's in contract order
            int32_t tcount[kWave];       // staged links per target
            int32_t tstart[kWave];       // first of each target's links in sorted
            uint32_t log_idx[(LOG || EDGE) ? CAPL : 1];  // each sorted link's index into the link table
            int32_t log_src[LOG ? CAPL : 1];            //   and its source slot (delivery log)
        };
        struct {                     // row intake
            unsigned long long bm[BM];   // the previous step's spike bitmap, then a zero word
            float wtab[256];             // compact rows: the weight table
        };
    };
    uint32_t param_words[sizeof(n2h_pop_params) / 4];  // the group's population parameters
    unsigned long long stamps[kTraceEvents];             // phase stamps (tracing)
};

// Bitmap words of the kernel variant for KW bitmap words per lane (n_groups <= 64 KW): the groups'
// words and the sentinel; KW = 0 (more than 2048 groups) has no row intake.
constexpr int bm_words(int kw) { return kw == 0 ? 1 : (64 * kw < kDenseWords - 1 ? 64 * kw : kDenseWords - 1) + 1; }

// One delivered link of weight w to delay cell b of this lane's target, with the lane's current cell held
// in a register (cb, cr): a change of cell writes the old one back and fetches the new one (or starts it
// at +0.0f if untouched), so each cell still sees exactly its links' adds, in order.
__device__ __forceinline__ void cell_add(float* acc, int lane, int b, float w, int& cb, float& cr,
                                         unsigned long long& t0, unsigned long long& t1) {
    if (b != cb) {
        if (cb >= 0) acc[cb * kWave + lane] = cr;
        bool seen;
        if (b < 64) {
            seen = (t0 >> b) & 1ull;
            t0 |= 1ull << b;
        } else {
            seen = (t1 >> (b - 64)) & 1ull;
            t1 |= 1ull << (b - 64);
        }
        cr = seen ? acc[b * kWave + lane] : 0.0f;
        cb = b;
    }
    cr = cr + w;
}

// Walk of the nl staged links.  A stable counting sort by target (ranks among equal targets from one
// ballot per target bit, batch by batch) regroups them so that target lane t then walks exactly its own
// links, in contract order, adding each weight to its (delay, t) partial sum: the target-lane walk of the
// reference's _walk (fused.py:939-991) without every lane reading every link.  Touched delays are
// recorded in the lane's own mask (t0: delays 1..64, t1: 65..128).
// Logging (St::kLog): each staged link's table index and source slot travel with it through the sort, and
// every target lane appends a record per link it walks, in walk order (records of this chunk at one base
// reserved by lane 0): {arrival step, emission step, connection, source neuron, target neuron}.
struct LogCtx {
    int64_t s;     // the step: the staged links are the spikes of step s - 1
    int first;     // the target group's first neuron
    const uint32_t* idx;
    const int32_t* src;
    float* acc_v;                // edge plans with ports: the potential port's partial sums [v_slots][64]
    unsigned long long* u;       //   and its touched delays (u[0]: 1..64, u[1]: 65..128)
};

// Edges (St::kEdge; the reference's D-6 receive variants, D-7 per-link attribute blocks and edge forward):
// the value a link delivers by its kind -- static w; gain w * g; plastic: w += f first (the stored weight
// too: the update persists), then the new w -- into the partial sums of its port (current or potential).
__device__ __forceinline__ float edge_value(const StepArgs& a, uint32_t x, float w, uint32_t idx) {
    const uint32_t kind = (x >> 8) & 3u;
    if (kind == 1u) return w * a.edge_attr[idx];
    if (kind == 2u) {
        const float nw = w + a.edge_attr[idx];
        a.links_rw[idx].y = __float_as_uint(nw);
        return nw;
    }
    return w;
}

template <class St>
__device__ __forceinline__ void walk_chunk(St& st, float* acc, int lane, int nl, const uint2 (&r)[St::kL / kWave],
                                           unsigned long long& t0, unsigned long long& t1, const StepArgs& a,
                                           const LogCtx& lc) {
    constexpr int kPer = St::kL / kWave;
    const unsigned long long below = (1ull << lane) - 1ull;
    st.tcount[lane] = 0;
    __syncthreads();
    uint2 ln[kPer];
    int off[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        off[u] = 0;
        ln[u] = make_uint2(0u, 0u);
        if (u * kWave >= nl) continue;
        const int e = u * kWave + lane;
        const bool valid = e < nl;
        int t = kWave;  // invalid lanes: a target value no valid lane has
        if (valid) {
            ln[u] = r[u];
            t = static_cast<int>(ln[u].x & (St::kEdge ? 0x3fu : 0xffffu));
        }
        unsigned long long same = ~0ull;
#pragma unroll
        for (int p = 0; p < 7; ++p) {
            const bool one = (t >> p) & 1;
            const unsigned long long bb = __ballot(one);
            same &= one ? bb : ~bb;
        }
        if (valid) {
            const int c = st.tcount[t];
            off[u] = c + __popcll(same & below);
            if (!(same & below)) st.tcount[t] = c + __popcll(same);  // the first lane of the target
        }
    }
    __syncthreads();
    const int cnt = st.tcount[lane];
    const int incl = wave_incl_scan(cnt, lane);
    st.tstart[lane] = incl - cnt;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int e = u * kWave + lane;
        if (e < nl) {
            const int pos = st.tstart[ln[u].x & (St::kEdge ? 0x3fu : 0xffffu)] + off[u];
            st.sorted[pos] = ln[u];
            if constexpr (St::kCarry) st.log_idx[pos] = lc.idx[u];
            if constexpr (St::kLog) st.log_src[pos] = lc.src[u];
        }
    }
    __syncthreads();
    const int j0 = incl - cnt;
    int64_t log_base = 0;
    if constexpr (St::kLog) {
        int b = 0;
        if (lane == 0) b = atomicAdd(a.log_count, nl);
        log_base = __builtin_amdgcn_readfirstlane(b);
    }
    int cb = -1, cbv = -1;
    float cr = 0.0f, crv = 0.0f;
    for (int j = j0; j < incl; ++j) {
        const uint2 l = st.sorted[j];
        if constexpr (St::kEdge) {
            const float v = edge_value(a, l.x, __uint_as_float(l.y), st.log_idx[j]);
            const int b = static_cast<int>(l.x >> 16) - 1;
            if constexpr (St::kPorts) {
                // the potential port's cells stop at v_slots (its links' largest delay, set by the host)
                if ((l.x >> 10) & 1u) {
                    if (b < a.v_slots) cell_add(lc.acc_v, lane, b, v, cbv, crv, lc.u[0], lc.u[1]);
                } else {
                    cell_add(acc, lane, b, v, cb, cr, t0, t1);
                }
            } else {
                cell_add(acc, lane, b, v, cb, cr, t0, t1);
            }
        } else {
            cell_add(acc, lane, static_cast<int>(l.x >> 16) - 1, __uint_as_float(l.y), cb, cr, t0, t1);
        }
        if constexpr (St::kLog) {
            const int64_t at = log_base + j;
            if (at < a.log_cap) {
                const int slot = st.log_src[j];
                const int32_t src = a.group_first[slot >> 6] + (slot & (kWave - 1));
                int32_t* rec = a.log_rec + at * 5;
                rec[0] = static_cast<int32_t>(lc.s - 1 + static_cast<int64_t>(l.x >> 16));
                rec[1] = static_cast<int32_t>(lc.s - 1);
                rec[2] = a.log_conn ? static_cast<int32_t>(a.log_conn[st.log_idx[j]]) : -1;
                rec[3] = src;
                rec[4] = lc.first + lane;
            }
        }
    }
    if (cb >= 0) acc[cb * kWave + lane] = cr;
    if constexpr (St::kPorts) {
        if (cbv >= 0) lc.acc_v[cbv * kWave + lane] = crv;
    }
}

// Row intake of one batch of St::kRU 16-byte cells of this lane's row (kRU * PER entries).  Padding
// entries name the sentinel source slot, whose bitmap bit is always clear, so no entry needs a bounds
// test.  The bitmap words and weights of all entries are read from LDS first (one wait).
//   UNIFORM (every link has delay cell b0): branch-free -- the hit bit, sign-extended to a mask, selects
//   either the weight or -0.0f, the exact identity of fp32 addition, into the register sum cr, so the sum
//   sees exactly the spiking sources' weights in row order; `any` records whether there was one.
//   Otherwise: each hit goes to its own delay cell through cell_add.
// The slot16 format's weights: w_lo for source slots below `split`, w_hi from it on (fp32 bits).
struct SlotWeights {
    uint32_t split, w_lo, w_hi;
};

template <int PER, bool UNIFORM, class St>
__device__ __forceinline__ void row_batch(const St& st, float* acc, int lane, const uint4 (&c)[St::kRU], int& cb,
                                          float& cr, uint32_t& any, unsigned long long& t0, unsigned long long& t1,
                                          const SlotWeights& sw) {
    if constexpr (PER == 8) {
        // slot16 (uniform delay only): a cell's eight 16-bit source slots, low half of each word first; the
        // weight is the slot's side of the split; branch-free as the UNIFORM walk below
        static_assert(UNIFORM, "slot16 rows hold no delay: every link has the uniform delay");
        const uint32_t* bm32 = reinterpret_cast<const uint32_t*>(st.bm);
#pragma unroll
        for (int u = 0; u < St::kRU; ++u) {
            const uint32_t cw[4] = {c[u].x, c[u].y, c[u].z, c[u].w};
            uint32_t src[8], bw[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                src[e] = (cw[e >> 1] >> (16 * (e & 1))) & 0xffffu;
                bw[e] = bm32[src[e] >> 5];
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(bw[e]), src[e], 1u));
                any |= m;
                const uint32_t wb = src[e] >= sw.split ? sw.w_hi : sw.w_lo;
                cr = cr + __uint_as_float((wb & m) | (0x80000000u & ~m));
            }
        }
        return;
    }
    constexpr int kHalf = St::kRU / 2;  // cells per LDS round (bounds the registers the decoded entries hold)
    constexpr int E = kHalf * PER;
    const uint32_t* bm32 = reinterpret_cast<const uint32_t*>(st.bm);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        uint32_t word[E], bw[E], wb[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t* cw = reinterpret_cast<const uint32_t*>(&c[h * kHalf + e / PER]);
            const uint32_t x = (PER == 4) ? cw[e % PER] : cw[2 * (e % PER)];
            const uint32_t src = (PER == 4) ? (x >> 15) : (x >> 7);
            word[e] = x;
            bw[e] = bm32[src >> 5];
            wb[e] = (PER == 4) ? __float_as_uint(st.wtab[x & 255u]) : cw[2 * (e % PER) + 1];
        }
        if (UNIFORM) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t src = (PER == 4) ? (word[e] >> 15) : (word[e] >> 7);
                const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(bw[e]), src, 1u));
                any |= m;
                cr = cr + __uint_as_float((wb[e] & m) | (0x80000000u & ~m));
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t src = (PER == 4) ? (word[e] >> 15) : (word[e] >> 7);
                if ((bw[e] >> (src & 31u)) & 1u) {
                    const int b = (PER == 4) ? static_cast<int>((word[e] >> 8) & 127u) : static_cast<int>(word[e] & 127u);
                    cell_add(acc, lane, b, __uint_as_float(wb[e]), cb, cr, t0, t1);
                }
            }
        }
    }
}

// Row intake: this lane walks its target's in-link cells row[64 q], q < depth, against the spike bitmap in
// LDS.  Two register batches alternate so that one batch's loads are in flight while the other is
// processed; the loads are unconditional (indices clamped into the group's block), so each stays one
// dwordx4 and the compiler waits only for the batch it processes.  `ready` runs once the first batch is
// issued (the bitmap's barrier).  b0: the uniform delay cell (UNIFORM).
template <int PER, bool UNIFORM, class St, typename Ready>
__device__ __forceinline__ void row_walk(const St& st, float* acc, int lane, const uint4* row, int depth, int b0,
                                         uint32_t sentinel, unsigned long long& t0, unsigned long long& t1,
                                         Ready ready, const SlotWeights& sw = SlotWeights{0u, 0u, 0u}) {
    constexpr int RU = St::kRU;
    if (depth <= 0) {
        ready();
        return;
    }
    const int last = depth - 1;
    // a batch past the end of the block: its clamped loads repeat the last cell, so those cells are
    // replaced by sentinel entries (uniform branch, once per walk)
    const uint32_t sx = (PER == 8) ? (sentinel | (sentinel << 16)) : (PER == 4) ? (sentinel << 15) : (sentinel << 7);
    auto trim = [&](uint4 (&c)[RU], int q_first) {
        if (q_first + RU > depth) {
#pragma unroll
            for (int u = 0; u < RU; ++u)
                if (q_first + u > last) c[u] = make_uint4(sx, sx, sx, sx);
        }
    };
    uint4 A[RU], B[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) A[u] = row[static_cast<int64_t>(min(u, last)) * kWave];
    ready();
    int cb = -1;
    float cr = 0.0f;
    uint32_t any = 0u;
    for (int q0 = 0; q0 < depth; q0 += 2 * RU) {
#pragma unroll
        for (int u = 0; u < RU; ++u) B[u] = row[static_cast<int64_t>(min(q0 + RU + u, last)) * kWave];
        trim(A, q0);
        row_batch<PER, UNIFORM>(st, acc, lane, A, cb, cr, any, t0, t1, sw);
        if (q0 + RU >= depth) break;
#pragma unroll
        for (int u = 0; u < RU; ++u) A[u] = row[static_cast<int64_t>(min(q0 + 2 * RU + u, last)) * kWave];
        trim(B, q0 + RU);
        row_batch<PER, UNIFORM>(st, acc, lane, B, cb, cr, any, t0, t1, sw);
    }
    if (UNIFORM) {
        if (any) {  // the row intake is the step's only intake: the cell starts untouched
            if (b0 < 64) t0 |= 1ull << b0;
            else t1 |= 1ull << (b0 - 64);
            acc[b0 * kWave + lane] = cr;
        }
    } else if (cb >= 0) {
        acc[cb * kWave + lane] = cr;
    }
}

template <class St>
__device__ __forceinline__ void deliver_links(const StepArgs& a, St& st, float* acc, int lane, int n_sp, int L,
                                              unsigned long long& t0, unsigned long long& t1, int64_t s, int first,
                                              float* acc_v, unsigned long long* u);

// Delivery of the n_sp spiking sources in st.sp (ascending): their segments into group g, staged in
// contract order, walked by the target lanes into the per-(delay, target) partial sums acc.
template <class St>
__device__ __forceinline__ void deliver_pass(const StepArgs& a, St& st, float* acc, int g, int lane, int n_sp,
                                             unsigned long long& t0, unsigned long long& t1, int64_t s, int first,
                                             float* acc_v, unsigned long long* u) {
    const int NG = a.n_groups;
    // segment bounds: every load of the pass issued before the first is used
    int64_t lo_r[St::kS / kWave], hi_r[St::kS / kWave];
#pragma unroll
    for (int u = 0; u < St::kS / kWave; ++u) {
        const int i = u * kWave + lane;
        lo_r[u] = hi_r[u] = 0;
        if (i < n_sp) {
            const int64_t idx = static_cast<int64_t>(st.sp[i]) * NG + g;
            lo_r[u] = a.seg_off[idx];
            hi_r[u] = a.seg_off[idx + 1];
        }
    }
    int carry = 0;
#pragma unroll
    for (int u = 0; u < St::kS / kWave; ++u) {
        if (u * kWave >= n_sp) break;
        const int i = u * kWave + lane;
        const int n = static_cast<int>(hi_r[u] - lo_r[u]);
        const int inc = wave_incl_scan(n, lane);
        if (i < n_sp) {
            st.seg_start[i] = carry + inc - n;
            st.seg_lo[i] = lo_r[u];
        }
        carry += last_lane(inc);
    }
    if (lane == 0) st.seg_start[n_sp] = carry;
    __syncthreads();
    N2H_STAMP(3);
    deliver_links(a, st, acc, lane, n_sp, carry, t0, t1, s, first, acc_v, u);
}

// The L links of the n_sp sources whose segment bounds are in st (seg_lo, seg_start): staged in chunks of
// the stage's kL in contract order and walked into acc.
template <class St>
__device__ __forceinline__ void deliver_links(const StepArgs& a, St& st, float* acc, int lane, int n_sp, int L,
                                              unsigned long long& t0, unsigned long long& t1, int64_t s, int first,
                                              float* acc_v, unsigned long long* u) {
    for (int e0 = 0; e0 < L; e0 += St::kL) {
        const int e1 = min(L, e0 + St::kL);
        // owner[j]: the source whose segment holds link e0 + j.  Short chunks: a binary search per link;
        // long ones: the segment starts scattered, then an inclusive max-scan (sources in ascending start
        // order).
        constexpr int kPer = St::kL / kWave;
        if (e1 - e0 <= 2 * kWave) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = e0 + u * kWave + lane;
                if (e < e1) {
                    int lo_i = 0, hi_i = n_sp;  // seg_start[lo_i] <= e < seg_start[hi_i]
                    while (hi_i - lo_i > 1) {
                        const int mid = (lo_i + hi_i) >> 1;
                        if (st.seg_start[mid] <= e) lo_i = mid;
                        else hi_i = mid;
                    }
                    st.owner[e - e0] = static_cast<int16_t>(lo_i);
                }
            }
        } else {
    #pragma unroll
            for (int q = 0; q < kPer; ++q) st.owner[lane * kPer + q] = -1;
            int c0 = 0, c1 = n_sp;  // the segment holding e0: seg_start[c0] <= e0 < seg_start[c1]
            while (c1 - c0 > 1) {
                const int mid = (c0 + c1) >> 1;
                if (st.seg_start[mid] <= e0) c0 = mid;
                else c1 = mid;
            }
            __syncthreads();
    #pragma unroll
            for (int u = 0; u < St::kS / kWave; ++u) {
                const int i = u * kWave + lane;
                if (i < n_sp) {
                    const int st0 = st.seg_start[i];
                    if (st0 > e0 && st0 < e1 && st.seg_start[i + 1] > st0) st.owner[st0 - e0] = i;
                }
            }
            if (lane == 0) st.owner[0] = c0;
            __syncthreads();
            int run_max = -1;
            int own[kPer];
    #pragma unroll
            for (int q = 0; q < kPer; ++q) {
                run_max = max(run_max, static_cast<int>(st.owner[lane * kPer + q]));
                own[q] = run_max;
            }
            // the running owner over the lanes before this one (lane 0's first entry is c0 >= 0)
            const int pre = wave_shift_up1(wave_incl_max_nonneg(run_max), -1);
    #pragma unroll
            for (int q = 0; q < kPer; ++q) st.owner[lane * kPer + q] = static_cast<int16_t>(max(pre, own[q]));
            __syncthreads();
        }
        uint2 r[kPer];
        uint32_t lidx[St::kCarry ? kPer : 1];
        int32_t lsrc[St::kLog ? kPer : 1];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = e0 + u * kWave + lane;
            r[u] = make_uint2(0u, 0u);
            if (e < e1) {
                const int i = st.owner[e - e0];
                const int64_t li = st.seg_lo[i] + (e - st.seg_start[i]);
                r[u] = a.links[li];
                if constexpr (St::kCarry) lidx[u] = static_cast<uint32_t>(li);
                if constexpr (St::kLog) lsrc[u] = st.sp[i];
            }
        }
        N2H_STAMP(4);
        const int nl = e1 - e0;
        walk_chunk(st, acc, lane, nl, r, t0, t1, a, LogCtx{s, first, lidx, lsrc, acc_v, u});
        __syncthreads();
    }
}

// One step s of the whole network for target group g (one wavefront).
// KW: bitmap words per lane held in registers on the burst path (0: read them in a loop).
template <int MODEL, int KW, class St>
__device__ __forceinline__ void step_group(const StepArgs& a, const int g, const int32_t k, St& st, float* acc_lds) {
    const int lane = threadIdx.x;
    // acc[b * 64 + lane]: the partial sum of delay cell b; the LDS holds cells acc_base onwards (one cell for
    // a uniform delay)
    float* acc = acc_lds - a.acc_base * kWave;
    // edge plans: the potential port's partial sums follow the current port's (acc_base is 0 there)
    float* acc_v = St::kPorts ? acc_lds + a.d_slots * kWave : nullptr;
    unsigned long long tv[2] = {0ull, 0ull};
    const int NG = a.n_groups;
    const int64_t N = a.n_neurons;
    const int D = a.d_slots;
    if (a.trace && lane < kTraceEvents) st.stamps[lane] = 0ull;
    N2H_STAMP(0);
    const int slot_prev = (a.phase + k + kListSlots - 1) & (kListSlots - 1);  // step s-1
    const int slot_cur = (a.phase + k) & (kListSlots - 1);                    // step s

    // the spikes of step s-1, first (the longest chain starts here): their count and, speculatively, the
    // first 64 source slots
    const int32_t* plist = a.spikes + slot_prev * a.list_stride;
    const int n_prev = plist[0];
    const int my_src = plist[1 + lane];

    // uniform values in scalar registers: later code then never waits on the vector memory counter for them
    // (that counter also covers the fire-and-forget ring atomics)
    const int first = __builtin_amdgcn_readfirstlane(a.group_first[g]);
    const int info = __builtin_amdgcn_readfirstlane(a.group_info[g]);
    const int count = info & 0xffff;
    const int pop = info >> 16;
    const bool live = lane < count;
    const int nid = first + lane;
    // the population's parameters, one 32-bit word per lane, kept in LDS once the delivery is done
    constexpr int kParamWords = sizeof(n2h_pop_params) / 4;
    const uint32_t pword = lane < kParamWords ? reinterpret_cast<const uint32_t*>(a.pops + pop)[lane] : 0u;

    // state; the step (a load in graph launches) only addresses the ring and the bit history
    float V = 0.0f, I = 0.0f, x = 0.0f;
    int32_t R = 0;
    if (live && !a.multirate) {  // (multi-rate plans: at update steps only, below)
        V = a.v[nid];
        R = a.ref[nid];
        if (MODEL == N2H_MODEL_LIF_EXP) I = a.isyn[nid];
    }
    int64_t s0 = a.step_imm;
    if (a.step_ctr) {
        const uint64_t v = static_cast<uint64_t>(*a.step_ctr);
        s0 = static_cast<int64_t>(
            (static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32)))) << 32) |
            static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v))));
    }
    const int64_t s = s0 + k;
    const int s_mod = static_cast<int>(s % D);  // ring slot of step s; delays stay in 32-bit arithmetic
    const int64_t cur = static_cast<int64_t>(s_mod) * N + nid;
    // multi-rate: a population with stride k > 1 updates at the steps s with (s + 1) % k == 0 only, over the
    // sum of its window's k ring slots; at its other steps it only takes its intake into the ring
    int stride = 1;
    if (a.multirate) stride = __builtin_amdgcn_readfirstlane(a.pops[pop].stride);
    const bool upd = stride <= 1 || static_cast<uint64_t>(s + 1) % static_cast<uint64_t>(stride) == 0ull;
    if (live && upd) x = a.ring[cur];
    float xv = 0.0f;  // edge plans with ports: the potential port's due input
    if constexpr (St::kPorts) {
        if (live && upd) xv = a.ring_v[cur];
    }
    // the background draw needs only the neuron and the step: computed here, under the latency of the
    // spike-list load, and kept in a register (the empty asm pins it here)
    uint32_t u_bg = philox_word0(static_cast<uint32_t>(nid), static_cast<uint32_t>(s),
                                 static_cast<uint32_t>(static_cast<uint64_t>(s) >> 32), a.instance, a.key0, a.key1);
    asm volatile("" : "+v"(u_bg));
    if (live && upd && a.multirate) {
        V = a.v[nid];
        R = a.ref[nid];
        if (MODEL == N2H_MODEL_LIF_EXP) I = a.isyn[nid];
    }

    unsigned long long t0 = 0ull, t1 = 0ull;  // delays 1..64 and 65..128 with a partial sum for this lane
    int path = 0;  // the intake this step took (status rows): 0 none, 1 sorted list, 2 bitmap scan, 3 rows
    if (KW > 0 && a.rows && n_prev >= a.dense_min) {
        path = 3;
        // row intake (a CSR-by-target pull): every target lane walks its own in-links in
        // contract order against the previous step's spike bitmap held in LDS -- the whole table, so this
        // path pays off when a large share of the sources spiked
        const int64_t prev_slot = ((s - 1) % a.bit_slots + a.bit_slots) % a.bit_slots;
        const unsigned long long* prev = a.bits + prev_slot * NG;
        for (int w = lane; w < NG; w += kWave) st.bm[w] = prev[w];
        if (lane == 0) st.bm[NG] = 0ull;  // the sentinel word of padding entries
        const bool compact = a.row_format == N2H_ROWS_COMPACT;
        if (compact) {
#pragma unroll
            for (int j = 0; j < 256 / kWave; ++j) st.wtab[j * kWave + lane] = a.row_wtab[j * kWave + lane];
        }
        const int depth = __builtin_amdgcn_readfirstlane(a.row_depth[g]);
        const uint4* row = a.rows + a.row_base[g] + lane;
        const int b0 = a.row_uniform;
        auto ready = [&]() {
            __syncthreads();
            N2H_STAMP(2);
        };
        const uint32_t sentinel = static_cast<uint32_t>(kWave) * static_cast<uint32_t>(NG);
        if (a.row_format == N2H_ROWS_SLOT16) {  // uniform delay (checked when bound)
            const SlotWeights sw{__float_as_uint(a.row_wtab[2]), __float_as_uint(a.row_wtab[0]),
                                 __float_as_uint(a.row_wtab[1])};
            row_walk<8, true>(st, acc, lane, row, depth, b0, sentinel, t0, t1, ready, sw);
        } else if (compact) {
            if (b0 >= 0) row_walk<4, true>(st, acc, lane, row, depth, b0, sentinel, t0, t1, ready);
            else row_walk<4, false>(st, acc, lane, row, depth, b0, sentinel, t0, t1, ready);
        } else {
            if (b0 >= 0) row_walk<2, true>(st, acc, lane, row, depth, b0, sentinel, t0, t1, ready);
            else row_walk<2, false>(st, acc, lane, row, depth, b0, sentinel, t0, t1, ready);
        }
        __syncthreads();
    } else if (n_prev <= kWave) {
        path = n_prev > 0 ? 1 : 0;
        // sparse path: sort the (atomically appended) source slots, then deliver them.  Each lane's
        // segment bounds are loaded before the ranking ballots, which hide their latency, and land in LDS
        // at the lane's rank.
        N2H_STAMP(1);
        const bool has = lane < n_prev;
        int64_t lo = 0, hi = 0;
        if (has) {
            const int64_t idx = static_cast<int64_t>(my_src) * NG + g;
            lo = a.seg_off[idx];
            hi = a.seg_off[idx + 1];
        }
        const int rank = wave_rank(my_src, has, a.slot_bits);
        if (has) {
            st.sp[rank] = my_src;
            st.seg_lo[rank] = lo;
            st.seg_start[rank] = static_cast<int>(hi - lo);  // the length, until the scan below
        }
        __syncthreads();
        N2H_STAMP(2);
        if (n_prev > 0) {
            const int n = has ? st.seg_start[lane] : 0;
            const int inc = wave_incl_scan(n, lane);
            const int total = last_lane(inc);
            if (has) st.seg_start[lane] = inc - n;
            if (lane == 0) st.seg_start[n_prev] = total;
            __syncthreads();
            N2H_STAMP(3);
            deliver_links(a, st, acc, lane, n_prev, total, t0, t1, s, first, acc_v, tv);
        }
    } else {
        path = 2;
        // burst path: the spike bitmap of step s-1, lane-contiguous words, ascending source slots
        const int64_t prev_slot = ((s - 1) % a.bit_slots + a.bit_slots) % a.bit_slots;
        const unsigned long long* prev = a.bits + prev_slot * NG;
        const int W = (NG + kWave - 1) / kWave;
        const int w_lo = lane * W;
        const int w_hi = min(NG, w_lo + W);
        unsigned long long wd[KW > 0 ? KW : 1];
        int cnt = 0;
        if (KW > 0) {
#pragma unroll
            for (int q = 0; q < (KW > 0 ? KW : 1); ++q) wd[q] = (w_lo + q < w_hi) ? prev[w_lo + q] : 0ull;
#pragma unroll
            for (int q = 0; q < (KW > 0 ? KW : 1); ++q) cnt += __popcll(wd[q]);
        } else {
            for (int w = w_lo; w < w_hi; ++w) cnt += __popcll(prev[w]);
        }
        const int incl = wave_incl_scan(cnt, lane);
        const int excl = incl - cnt;
        const int total = last_lane(incl);
        for (int r0 = 0; r0 < total; r0 += St::kS) {
            const int r1 = min(total, r0 + St::kS);
            if (excl < r1 && excl + cnt > r0) {
                // this lane's words in order: each word's set bits ranked from the lane's scan offset, the
                // pass's [r0, r1) stored
                int rank = excl;
                auto take = [&](unsigned long long m, int q) {
                    const int pc = __popcll(m);
                    if (rank + pc <= r0) {
                        rank += pc;
                        return;
                    }
                    const int base = (w_lo + q) * kWave;
                    while (m) {
                        const int b = __ffsll(static_cast<long long>(m)) - 1;
                        if (rank >= r0 && rank < r1) st.sp[rank - r0] = base + b;
                        ++rank;
                        m &= m - 1ull;
                    }
                };
                if constexpr (KW > 0 && KW <= 16) {
                    // q is wave-uniform: the word is read by indexed register moves, no per-word select
#pragma unroll 1
                    for (int q = 0; q < w_hi - w_lo && rank < r1; ++q) take(wd[q], q);
                } else if constexpr (KW > 0) {
                    // wider arrays would go to scratch memory when indexed: the word is selected
                    for (int q = 0; q < w_hi - w_lo && rank < r1; ++q) {
                        unsigned long long m = 0ull;
#pragma unroll
                        for (int z = 0; z < KW; ++z) m = (z == q) ? wd[z] : m;
                        take(m, q);
                    }
                } else {
                    for (int q = 0; q < w_hi - w_lo && rank < r1; ++q) take(prev[w_lo + q], q);
                }
            }
            __syncthreads();
            deliver_pass(a, st, acc, g, lane, r1 - r0, t0, t1, s, first, acc_v, tv);
        }
    }

    N2H_STAMP(5);
    // the parameters into LDS before any store or atomic is issued: waiting for their load then waits for
    // loads only (the memory counter also counts stores and atomics)
    if (lane < kParamWords) st.param_words[lane] = pword;
    // the delay-1 cell's partial sum is the one due now: ring[s][i] + p, read in registers
    if (live && (t0 & 1ull)) x = x + acc[lane];
    if constexpr (St::kPorts) {
        if (live && (tv[0] & 1ull)) xv = xv + acc_v[lane];
    }

    // ---- background, LIF update, threshold ----
    __syncthreads();
    const n2h_pop_params& P = *reinterpret_cast<const n2h_pop_params*>(st.param_words);
    bool fired = false;
    float vpre = __builtin_inff();  // the potential the threshold sees (+inf: refractory)
    float vpre_lo = 0.0f;           // its remainder (compensated LIF_EXP only)
    if (live && upd) {
        if (stride > 1) {
            // the window's earlier slots, oldest first: ((r[s-k+1] + r[s-k+2]) + ... + r[s-1]) + r[s]
            int q = s_mod - (stride - 1);
            if (q < 0) q += D;
            float xw = a.ring[static_cast<int64_t>(q) * N + nid];
            for (int j = 2; j < stride; ++j) {
                if (++q == D) q = 0;
                xw = xw + a.ring[static_cast<int64_t>(q) * N + nid];
            }
            x = xw + x;
        }
        const uint32_t u = u_bg;
        int c = 0;
#pragma unroll
        for (int j = 0; j < N2H_POISSON_K; ++j) c += (u >= P.poisson_thr[j]) ? 1 : 0;
        const float in = x + P.j_ext * static_cast<float>(c);
        if (MODEL == N2H_MODEL_LIF_DELTA) {
            if (R > 0) {
                R -= 1;
            } else {
                const float vd = V * P.decay;
                V = vd + in;
                vpre = V;
                if (V >= P.v_th) {
                    fired = true;
                    V = P.v_reset;
                    R = P.ref_steps;
                }
            }
        } else if (!a.vlo) {
            I = I + in;
            if (R > 0) {
                V = P.v_reset;
                R -= 1;
            } else {
                const float dv = V - P.v_rest;
                const float pv = P.decay * dv;
                const float vr = P.v_rest + pv;
                const float pvi = P.p_vi * I;
                V = vr + pvi;
                if constexpr (St::kPorts) V = V + xv;  // the potential port: a jump after the propagation
                vpre = V;
            }
            I = P.p_i * I;
            if (V > P.v_th) {
                fired = true;
                V = P.v_reset;
                R = P.ref_steps;
            }
        } else {
            // The same operations in the same order, each with its exact rounding error: V, I and the spikes
            // are bit-identical to the branch above; vl / il carry the remainder of the trajectory that exact
            // arithmetic on the same inputs would follow (second-order terms rounded once more).
            using namespace n2h_twofloat;
            float vl = a.vlo[nid], il = a.ilo[nid];
            const df jc = two_prod(P.j_ext, static_cast<float>(c));
            const df sx = two_sum(x, jc.hi);  // sx.hi == in
            const df i1 = two_sum(I, sx.hi);
            I = i1.hi;
            il = i1.lo + (il + (sx.lo + jc.lo));
            if (R > 0) {
                V = P.v_reset;
                vl = 0.0f;
                R -= 1;
            } else {
                const df dv = two_sum(V, -P.v_rest);
                const float dv_lo = dv.lo + vl;
                const df pv = two_prod(P.decay, dv.hi);
                const float pv_lo = pv.lo + P.decay * dv_lo;
                const df vr = two_sum(P.v_rest, pv.hi);
                const float vr_lo = vr.lo + pv_lo;
                const df pvi = two_prod(P.p_vi, I);
                const float pvi_lo = pvi.lo + P.p_vi * il;
                const df vn = two_sum(vr.hi, pvi.hi);
                V = vn.hi;
                vl = vn.lo + (vr_lo + pvi_lo);
                vpre = V;
                vpre_lo = vl;
            }
            const df i2 = two_prod(P.p_i, I);
            I = i2.hi;
            il = i2.lo + P.p_i * il;
            if (V > P.v_th) {
                fired = true;
                V = P.v_reset;
                vl = 0.0f;
                R = P.ref_steps;
            }
            a.vlo[nid] = vl;
            a.ilo[nid] = il;
        }
    }

    N2H_STAMP(6);
    // ---- emission: the source slots appended to the step's list (the one returning atomic, first, so that
    // waiting for it waits for no other memory operation), then the bitmap word ----
    const unsigned long long mask = __ballot(fired);
    int32_t* clist = a.spikes + slot_cur * a.list_stride;
    if (mask) {
        int base = 0;
        if (lane == 0) base = atomicAdd(clist, __popcll(mask));
        base = __builtin_amdgcn_readlane(base, 0);
        if (fired) {
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            clist[1 + base + rank] = g * kWave + lane;
        }
    }
    if (lane == 0) a.bits[(s % a.bit_slots) * NG + g] = mask;
    if (a.status && lane == 0) {  // this group's status row: {step, spikes emitted, intake path, sources seen}
        int4* row = reinterpret_cast<int4*>(a.status) + g;
        *row = make_int4(static_cast<int32_t>(s), __popcll(mask), path, n_prev);
    }
    // the list of step s+1 starts empty (its slot was last read during step s-2)
    if (g == 0 && lane == 0) a.spikes[((slot_cur + 1) & (kListSlots - 1)) * a.list_stride] = 0;

    // ---- every remaining write last: nothing after them waits on the memory counter ----
    if (live) {
        if (upd) {
            a.v[nid] = V;
            a.ref[nid] = R;
            if (MODEL == N2H_MODEL_LIF_EXP) a.isyn[nid] = I;
            a.ring[cur] = 0.0f;
            if constexpr (St::kPorts) a.ring_v[cur] = 0.0f;
            int q = s_mod;
            for (int j = 1; j < stride; ++j) {  // the window's earlier slots, read above
                q = (q == 0) ? D - 1 : q - 1;
                a.ring[static_cast<int64_t>(q) * N + nid] = 0.0f;
            }
        }
        if (a.vtrace) {
            const int64_t q = s - a.vtrace_step0;
            if (q >= 0 && q < a.vtrace_steps) a.vtrace[q * N + nid] = (vpre - P.v_th) + vpre_lo;
        }
        // ring: add each touched cell's partial sum (arrival s - 1 + (b + 1), b < D); the cell's one add.
        // Without an update the delay-1 cell (arrival s) stays in the ring for the window's sum.
        unsigned long long m = upd ? (t0 & ~1ull) : t0;
        while (m) {
            const int b = __ffsll(static_cast<long long>(m)) - 1;
            m &= m - 1ull;
            int slot = s_mod + b;
            if (slot >= D) slot -= D;
            unsafeAtomicAdd(&a.ring[static_cast<int64_t>(slot) * N + nid], acc[b * kWave + lane]);
        }
        m = t1;
        while (m) {
            const int b = __ffsll(static_cast<long long>(m)) - 1 + 64;
            m &= m - 1ull;
            int slot = s_mod + b;
            if (slot >= D) slot -= D;
            unsafeAtomicAdd(&a.ring[static_cast<int64_t>(slot) * N + nid], acc[b * kWave + lane]);
        }
        if constexpr (St::kPorts) {  // the potential port's cells into its own ring, as above
            for (int h = 0; h < 2; ++h) {
                m = h == 0 ? (upd ? (tv[0] & ~1ull) : tv[0]) : tv[1];
                while (m) {
                    const int b = __ffsll(static_cast<long long>(m)) - 1 + 64 * h;
                    m &= m - 1ull;
                    int slot = s_mod + b;
                    if (slot >= D) slot -= D;
                    unsafeAtomicAdd(&a.ring_v[static_cast<int64_t>(slot) * N + nid], acc_v[b * kWave + lane]);
                }
            }
        }
    }
    N2H_STAMP(7);
    if (a.trace && lane < kTraceEvents) a.trace[static_cast<int64_t>(g) * kTraceEvents + lane] = st.stamps[lane];
}

template <int MODEL, int KW>
__global__ __launch_bounds__(kWave) void step_kernel(StepArgs a, int32_t k) {
    __shared__ Stage<bm_words(KW)> st;
    extern __shared__ float acc[];  // [d_slots][64] partial sums of this emission step
    const int g = plan_group(a);
    if (g < 0) return;
    step_group<MODEL, KW>(a, g, k, st, acc);
}

// The edge variant (n2h_bind_edges): the step without the row intake, links delivering by their edge kind into
// their port's partial sums and ring.
// With ports the LDS holds a second set of partial sums (v_slots cells): the stage then takes the batched
// kernel's smaller passes (kBatchCapS sources, kBatchCapL links), so that a PD14-sized plan (60 slots, ~30
// potential-port slots) still fits five groups per CU.
template <int MODEL, bool LOG, bool PORTS>
__global__ __launch_bounds__(kWave) void step_edge_kernel(StepArgs a, int32_t k) {
    __shared__ Stage<bm_words(0), PORTS ? kBatchCapS : kCapS, PORTS ? kBatchCapL : kCapL, kRowU, LOG, true, PORTS> st;
    extern __shared__ float acc[];
    const int g = plan_group(a);
    if (g < 0) return;
    step_group<MODEL, 0>(a, g, k, st, acc);
}

// The logging variant (n2h_bind_log): the step without the row intake (KW = 0: segment intakes only, which
// give the same results), every delivered link recorded.
template <int MODEL>
__global__ __launch_bounds__(kWave) void step_log_kernel(StepArgs a, int32_t k) {
    __shared__ Stage<bm_words(0), kCapS, kCapL, kRowU, true> st;
    extern __shared__ float acc[];
    const int g = plan_group(a);
    if (g < 0) return;
    step_group<MODEL, 0>(a, g, k, st, acc);
}

// Batched step: the target groups of many independent networks (plans) in one grid.  Tiles are the
// plans' groups in plan order; block b runs tile xcd_tile(b, T, xcd, run): with xcd = k > 1 consecutive
// tiles -- neighbouring groups of one plan -- land on the same XCD and share its L2.
struct BatchArgs {
    const StepArgs* plans;     // device [n_plans]
    const int64_t* step_ctr;   // graph launches: the window's first step; eager launches: null
    int64_t step_imm;
    int32_t phase;
    int32_t n_plans;
    int32_t n_tiles;
    int32_t xcd;
    int32_t run;
    int32_t first[kMaxBatch + 1];  // first tile of each plan
};

// MODEL = kModelMixed: a batch of networks of both models (a heterogeneous fused round, naer2/fusion.py:61-116
// over runtimes of different node types); each tile takes its own network's model, a branch uniform over
// the workgroup.
constexpr int kModelMixed = -1;

// WIDE (n2h_tuning.stage 1): the single-plan kernel's stage (kCapS sources, kCapL links a pass, kRowU row cells
// in flight) instead of the lean batched one -- fewer passes and more row cells in flight a group, more LDS
// and registers a group.
template <int MODEL, int KW, bool WIDE = false>
__global__ __launch_bounds__(kWave) void batch_kernel(BatchArgs ba, int32_t k) {
    __shared__ Stage<bm_words(KW), WIDE ? kCapS : kBatchCapS, WIDE ? kCapL : kBatchCapL, WIDE ? kRowU : kBatchRowU> st;
    extern __shared__ float acc[];
    const int tile = xcd_tile(static_cast<int>(blockIdx.x), ba.n_tiles, ba.xcd, ba.run);
    if (tile < 0) return;
    int lo = 0, hi = ba.n_plans;  // first[lo] <= tile < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ba.first[mid] <= tile) lo = mid;
        else hi = mid;
    }
    StepArgs a = ba.plans[lo];
    a.step_ctr = ba.step_ctr;
    a.step_imm = ba.step_imm;
    a.phase = ba.phase;
    if constexpr (MODEL == kModelMixed) {
        if (a.model == N2H_MODEL_LIF_DELTA) step_group<N2H_MODEL_LIF_DELTA, KW>(a, tile - ba.first[lo], k, st, acc);
        else step_group<N2H_MODEL_LIF_EXP, KW>(a, tile - ba.first[lo], k, st, acc);
    } else {
        step_group<MODEL, KW>(a, tile - ba.first[lo], k, st, acc);
    }
}

__global__ void advance_kernel(int64_t* ctr, int32_t n) { *ctr += n; }

// n2h_batch_sync: up to kSyncChunk plans' launch arguments by value (kernel arguments, so nothing on the host
// has to outlive the call) written over the batch's device copy, and the batch's step counter.
constexpr int kSyncChunk = 8;
struct SyncChunk {
    StepArgs args[kSyncChunk];
    int64_t* step_ctr;  // null: leave it
    int64_t step;
    int32_t first;
    int32_t n;
};
static_assert(sizeof(SyncChunk) <= 4000, "kernel arguments are limited to 4 KB");
__global__ void batch_sync_kernel(StepArgs* dst, SyncChunk c) {
    const int i = static_cast<int>(threadIdx.x);
    if (i < c.n) dst[c.first + i] = c.args[i];
    if (i == 0 && c.step_ctr) *c.step_ctr = c.step;
}
__global__ void advance_many_kernel(int64_t* const* ctrs, int32_t n_ctrs, int32_t n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_ctrs) *ctrs[i] += n;
}
__global__ void set_kernel(int64_t* ctr, int64_t v) { *ctr = v; }

}  // namespace

struct n2h_plan {
    n2h_model_desc desc;
    n2h_tables tables;
    int32_t* pop_first_host;
    n2h_pop_params* pops_dev;
    int64_t* step_ctr;
    float* v;
    float* isyn;
    int32_t* ref;
    float* ring;
    unsigned long long* bits;
    int32_t* spikes;
    int64_t bit_slots;
    int64_t host_step;  // the step the next enqueued launch computes (mirror of *step_ctr)
    unsigned long long* trace;
    const uint4* rows;
    const int64_t* row_base;
    const int32_t* row_depth;
    const float* row_wtab;
    int32_t row_format;
    int32_t dense_min;
    float* vtrace;
    int64_t vtrace_step0;
    int32_t vtrace_steps;
    float* vlo;
    float* ilo;
    int32_t* log_rec;
    int32_t* log_count;
    const int16_t* log_conn;
    int64_t log_cap;
    int32_t* status;
    const float* edge_attr;
    float* ring_v;
    int32_t v_slots;        // edge plans: LDS cells of the potential port (0: no potential-port links)
    bool edges;             // n2h_bind_edges bound kinds or ports
    int32_t multirate;      // some population's stride > 1
    int32_t xcd;            // launch mapping (n2h_set_tuning): xcd_tile's k and run
    int32_t xcd_run;
    size_t lds_bytes;
    int device;
};

struct n2h_batch {
    n2h_plan** plans;        // in launch order
    int32_t n_plans;
    int32_t model;
    int32_t words;           // bitmap words per lane on the burst path (kernel variant)
    n2h_tuning tuning;
    StepArgs* args_dev;      // device [n_plans]
    int64_t** ctrs_dev;      // device [n_plans]: the plans' step counters
    int64_t* step_ctr;       // the batch's own counter (graph windows)
    int64_t host_step;
    int32_t n_tiles;
    size_t lds_bytes;
    int32_t first[kMaxBatch + 1];
};

struct n2h_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    n2h_plan* plan;          // a plan's window, or
    n2h_batch* batch;        // a batch's window
    int32_t n_steps;
    int32_t phase;  // first step mod 4 of every window it replays
};

namespace {

// Enqueue n_steps step kernels.  graph: the steps come from the device counter (the window replays
// anywhere with the same phase); eager: from the host mirror.  Either way the counter and the mirror
// advance by n_steps.
StepArgs make_args(const n2h_plan* p) {
    StepArgs a;
    memset(&a, 0, sizeof(a));
    a.n_neurons = p->desc.n_neurons;
    a.model = p->desc.model;
    a.n_groups = p->tables.n_groups;
    a.d_slots = p->desc.d_slots;
    a.instance = p->desc.instance;
    a.bit_slots = p->bit_slots;
    a.key0 = static_cast<uint32_t>(p->desc.seed);
    a.key1 = static_cast<uint32_t>(p->desc.seed >> 32);
    a.group_first = p->tables.group_first;
    a.group_info = p->tables.group_info;
    a.seg_off = p->tables.seg_off;
    a.links = reinterpret_cast<const uint2*>(p->tables.links);
    a.pops = p->pops_dev;
    a.v = p->v;
    a.isyn = p->isyn;
    a.ref = p->ref;
    a.ring = p->ring;
    a.bits = p->bits;
    a.step_ctr = nullptr;
    a.step_imm = p->host_step;
    a.phase = static_cast<int32_t>(p->host_step & (kListSlots - 1));
    a.slot_bits = 6;
    while ((1ll << a.slot_bits) < static_cast<int64_t>(kWave) * a.n_groups) ++a.slot_bits;
    a.spikes = p->spikes;
    a.list_stride = 1 + static_cast<int64_t>(kWave) * a.n_groups;
    a.trace = p->trace;
    a.rows = p->rows;
    a.row_base = p->row_base;
    a.row_depth = p->row_depth;
    a.row_wtab = p->row_wtab;
    a.row_format = p->row_format;
    a.row_uniform = p->tables.uniform_delay - 1;
    a.acc_base = p->tables.uniform_delay > 0 ? p->tables.uniform_delay - 1 : 0;
    a.dense_min = p->dense_min;
    a.vtrace = p->vtrace;
    a.vtrace_step0 = p->vtrace_step0;
    a.vtrace_steps = p->vtrace_steps;
    a.vlo = p->vlo;
    a.ilo = p->ilo;
    a.log_rec = p->log_rec;
    a.log_count = p->log_count;
    a.log_conn = p->log_conn;
    a.log_cap = p->log_cap;
    a.status = p->status;
    a.edge_attr = p->edge_attr;
    a.ring_v = p->ring_v;
    a.v_slots = p->v_slots;
    a.links_rw = const_cast<uint2*>(reinterpret_cast<const uint2*>(p->tables.links));
    a.multirate = p->multirate;
    a.xcd = p->xcd;
    a.xcd_run = p->xcd_run;
    return a;
}

// The kernel variant for `words` bitmap words per lane (n_groups <= 64 words).
template <int MODEL>
void launch_step_model(int words, dim3 grid, size_t lds, hipStream_t stream, const StepArgs& a, int32_t k) {
    const dim3 block(kWave);
    if (words <= 8) hipLaunchKernelGGL((step_kernel<MODEL, 8>), grid, block, lds, stream, a, k);
    else if (words <= 16) hipLaunchKernelGGL((step_kernel<MODEL, 16>), grid, block, lds, stream, a, k);
    else if (words <= 32) hipLaunchKernelGGL((step_kernel<MODEL, 32>), grid, block, lds, stream, a, k);
    else hipLaunchKernelGGL((step_kernel<MODEL, 0>), grid, block, lds, stream, a, k);
}

int launch_steps(n2h_plan* p, int32_t n_steps, hipStream_t stream, bool graph) {
    StepArgs a = make_args(p);
    a.step_ctr = graph ? p->step_ctr : nullptr;
    const dim3 grid(static_cast<unsigned>(xcd_blocks(a.n_groups, a.xcd, a.xcd_run))), block(kWave);
    const int words = (a.n_groups + kWave - 1) / kWave;  // bitmap words per lane on the burst path
    for (int32_t k = 0; k < n_steps; ++k) {
        if (p->edges) {  // edge plans (LIF_EXP only)
            constexpr int E = N2H_MODEL_LIF_EXP;
            if (p->log_rec && p->ring_v)
                hipLaunchKernelGGL((step_edge_kernel<E, true, true>), grid, block, p->lds_bytes, stream, a, k);
            else if (p->log_rec)
                hipLaunchKernelGGL((step_edge_kernel<E, true, false>), grid, block, p->lds_bytes, stream, a, k);
            else if (p->ring_v)
                hipLaunchKernelGGL((step_edge_kernel<E, false, true>), grid, block, p->lds_bytes, stream, a, k);
            else
                hipLaunchKernelGGL((step_edge_kernel<E, false, false>), grid, block, p->lds_bytes, stream, a, k);
        } else if (p->log_rec) {
            if (p->desc.model == N2H_MODEL_LIF_DELTA)
                hipLaunchKernelGGL(step_log_kernel<N2H_MODEL_LIF_DELTA>, grid, block, p->lds_bytes, stream, a, k);
            else
                hipLaunchKernelGGL(step_log_kernel<N2H_MODEL_LIF_EXP>, grid, block, p->lds_bytes, stream, a, k);
        } else if (p->desc.model == N2H_MODEL_LIF_DELTA) {
            launch_step_model<N2H_MODEL_LIF_DELTA>(words, grid, p->lds_bytes, stream, a, k);
        } else {
            launch_step_model<N2H_MODEL_LIF_EXP>(words, grid, p->lds_bytes, stream, a, k);
        }
    }
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, stream, p->step_ctr, n_steps);
    N2H_HIP(hipGetLastError());
    p->host_step += n_steps;
    return N2H_OK;
}

// Capture the launches fn enqueues on a private stream into an instantiated graph.
template <typename F>
int capture(F fn, hipGraph_t* graph_out, hipGraphExec_t* exec_out) {
    hipStream_t cap;
    N2H_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cap);
        return fail(N2H_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    int rc = fn(cap);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cap, &graph);
    (void)hipStreamDestroy(cap);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(N2H_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(N2H_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *graph_out = graph;
    *exec_out = exec;
    return N2H_OK;
}

// Every plan of a batch must still be at the batch's step (none stepped on its own since).
int batch_in_step(const n2h_batch* bt) {
    for (int32_t i = 0; i < bt->n_plans; ++i)
        if (bt->plans[i]->host_step != bt->host_step)
            return fail(N2H_ESTATE, "plan %d is at step %lld, the batch at %lld", i,
                        static_cast<long long>(bt->plans[i]->host_step), static_cast<long long>(bt->host_step));
    return N2H_OK;
}

template <int MODEL>
void launch_batch_model(const n2h_batch* bt, const BatchArgs& ba, dim3 grid, hipStream_t stream, int32_t k) {
    const dim3 block(kWave);
    if (MODEL != kModelMixed && bt->tuning.stage == 1 && bt->words <= 16) {  // the wide stage (single-model batches)
        if (bt->words <= 8)
            hipLaunchKernelGGL((batch_kernel<MODEL, 8, true>), grid, block, bt->lds_bytes, stream, ba, k);
        else if (bt->words <= 13)
            hipLaunchKernelGGL((batch_kernel<MODEL, 13, true>), grid, block, bt->lds_bytes, stream, ba, k);
        else
            hipLaunchKernelGGL((batch_kernel<MODEL, 16, true>), grid, block, bt->lds_bytes, stream, ba, k);
        return;
    }
    if (bt->words <= 8)
        hipLaunchKernelGGL((batch_kernel<MODEL, 8>), grid, block, bt->lds_bytes, stream, ba, k);
    else if (bt->words <= 13)  // up to 832 groups a plan (Brunel networks up to 53k neurons) at 20 groups per CU
        hipLaunchKernelGGL((batch_kernel<MODEL, 13>), grid, block, bt->lds_bytes, stream, ba, k);
    else if (bt->words <= 16)
        hipLaunchKernelGGL((batch_kernel<MODEL, 16>), grid, block, bt->lds_bytes, stream, ba, k);
    else if (bt->words <= 32)
        hipLaunchKernelGGL((batch_kernel<MODEL, 32>), grid, block, bt->lds_bytes, stream, ba, k);
    else
        hipLaunchKernelGGL((batch_kernel<MODEL, 0>), grid, block, bt->lds_bytes, stream, ba, k);
}

// Enqueue n_steps batched step kernels and advance the batch's and every plan's device counter.
int launch_batch(n2h_batch* bt, int32_t n_steps, hipStream_t stream, bool graph) {
    BatchArgs ba;
    memset(&ba, 0, sizeof(ba));
    ba.plans = bt->args_dev;
    ba.step_ctr = graph ? bt->step_ctr : nullptr;
    ba.step_imm = bt->host_step;
    ba.phase = static_cast<int32_t>(bt->host_step & (kListSlots - 1));
    ba.n_plans = bt->n_plans;
    ba.n_tiles = bt->n_tiles;
    ba.xcd = bt->tuning.xcd;
    ba.run = bt->tuning.run;
    for (int32_t i = 0; i <= bt->n_plans; ++i) ba.first[i] = bt->first[i];
    const dim3 grid(static_cast<unsigned>(xcd_blocks(bt->n_tiles, ba.xcd, ba.run)));
    for (int32_t k = 0; k < n_steps; ++k) {
        if (bt->model == N2H_MODEL_LIF_DELTA) launch_batch_model<N2H_MODEL_LIF_DELTA>(bt, ba, grid, stream, k);
        else if (bt->model == N2H_MODEL_LIF_EXP) launch_batch_model<N2H_MODEL_LIF_EXP>(bt, ba, grid, stream, k);
        else launch_batch_model<kModelMixed>(bt, ba, grid, stream, k);
    }
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, stream, bt->step_ctr, n_steps);
    hipLaunchKernelGGL(advance_many_kernel, dim3((bt->n_plans + 63) / 64), dim3(64), 0, stream, bt->ctrs_dev,
                       bt->n_plans, n_steps);
    N2H_HIP(hipGetLastError());
    bt->host_step += n_steps;
    return N2H_OK;
}

int check_ready(n2h_plan* p) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (!p->v || !p->ref || (p->desc.model == N2H_MODEL_LIF_EXP && !p->isyn))
        return fail(N2H_ESTATE, "state not bound (n2h_bind_state)");
    if (!p->ring || !p->bits || !p->spikes) return fail(N2H_ESTATE, "ring not bound (n2h_bind_ring)");
    return N2H_OK;
}

}  // namespace

extern "C" {

int n2h_version(void) { return N2H_ABI_VERSION; }

int n2h_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return N2H_EINVAL;
    strncpy(buf, g_err, len - 1);
    buf[len - 1] = '\0';
    return N2H_OK;
}

int n2h_plan_create(const n2h_model_desc* desc, const n2h_tables* tables, n2h_plan** out) {
    if (!desc || !tables || !out) return fail(N2H_EINVAL, "null argument");
    *out = nullptr;
    if (desc->model != N2H_MODEL_LIF_DELTA && desc->model != N2H_MODEL_LIF_EXP)
        return fail(N2H_EINVAL, "unknown model %d", desc->model);
    if (desc->n_neurons <= 0) return fail(N2H_EINVAL, "n_neurons must be positive");
    if (desc->n_pops < 1 || desc->n_pops > N2H_MAX_POPS)
        return fail(N2H_EINVAL, "n_pops %d outside [1, %d]", desc->n_pops, N2H_MAX_POPS);
    if (desc->d_slots < 1 || desc->d_slots > N2H_MAX_SLOTS)
        return fail(N2H_EINVAL, "d_slots %d outside [1, %d]", desc->d_slots, N2H_MAX_SLOTS);
    if (!desc->pop_first || !desc->pops) return fail(N2H_EINVAL, "null population arrays");
    if (desc->pop_first[0] != 0 || desc->pop_first[desc->n_pops] != desc->n_neurons)
        return fail(N2H_EINVAL, "pop_first must run from 0 to n_neurons");
    for (int p = 0; p < desc->n_pops; ++p) {
        if (desc->pop_first[p + 1] < desc->pop_first[p]) return fail(N2H_EINVAL, "pop_first not monotone");
        for (int j = 1; j < N2H_POISSON_K; ++j)
            if (desc->pops[p].poisson_thr[j] < desc->pops[p].poisson_thr[j - 1])
                return fail(N2H_EINVAL, "population %d: poisson_thr not non-decreasing", p);
        if (desc->pops[p].ref_steps < 0) return fail(N2H_EINVAL, "population %d: negative ref_steps", p);
        const int stride = desc->pops[p].stride;
        if (stride < 0 || stride > N2H_MAX_STRIDE || stride > desc->d_slots)
            return fail(N2H_EINVAL, "population %d: stride %d outside [0, min(%d, d_slots)]", p, stride,
                        N2H_MAX_STRIDE);
    }
    if (tables->n_groups <= 0) return fail(N2H_EINVAL, "n_groups must be positive");
    if (tables->uniform_delay < 0 || tables->uniform_delay > desc->d_slots)
        return fail(N2H_EINVAL, "uniform_delay %d outside [0, d_slots]", tables->uniform_delay);
    if (!tables->group_first || !tables->group_info || !tables->seg_off || (!tables->links && tables->n_links))
        return fail(N2H_EINVAL, "null table pointer");
    int64_t min_groups = 0;
    for (int p = 0; p < desc->n_pops; ++p)
        min_groups += (desc->pop_first[p + 1] - desc->pop_first[p] + kWave - 1) / kWave;
    if (tables->n_groups < min_groups) return fail(N2H_EINVAL, "n_groups %d below the populations' %lld",
                                                   tables->n_groups, static_cast<long long>(min_groups));

    n2h_plan* p = new (std::nothrow) n2h_plan();
    if (!p) return fail(N2H_ENOMEM, "host allocation");
    p->desc = *desc;
    p->tables = *tables;
    p->pop_first_host = new (std::nothrow) int32_t[desc->n_pops + 1];
    if (!p->pop_first_host) {
        delete p;
        return fail(N2H_ENOMEM, "host allocation");
    }
    memcpy(p->pop_first_host, desc->pop_first, sizeof(int32_t) * (desc->n_pops + 1));
    p->desc.pop_first = p->pop_first_host;
    p->desc.pops = nullptr;
    for (int q = 0; q < desc->n_pops; ++q) p->multirate |= desc->pops[q].stride > 1 ? 1 : 0;
    // partial sums: one cell per delay, or the one cell of a uniform delay
    p->lds_bytes = sizeof(float) * kWave * static_cast<size_t>(tables->uniform_delay > 0 ? 1 : desc->d_slots);
    p->xcd = kDefaultXcds;
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipMalloc(&p->pops_dev, sizeof(n2h_pop_params) * desc->n_pops);
    if (e == hipSuccess) e = hipMalloc(&p->step_ctr, sizeof(int64_t));
    if (e == hipSuccess)
        e = hipMemcpy(p->pops_dev, desc->pops, sizeof(n2h_pop_params) * desc->n_pops, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p->step_ctr, 0, sizeof(int64_t));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        n2h_plan_destroy(p);
        return fail(N2H_EHIP, "plan setup: %s", hipGetErrorString(e));
    }
    *out = p;
    return N2H_OK;
}

int n2h_plan_destroy(n2h_plan* p) {
    if (!p) return N2H_OK;
    if (p->pops_dev) (void)hipFree(p->pops_dev);
    if (p->step_ctr) (void)hipFree(p->step_ctr);
    delete[] p->pop_first_host;
    delete p;
    return N2H_OK;
}

int n2h_bind_state(n2h_plan* p, float* v, float* i_syn, int32_t* ref) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (!v || !ref) return fail(N2H_EINVAL, "null state pointer");
    if (p->desc.model == N2H_MODEL_LIF_EXP && !i_syn) return fail(N2H_EINVAL, "LIF_EXP needs i_syn");
    p->v = v;
    p->isyn = i_syn;
    p->ref = ref;
    return N2H_OK;
}

int n2h_bind_ring(n2h_plan* p, float* ring, uint64_t* spike_bits, int64_t bit_slots, int32_t* spike_list) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (!ring || !spike_bits || !spike_list) return fail(N2H_EINVAL, "null ring pointer");
    if (bit_slots < 2) return fail(N2H_EINVAL, "bit_slots must be at least 2");
    p->ring = ring;
    p->spikes = spike_list;
    p->bits = reinterpret_cast<unsigned long long*>(spike_bits);
    p->bit_slots = bit_slots;
    return N2H_OK;
}

int n2h_set_step(n2h_plan* p, int64_t step, void* stream) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (step < 0) return fail(N2H_EINVAL, "negative step");
    hipLaunchKernelGGL(set_kernel, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), p->step_ctr, step);
    N2H_HIP(hipGetLastError());
    p->host_step = step;
    return N2H_OK;
}

int n2h_get_step(n2h_plan* p, int64_t* step) {
    if (!p || !step) return fail(N2H_EINVAL, "null argument");
    N2H_HIP(hipMemcpy(step, p->step_ctr, sizeof(int64_t), hipMemcpyDeviceToHost));
    return N2H_OK;
}

int n2h_step(n2h_plan* p, int32_t n_steps, void* stream) {
    int rc = check_ready(p);
    if (rc) return rc;
    if (n_steps < 0) return fail(N2H_EINVAL, "negative n_steps");
    if (n_steps == 0) return N2H_OK;
    return launch_steps(p, n_steps, static_cast<hipStream_t>(stream), false);
}

int n2h_graph_create(n2h_plan* p, int32_t n_steps, void* stream, n2h_graph** out) {
    (void)stream;
    if (!out) return fail(N2H_EINVAL, "null out");
    *out = nullptr;
    int rc = check_ready(p);
    if (rc) return rc;
    if (n_steps <= 0) return fail(N2H_EINVAL, "n_steps must be positive");
    const int64_t at = p->host_step;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    rc = capture([&](hipStream_t cap) { return launch_steps(p, n_steps, cap, true); }, &graph, &exec);
    p->host_step = at;  // captured, not run
    if (rc) return rc;
    n2h_graph* gr = new (std::nothrow)
        n2h_graph{graph, exec, p, nullptr, n_steps, static_cast<int32_t>(at & (kListSlots - 1))};
    if (!gr) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        return fail(N2H_ENOMEM, "host allocation");
    }
    *out = gr;
    return N2H_OK;
}

int n2h_graph_launch(n2h_graph* gr, void* stream) {
    if (!gr) return fail(N2H_EINVAL, "null graph");
    if (gr->batch) {
        int rc = batch_in_step(gr->batch);
        if (rc) return rc;
    }
    const int64_t at = gr->batch ? gr->batch->host_step : gr->plan->host_step;
    if ((at & (kListSlots - 1)) != gr->phase)
        return fail(N2H_EINVAL, "graph captured at step phase %d (mod %d), launched at step %lld", gr->phase,
                    kListSlots, static_cast<long long>(at));
    N2H_HIP(hipGraphLaunch(gr->exec, static_cast<hipStream_t>(stream)));
    if (gr->batch) {
        gr->batch->host_step += gr->n_steps;
        for (int32_t i = 0; i < gr->batch->n_plans; ++i) gr->batch->plans[i]->host_step += gr->n_steps;
    } else {
        gr->plan->host_step += gr->n_steps;
    }
    return N2H_OK;
}

int n2h_batch_create(n2h_plan* const* plans, int32_t n_plans, const n2h_tuning* tuning, n2h_batch** out) {
    if (!plans || !out) return fail(N2H_EINVAL, "null argument");
    *out = nullptr;
    if (n_plans < 1 || n_plans > kMaxBatch) return fail(N2H_EINVAL, "n_plans %d outside [1, %d]", n_plans, kMaxBatch);
    for (int32_t i = 0; i < n_plans; ++i)
        if (plans[i] && plans[i]->log_rec) return fail(N2H_EINVAL, "plan %d logs its deliveries: a logging plan steps alone", i);
    for (int32_t i = 0; i < n_plans; ++i)
        if (plans[i] && plans[i]->edges) return fail(N2H_EINVAL, "plan %d has edge kinds or ports: it steps alone", i);
    n2h_tuning tu{};
    if (tuning) tu = *tuning;
    if (tu.order < 0 || tu.order > 1) return fail(N2H_EINVAL, "tuning.order must be 0 or 1");
    if (tu.xcd < 0 || tu.xcd > 64) return fail(N2H_EINVAL, "tuning.xcd must be in [0, 64]");
    if (tu.stage < 0 || tu.stage > 1) return fail(N2H_EINVAL, "tuning.stage must be 0 or 1");
    if (tu.run < 0 || tu.run > 4096) return fail(N2H_EINVAL, "tuning.run must be in [0, 4096]");
    for (int32_t i = 0; i < n_plans; ++i) {
        int rc = check_ready(plans[i]);
        if (rc) return rc;
        if (plans[i]->host_step != plans[0]->host_step) return fail(N2H_EINVAL, "plans at different steps");
        for (int32_t j = 0; j < i; ++j)
            if (plans[j] == plans[i]) return fail(N2H_EINVAL, "plan %d listed twice", i);
    }
    n2h_batch* bt = new (std::nothrow) n2h_batch();
    if (!bt) return fail(N2H_ENOMEM, "host allocation");
    bt->plans = new (std::nothrow) n2h_plan*[n_plans];
    if (!bt->plans) {
        delete bt;
        return fail(N2H_ENOMEM, "host allocation");
    }
    for (int32_t i = 0; i < n_plans; ++i) bt->plans[i] = plans[i];
    if (tu.order == 1) {  // the most links first: the heaviest tiles start first, light ones fill the tail
        for (int32_t i = 1; i < n_plans; ++i)
            for (int32_t j = i; j > 0 && bt->plans[j - 1]->tables.n_links < bt->plans[j]->tables.n_links; --j) {
                n2h_plan* t = bt->plans[j];
                bt->plans[j] = bt->plans[j - 1];
                bt->plans[j - 1] = t;
            }
    }
    bt->n_plans = n_plans;
    bt->model = plans[0]->desc.model;
    for (int32_t i = 1; i < n_plans; ++i)
        if (plans[i]->desc.model != bt->model) bt->model = kModelMixed;  // both models: the mixed kernel
    bt->tuning = tu;
    bt->host_step = plans[0]->host_step;
    int max_groups = 0;
    int64_t tiles = 0;
    bt->lds_bytes = 0;
    bt->first[0] = 0;
    StepArgs* host_args = new (std::nothrow) StepArgs[n_plans];
    int64_t** host_ctrs = new (std::nothrow) int64_t*[n_plans];
    if (!host_args || !host_ctrs) {
        delete[] host_args;
        delete[] host_ctrs;
        n2h_batch_destroy(bt);
        return fail(N2H_ENOMEM, "host allocation");
    }
    for (int32_t i = 0; i < n_plans; ++i) {
        const n2h_plan* p = bt->plans[i];
        host_args[i] = make_args(p);
        host_ctrs[i] = p->step_ctr;
        tiles += p->tables.n_groups;
        bt->first[i + 1] = static_cast<int32_t>(tiles);
        if (p->tables.n_groups > max_groups) max_groups = p->tables.n_groups;
        if (p->lds_bytes > bt->lds_bytes) bt->lds_bytes = p->lds_bytes;
    }
    if (tiles > (1ll << 30)) {
        delete[] host_args;
        delete[] host_ctrs;
        n2h_batch_destroy(bt);
        return fail(N2H_EINVAL, "too many groups in one batch");
    }
    bt->n_tiles = static_cast<int32_t>(tiles);
    bt->words = (max_groups + kWave - 1) / kWave;
    hipError_t e = hipMalloc(&bt->args_dev, sizeof(StepArgs) * n_plans);
    if (e == hipSuccess) e = hipMalloc(&bt->ctrs_dev, sizeof(int64_t*) * n_plans);
    if (e == hipSuccess) e = hipMalloc(&bt->step_ctr, sizeof(int64_t));
    if (e == hipSuccess) e = hipMemcpy(bt->args_dev, host_args, sizeof(StepArgs) * n_plans, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(bt->ctrs_dev, host_ctrs, sizeof(int64_t*) * n_plans, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(bt->step_ctr, &bt->host_step, sizeof(int64_t), hipMemcpyHostToDevice);
    delete[] host_args;
    delete[] host_ctrs;
    if (e != hipSuccess) {
        n2h_batch_destroy(bt);
        return fail(N2H_EHIP, "batch setup: %s", hipGetErrorString(e));
    }
    *out = bt;
    return N2H_OK;
}

int n2h_batch_destroy(n2h_batch* bt) {
    if (!bt) return N2H_OK;
    if (bt->args_dev) (void)hipFree(bt->args_dev);
    if (bt->ctrs_dev) (void)hipFree(bt->ctrs_dev);
    if (bt->step_ctr) (void)hipFree(bt->step_ctr);
    delete[] bt->plans;
    delete bt;
    return N2H_OK;
}

int n2h_batch_sync(n2h_batch* bt, void* stream) {
    if (!bt) return fail(N2H_EINVAL, "null batch");
    const int64_t at = bt->plans[0]->host_step;
    size_t lds = 0;
    for (int32_t i = 0; i < bt->n_plans; ++i) {
        const n2h_plan* p = bt->plans[i];
        if (p->host_step != at) return fail(N2H_ESTATE, "plan %d is at step %lld, plan 0 at %lld", i,
                                            static_cast<long long>(p->host_step), static_cast<long long>(at));
        if (p->log_rec) return fail(N2H_EINVAL, "plan %d logs its deliveries: a logging plan steps alone", i);
        if (p->edges) return fail(N2H_EINVAL, "plan %d has edge kinds or ports: it steps alone", i);
        if (p->lds_bytes > lds) lds = p->lds_bytes;
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    for (int32_t i0 = 0; i0 < bt->n_plans; i0 += kSyncChunk) {
        SyncChunk c;
        memset(&c, 0, sizeof(c));
        c.first = i0;
        c.n = bt->n_plans - i0 < kSyncChunk ? bt->n_plans - i0 : kSyncChunk;
        for (int32_t j = 0; j < c.n; ++j) c.args[j] = make_args(bt->plans[i0 + j]);
        c.step_ctr = i0 == 0 ? bt->step_ctr : nullptr;
        c.step = at;
        hipLaunchKernelGGL(batch_sync_kernel, dim3(1), dim3(kSyncChunk), 0, st, bt->args_dev, c);
    }
    N2H_HIP(hipGetLastError());
    bt->host_step = at;
    bt->lds_bytes = lds;
    return N2H_OK;
}

int n2h_batch_step(n2h_batch* bt, int32_t n_steps, void* stream) {
    if (!bt) return fail(N2H_EINVAL, "null batch");
    if (n_steps < 0) return fail(N2H_EINVAL, "negative n_steps");
    int rc = batch_in_step(bt);
    if (rc) return rc;
    if (n_steps == 0) return N2H_OK;
    rc = launch_batch(bt, n_steps, static_cast<hipStream_t>(stream), false);
    if (rc) return rc;
    for (int32_t i = 0; i < bt->n_plans; ++i) bt->plans[i]->host_step += n_steps;
    return N2H_OK;
}

int n2h_batch_graph_create(n2h_batch* bt, int32_t n_steps, void* stream, n2h_graph** out) {
    (void)stream;
    if (!bt || !out) return fail(N2H_EINVAL, "null argument");
    *out = nullptr;
    if (n_steps <= 0) return fail(N2H_EINVAL, "n_steps must be positive");
    int rc = batch_in_step(bt);
    if (rc) return rc;
    const int64_t at = bt->host_step;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    rc = capture([&](hipStream_t cap) { return launch_batch(bt, n_steps, cap, true); }, &graph, &exec);
    bt->host_step = at;  // captured, not run
    if (rc) return rc;
    n2h_graph* gr = new (std::nothrow)
        n2h_graph{graph, exec, nullptr, bt, n_steps, static_cast<int32_t>(at & (kListSlots - 1))};
    if (!gr) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        return fail(N2H_ENOMEM, "host allocation");
    }
    *out = gr;
    return N2H_OK;
}

int n2h_graph_destroy(n2h_graph* gr) {
    if (!gr) return N2H_OK;
    (void)hipGraphExecDestroy(gr->exec);
    (void)hipGraphDestroy(gr->graph);
    delete gr;
    return N2H_OK;
}

int n2h_bind_rows(n2h_plan* p, const uint32_t* rows, const int64_t* row_base, const int32_t* row_depth,
                  int32_t format, const float* wtab, int32_t dense_min_spikes) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (rows && (!row_base || !row_depth)) return fail(N2H_EINVAL, "null row table");
    if (rows && p->tables.n_groups > kDenseWords - 1)
        return fail(N2H_EINVAL, "row intake holds the spike bitmap in LDS: at most %d groups", kDenseWords - 1);
    if (rows && format != N2H_ROWS_WIDE && format != N2H_ROWS_COMPACT && format != N2H_ROWS_SLOT16)
        return fail(N2H_EINVAL, "unknown row format %d", format);
    if (rows && format == N2H_ROWS_SLOT16) {
        if (!wtab) return fail(N2H_EINVAL, "slot16 rows need their weights and split (wtab[3])");
        if (p->tables.uniform_delay <= 0) return fail(N2H_EINVAL, "slot16 rows need one delay for every link");
        if (static_cast<int64_t>(kWave) * p->tables.n_groups > 0xffff)
            return fail(N2H_EINVAL, "slot16 rows name at most 65,535 source slots (the sentinel included)");
    }
    if (rows && format == N2H_ROWS_COMPACT && !wtab) return fail(N2H_EINVAL, "compact rows need a weight table");
    if (rows && format == N2H_ROWS_COMPACT && static_cast<int64_t>(kWave) * p->tables.n_groups > (1 << 17))
        return fail(N2H_EINVAL, "compact rows name at most 2^17 source slots");
    if (dense_min_spikes < 0) return fail(N2H_EINVAL, "negative dense_min_spikes");
    p->rows = reinterpret_cast<const uint4*>(rows);
    p->row_base = row_base;
    p->row_depth = row_depth;
    p->row_format = format;
    p->row_wtab = wtab;
    p->dense_min = dense_min_spikes;
    return N2H_OK;
}

int n2h_set_vtrace(n2h_plan* p, float* buffer, int64_t step0, int32_t n_steps, float* v_lo, float* i_lo) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (buffer && n_steps <= 0) return fail(N2H_EINVAL, "n_steps must be positive");
    if ((buffer || v_lo) && p->multirate)
        return fail(N2H_EINVAL, "potential traces need every population updating every step (stride 1)");
    if ((v_lo == nullptr) != (i_lo == nullptr)) return fail(N2H_EINVAL, "v_lo and i_lo go together");
    if (v_lo && p->desc.model != N2H_MODEL_LIF_EXP)
        return fail(N2H_EINVAL, "the compensated state is tracked for the exponential-synapse model only");
    if (v_lo && p->edges) return fail(N2H_EINVAL, "edge plans take no compensated trace");
    p->vtrace = buffer;
    p->vtrace_step0 = step0;
    p->vtrace_steps = buffer ? n_steps : 0;
    p->vlo = v_lo;
    p->ilo = i_lo;
    return N2H_OK;
}

int n2h_bind_log(n2h_plan* p, const int16_t* link_conn, int32_t* records, int32_t* count, int64_t capacity) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (records && (!count || capacity < 0)) return fail(N2H_EINVAL, "a log needs its count and a capacity >= 0");
    if (records && p->tables.n_links > static_cast<int64_t>(UINT32_MAX))
        return fail(N2H_EINVAL, "logging plans hold at most 2^32 - 1 links (%lld)", static_cast<long long>(p->tables.n_links));
    p->log_rec = records;
    p->log_count = records ? count : nullptr;
    p->log_conn = records ? link_conn : nullptr;
    p->log_cap = records ? capacity : 0;
    return N2H_OK;
}

int n2h_bind_edges(n2h_plan* p, const float* edge_attr, float* ring_v, int32_t v_slots) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (v_slots < 0 || v_slots > p->desc.d_slots) return fail(N2H_EINVAL, "v_slots %d outside [0, d_slots]", v_slots);
    if ((v_slots > 0) != (ring_v != nullptr)) return fail(N2H_EINVAL, "ring_v and v_slots > 0 go together");
    const size_t cell = sizeof(float) * kWave;
    if (!edge_attr && !ring_v) {
        p->edges = false;
        p->edge_attr = nullptr;
        p->ring_v = nullptr;
        p->v_slots = 0;
        p->lds_bytes = cell * static_cast<size_t>(p->tables.uniform_delay > 0 ? 1 : p->desc.d_slots);
        return N2H_OK;
    }
    if (p->desc.model != N2H_MODEL_LIF_EXP) return fail(N2H_EINVAL, "edge kinds and ports need the exponential-synapse model");
    if (p->tables.uniform_delay > 0) return fail(N2H_EINVAL, "edge plans keep every delay cell (uniform_delay 0)");
    if (p->multirate) return fail(N2H_EINVAL, "edge plans update every population every step");
    if (p->vlo) return fail(N2H_EINVAL, "edge plans take no compensated trace");
    if (p->tables.n_links > static_cast<int64_t>(UINT32_MAX))
        return fail(N2H_EINVAL, "edge plans hold at most 2^32 - 1 links (%lld)", static_cast<long long>(p->tables.n_links));
    p->edges = true;
    p->edge_attr = edge_attr;
    p->ring_v = ring_v;
    p->v_slots = v_slots;
    p->lds_bytes = cell * static_cast<size_t>(p->desc.d_slots + v_slots);
    return N2H_OK;
}

int n2h_set_status(n2h_plan* p, int32_t* rows) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    p->status = rows;
    return N2H_OK;
}

int n2h_set_tuning(n2h_plan* p, const n2h_tuning* t) {
    if (!p || !t) return fail(N2H_EINVAL, "null argument");
    if (t->xcd < 0 || t->xcd > 64) return fail(N2H_EINVAL, "tuning.xcd must be in [0, 64]");
    if (t->run < 0 || t->run > 4096) return fail(N2H_EINVAL, "tuning.run must be in [0, 4096]");
    p->xcd = t->xcd;
    p->xcd_run = t->run;
    return N2H_OK;
}

int n2h_set_trace(n2h_plan* p, uint64_t* buffer) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    p->trace = reinterpret_cast<unsigned long long*>(buffer);
    return N2H_OK;
}

int n2h_status(n2h_plan* p, int32_t* info, int32_t n_info) {
    if (!p || !info) return fail(N2H_EINVAL, "null argument");
    const int32_t vals[4] = {p->tables.n_groups, p->desc.d_slots, static_cast<int32_t>(p->lds_bytes),
                             p->desc.model};
    for (int32_t i = 0; i < n_info && i < 4; ++i) info[i] = vals[i];
    return N2H_OK;
}

}  // extern "C"
