// Surrogate-gradient adjoint of the exponential-synapse LIF step (the LIF_EXP model of naer2_hip.hip) for a
// per-population rate loss, batched over networks: one launch per reverse step (adj_step_kernel: short links
// and every neuron's update), and per window of long links a clear, a mark, the long-link gather and the
// combine of its target ranges (or, for range counts that are not a multiple of 8, one gather launch a range).
//
// The reference is forward-only; this adjoint is added beside its plain path.  Forward step t of neuron i (naer2_hip.hip, step_group):
//     I'_t = I_{t-1} + x_t + ext_t
//     V^pre_t = V_rest + P_V (V_{t-1} - V_rest) + P_VI I'_t   (r_t = 1), or V_reset while refractory (r_t = 0)
//     I_t = P_I I'_t;  s_t = [V^pre_t > theta];  V_t = V^pre_t (1 - s_t) + V_reset s_t
//     x_t[k] = sum over links j -> k of alpha[pop k][pop j] w s_{t-d}[j]
// Surrogate: ds_t/dV^pre_t = r_t h(V^pre_t), h(v) = gamma (kappa / 2) / dv / (1 + kappa |u|)^2 for |u| <= u_c
// and 0 beyond, u = (v - theta) / dv, dv = theta - V_reset: the derivative of the smooth spike
// 1/2 + gamma (kappa/2) u' / (1 + kappa |u'|), u' = u clamped to [-u_c, u_c].  The cutoff makes the
// gradient sparse (Perez-Nieves & Goodman, NeurIPS 2021): a neuron farther than u_c dv from the threshold
// that did not spike skips its out-link gather.  The refractory gate r_t and the reset carry no gradient (through the hard reset the
// adjoint of V grows by ~h (V^pre - V_reset) ~ kappa / 2 at every spike and diverges; detaching it is the
// usual remedy, Zenke & Vogels 2021).  Reverse step t, given the adjoints gV, gI of step t's outputs and gx
// of the inputs of steps t + 1 .. t + D:
//     gs     = c[pop i] + sum over out-links i -> k of alpha w gx_{t+d}[k]
//     gVpre  = (1 - s_t) gV + r_t h(V^pre_t) gs
//     gI'    = r_t P_VI gVpre + P_I gI;   gV <- r_t P_V gVpre;   gI <- gI';   gx_t[i] = gI'
//     dL/dalpha[a][pop i] += s_t[i] * sum over out-links i -> k in pop a of w gx_{t+d}[k]
// (accumulated per source in compensated fp32: a hi + lo pair, every addition's exact error kept in lo).
// The surrogate takes the traced distance to the threshold, V^pre - theta, which the recompute forward
// computes compensated (n2h_set_vtrace with v_lo / i_lo): the potential as exact arithmetic on the fp32
// inputs would give it, to ~2^-44, rather than the fp32 potential whose rounding the surrogate's slope at
// the threshold (kappa / dv per unit of V) would amplify.
// A wavefront gathers the out-links of one neuron that needs them (lanes stride the list; per-lane sums per
// target population, then a fixed butterfly): every sum has a fixed order, so results are deterministic.
//
// Windows.  The out-links are split by delay: short (d < W) and long (d >= W), W = long_delay.  For a window
// of reverse steps [t0, t0 + W) every long link's input adjoints gx_{t0+w+d} (w < W) lie past the window, so
// adj_long_kernel reads each gathering source's long links once per window and gathers a link's W values
// in one contiguous read: the ring is target-major, gx[k][slot], step s in slot s mod R, with the first W - 1
// slots mirrored past slot R - 1 so that W consecutive steps never wrap.  The long links are walked in
// long_parts target ranges so that the L2s hold the rows being read (adj_long_pipek_kernel: all ranges in one
// launch over the XCDs, adj_combine_kernel adding their sums in order; adj_long_kernel: one launch a
// range).  Each step, adj_step_kernel walks only the short links, adds the window's long sums (bins[a] =
// short sum + long sum) and updates every neuron.  The split depends on the link alone, so the sums do not
// depend on where windows, checkpoints or batches start.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "naer2_hip.h"
#include "twofloat.h"

namespace {

constexpr int kWave = 64;
constexpr int kAdjPops = 8;      // target-population bins per lane
constexpr int kAdjNets = 16;     // networks in one launch
constexpr int kWavesPerBlock = 4;
constexpr int kMaxWindow = 8;
constexpr int kMaxParts = 64;      // target ranges of the long links
constexpr int kUnroll = 4;        // links a lane loads per pass of a gather loop

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

struct AdjBatch {
    n2h_adj_net net[kAdjNets];
    int32_t first_block[kAdjNets + 1];
    int32_t n_nets;
    int32_t ring_pos;  // step (window: its first step) mod ring_slots
    int32_t n_win;     // window: its steps
    int32_t win_step;  // step: its place in the window
    int32_t recent_pos;  // step mod long_delay: this step's slot of the short ring
    int32_t sub_cap;   // entries a work sub-list holds
    int32_t part;      // long kernel, one range a launch: the target range of this launch
    float* partial;    // long kernel, ranges split over the XCDs: [window entry][range][8 W'] sums
    int32_t* work;     // counters and work lists
    n2h_surrogate sur;
};

// The network of this block and the first neuron of this wavefront: a block covers kWavesPerBlock * 64
// consecutive neurons of one network, a lane one neuron.
__device__ inline int adj_wave_base_at(const AdjBatch& b, int blk, int& q) {
    q = 0;
    while (q + 1 < b.n_nets && b.first_block[q + 1] <= blk) ++q;
    return ((blk - b.first_block[q]) * kWavesPerBlock + static_cast<int>(threadIdx.x >> 6)) * kWave;
}
__device__ inline int adj_wave_base(const AdjBatch& b, int& q) { return adj_wave_base_at(b, blockIdx.x, q); }

// j's population: the count of population starts at or below j, the starts loaded all at once (past the last
// population the end, n > j) rather than one after another
__device__ inline int adj_pop(const n2h_adj_net& a, int j) {
    int pj = 0;
#pragma unroll
    for (int p = 1; p < kAdjPops; ++p) pj += a.pop_first[min(p, a.n_pops)] <= j ? 1 : 0;
    return pj;
}

// |V - theta| / dv from the traced distance to the threshold V - theta (n2h_set_vtrace): +inf on refractory
// steps (the trace holds +inf there)
__device__ inline float adj_u(float dist, float dv) { return fabsf(dist) / dv; }

// A link record, read once per window: a non-temporal load, so that the stream of records does not evict
// the L2's ring rows of the range being walked, which every gatherer's links read again.
__device__ __forceinline__ uint2 adj_load_link(const uint2* __restrict__ p) {
    typedef uint32_t u2v __attribute__((ext_vector_type(2)));
    const u2v v = __builtin_nontemporal_load(reinterpret_cast<const u2v*>(p));
    return make_uint2(v[0], v[1]);
}

// W consecutive floats from a dword-aligned address: 16-byte vector loads (gfx950 takes dword-aligned
// global_load_dwordx4) for each full four, single loads for the rest -- one request per link and four
// adjoints instead of four, which the texture path serves one lane address at a time
template <int W>
__device__ __forceinline__ void adj_load_window(const float* __restrict__ g, float (&out)[W]) {
    typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
#pragma unroll
    for (int w = 0; w + 4 <= W; w += 4) {
        const f4u v = *reinterpret_cast<const f4u*>(g + w);
        out[w] = v[0];
        out[w + 1] = v[1];
        out[w + 2] = v[2];
        out[w + 3] = v[3];
    }
#pragma unroll
    for (int w = W & ~3; w < W; ++w) out[w] = g[w];
}

// one link {target | (d - 1) << 17 | target pop << 24, base weight}: the ring offset of gx_{t+d}[target]
// (pos = t mod R; d < R) and the target population
__device__ inline int64_t adj_link_slot(uint32_t x, int pos, int R, int RS, uint32_t& tp) {
    const uint32_t tgt = x & 0x1ffffu;
    const int d = static_cast<int>((x >> 17) & 127u) + 1;
    tp = x >> 24;
    int slot = pos + d;
    slot -= (slot >= R) ? R : 0;
    return static_cast<int64_t>(tgt) * RS + slot;
}

// Sums v[0..V) (V a power of 2, 1 to 64) over the 64 lanes of the wavefront by recursive halving, a fixed
// tree: at the stage of lane offset o, while more than one element is left, a lane keeps one half of its
// elements (the upper half if its lane bit o is set) and adds its partner's copy of that half.  Lane l ends
// with element l >> (6 - log2 V) in v[0]; 2 V - 2 + log2 (64 / V) lane exchanges instead of a butterfly's 6 V.
template <int L, int O, int V>
__device__ inline void adj_halve(float (&v)[V], int lane) {  // L elements left, lane offset O
    if constexpr (O > 0) {
        if constexpr (L > 1) {
            constexpr int H = L / 2;
            const bool upper = (lane & O) != 0;
#pragma unroll
            for (int i = 0; i < H; ++i) {
                const float send = upper ? v[i] : v[i + H];
                const float keep = upper ? v[i + H] : v[i];
                v[i] = keep + __shfl_xor(send, O, kWave);
            }
            adj_halve<H, O / 2, V>(v, lane);
        } else {
            v[0] += __shfl_xor(v[0], O, kWave);
            adj_halve<1, O / 2, V>(v, lane);
        }
    }
}
template <int V>
__device__ inline float adj_reduce_halving(float (&v)[V], int lane) {
    adj_halve<V, kWave / 2, V>(v, lane);
    return v[0];
}
template <int V>
constexpr int adj_halving_shift() {  // 6 - log2 V: lane l ends with element l >> shift
    return V == 64 ? 0 : V == 32 ? 1 : V == 16 ? 2 : V == 8 ? 3 : V == 4 ? 4 : V == 2 ? 5 : 6;
}

// Work lists.  Few neurons gather (a few percent a step), and they cluster in the more active populations,
// so a lane-per-neuron wavefront walking its own gatherers in turn leaves most wavefronts idle behind a few.
// Instead adj_mark_kernel appends every gatherer to a list (the window's, entry mask << 21 | net << 17 |
// neuron, and each window step's, net << 17 | neuron), and the gather kernels give each list entry a whole
// wavefront in grid-stride order.  List order depends on the schedule, but every entry's sums are its own,
// in a fixed order, so results do not.  Each list (which = 0: the window's, 1 + w: step w's) is kept as
// kSubLists sub-lists, each counter on a cache line of its own: appends to one counter serialize in the L2.
// Every neuron has an XCD of its own: network q's neurons are dealt, by mark workgroup, over the XCDs of
// network q (with at most 8 networks a contiguous share of the 8, else XCD q mod 8).  The wavefronts of a
// mark workgroup append to sub-lists x and 8 + x of its XCD x (by their index's parity), and the step
// kernel's workgroups on XCD x walk sub-lists x and 8 + x.  A neuron is then gathered on the same XCD in
// every step, so that its short links, read again in the steps it stays near the threshold, are in that
// XCD's L2, and an XCD's gatherers read the short ring of its own networks only (one of 4 PD14s: 0.9 MB
// of adjoints a step instead of 3.7).  (Workgroup b runs on XCD b mod 8: the hardware deals a launch's
// workgroups to the 8 XCDs in turn.)
constexpr int kLongBlocks = 1536;    // 6 blocks of 4 wavefronts per CU (82 VGPRs: 6 wavefronts per SIMD fit)
constexpr int kShortBlocks = 2048;   // 8 blocks of 4 wavefronts per CU
constexpr int kSplitBlocks = 1280;   // the split long kernel: 5 blocks per CU (93 VGPRs)
constexpr int kPipeBlocks = 768;     // its pipelined form: 3 blocks per CU (108 VGPRs at W = 4 would fit
                                     // 4; with 3 fewer items' ring rows share each L2, measured faster)
constexpr int kPipeBins = 4;         // its population bins a lane per window step
constexpr int kGroupLanes = 8;       // the population-pure long kernel: lanes a (gatherer, range) task
constexpr int kGroupBlocks = 1536;   // its workgroups: 6 blocks per CU (74 VGPRs at W = 4: 6 wavefronts per SIMD)
constexpr int kXcds = 8;
static_assert(kSplitBlocks % kXcds == 0 && kGroupBlocks % kXcds == 0 && kPipeBlocks % kXcds == 0 &&
                  kShortBlocks % kXcds == 0,
              "the split long kernels and the step kernel deal workgroups to the XCDs in turn");
constexpr int kSubLists = 2 * kXcds;
constexpr int kCounterStride = 32;   // int32 words: 128 B
constexpr int kAdjListHead = (1 + kMaxWindow) * kSubLists * kCounterStride;
// network q's XCDs: the first, and their count
__host__ __device__ inline int adj_net_xcd0(int q, int n_nets) { return n_nets <= kXcds ? q * kXcds / n_nets : q % kXcds; }
__host__ __device__ inline int adj_net_xcds(int q, int n_nets) {
    return n_nets <= kXcds ? (q + 1) * kXcds / n_nets - q * kXcds / n_nets : 1;
}
// entries a sub-list holds: sub-list c takes the appends of the wavefronts of parity c / kXcds (two a
// workgroup) of the mark workgroups dealt to XCD c mod kXcds (workgroup bq of network q to XCD
// adj_net_xcd0(q) + bq mod adj_net_xcds(q)), at most 64 each: sized for the XCD dealt the most
static_assert(kWavesPerBlock % 2 == 0, "mark workgroups append to two sub-lists by wavefront parity");
inline int64_t adj_sub_cap(const int32_t* n_neurons, int n_nets) {
    int64_t per[kXcds] = {};
    for (int q = 0; q < n_nets; ++q) {
        const int64_t nb = (n_neurons[q] + kWave * kWavesPerBlock - 1) / (kWave * kWavesPerBlock);
        const int x0 = adj_net_xcd0(q, n_nets), m = adj_net_xcds(q, n_nets);
        for (int k = 0; k < m; ++k) per[x0 + k] += (nb + m - 1) / m;
    }
    int64_t most = 0;
    for (int x = 0; x < kXcds; ++x) most = per[x] > most ? per[x] : most;
    return static_cast<int64_t>(kWave) * (kWavesPerBlock / 2) * most;
}

__device__ inline int64_t adj_counter(int which, int c) {
    return static_cast<int64_t>(which * kSubLists + c) * kCounterStride;
}
__device__ inline int64_t adj_sublist(const AdjBatch& b, int which, int c) {
    return kAdjListHead + static_cast<int64_t>(which * kSubLists + c) * b.sub_cap;
}

// entry i of list `which` over its sub-lists in order (i < the list's total)
struct AdjList {
    int cnt[kSubLists];
    int total;
};
__device__ inline AdjList adj_list_counts(const AdjBatch& b, int which) {
    AdjList l;
    l.total = 0;
#pragma unroll
    for (int c = 0; c < kSubLists; ++c) {
        l.cnt[c] = b.work[adj_counter(which, c)];
        l.total += l.cnt[c];
    }
    return l;
}
// (per lane: each lane's own entry i; adj_list_entry below takes a wavefront-uniform i)
__device__ inline uint32_t adj_list_entry_lane(const AdjBatch& b, const AdjList& l, int which, int i) {
    int k = i, sel = kSubLists - 1;
    bool found = false;
#pragma unroll
    for (int c = 0; c < kSubLists; ++c) {
        if (!found) {
            if (k < l.cnt[c]) {
                sel = c;
                found = true;
            } else {
                k -= l.cnt[c];
            }
        }
    }
    return static_cast<uint32_t>(b.work[adj_sublist(b, which, sel) + k]);
}
__device__ inline uint32_t adj_list_entry(const AdjBatch& b, const AdjList& l, int which, int i) {
    int k = i, sel = kSubLists - 1;
    bool found = false;
#pragma unroll
    for (int c = 0; c < kSubLists; ++c) {
        if (!found) {
            if (k < l.cnt[c]) {
                sel = c;
                found = true;
            } else {
                k -= l.cnt[c];
            }
        }
    }
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(b.work[adj_sublist(b, which, sel) + k]));
}

// one short link {target | (d - 1) << 17 | ...}, d < long_delay: the offset of gx_{t+d}[target] in the
// short ring [long_delay][n] (pos = t mod long_delay; clamped into the ring whatever the link holds)
__device__ inline int64_t adj_recent_slot(uint32_t x, int pos, int W, int n, uint32_t& tp) {
    const uint32_t tgt = x & 0x1ffffu;
    const int d = static_cast<int>((x >> 17) & 127u) + 1;
    tp = x >> 24;
    int slot = pos + d;
    slot -= (slot >= W) ? W : 0;
    slot = slot < W ? slot : W - 1;
    return static_cast<int64_t>(slot) * n + tgt;
}

// lane per neuron: its window mask (steps where it is within the support or spiked), appended to the lists
// the list counters to zero before a window's mark kernel (a kernel rather than hipMemsetAsync, so that a
// caller may capture the reverse pass into a graph)
__global__ __launch_bounds__(256) void adj_clear_kernel(int32_t* __restrict__ work) {
    for (int i = threadIdx.x; i < kAdjListHead; i += blockDim.x) work[i] = 0;
}

__global__ __launch_bounds__(kWave * kWavesPerBlock) void adj_mark_kernel(AdjBatch b) {
    int q;
    const int base = adj_wave_base(b, q);
    const n2h_adj_net& a = b.net[q];
    const int n = a.n_neurons;
    if (base >= n) return;  // uniform over the wavefront
    const int lane = threadIdx.x & (kWave - 1);
    const int j = base + lane;
    uint32_t mask = 0;
    if (j < n) {
        const float* C = a.pop_coef + adj_pop(a, j) * 8;
        const float dv = C[5];
        // the window's traces and spike flags loaded all at once (unrolled to the longest window)
        uint8_t s[kMaxWindow];
        float vt[kMaxWindow];
#pragma unroll
        for (int w = 0; w < kMaxWindow; ++w) {
            s[w] = w < b.n_win ? a.spikes[static_cast<int64_t>(w) * n + j] : uint8_t{0};
            vt[w] = w < b.n_win ? a.vtrace[static_cast<int64_t>(w) * n + j] : __builtin_inff();
        }
#pragma unroll
        for (int w = 0; w < kMaxWindow; ++w) {
            const bool near = adj_u(vt[w], dv) <= b.sur.cutoff;
            mask |= (w < b.n_win && (s[w] != 0 || near)) ? (1u << w) : 0u;
        }
    }
    const uint64_t below = (1ull << lane) - 1ull;
    const int bq = static_cast<int>(blockIdx.x) - b.first_block[q];  // this workgroup's XCD among its network's
    const int x = adj_net_xcd0(q, b.n_nets) + bq % adj_net_xcds(q, b.n_nets);
    const int sub = x + kXcds * static_cast<int>((threadIdx.x >> 6) & 1u);
    // every list's ballot, then the entries of all lists reserved at once: lane `which` adds list which's
    // count to its counter (one returning atomic a lane, one round trip for all lists), then the entries
    uint64_t bal[1 + kMaxWindow];
    int cnt_l = 0;  // lane which: list which's entries from this wavefront
#pragma unroll
    for (int which = 0; which <= kMaxWindow; ++which) {
        const bool in = which <= b.n_win && (which == 0 ? mask != 0 : ((mask >> (which - 1)) & 1u) != 0);
        bal[which] = __ballot(in);
        cnt_l = lane == which ? __popcll(bal[which]) : cnt_l;
    }
    int first_l = 0;
    if (lane <= b.n_win && cnt_l > 0) first_l = atomicAdd(b.work + adj_counter(lane, sub), cnt_l);
#pragma unroll
    for (int which = 0; which <= kMaxWindow; ++which) {
        if (bal[which] == 0) continue;  // uniform (and every list past the window's steps)
        const int first = __shfl(first_l, which, kWave);
        const uint32_t ent = (which == 0 ? mask << 21 : 0u) | (static_cast<uint32_t>(q) << 17) | static_cast<uint32_t>(j);
        if ((bal[which] >> lane) & 1ull)
            b.work[adj_sublist(b, which, sub) + first + __popcll(bal[which] & below)] = static_cast<int32_t>(ent);
    }
}

template <int W>
constexpr int adj_win_pad() { return W <= 1 ? 1 : W <= 2 ? 2 : W <= 4 ? 4 : 8; }  // W padded to a power of 2

// The window's gatherers' long links into one target range (b.part), for range counts that are not a
// multiple of 8 (one launch a range): a wavefront a gatherer, a link's W input adjoints in one contiguous
// read; lanes stride the list with per-lane sums per (target population, window step), then a fixed
// reduction; each launch adds its sums to the long_bins rows of the steps in the entry's mask, in launch
// (range) order.  Range counts that are a multiple of 8 take one launch for all ranges instead
// (adj_long_pipek_kernel, or adj_long_group_kernel for population-pure ranges): workgroup b runs on XCD
// b mod kXcds (the hardware deals workgroups to the XCDs in turn), whose wavefronts walk ranges x,
// x + kXcds, ... in that order, so each XCD's L2 holds the ring rows of one range at a time and no two L2s
// hold the same rows; the range sums go to `partial`, which adj_combine_kernel adds in range order (the
// same sums, in the same order, as here).
template <int W>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void adj_long_kernel(AdjBatch b) {
    constexpr int WP = adj_win_pad<W>();
    constexpr int V = kAdjPops * WP;
    constexpr int SH = adj_halving_shift<V>();
    const int lane = threadIdx.x & (kWave - 1);
    const AdjList list = adj_list_counts(b, 0);
    const int idx = lane >> SH;  // the element this lane ends with: w * 8 + p
    const int ww = idx >> 3, pp = idx & 7;
    const bool writer = (lane & ((1 << SH) - 1)) == 0 && ww < W;
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int part = b.part;
    for (int i = static_cast<int>(blockIdx.x) * kWavesPerBlock + wave; i < list.total;
         i += static_cast<int>(gridDim.x) * kWavesPerBlock) {
        const uint32_t ent = adj_list_entry(b, list, 0, i);
        const int q = static_cast<int>((ent >> 17) & 15u), jj = static_cast<int>(ent & 0x1ffffu);
        const uint32_t mk = ent >> 21;
        const n2h_adj_net& a = b.net[q];
        const int n = a.n_neurons, R = a.ring_slots, RS = R + W - 1;
        float bins[V];
#pragma unroll
        for (int k = 0; k < V; ++k) bins[k] = 0.0f;
        const int64_t seg = static_cast<int64_t>(jj) * a.long_parts + part;
        const int64_t e0 = a.long_ptr[seg], e1 = a.long_ptr[seg + 1];
        // kUnroll links a lane per pass, every load issued before the sums; the lane's links are added in
        // list order (e, e + 64, ...) as in a plain loop
        const uint2* links = reinterpret_cast<const uint2*>(a.long_links);
        for (int64_t e = e0 + lane; e < e1; e += kUnroll * kWave) {
            uint2 l[kUnroll];
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) {
                const int64_t er = e + r * kWave;
                l[r] = er < e1 ? links[er] : make_uint2(0u, 0u);  // padding: target 0, delay 1, weight +0
            }
            float gw[kUnroll][W];
            uint32_t tp[kUnroll];
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) adj_load_window<W>(a.gx_ring + adj_link_slot(l[r].x, b.ring_pos, R, RS, tp[r]), gw[r]);
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) {
                const bool live = e + r * kWave < e1;
                const float wt = __uint_as_float(l[r].y);
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    const float c = wt * gw[r][w];
#pragma unroll
                    for (int p = 0; p < kAdjPops; ++p)
                        bins[w * kAdjPops + p] = (live && tp[r] == static_cast<uint32_t>(p)) ? bins[w * kAdjPops + p] + c
                                                                                              : bins[w * kAdjPops + p];
                }
            }
        }
        const float v = (e1 > e0) ? adj_reduce_halving<V>(bins, lane) : 0.0f;  // uniform branch
        if (writer && ((mk >> ww) & 1u)) {
            // the rows of the steps in the mask: target ranges added in launch order (the first one stores)
            float* row = a.long_bins + (static_cast<int64_t>(ww) * n + jj) * kAdjPops + pp;
            *row = part == 0 ? v : *row + v;
        }
    }
}

// The split long kernel (all ranges in one launch over the XCDs, equal ranges), with its items pipelined in
// each wavefront: the wavefront's next 64 items' list entries, run bounds and populations are loaded at once
// (an item a lane, read back uniform), and the first pass of the next item's link records is issued right
// after the current item's ring-row loads, so that its HBM round trip runs under the current item's sums and
// reduction.  A lane keeps K population bins per window step instead of 8: a target range (range k of a
// network: targets t with t * long_parts / n = k) spans few populations, populations being contiguous in the
// neuron order, so an item sums into the bins of populations [g0, g0 + K) of its range's span, one sweep over
// its links for each K populations of the span (one sweep for every range of PD14 at 24 ranges).  A bin's
// sum takes its population's links in link order, and the partial row holds +0 for populations outside the
// range, as the launch-per-range adj_long_kernel's sums are.
template <int W, int K>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void adj_long_pipek_kernel(AdjBatch b) {
    constexpr int WP = adj_win_pad<W>();
    constexpr int V = K * WP;
    constexpr int SH = adj_halving_shift<V>();
    const int lane = threadIdx.x & (kWave - 1);
    const AdjList list = adj_list_counts(b, 0);
    // this lane's slot of a partial row: window step ws, population ps (lanes below 8 W write)
    const int ws = lane >> 3, ps = lane & 7;
    const bool writer = ws < W;
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int xcd = static_cast<int>(blockIdx.x % kXcds);
    const int64_t f0 = static_cast<int64_t>(blockIdx.x / kXcds) * kWavesPerBlock + wave;
    const int64_t fs = static_cast<int64_t>(gridDim.x / kXcds) * kWavesPerBlock;
    const int P = b.net[0].long_parts;
    const int64_t total = list.total;
    const int64_t F = static_cast<int64_t>(P / kXcds) * total;
    // the items of a range are walked network by network: the window list's sub-lists in the order of their
    // XCDs' networks (each network's XCDs' sub-lists together), an entry keeping its place in the whole list
    int ord[kSubLists], ocum[kSubLists + 1], ooff[kSubLists];
    {
        int k = 0;
#pragma unroll
        for (int x = 0; x < kXcds; ++x) {  // XCD x's two sub-lists; XCDs of one network are adjacent
            ord[k++] = x;
            ord[k++] = x + kXcds;
        }
        int acc = 0;
        ocum[0] = 0;
#pragma unroll
        for (int kk = 0; kk < kSubLists; ++kk) {
            int off = 0;
#pragma unroll
            for (int c = 0; c < kSubLists; ++c) off += c < ord[kk] ? list.cnt[c] : 0;
            ooff[kk] = off;
            acc += list.cnt[ord[kk]];
            ocum[kk + 1] = acc;
        }
    }
    for (int64_t fb = f0; fb < F; fb += kWave * fs) {
        // lane l: item fb + l fs -- its list index, range, network, source, run bounds and the range's
        // populations [plo_l, phi_l]
        const int64_t fl = fb + static_cast<int64_t>(lane) * fs;
        int i_l = 0, part_l = 0, q_l = 0, plo_l = 0, phi_l = 0;
        int64_t e0_l = 0, e1_l = 0;
        if (fl < F) {
            const int m = static_cast<int>(fl / total);
            const int e = static_cast<int>(fl - static_cast<int64_t>(m) * total);
            part_l = xcd + kXcds * m;
            int k = 0;
#pragma unroll
            for (int kk = 1; kk < kSubLists; ++kk) k += ocum[kk] <= e ? 1 : 0;
            int sub = 0, cum = 0, off = 0;
#pragma unroll
            for (int kk = 0; kk < kSubLists; ++kk) {
                sub = kk == k ? ord[kk] : sub;
                cum = kk == k ? ocum[kk] : cum;
                off = kk == k ? ooff[kk] : off;
            }
            i_l = off + (e - cum);
            const uint32_t ent = static_cast<uint32_t>(b.work[adj_sublist(b, 0, sub) + (e - cum)]);
            q_l = static_cast<int>((ent >> 17) & 15u);
            const int jj = static_cast<int>(ent & 0x1ffffu);
            const n2h_adj_net& aq = b.net[q_l];
            const int64_t seg = static_cast<int64_t>(jj) * aq.long_parts + part_l;
            e0_l = aq.long_ptr[seg];
            e1_l = aq.long_ptr[seg + 1];
            const int64_t n = aq.n_neurons;
            const int t0 = static_cast<int>((part_l * n + P - 1) / P);
            const int t1 = static_cast<int>(((part_l + 1) * n + P - 1) / P) - 1;  // last target (t1 < t0: none)
#pragma unroll
            for (int p = 1; p < kAdjPops; ++p) {
                if (p < aq.n_pops) {
                    const int first = aq.pop_first[p];
                    plo_l += first <= t0 ? 1 : 0;
                    phi_l += first <= t1 ? 1 : 0;
                }
            }
            phi_l = max(phi_l, plo_l);
        }
        const int n_items = static_cast<int>(min(static_cast<int64_t>(kWave), (F - fb + fs - 1) / fs));
        auto rd64 = [](int64_t v, int k) {
            const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), k));
            const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), k));
            return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
        };
        // item k's first pass of link records, kUnroll a lane, as loaded: the loads are unconditional -- a lane
        // past the run reads the network's first run bound, 8 valid bytes -- so that no branch hides from the
        // compiler how many loads are in flight, and nothing uses them before `pass` (which pads the records
        // past the run) so that they stay in flight meanwhile.
        auto first_pass = [&](int k, uint2 (&l)[kUnroll]) {
            const int q = __builtin_amdgcn_readlane(q_l, k);
            const int64_t e0 = rd64(e0_l, k), e1 = rd64(e1_l, k);
            const uint2* links = reinterpret_cast<const uint2*>(b.net[q].long_links);
            const uint2* valid = reinterpret_cast<const uint2*>(b.net[q].long_ptr);
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) {
                const int64_t er = e0 + lane + r * kWave;
                l[r] = adj_load_link(er < e1 ? links + er : valid);
            }
        };
        uint2 cur[kUnroll], nxt[kUnroll];
        first_pass(0, cur);
        for (int k = 0; k < n_items; ++k) {
            const int q = __builtin_amdgcn_readlane(q_l, k);
            const int i = __builtin_amdgcn_readlane(i_l, k);
            const int part = __builtin_amdgcn_readlane(part_l, k);
            const int plo = __builtin_amdgcn_readlane(plo_l, k), phi = __builtin_amdgcn_readlane(phi_l, k);
            const int64_t e0 = rd64(e0_l, k), e1 = rd64(e1_l, k);
            const n2h_adj_net& a = b.net[q];
            const int R = a.ring_slots, RS = R + W - 1;
            const uint2* links = reinterpret_cast<const uint2*>(a.long_links);
            float out = 0.0f;  // this lane's slot (ws, ps); +0 for populations outside the range
            // one sweep over the item's links, the sums of populations [g0, g0 + K); FIRST: the first sweep,
            // whose first pass takes the records loaded ahead and issues the next item's
            auto sweep = [&](int g0, auto first) {
                float bins[V];
#pragma unroll
                for (int t = 0; t < V; ++t) bins[t] = 0.0f;
                // one pass of kUnroll records a lane: ring rows, then the sums (records at or past e1 become
                // the padding {0, 0}, never added)
                auto pass = [&](int64_t e, const uint2 (&raw)[kUnroll], auto prefetch) {
                    uint2 l[kUnroll];
#pragma unroll
                    for (int r = 0; r < kUnroll; ++r) l[r] = e + r * kWave < e1 ? raw[r] : make_uint2(0u, 0u);
                    float gw[kUnroll][W];
                    uint32_t tp[kUnroll];
#pragma unroll
                    for (int r = 0; r < kUnroll; ++r)
                        adj_load_window<W>(a.gx_ring + adj_link_slot(l[r].x, b.ring_pos, R, RS, tp[r]), gw[r]);
                    // the next item's first pass, issued behind these ring-row loads by every lane and on
                    // every path (the last item loads its own again), so that the compiler counts it in flight
                    if constexpr (decltype(prefetch)::value) {
                        first_pass(min(k + 1, n_items - 1), nxt);
                        __builtin_amdgcn_sched_barrier(0);  // the scheduler would sink these loads below the sums
                    }
#pragma unroll
                    for (int r = 0; r < kUnroll; ++r) {
                        const bool live = e + r * kWave < e1;
                        const uint32_t kk = tp[r] - static_cast<uint32_t>(g0);  // >= K: another sweep's
                        const float wt = __uint_as_float(l[r].y);
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            const float c = wt * gw[r][w];
#pragma unroll
                            for (int p = 0; p < K; ++p)
                                bins[w * K + p] = (live && kk == static_cast<uint32_t>(p)) ? bins[w * K + p] + c
                                                                                          : bins[w * K + p];
                        }
                    }
                };
                if constexpr (decltype(first)::value) {
                    pass(e0 + lane, cur, std::true_type{});
                } else {
                    uint2 l[kUnroll];
#pragma unroll
                    for (int r = 0; r < kUnroll; ++r) {
                        const int64_t er = e0 + lane + r * kWave;
                        l[r] = er < e1 ? adj_load_link(links + er) : make_uint2(0u, 0u);
                    }
                    pass(e0 + lane, l, std::false_type{});
                }
                for (int64_t e = e0 + lane + kUnroll * kWave; e < e1; e += kUnroll * kWave) {
                    uint2 l[kUnroll];
#pragma unroll
                    for (int r = 0; r < kUnroll; ++r) {
                        const int64_t er = e + r * kWave;
                        l[r] = er < e1 ? adj_load_link(links + er) : make_uint2(0u, 0u);
                    }
                    pass(e, l, std::false_type{});
                }
                const float v = adj_reduce_halving<V>(bins, lane);  // lane l: element l >> SH = w K + p
                const int kp = ps - g0;
                const float mine = __shfl(v, ((ws * K + min(max(kp, 0), K - 1)) << SH) & (kWave - 1), kWave);
                if (kp >= 0 && kp < K) out = mine;
            };
            sweep(plo, std::true_type{});
            for (int g0 = plo + K; g0 <= phi; g0 += K) sweep(g0, std::false_type{});  // spans wider than K
            if (writer) b.partial[(static_cast<int64_t>(i) * P + part) * (kAdjPops * WP) + lane] = out;
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) cur[r] = nxt[r];
        }
    }
}

// Population-pure ranges, LANES lanes a task: a wavefront takes 64 / LANES consecutive gatherers of one
// range at once (their list entries, run bounds and links all in flight together), each walked by its lane
// group with the group's lanes striding the run (e, e + LANES, ...), kUnroll links a lane per pass, then a
// fixed butterfly over the group.  The partial row of (entry, range) is [W']: the range's one population
// only, which adj_combine_kernel<W, true> adds over that population's ranges.  The XCD of workgroup b
// (b mod kXcds) walks ranges x, x + kXcds, ... in turn, as in adj_long_pipek_kernel.
template <int W, int LANES>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void adj_long_group_kernel(AdjBatch b) {
    constexpr int WP = adj_win_pad<W>();
    constexpr int G = kWave / LANES;  // tasks a wavefront
    constexpr int U = 2 * kUnroll;    // links a lane per pass
    const int lane = threadIdx.x & (kWave - 1);
    const int sub = lane % LANES, gq = lane / LANES;
    const AdjList list = adj_list_counts(b, 0);
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int xcd = static_cast<int>(blockIdx.x % kXcds);
    const int f0 = static_cast<int>(blockIdx.x / kXcds) * kWavesPerBlock + wave;
    const int fs = static_cast<int>(gridDim.x / kXcds) * kWavesPerBlock;
    const int P = b.net[0].long_parts;
    const int chunks = (list.total + G - 1) / G;  // wavefront tasks per range
    const int64_t F = static_cast<int64_t>(P / kXcds) * chunks;
    for (int64_t f = f0; f < F; f += fs) {
        const int m = static_cast<int>(f / chunks);
        const int i = static_cast<int>(f - static_cast<int64_t>(m) * chunks) * G + gq;
        const int part = xcd + kXcds * m;
        const bool task = i < list.total;
        float bins[WP];
#pragma unroll
        for (int k = 0; k < WP; ++k) bins[k] = 0.0f;
        if (task) {
            const uint32_t ent = adj_list_entry_lane(b, list, 0, i);
            const int q = static_cast<int>((ent >> 17) & 15u), jj = static_cast<int>(ent & 0x1ffffu);
            const n2h_adj_net& a = b.net[q];
            const int R = a.ring_slots, RS = R + W - 1;
            const int64_t seg = static_cast<int64_t>(jj) * a.long_parts + part;
            const int64_t e0 = a.long_ptr[seg], e1 = a.long_ptr[seg + 1];
            const uint2* links = reinterpret_cast<const uint2*>(a.long_links);
            for (int64_t e = e0 + sub; e < e1; e += U * LANES) {
                uint2 l[U];
#pragma unroll
                for (int r = 0; r < U; ++r) {
                    const int64_t er = e + r * LANES;
                    l[r] = er < e1 ? adj_load_link(links + er) : make_uint2(0u, 0u);  // padding: target 0, delay 1, weight +0
                }
                float gw[U][W];
#pragma unroll
                for (int r = 0; r < U; ++r) {
                    uint32_t tp;
                    adj_load_window<W>(a.gx_ring + adj_link_slot(l[r].x, b.ring_pos, R, RS, tp), gw[r]);
                }
#pragma unroll
                for (int r = 0; r < U; ++r) {
                    const bool live = e + r * LANES < e1;
                    const float wt = __uint_as_float(l[r].y);
#pragma unroll
                    for (int w = 0; w < W; ++w) bins[w] = live ? bins[w] + wt * gw[r][w] : bins[w];
                }
            }
        }
#pragma unroll
        for (int o = LANES / 2; o > 0; o /= 2) {  // every lane of the group takes part (task uniform per group)
#pragma unroll
            for (int w = 0; w < W; ++w) bins[w] += __shfl_xor(bins[w], o, kWave);
        }
        if (task && sub < W) {
            float v = bins[0];
#pragma unroll
            for (int w = 1; w < W; ++w) v = sub == w ? bins[w] : v;
            b.partial[(static_cast<int64_t>(i) * P + part) * WP + sub] = v;
        }
    }
}

// The window's gatherers, a wavefront each: lane w * 8 + p adds the entry's range sums of (window step w,
// target population p) in range order into its long_bins row (steps in the entry's mask).  PURE: the partial
// rows are [W'] per range, and population p's sum takes the ranges whose range_pop is p, in range order.
template <int W, bool PURE>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void adj_combine_kernel(AdjBatch b) {
    constexpr int WP = adj_win_pad<W>();
    constexpr int V = PURE ? WP : kAdjPops * WP;
    const int lane = threadIdx.x & (kWave - 1);
    const int n_waves = gridDim.x * kWavesPerBlock;
    const AdjList list = adj_list_counts(b, 0);
    const int P = b.net[0].long_parts;
    const int ww = lane >> 3, pp = lane & 7;
    for (int i = blockIdx.x * kWavesPerBlock + static_cast<int>(threadIdx.x >> 6); i < list.total; i += n_waves) {
        const uint32_t ent = adj_list_entry(b, list, 0, i);
        const int q = static_cast<int>((ent >> 17) & 15u), jj = static_cast<int>(ent & 0x1ffffu);
        const uint32_t mk = ent >> 21;
        if (lane >= W * kAdjPops || !((mk >> ww) & 1u)) continue;
        const n2h_adj_net& a = b.net[q];
        float sum = 0.0f;
        // the ranges' sums loaded kBatch at a time, all in flight together, then added in range order (a load
        // and its wait per range would make P dependent round trips)
        constexpr int kBatch = 32;
        if constexpr (PURE) {
            const float* src = b.partial + static_cast<int64_t>(i) * P * V + ww;
            bool first = true;
            for (int r0 = 0; r0 < P; r0 += kBatch) {
                float x[kBatch];
                bool mine[kBatch];
#pragma unroll
                for (int k = 0; k < kBatch; ++k) {
                    mine[k] = r0 + k < P && a.range_pop[r0 + k] == pp;
                    x[k] = mine[k] ? src[static_cast<int64_t>(r0 + k) * V] : 0.0f;
                }
#pragma unroll
                for (int k = 0; k < kBatch; ++k) {
                    if (!mine[k]) continue;
                    sum = first ? x[k] : sum + x[k];
                    first = false;
                }
            }
        } else {
            const float* src = b.partial + static_cast<int64_t>(i) * P * V + lane;
            for (int r0 = 0; r0 < P; r0 += kBatch) {
                float x[kBatch];
#pragma unroll
                for (int k = 0; k < kBatch; ++k) x[k] = r0 + k < P ? src[static_cast<int64_t>(r0 + k) * V] : 0.0f;
#pragma unroll
                for (int k = 0; k < kBatch; ++k)
                    if (r0 + k < P) sum = r0 + k == 0 ? x[k] : sum + x[k];
            }
        }
        a.long_bins[(static_cast<int64_t>(ww) * a.n_neurons + jj) * kAdjPops + pp] = sum;
    }
}

__device__ inline void adj_update_quiet(const AdjBatch& b, const n2h_adj_net& a, int j);

// The update of a gatherer j (the reverse step's update above: gs = c + sum over p of alpha[p][pj] bins[p],
// gVpre = (1 - s) gV + h gs, ...) from inputs its wavefront loaded ahead: lane p < 8 holds its population's
// total tot_l (short + long sums), lane c < 8 the coefficient C[c] of j's population pj, lane p < n_pops
// alpha[p][pj] and j's compensated dL/dalpha pair; vp, spk, gv, gi are uniform.  The coupling sum runs over
// p in order; lane p updates its dL/dalpha pair; lane 0 writes the state and the adjoint rings.
__device__ inline void adj_update_gatherer(const AdjBatch& b, const n2h_adj_net& a, int j, float tot_l, float coef,
                                           float alpha_l, float2 gs_l, float vp, bool spk, float gv, float gi,
                                           int lane) {
    const int np = a.n_pops;
    float coup = 0.0f;
    for (int p = 0; p < np; ++p) coup += __shfl(alpha_l, p, kWave) * __shfl(tot_l, p, kWave);
    if (spk && lane < np) {  // compensated fp32 accumulation (hi, lo) over the reverse steps
        const n2h_twofloat::df t = n2h_twofloat::two_sum(gs_l.x, tot_l);
        gs_l.x = t.hi;
        gs_l.y = gs_l.y + t.lo;
        reinterpret_cast<float2*>(a.g_sum)[static_cast<int64_t>(j) * np + lane] = gs_l;
    }
    const float P_V = __shfl(coef, 0, kWave), P_I = __shfl(coef, 1, kWave), P_VI = __shfl(coef, 2, kWave);
    const float dv = __shfl(coef, 5, kWave), c_loss = __shfl(coef, 6, kWave);
    if (lane != 0) return;
    const float u = adj_u(vp, dv);
    const bool near = u <= b.sur.cutoff;
    float gip, gv_new;
    if (isfinite(vp)) {
        const float gsum = c_loss + coup;
        const float den = 1.0f + b.sur.kappa * u;
        const float h = near ? (b.sur.gamma * 0.5f * b.sur.kappa / dv) / (den * den) : 0.0f;
        const float gvpre = (spk ? 0.0f : gv) + h * gsum;
        gip = P_VI * gvpre + P_I * gi;
        gv_new = P_V * gvpre;
    } else {
        gip = P_I * gi;
        gv_new = 0.0f;
    }
    const int R = a.ring_slots, RS = R + a.long_delay - 1;
    a.g_v[j] = gv_new;
    a.g_i[j] = gip;
    float* row = a.gx_ring + static_cast<int64_t>(j) * RS;
    row[b.ring_pos] = gip;
    if (b.ring_pos < a.long_delay - 1) row[b.ring_pos + R] = gip;  // the mirror
    a.gx_recent[static_cast<int64_t>(b.recent_pos) * a.n_neurons + j] = gip;
}

// One reverse step in one launch.  Workgroups [0, kShortBlocks) take this step's gatherers, a wavefront
// each: their short links (lanes stride the list, per-lane sums per target population, then a fixed
// reduction), plus the window's long sums (short + long), then lane 0 updates the neuron.  The other
// workgroups update every neuron that did not gather, a lane each.  Neither half waits for the other: a
// gather reads the adjoints of steps t + 1 .. t + W - 1 only, never those this step writes.
__global__ __launch_bounds__(kWave * kWavesPerBlock) void adj_step_kernel(AdjBatch b) {
    const int lane = threadIdx.x & (kWave - 1);
    if (static_cast<int>(blockIdx.x) >= kShortBlocks) {
        // the quiet neurons, a lane each: with at most 8 networks, workgroup u (on XCD x = u mod 8) takes the
        // (u / 8)-th neuron block that the mark kernel deals to XCD x, so that a neuron is updated on the XCD
        // that gathers it (its adjoint, written here, is read there in the next steps)
        const int u = static_cast<int>(blockIdx.x) - kShortBlocks;
        int q, base;
        if (b.n_nets <= kXcds) {
            const int x = u % kXcds, k = u / kXcds;
            q = 0;
            while (q + 1 < b.n_nets && x >= adj_net_xcd0(q + 1, b.n_nets)) ++q;
            const int bq = (x - adj_net_xcd0(q, b.n_nets)) + adj_net_xcds(q, b.n_nets) * k;
            if (bq >= b.first_block[q + 1] - b.first_block[q]) return;
            base = (bq * kWavesPerBlock + static_cast<int>(threadIdx.x >> 6)) * kWave;
        } else {
            base = adj_wave_base_at(b, u, q);
        }
        const n2h_adj_net& a = b.net[q];
        const int j = base + lane;
        if (j >= a.n_neurons) return;
        adj_update_quiet(b, a, j);
        return;
    }
    // this XCD's sub-lists x and 8 + x of the step's list, in that order
    const int x = static_cast<int>(blockIdx.x % kXcds);
    const int n_waves = kShortBlocks / kXcds * kWavesPerBlock;
    const int which = 1 + b.win_step;
    const int c0 = b.work[adj_counter(which, x)], c1 = b.work[adj_counter(which, x + kXcds)];
    for (int i = static_cast<int>(blockIdx.x / kXcds) * kWavesPerBlock + static_cast<int>(threadIdx.x >> 6);
         i < c0 + c1; i += n_waves) {
        const uint32_t ent = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(
            i < c0 ? b.work[adj_sublist(b, which, x) + i] : b.work[adj_sublist(b, which, x + kXcds) + (i - c0)]));
        const int q = static_cast<int>((ent >> 17) & 15u), jj = static_cast<int>(ent & 0x1ffffu);
        const n2h_adj_net& a = b.net[q];
        const int W = a.long_delay, n = a.n_neurons, np = a.n_pops;
        float bins[kAdjPops];
#pragma unroll
        for (int p = 0; p < kAdjPops; ++p) bins[p] = 0.0f;
        // The run bounds and, in the same round, every input of the neuron's update, spread over the lanes,
        // so that after the sums only arithmetic and stores are left: lane p's population start (the
        // population is the count of starts <= jj), then the population's coefficients (lane c < 8) and
        // its alpha column (lane p: alpha[p][pj]); the long sums' row, the compensated dL/dalpha pair
        // (lane p), and the trace, spike flag and adjoint state.
        const int64_t e0 = a.out_ptr[jj], e1 = a.out_ptr[jj + 1];
        const bool starts_before = lane >= 1 && lane < np && a.pop_first[lane] <= jj;
        const float lb = a.long_bins[static_cast<int64_t>(jj) * kAdjPops + (lane & 7)];
        const float vp = a.vtrace[jj];
        const bool spk = a.spikes[jj] != 0;
        const float gv = a.g_v[jj], gi = a.g_i[jj];
        float2 gs_l = make_float2(0.0f, 0.0f);
        if (lane < np) gs_l = reinterpret_cast<const float2*>(a.g_sum)[static_cast<int64_t>(jj) * np + lane];
        const int pj = __popcll(__ballot(starts_before));
        const float coef = a.pop_coef[pj * 8 + (lane & 7)];
        const float alpha_l = lane < np ? a.alpha[lane * np + pj] : 0.0f;
        const uint2* links = reinterpret_cast<const uint2*>(a.out_links);
        for (int64_t e = e0 + lane; e < e1; e += kUnroll * kWave) {  // as in adj_long_kernel
            uint2 l[kUnroll];
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) {
                const int64_t er = e + r * kWave;
                l[r] = er < e1 ? links[er] : make_uint2(0u, 0u);
            }
            float g[kUnroll];
            uint32_t tp[kUnroll];
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) g[r] = a.gx_recent[adj_recent_slot(l[r].x, b.recent_pos, W, n, tp[r])];
#pragma unroll
            for (int r = 0; r < kUnroll; ++r) {
                const bool live = e + r * kWave < e1;
                const float c = __uint_as_float(l[r].y) * g[r];
#pragma unroll
                for (int p = 0; p < kAdjPops; ++p)
                    bins[p] = (live && tp[r] == static_cast<uint32_t>(p)) ? bins[p] + c : bins[p];
            }
        }
        const float v = (e1 > e0) ? adj_reduce_halving<kAdjPops>(bins, lane) : 0.0f;  // lane l: population l >> 3
        // lane p < 8: its population's total, short + long (the long row prefetched)
        const float tot_l = __shfl(v, (lane & 7) * 8, kWave) + lb;
        adj_update_gatherer(b, a, jj, tot_l, coef, alpha_l, gs_l, vp, spk, gv, gi, lane);
    }
}

// The update of a quiet neuron j in this reverse step: it did not spike and lies outside the surrogate's
// support (h = 0), so it gathered nothing.  This is the update of adj_update_gatherer with those values --
// gs = c + sum over a of alpha[a][pj] * 0 = c + 0 (the sum is +0 for finite alpha), gVpre = gV + 0 gs -- from
// inputs loaded in one round (the population starts, then its coefficients; trace, spike flag and state).
__device__ inline void adj_update_quiet(const AdjBatch& b, const n2h_adj_net& a, int j) {
    const int R = a.ring_slots, RS = R + a.long_delay - 1;
    const float* C = a.pop_coef + adj_pop(a, j) * 8;
    const float P_V = C[0], P_I = C[1], P_VI = C[2], dv = C[5], c_loss = C[6];
    const float vp = a.vtrace[j];
    const bool s = a.spikes[j] != 0;
    const float gv = a.g_v[j], gi = a.g_i[j];
    if (s || adj_u(vp, dv) <= b.sur.cutoff) return;  // a gatherer: a wavefront of the gather half updates it
    float gip, gv_new;
    if (isfinite(vp)) {
        const float gsum = c_loss + 0.0f;
        const float gvpre = gv + 0.0f * gsum;
        gip = P_VI * gvpre + P_I * gi;
        gv_new = P_V * gvpre;
    } else {
        gip = P_I * gi;
        gv_new = 0.0f;
    }
    a.g_v[j] = gv_new;
    a.g_i[j] = gip;
    float* row = a.gx_ring + static_cast<int64_t>(j) * RS;
    row[b.ring_pos] = gip;
    if (b.ring_pos < a.long_delay - 1) row[b.ring_pos + R] = gip;  // the mirror
    a.gx_recent[static_cast<int64_t>(b.recent_pos) * a.n_neurons + j] = gip;
}

}  // namespace

extern "C" {

int n2h_adj_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return N2H_EINVAL;
    snprintf(buf, len, "%s", g_err);
    return N2H_OK;
}

}  // extern "C"

namespace {

// checks the networks and fills the launch block (first blocks, ring position of `step`)
int adj_prepare(const n2h_adj_net* nets, int32_t n_nets, int64_t step, const n2h_surrogate* sur, AdjBatch& b,
                int64_t& blocks) {
    if (!nets || n_nets < 1 || n_nets > kAdjNets) return fail(N2H_EINVAL, "1 to %d networks per launch", kAdjNets);
    if (step < 0) return fail(N2H_EINVAL, "negative step");
    if (!sur || !(sur->kappa > 0.0f) || !(sur->cutoff >= 0.0f) || !(sur->gamma >= 0.0f))
        return fail(N2H_EINVAL, "surrogate: kappa > 0, cutoff >= 0, gamma >= 0");
    memset(&b, 0, sizeof(b));
    b.n_nets = n_nets;
    b.sur = *sur;
    blocks = 0;
    int ring = -1, win = -1;
    for (int i = 0; i < n_nets; ++i) {
        const n2h_adj_net& a = nets[i];
        if (a.n_neurons <= 0 || a.n_neurons > (1 << 17)) return fail(N2H_EINVAL, "net %d: n_neurons outside [1, 2^17]", i);
        if (a.n_pops < 1 || a.n_pops > kAdjPops) return fail(N2H_EINVAL, "net %d: 1 to %d populations", i, kAdjPops);
        if (a.ring_slots < 2 || a.ring_slots > N2H_MAX_SLOTS + 1) return fail(N2H_EINVAL, "net %d: ring_slots", i);
        if (a.long_delay < 1 || a.long_delay > kMaxWindow)
            return fail(N2H_EINVAL, "net %d: long_delay outside [1, %d]", i, kMaxWindow);
        if (a.long_parts < 1 || a.long_parts > kMaxParts)
            return fail(N2H_EINVAL, "net %d: long_parts outside [1, %d]", i, kMaxParts);
        if (!a.pop_first || !a.out_ptr || !a.out_links || !a.long_ptr || !a.long_links || !a.pop_coef || !a.alpha ||
            !a.vtrace || !a.spikes || !a.gx_ring || !a.gx_recent || !a.long_bins || !a.g_v || !a.g_i || !a.g_sum)
            return fail(N2H_EINVAL, "net %d: null pointer", i);
        if (ring >= 0 && ring != a.ring_slots) return fail(N2H_EINVAL, "every network needs the same ring_slots");
        if (win >= 0 && win != a.long_delay) return fail(N2H_EINVAL, "every network needs the same long_delay");
        if (i > 0 && a.long_parts != nets[0].long_parts) return fail(N2H_EINVAL, "every network needs the same long_parts");
        if ((a.range_pop == nullptr) != (nets[0].range_pop == nullptr))
            return fail(N2H_EINVAL, "range_pop given for some networks only");
        if (a.range_pop && a.long_parts % kXcds != 0)
            return fail(N2H_EINVAL, "net %d: population-pure ranges need long_parts a multiple of %d", i, kXcds);
        ring = a.ring_slots;
        win = a.long_delay;
        b.net[i] = a;
        b.first_block[i] = static_cast<int32_t>(blocks);
        blocks += (a.n_neurons + kWave * kWavesPerBlock - 1) / (kWave * kWavesPerBlock);
    }
    b.first_block[n_nets] = static_cast<int32_t>(blocks);
    b.ring_pos = static_cast<int32_t>(step % ring);
    b.recent_pos = static_cast<int32_t>(step % win);
    int32_t sizes[kAdjNets];
    for (int i = 0; i < n_nets; ++i) sizes[i] = nets[i].n_neurons;
    b.sub_cap = static_cast<int32_t>(adj_sub_cap(sizes, n_nets));
    return N2H_OK;
}

// the work buffer: the counters, then 1 + long_delay lists of kSubLists sub-lists of sub_cap entries, then
// (long_parts a multiple of kXcds) the range sums of every possible window entry
int64_t adj_lists_words(int64_t sub_cap, int32_t long_delay) {
    return kAdjListHead + (1 + static_cast<int64_t>(long_delay)) * kSubLists * sub_cap;
}
int64_t adj_partial_floats(int64_t n_total, int32_t long_delay, int32_t long_parts) {
    if (long_parts % kXcds != 0) return 0;
    const int wp = long_delay <= 1 ? 1 : long_delay <= 2 ? 2 : long_delay <= 4 ? 4 : 8;
    return n_total * long_parts * kAdjPops * wp;
}
int64_t adj_work_words(int64_t sub_cap, int64_t n_total, int32_t long_delay, int32_t long_parts) {
    return adj_lists_words(sub_cap, long_delay) + adj_partial_floats(n_total, long_delay, long_parts);
}

int adj_bind_work(AdjBatch& b, const n2h_adj_net* nets, int32_t n_nets, int32_t* work, int64_t work_words) {
    int64_t n_total = 0;
    for (int i = 0; i < n_nets; ++i) n_total += nets[i].n_neurons;
    const int32_t W = nets[0].long_delay, P = nets[0].long_parts;
    const int64_t need = adj_work_words(b.sub_cap, n_total, W, P);
    if (!work || work_words < need)
        return fail(N2H_EINVAL, "work: %lld int32 words needed (n2h_adj_work_words), %lld given",
                    static_cast<long long>(need), static_cast<long long>(work_words));
    b.work = work;
    b.partial = adj_partial_floats(n_total, W, P) ? reinterpret_cast<float*>(work + adj_lists_words(b.sub_cap, W)) : nullptr;
    return N2H_OK;
}

template <int W>
void launch_window(int64_t blocks, hipStream_t stream, AdjBatch& b) {
    const dim3 block(kWave * kWavesPerBlock);
    hipLaunchKernelGGL(adj_mark_kernel, dim3(static_cast<unsigned>(blocks)), block, 0, stream, b);
    if (b.partial && b.net[0].range_pop) {  // population-pure ranges split over the XCDs, then added in order
        hipLaunchKernelGGL((adj_long_group_kernel<W, kGroupLanes>), dim3(kGroupBlocks), block, 0, stream, b);
        hipLaunchKernelGGL((adj_combine_kernel<W, true>), dim3(kShortBlocks), block, 0, stream, b);
        return;
    }
    if (b.partial) {  // the ranges split over the XCDs, then added in order
        hipLaunchKernelGGL((adj_long_pipek_kernel<W, kPipeBins>), dim3(kPipeBlocks), block, 0, stream, b);
        hipLaunchKernelGGL((adj_combine_kernel<W, false>), dim3(kShortBlocks), block, 0, stream, b);
        return;
    }
    for (int part = 0; part < b.net[0].long_parts; ++part) {  // one target range a launch
        b.part = part;
        hipLaunchKernelGGL((adj_long_kernel<W>), dim3(kLongBlocks), block, 0, stream, b);
    }
}

}  // namespace

extern "C" {

int64_t n2h_adj_work_words(const int32_t* n_neurons, int32_t n_nets, int32_t long_delay, int32_t long_parts) {
    if (!n_neurons || n_nets < 1 || n_nets > kAdjNets || long_delay < 1 || long_delay > kMaxWindow) return -1;
    if (long_parts < 1 || long_parts > kMaxParts) return -1;
    int64_t n_total = 0;
    for (int i = 0; i < n_nets; ++i) {
        if (n_neurons[i] < 1) return -1;
        n_total += n_neurons[i];
    }
    return adj_work_words(adj_sub_cap(n_neurons, n_nets), n_total, long_delay, long_parts);
}

int n2h_adj_step(const n2h_adj_net* nets, int32_t n_nets, int64_t step, int32_t win_step, const n2h_surrogate* sur,
                 int32_t* work, int64_t work_words, void* stream) {
    AdjBatch b;
    int64_t blocks;
    int rc = adj_prepare(nets, n_nets, step, sur, b, blocks);
    if (rc != N2H_OK) return rc;
    if ((rc = adj_bind_work(b, nets, n_nets, work, work_words)) != N2H_OK) return rc;
    if (win_step < 0 || win_step >= nets[0].long_delay)
        return fail(N2H_EINVAL, "win_step %d outside [0, long_delay = %d)", win_step, nets[0].long_delay);
    b.win_step = win_step;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // the quiet neurons' workgroups: with at most 8 networks, 8 times the most neuron blocks an XCD is dealt
    int64_t quiet = blocks;
    if (n_nets <= kXcds) {
        int64_t most = 0;
        for (int q = 0; q < n_nets; ++q) {
            const int64_t nb = b.first_block[q + 1] - b.first_block[q];
            const int m = adj_net_xcds(q, n_nets);
            most = std::max(most, (nb + m - 1) / m);
        }
        quiet = kXcds * most;
    }
    hipLaunchKernelGGL(adj_step_kernel, dim3(static_cast<unsigned>(kShortBlocks + quiet)), dim3(kWave * kWavesPerBlock), 0,
                       s, b);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(N2H_EHIP, "adjoint step kernels: %s", hipGetErrorString(e));
    return N2H_OK;
}

int n2h_adj_window(const n2h_adj_net* nets, int32_t n_nets, int64_t step0, int32_t n_win, const n2h_surrogate* sur,
                   int32_t* work, int64_t work_words, void* stream) {
    AdjBatch b;
    int64_t blocks;
    int rc = adj_prepare(nets, n_nets, step0, sur, b, blocks);
    if (rc != N2H_OK) return rc;
    if ((rc = adj_bind_work(b, nets, n_nets, work, work_words)) != N2H_OK) return rc;
    const int W = nets[0].long_delay;
    if (n_win < 1 || n_win > W) return fail(N2H_EINVAL, "n_win %d outside [1, long_delay = %d]", n_win, W);
    b.n_win = n_win;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(adj_clear_kernel, dim3(1), dim3(256), 0, s, work);  // the counters
    switch (W) {
        case 1: launch_window<1>(blocks, s, b); break;
        case 2: launch_window<2>(blocks, s, b); break;
        case 3: launch_window<3>(blocks, s, b); break;
        case 4: launch_window<4>(blocks, s, b); break;
        case 5: launch_window<5>(blocks, s, b); break;
        case 6: launch_window<6>(blocks, s, b); break;
        case 7: launch_window<7>(blocks, s, b); break;
        default: launch_window<8>(blocks, s, b); break;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(N2H_EHIP, "adjoint window kernels: %s", hipGetErrorString(e));
    return N2H_OK;
}

}  // extern "C"
