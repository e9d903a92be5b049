// event.hip -- the event-driven, exact-time spiking network on MI355X (gfx950), and its C-ABI (n2h_ev_*).
//
// Brunel's model A as the reference runs it (test/networks/Brunel): delta-synapse LIF neurons whose potential
// changes only at an input's exact arrival time e + D, firing inside delivery at that time, driven by Poisson
// sources with exact spike times; one homogeneous delay D on every connection.  Since every delay is D, the
// spikes emitted in one window of length D are exactly the inputs of the next window, so a window is:
//   1. ev_deliver_kernel, one 64-lane wavefront per group: a neuron group walks the window's input stream
//      (the previous window's spikes of its source populations, merged by arrival), gathers each spike's links
//      into the group (the group segments of n2h_tables), sorts them by target lane (stable), and every lane
//      runs its neuron's inputs in order -- receive (decay to the arrival, add the weight, unless refractory),
//      and the threshold once per distinct arrival time; a source group draws its spikes for the window.
//      Spikes go to the group's region of the window.
//   2. ev_sort_kernel, one workgroup of 1024 threads: one merged stream per distinct set of source
//      populations, in (arrival, connection, neuron) order -- the input of the next window -- by a counting
//      sort into arrival-time bins; with a record bound, also the window's spikes in (population, time,
//      neuron) order.
// Both read the window index from a device counter (the sort advances it), so windows replay in hipGraphs.
//
// Reference behaviour followed (MatteoSchmider/naer2):
//   naer2/transpiler/templates/fused.py:2084-2301 (_pull_queues): a source run's spikes in (e, node) order;
//     fused.py:1300-1370 (_merge_segments): the connections' queues merged by arrival e + D_c, ties to the
//     lower connection (here a connection's spikes of one arrival time by source neuron: the reference's
//     (e, node) order differs only when two emission times round to the same fp32 arrival); fused.py:1421-1526 (_walk_batch): links unpacked per segment, sorted stably into one
//     inbox per neuron, each neuron's inputs at one time received one by one, then its send once;
//   fused.py:2667-2702 (_emit_delivered) and 2715-2758 (_sort_region): spikes inside delivery at the arrival
//     time, a region per program, sorted by (e, node);
//   test/networks/Brunel/neuron.py:34-45 (send / receive) and poisson.py:21-24 (update).
// The fp32 exp and log here are this build's own (Cody-Waite + Taylor; a least-squares log1p fit), computed
// with exactly the operations of the oracle, so that spike times match the oracle's bit for bit.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "naer2_hip.h"
#include "philox.h"

namespace {

constexpr int kWave = 64;
constexpr int kEnt = 256;          // input-stream entries a group takes per round (one per thread)
constexpr int kCapL = 2048;        // links a group stages per round (LDS: 8 B staged + 8 B inbox each)
constexpr int kLinkU = 8;          // link loads per lane in flight while staging
constexpr int kWaves = 4;          // wavefronts per group's workgroup (staging); wave 0 walks
constexpr int kThreads = kWaves * kWave;
constexpr int kXcdRun = 16;        // consecutive groups one XCD runs in a row (one 128-byte line of offsets)
constexpr int kSortThreads = 1024;
constexpr int kPerThread = 2;                          // keys per thread in the register sort
constexpr int kSmallSort = kPerThread * kSortThreads;  // key counts up to this sort in registers (bitonic_small)
constexpr int kBins = 2048;                            // arrival-time bins of a window's stream sort
constexpr int kBinCap = 2048;                          // spikes a stream bins in LDS (more: the bitonic network)
constexpr int kBinRankMax = 256;                       // spikes of one bin placed by rank (more: the bitonic network)
constexpr int kPrefixLds = 4096;                       // group prefix held in LDS (networks of < 4096 groups)

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define EV_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(N2H_EHIP, "%s: %s", #call, hipGetErrorString(e_));   \
    } while (0)

struct EvPopDev {
    int32_t kind;
    int32_t cap;        // region entries per group
    float tau_inv, theta, v_reset, t_ref, rate_inv;
    int32_t stream;     // neuron populations: their input stream, -1: none
    int32_t group0;     // first group
    int32_t n_groups;
    int32_t first;      // first neuron
    int32_t n;
    int64_t region0;    // first region entry of group0
};

struct EvStreamDev {
    int32_t n_src;
    int32_t src[N2H_EV_MAX_SOURCES];
    float delay[N2H_EV_MAX_SOURCES];   // per source: its connection's delay (ms) ...
    int32_t lag[N2H_EV_MAX_SOURCES];   // ... = lag windows
    int64_t base;       // first entry in the stream arrays
};

struct EvArgs {
    int32_t n_groups;
    int32_t n_pops;
    int32_t n_streams;
    int32_t rec_windows;
    const int32_t* group_first;
    const int32_t* group_info;
    const int64_t* seg_off;
    const uint2* links;
    const int64_t* train_ptr;
    const float* train_t;
    const EvPopDev* pops;
    const EvStreamDev* streams;
    float* v;
    float* last;
    float* tre;
    float* nxt;
    int32_t* reg_lane;
    float* reg_e;
    int32_t* reg_count;
    int32_t* stream_count;
    int32_t* stream_slot;
    float* stream_a;
    unsigned long long* keys;     // [keys_cap]: sort scratch beyond LDS
    int64_t keys_cap;
    int32_t* group_prefix;        // [n_groups + 1]
    int64_t* window;              // the device window counter
    int32_t* flags;               // [N2H_EV_FLAGS]
    int32_t* rec_node;
    float* rec_e;
    int32_t* rec_count;
    int32_t* rec_wcount;
    int64_t rec_cap;
    unsigned long long* counters; // [2]: inputs delivered, spikes emitted (since the plan was made)
    unsigned long long* trace;    // null, or [n_groups + 1][8] phase stamps of the last window (100 MHz clock)
    int32_t xcd;                  // 0: block b runs group b; k > 1: runs of kXcdRun consecutive groups per XCD
    int32_t* log_rec;             // delivery log (the logging kernel variant): [log_cap][4] {window, connection,
    float* log_a;                 //   source, target} and [log_cap] arrival, appended at *log_count
    int32_t* log_count;
    const int32_t* log_conn;      // [n_pops][n_pops]: connection index from source population s into t at t * n + s
    int64_t log_cap;
    int32_t n_hist;               // windows of sorted spikes kept (the largest lag; 1: streams from the regions)
    int32_t pad_h;
    unsigned long long* hist_key; // [n_hist][keys_cap]: a window's spikes, (population, time, neuron) keys
    int32_t* hist_off;            // [n_hist][n_pops + 1]: each population's first key in its window's slot
    int64_t* hist_win;            // [n_hist]: the window a slot holds (-1: none)
    float W;
    uint32_t key0, key1, instance;
};

// ---------------------------------------------------------------- the contract's fp32 math

// exp(x) for x <= 88: x = n ln2 + r (Cody-Waite: n * 0.693145751953125 is exact), the degree-7 Taylor
// polynomial of exp(r) by Estrin's scheme (a short dependency chain for the walk), times 2^n; 0 below -87.
__device__ __forceinline__ float ev_exp(float x) {
    if (x < -87.0f) return 0.0f;
    const float n = rintf(x * 1.44269502f);
    float r = x - n * 0.693145751953125f;
    r = r - n * 1.42860677e-06f;
    const float r2 = r * r;
    const float r4 = r2 * r2;
    const float q01 = 1.0f + 1.0f * r;
    const float q23 = 0.5f + 1.66666672e-01f * r;
    const float q45 = 4.16666679e-02f + 8.33333377e-03f * r;
    const float q67 = 1.38888892e-03f + 1.98412701e-04f * r;
    const float q03 = q01 + r2 * q23;
    const float q47 = q45 + r2 * q67;
    return ldexpf(q03 + r4 * q47, static_cast<int>(n));
}

// ev_exp without a branch: the same value for every finite x (below -87 the reduction still runs, then 0 is
// selected), for the walk's straight-line loop.
__device__ __forceinline__ float ev_exp_sel(float x) {
    const float n = rintf(x * 1.44269502f);
    float r = x - n * 0.693145751953125f;
    r = r - n * 1.42860677e-06f;
    const float r2 = r * r;
    const float r4 = r2 * r2;
    const float q01 = 1.0f + 1.0f * r;
    const float q23 = 0.5f + 1.66666672e-01f * r;
    const float q45 = 4.16666679e-02f + 8.33333377e-03f * r;
    const float q67 = 1.38888892e-03f + 1.98412701e-04f * r;
    const float q03 = q01 + r2 * q23;
    const float q47 = q45 + r2 * q67;
    const float y = ldexpf(q03 + r4 * q47, static_cast<int>(fmaxf(n, -1000.0f)));
    return x < -87.0f ? 0.0f : y;
}

// log(u) for finite u > 0: u = 2^e m, m in [sqrt(1/2), sqrt(2)), log m = t P(t) with t = m - 1 exact and P a
// degree-10 least-squares fit of log1p(t)/t on [sqrt(1/2) - 1, sqrt(2) - 1].
__device__ __forceinline__ float ev_log(float u) {
    const uint32_t b = __float_as_uint(u);
    int e = static_cast<int>((b >> 23) & 0xFF) - 127;
    float m = __uint_as_float((b & 0x7FFFFFu) | 0x3F800000u);
    if (m > 1.41421354f) {
        m = m * 0.5f;
        e += 1;
    }
    const float t = m - 1.0f;
    float p = 0.0674023256f;
    p = p * t + -0.117272295f;
    p = p * t + 0.119239263f;
    p = p * t + -0.123992212f;
    p = p * t + 0.142126068f;
    p = p * t + -0.166680425f;
    p = p * t + 0.200025767f;
    p = p * t + -0.250000238f;
    p = p * t + 0.333333015f;
    p = p * t + -0.5f;
    p = p * t + 1.0f;
    const float fe = static_cast<float>(e);
    float r = t * p;
    r = r + fe * 1.42860677e-06f;
    return r + fe * 0.693145751953125f;
}

// One exponential interval of a Poisson source (its m-th draw in window k), in ms.
__device__ __forceinline__ float ev_interval(const EvArgs& a, uint32_t node, int64_t k, uint32_t m, float rate_inv) {
    const uint32_t x = philox_word0(node, static_cast<uint32_t>(k), m, a.instance, a.key0, a.key1);
    const float u = (static_cast<float>(x >> 9) + 0.5f) * 1.1920928955078125e-07f;
    return (-ev_log(u)) * rate_inv;
}

__device__ __forceinline__ float window_time(int64_t k, float W) {
    return static_cast<float>(static_cast<double>(k) * static_cast<double>(W));
}

// Inclusive prefix sum over the 64 lanes of a wavefront in DPP lane moves: a scan inside each row of 16
// lanes (row_shr 1, 2, 4, 8; lanes shifted in from outside the row add 0), then row 0's total into rows 1
// and 3 (row_bcast:15) and row 1's running total into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ int wave_incl_scan(int x, int /*lane*/) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return x;
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = max(x, __shfl_xor(x, off, kWave));
    return x;
}

// The lanes in `fire` spike at t[lane]: appended to the group's region in lane order.  emitted is uniform.
__device__ __forceinline__ void emit(const EvArgs& a, int64_t rbase, int cap, int& emitted, bool fire, int lane,
                                     float t) {
    const uint64_t m = __ballot(fire);
    if (m) {
        const int pos = emitted + __popcll(m & ((1ull << lane) - 1ull));
        if (fire && pos < cap) {
            a.reg_lane[rbase + pos] = lane;
            a.reg_e[rbase + pos] = t;
        }
        emitted += __popcll(m);
    }
}

template <bool LOG>
struct DeliverLds {
    int32_t off[kEnt + 1];          // the round's entries: first staged link (exclusive scan), off[r] = links
    int32_t slot[LOG ? kEnt : 1];   // (delivery log) the entry's source slot
    int32_t in_e[LOG ? kCapL : 1];  // (delivery log) the inbox input's entry
    int64_t lo[kEnt];               // first link of the entry's segment
    float a[kEnt];                  // arrival time
    uint32_t st_t[kCapL];           // staged link: target lane | entry << 8
    float st_w[kCapL];
    float in_w[kCapL];              // inboxes: the links by target lane, stable
    float in_a[kCapL];
    int32_t cnt[kWaves][kWave];     // links per (staging wave, target lane)
    int32_t wstart[kWaves][kWave];  // a wave's first inbox place per target lane
    int32_t wrun[kWaves][kWave];    // links a wave has placed per target lane
    int32_t scan[kWaves + 1];
    int32_t T;
    int32_t depth;
};

// One group of a window, one workgroup of kWaves wavefronts.  Neuron groups: all waves stage each round of the
// input stream (every wave a contiguous quarter of the round's links, so the per-target order stays the link
// order), wave 0 -- one lane per neuron -- walks the inboxes; then the final threshold.  Source groups: wave 0
// draws the window's spikes.  Region count and state written last.
// Block b's group: b itself, or with k XCDs (blocks are dealt to the XCDs in turn) runs of kXcdRun
// consecutive groups per XCD, so that neighbouring groups -- whose segments of one source and whose offset
// entries lie side by side in memory -- share an L2.
__device__ __forceinline__ int group_of_block(int b, int k) {
    if (k <= 1) return b;
    const int x = b % k, q = b / k;
    return (q / kXcdRun) * (k * kXcdRun) + x * kXcdRun + q % kXcdRun;
}

// The delivery log (the reference's deliveries(): (slot, e, connection, source, target) in delivery order per
// target): the lanes in `has` append {window, connection, source, target} and the arrival, one wave-wide
// atomic per call; past the capacity the count still grows and nothing is written.
__device__ __forceinline__ void log_input(const EvArgs& a, const int32_t* ent_slot, const int32_t* in_e, int at,
                                          bool has, int lane, int64_t k, float ar, int tpop, int target) {
    const uint64_t m = __ballot(has);
    if (!m) return;
    int base = 0;
    if (lane == 0) base = atomicAdd(a.log_count, __popcll(m));
    base = __builtin_amdgcn_readlane(base, 0);
    if (!has) return;
    const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
    if (pos >= a.log_cap) return;
    const int32_t slot = ent_slot[in_e[at]];
    const int sg = slot / kWave;
    const int spop = a.group_info[sg] >> 16;
    const int source = (sg - a.pops[spop].group0) * kWave + (slot & (kWave - 1));
    int32_t* rec = a.log_rec + 4 * static_cast<int64_t>(pos);
    rec[0] = static_cast<int32_t>(k);
    rec[1] = a.log_conn[tpop * a.n_pops + spop];
    rec[2] = source;
    rec[3] = target;
    a.log_a[pos] = ar;
}

template <bool LOG>
__global__ __launch_bounds__(kThreads) void ev_deliver_kernel(EvArgs a) {
    __shared__ DeliverLds<LOG> s;
    const int g = group_of_block(static_cast<int>(blockIdx.x), a.xcd);
    if (g >= a.n_groups) return;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;
    const int info = a.group_info[g];
    const int cnt = info & 0xFFFF;
    const EvPopDev P = a.pops[info >> 16];
    const int64_t k = *a.window;
    const int64_t rbase = P.region0 + static_cast<int64_t>(g - P.group0) * P.cap;
    const bool act = wv == 0 && lane < cnt;
    const int node = a.group_first[g] + lane;
    int emitted = 0;

    if (P.kind == N2H_EV_POISSON) {
        if (wv) return;
        const float t_end = window_time(k + 1, a.W);
        float next = act ? a.nxt[node] : 0.0f;
        uint32_t m = 0;
        if (act && next < 0.0f) next = ev_interval(a, static_cast<uint32_t>(node), k, m++, P.rate_inv);
        bool want = act && next < t_end;
        while (__ballot(want)) {
            emit(a, rbase, P.cap, emitted, want, lane, next);
            if (want) next = next + ev_interval(a, static_cast<uint32_t>(node), k, m++, P.rate_inv);
            want = act && next < t_end;
        }
        if (act) a.nxt[node] = next;
    } else if (P.kind == N2H_EV_TRAIN) {
        if (wv) return;
        const float t_start = window_time(k, a.W), t_end = window_time(k + 1, a.W);
        int64_t j = 0, hi = 0;
        if (act) {
            int64_t lo = a.train_ptr[node];
            hi = a.train_ptr[node + 1];
            int64_t h = hi;
            while (lo < h) {  // the first listed time >= t_start
                const int64_t mid = (lo + h) >> 1;
                if (a.train_t[mid] < t_start) lo = mid + 1;
                else h = mid;
            }
            j = lo;
        }
        float t = (act && j < hi) ? a.train_t[j] : 0.0f;
        bool want = act && j < hi && t < t_end;
        while (__ballot(want)) {
            emit(a, rbase, P.cap, emitted, want, lane, t);
            if (want) {
                ++j;
                t = j < hi ? a.train_t[j] : 0.0f;
            }
            want = want && j < hi && t < t_end;
        }
    } else {
        float V = act ? a.v[node] : 0.0f;
        float last = act ? a.last[node] : 0.0f;
        float tre = act ? a.tre[node] : 0.0f;
        float aprev = 0.0f;
        bool pend = false;
        if (P.stream >= 0) {
            const EvStreamDev& S = a.streams[P.stream];
            const int32_t n_s = a.stream_count[P.stream];
            const int32_t* sslot = a.stream_slot + S.base;
            const float* sa = a.stream_a + S.base;
            s.cnt[wv][lane] = 0;
            s.wrun[wv][lane] = 0;
            const uint64_t lt = (1ull << lane) - 1ull;
            const bool tr = a.trace != nullptr && tid == 0;
            unsigned long long t_in = tr ? wall_clock64() : 0ull, t_ph = t_in;
            unsigned long long ph[5] = {0ull, 0ull, 0ull, 0ull, 0ull};
            int rounds = 0;
#define EV_PHASE(i)                                 \
    do {                                            \
        if (tr) {                                   \
            const unsigned long long t_ = wall_clock64(); \
            ph[(i)] += t_ - t_ph;                   \
            t_ph = t_;                              \
        }                                           \
    } while (0)
            // the entries of a round, one per thread (a spike of the stream and its segment into the group),
            // loaded a round ahead: the stream entry while this round stages, its segment bounds after that
            bool pf_ok = tid < n_s;
            int32_t pf_slot = pf_ok ? sslot[tid] : 0;
            float pf_a = pf_ok ? sa[tid] : 0.0f;
            int64_t pf_lo = 0, pf_hi = 0;
            if (pf_ok) {
                const int64_t seg = static_cast<int64_t>(pf_slot) * a.n_groups + g;
                pf_lo = a.seg_off[seg];
                pf_hi = a.seg_off[seg + 1];
            }
            for (int base = 0; base < n_s;) {
                ++rounds;
                const bool ok = pf_ok;
                const float ae = pf_a;
                const int32_t cur_slot = pf_slot;
                const int64_t lo = pf_lo;
                const int nl = ok ? static_cast<int>(pf_hi - pf_lo) : 0;
                // inclusive scan over the workgroup (entry order = thread order)
                const int x = wave_incl_scan(nl, lane);
                if (lane == kWave - 1) s.scan[wv] = x;
                __syncthreads();
                int carry = 0;
#pragma unroll
                for (int w = 0; w < kWaves; ++w)
                    if (w < wv) carry += s.scan[w];
                const int incl = x + carry;
                // the longest prefix of entries whose links fit in LDS (incl is non-decreasing in entry order)
                int r = __syncthreads_count(ok && incl <= kCapL);
                if (r > 0 && tid == r - 1) s.T = incl;
                if (r > 0 && tid < r) {
                    s.off[tid] = incl - nl;
                    s.lo[tid] = lo;
                    s.a[tid] = ae;
                    if constexpr (LOG) s.slot[tid] = cur_slot;
                }
                __syncthreads();
                int T;
                if (r == 0) {  // one segment longer than the stage: its first kCapL links only, flagged
                    r = 1;
                    T = kCapL;
                    if (tid == 0) {
                        s.off[0] = 0;
                        s.lo[0] = lo;
                        s.a[0] = ae;
                        if constexpr (LOG) s.slot[0] = cur_slot;
                        atomicOr(&a.flags[N2H_EV_FLAG_SEGMENT], 1);
                    }
                } else {
                    T = s.T;
                }
                if (tid == 0) {
                    s.off[r] = T;
                    if (T) atomicAdd(&a.counters[0], static_cast<unsigned long long>(T));
                }
                __syncthreads();
                EV_PHASE(0);
                {
                    const int nidx = base + r + tid;
                    pf_ok = nidx < n_s;
                    pf_slot = pf_ok ? sslot[nidx] : 0;
                    pf_a = pf_ok ? sa[nidx] : 0.0f;
                }
                // pass 1: each wave stages its quarter of the links, counts per target lane; kLinkU loads per
                // lane in flight (indices clamped into the quarter, results past its end ignored)
                const int Q = ((T + kThreads - 1) / kThreads) * kWave;
                const int q0 = wv * Q, q1 = min(q0 + Q, T);
                for (int f0 = q0; f0 < q1; f0 += kWave * kLinkU) {
                    int pos[kLinkU];
                    uint2 lk[kLinkU];
#pragma unroll
                    for (int u = 0; u < kLinkU; ++u) {
                        const int f = min(f0 + u * kWave + lane, q1 - 1);
                        int p = 0;
#pragma unroll
                        for (int step = kEnt / 2; step > 0; step >>= 1)
                            if (p + step < r && s.off[p + step] <= f) p += step;
                        pos[u] = p;
                    }
#pragma unroll
                    for (int u = 0; u < kLinkU; ++u) {
                        const int f = min(f0 + u * kWave + lane, q1 - 1);
                        lk[u] = a.links[s.lo[pos[u]] + (f - s.off[pos[u]])];
                    }
#pragma unroll
                    for (int u = 0; u < kLinkU; ++u) {
                        const int f = f0 + u * kWave + lane;
                        if (f < q1) {
                            const uint32_t t = lk[u].x & 63u;  // (the target lane: x & 0xffff < 64 by construction)
                            s.st_t[f] = t | (static_cast<uint32_t>(pos[u]) << 8);
                            s.st_w[f] = __uint_as_float(lk[u].y);
                            atomicAdd(&s.cnt[wv][t], 1);
                        }
                    }
                }
                __syncthreads();
                EV_PHASE(1);
                if (pf_ok) {
                    const int64_t seg = static_cast<int64_t>(pf_slot) * a.n_groups + g;
                    pf_lo = a.seg_off[seg];
                    pf_hi = a.seg_off[seg + 1];
                }
                // inbox places: target lane t's inbox starts at the links to lanes < t; within it, wave w's
                // links follow those of waves < w
                int c = 0;
                if (wv == 0) {
                    int tot = 0;
#pragma unroll
                    for (int w = 0; w < kWaves; ++w) tot += s.cnt[w][lane];
                    int st = wave_incl_scan(tot, lane) - tot;
#pragma unroll
                    for (int w = 0; w < kWaves; ++w) {
                        s.wstart[w][lane] = st;
                        st += s.cnt[w][lane];
                    }
                    c = tot;
                    const int d = wave_max(tot);
                    if (lane == 0) s.depth = d;
                }
                __syncthreads();
                EV_PHASE(2);
                // pass 2: each wave places its quarter in order, stable ranks among equal targets per chunk
                for (int f0 = 0; f0 < Q; f0 += kWave) {
                    const int f = q0 + f0 + lane;
                    const bool fo = f < q1;
                    const uint32_t tw = fo ? s.st_t[f] : 0u;
                    const uint32_t t = tw & 0xFFu;
                    uint64_t match = __ballot(fo);
#pragma unroll
                    for (int b = 0; b < 6; ++b) {
                        const uint64_t bb = __ballot(fo && ((t >> b) & 1u));
                        match &= ((t >> b) & 1u) ? bb : ~bb;
                    }
                    if (fo) {
                        const int p = s.wstart[wv][t] + s.wrun[wv][t] + __popcll(match & lt);
                        s.in_w[p] = s.st_w[f];
                        s.in_a[p] = s.a[tw >> 8];
                        if constexpr (LOG) s.in_e[p] = static_cast<int32_t>(tw >> 8);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (fo && (match >> lane) == 1ull) s.wrun[wv][t] += __popcll(match);  // the last lane of t
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                __syncthreads();
                EV_PHASE(3);
                // the walk: wave 0, every lane its neuron's inputs of the round, in order
                if (wv == 0) {
                    const int depth = s.depth;
                    const int st = s.wstart[0][lane];
                    // (the next input read one iteration ahead: the LDS latency off the chain through V;
                    // branch-free: every lane computes both outcomes and selects)
                    float w_n = c > 0 ? s.in_w[st] : 0.0f;
                    float a_n = c > 0 ? s.in_a[st] : 0.0f;
                    for (int q = 0; q < depth; ++q) {
                        const bool has = act && q < c;
                        const float w = w_n;
                        const float ar = a_n;
                        if (q + 1 < c) {
                            w_n = s.in_w[st + q + 1];
                            a_n = s.in_a[st + q + 1];
                        }
                        const bool fire = has && pend && ar != aprev && V >= P.theta;
                        const float t_end_ref = aprev + P.t_ref;
                        V = fire ? P.v_reset : V;
                        tre = fire ? t_end_ref : tre;
                        last = fire ? t_end_ref : last;
                        emit(a, rbase, P.cap, emitted, fire, lane, aprev);
                        if constexpr (LOG) log_input(a, s.slot, s.in_e, st + q, has, lane, k, ar, info >> 16,
                                                     (g - P.group0) * kWave + lane);
                        const bool recv = has && ar >= tre;
                        const float vn = V * ev_exp_sel((last - ar) * P.tau_inv) + w;
                        V = recv ? vn : V;
                        last = recv ? ar : last;
                        aprev = has ? ar : aprev;
                        pend = pend || has;
                    }
                }
                s.cnt[wv][lane] = 0;
                s.wrun[wv][lane] = 0;
                __syncthreads();
                EV_PHASE(4);
                base += r;
            }
#undef EV_PHASE
            if (tr) {
                unsigned long long* row = a.trace + static_cast<int64_t>(g) * 8;
                row[0] = t_in;
                for (int i = 0; i < 5; ++i) row[1 + i] = ph[i];
                row[6] = static_cast<unsigned long long>(rounds);
                row[7] = wall_clock64();
            }
        }
        if (wv) return;
        const bool fire = act && pend && V >= P.theta;
        if (fire) {
            V = P.v_reset;
            tre = aprev + P.t_ref;
            last = tre;
        }
        emit(a, rbase, P.cap, emitted, fire, lane, aprev);
        if (act) {
            a.v[node] = V;
            a.last[node] = last;
            a.tre[node] = tre;
        }
    }
    if (tid == 0) {
        a.reg_count[g] = min(emitted, P.cap);
        if (emitted > P.cap) atomicOr(&a.flags[N2H_EV_FLAG_REGION], 1);
    }
}

// ---------------------------------------------------------------- the window's sort (one workgroup)

// Bitonic network over keys[0, n) (n a power of two) in place, 1024 threads.  LDS: n <= L keys in s_k.
// Beyond L the keys stay in global memory: stages whose partners are L or more apart run there, the
// others on chunks of L keys staged in LDS.  The direction of a pair at stage `size` is ascending iff
// (index & size) == 0, whichever memory holds it.
template <int L>
__device__ void bitonic_local(unsigned long long* s_k, int n, int64_t off, int size_lo, int size_hi, int stride_hi) {
    // stages (size, stride) for size in [size_lo, size_hi], stride from min(size/2, stride_hi) down to 1, on
    // the n keys at global index off..off+n-1 held in s_k
    const int tid = threadIdx.x;
    for (int size = size_lo; size <= size_hi; size <<= 1) {
        for (int stride = min(size >> 1, stride_hi); stride > 0; stride >>= 1) {
            for (int p = tid; p < (n >> 1); p += kSortThreads) {
                const int i = 2 * p - (p & (stride - 1));
                const int j = i + stride;
                const bool asc = ((off + i) & size) == 0;
                const unsigned long long x = s_k[i], y = s_k[j];
                if ((x > y) == asc) {
                    s_k[i] = y;
                    s_k[j] = x;
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long v, int m) {
    const int lo = __shfl_xor(static_cast<int>(static_cast<uint32_t>(v)), m, kWave);
    const int hi = __shfl_xor(static_cast<int>(static_cast<uint32_t>(v >> 32)), m, kWave);
    return (static_cast<unsigned long long>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo);
}

// One in-register stage of the bitonic network on a thread's kPerThread consecutive keys (elements
// kPerThread t .. + kPerThread - 1): pairs STRIDE apart (STRIDE < kPerThread), ascending where
// (index & size) == 0.
template <int STRIDE>
__device__ __forceinline__ void ce_local(unsigned long long (&x)[kPerThread], int i0, int size) {
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        if (j & STRIDE) continue;
        const bool asc = ((i0 + j) & size) == 0;
        if ((x[j] > x[j + STRIDE]) == asc) {
            const unsigned long long t = x[j];
            x[j] = x[j + STRIDE];
            x[j + STRIDE] = t;
        }
    }
}

// The bitonic network over kSmallSort keys in s_k, in registers: thread t holds elements kPerThread t ..
// kPerThread t + kPerThread - 1; pairs less than kPerThread apart meet in the thread, pairs up to
// 64 kPerThread / 2 apart by lane shuffles within a wavefront, farther pairs through LDS.
__device__ void bitonic_small(unsigned long long* s_k) {
    const int t = static_cast<int>(threadIdx.x);
    const int i0 = kPerThread * t;
    unsigned long long x[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) x[j] = s_k[i0 + j];
    for (int size = 2; size <= kSmallSort; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride < kPerThread) {
                if (stride == 1) ce_local<1>(x, i0, size);
                if constexpr (kPerThread > 2) {
                    if (stride == 2) ce_local<2>(x, i0, size);
                }
                if constexpr (kPerThread > 4) {
                    if (stride == 4) ce_local<4>(x, i0, size);
                }
            } else {
                unsigned long long y[kPerThread];
                if (stride / kPerThread < kWave) {
#pragma unroll
                    for (int j = 0; j < kPerThread; ++j) y[j] = shfl_xor64(x[j], stride / kPerThread);
                } else {
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < kPerThread; ++j) s_k[i0 + j] = x[j];
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < kPerThread; ++j) y[j] = s_k[(i0 + j) ^ stride];
                }
                const bool take_min = ((i0 & stride) == 0) == ((i0 & size) == 0);
#pragma unroll
                for (int j = 0; j < kPerThread; ++j) x[j] = take_min ? min(x[j], y[j]) : max(x[j], y[j]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) s_k[i0 + j] = x[j];
    __syncthreads();
}

template <int L>
__device__ void block_sort(unsigned long long* keys, int64_t n, unsigned long long* s_k, bool in_lds) {
    const int tid = threadIdx.x;
    if (in_lds && n == kSmallSort) {
        bitonic_small(s_k);
        return;
    }
    if (in_lds) {
        bitonic_local<L>(s_k, static_cast<int>(n), 0, 2, static_cast<int>(n), static_cast<int>(n));
        return;
    }
    // 1. every chunk of L sorted in LDS (stages of size <= L)
    for (int64_t c0 = 0; c0 < n; c0 += L) {
        for (int i = tid; i < L; i += kSortThreads) s_k[i] = keys[c0 + i];
        __syncthreads();
        bitonic_local<L>(s_k, L, c0, 2, L, L);
        for (int i = tid; i < L; i += kSortThreads) keys[c0 + i] = s_k[i];
        __syncthreads();
    }
    // 2. larger stages: partners >= L apart in global memory, then the rest per chunk in LDS
    for (int64_t size = 2 * static_cast<int64_t>(L); size <= n; size <<= 1) {
        for (int64_t stride = size >> 1; stride >= L; stride >>= 1) {
            for (int64_t p = tid; p < (n >> 1); p += kSortThreads) {
                const int64_t i = 2 * p - (p & (stride - 1));
                const int64_t j = i + stride;
                const bool asc = (i & size) == 0;
                const unsigned long long x = keys[i], y = keys[j];
                if ((x > y) == asc) {
                    keys[i] = y;
                    keys[j] = x;
                }
            }
            __syncthreads();
        }
        for (int64_t c0 = 0; c0 < n; c0 += L) {
            for (int i = tid; i < L; i += kSortThreads) s_k[i] = keys[c0 + i];
            __syncthreads();
            // stages of this size with stride < L: (c0 + i) & size gives the direction
            const int tid2 = threadIdx.x;
            for (int stride = L >> 1; stride > 0; stride >>= 1) {
                for (int p = tid2; p < (L >> 1); p += kSortThreads) {
                    const int i = 2 * p - (p & (stride - 1));
                    const int j = i + stride;
                    const bool asc = ((c0 + i) & size) == 0;
                    const unsigned long long x = s_k[i], y = s_k[j];
                    if ((x > y) == asc) {
                        s_k[i] = y;
                        s_k[j] = x;
                    }
                }
                __syncthreads();
            }
            for (int i = tid; i < L; i += kSortThreads) keys[c0 + i] = s_k[i];
            __syncthreads();
        }
    }
}

// Exclusive scan of one value per thread over the workgroup; returns the thread's prefix, *total the sum.
__device__ int block_excl_scan(int x, int* s_w, int* total) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const int incl = wave_incl_scan(x, lane);
    if (lane == kWave - 1) s_w[wv] = incl;
    __syncthreads();
    if (wv == 0) {
        const int y = lane < kSortThreads / kWave ? s_w[lane] : 0;
        const int yi = wave_incl_scan(y, lane);
        if (lane < kSortThreads / kWave) s_w[lane] = yi - y;
        if (lane == kSortThreads / kWave - 1) s_w[kSortThreads / kWave] = yi;
    }
    __syncthreads();
    const int r = s_w[wv] + incl - x;
    *total = s_w[kSortThreads / kWave];
    __syncthreads();
    return r;
}

__device__ __forceinline__ int64_t pow2_at_least(int64_t n) {
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

// An arrival time's bin among kBins over the window's arrivals: non-decreasing in arr (a subtraction and a
// multiplication by a positive constant round monotonically), clamped.
__device__ __forceinline__ int arrival_bin(float arr, float a0, float inv_bw) {
    const float x = (arr - a0) * inv_bw;
    return x <= 0.0f ? 0 : (x >= static_cast<float>(kBins - 1) ? kBins - 1 : static_cast<int>(x));
}

template <int L>
__global__ __launch_bounds__(kSortThreads) void ev_sort_kernel(EvArgs a) {
    extern __shared__ unsigned long long s_k[];
    __shared__ int s_w[kSortThreads / kWave + 1];
    __shared__ EvPopDev s_pops[N2H_MAX_POPS];
    __shared__ int s_bin[kBins + 1];     // a stream's spikes per arrival-time bin, then the bins' offsets
    __shared__ int s_fill[kBins];
    __shared__ int s_c[N2H_MAX_POPS];    // a stream's connection index of each population (-1: not a source)
    __shared__ int s_n;
    __shared__ int s_gp[kPrefixLds];     // the group prefix (G < kPrefixLds): a spike's group by binary search
    const int tid = threadIdx.x;
    const int64_t k = *a.window;
    if (tid < a.n_pops) s_pops[tid] = a.pops[tid];
    unsigned long long* trow = (a.trace != nullptr && tid == 0) ? a.trace + static_cast<int64_t>(a.n_groups) * 8 : nullptr;
    if (trow) trow[0] = wall_clock64();
    // 1. every group's region: its place in the window's list
    const int G = a.n_groups;
    const int per = (G + kSortThreads - 1) / kSortThreads;
    const int g0 = min(tid * per, G), g1 = min(g0 + per, G);
    int part = 0;
    for (int g = g0; g < g1; ++g) part += a.reg_count[g];
    int S;
    int pre = block_excl_scan(part, s_w, &S);
    const bool lds_prefix = G < kPrefixLds;
    for (int g = g0; g < g1; ++g) {
        a.group_prefix[g] = pre;
        if (lds_prefix) s_gp[g] = pre;
        pre += a.reg_count[g];
    }
    if (tid == 0) {
        a.group_prefix[G] = S;
        if (lds_prefix) s_gp[G] = S;
    }
    __syncthreads();
    // the group holding entry i of the window's list, among groups [lo, hi) with s_gp[lo] <= i < s_gp[hi]
    auto group_of_entry = [&](int i, int lo, int hi) {
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_gp[mid] <= i) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    if (trow) trow[1] = wall_clock64();
    if (S > a.keys_cap) {  // cannot happen: the keys hold every region entry
        if (tid == 0) atomicOr(&a.flags[N2H_EV_FLAG_SORT], 1);
        S = static_cast<int>(a.keys_cap);
    }
    // 2. the window's spikes in (population, time, neuron) order, keys population << 56 | time bits << 24 |
    //    neuron within its population: for the record (when one is bound) and the history of windows (delays
    //    of several windows)
    const bool multi = a.n_hist > 1;
    if (a.rec_node != nullptr || multi) {
        int64_t n2 = pow2_at_least(S > 0 ? S : 1);
        if (n2 < kSmallSort) n2 = kSmallSort;  // (padding sorts in registers)
        const bool in_lds = n2 <= L;
        unsigned long long* keys = in_lds ? s_k : a.keys;
        if (lds_prefix) {  // one thread per spike: every region load in flight at once
            for (int i = tid; i < S; i += kSortThreads) {
                const int g = group_of_entry(i, 0, G);
                const int j = i - s_gp[g];
                const int pop = a.group_info[g] >> 16;
                const EvPopDev& P = s_pops[pop];
                const int64_t rb = P.region0 + static_cast<int64_t>(g - P.group0) * P.cap;
                const uint32_t local = static_cast<uint32_t>(g - P.group0) * kWave + static_cast<uint32_t>(a.reg_lane[rb + j]);
                keys[i] = (static_cast<unsigned long long>(pop) << 56) |
                          (static_cast<unsigned long long>(__float_as_uint(a.reg_e[rb + j])) << 24) | local;
            }
        } else {
            for (int g = tid; g < G; g += kSortThreads) {  // one thread per group: a group holds few spikes
                const int c = a.reg_count[g];
                if (c == 0) continue;
                const int pop = a.group_info[g] >> 16;
                const int b = a.group_prefix[g];
                const EvPopDev& P = s_pops[pop];
                const int64_t rb = P.region0 + static_cast<int64_t>(g - P.group0) * P.cap;
                const uint32_t local0 = static_cast<uint32_t>(g - P.group0) * kWave;
                for (int j = 0; j < c && b + j < S; ++j) {
                    const uint32_t local = local0 + static_cast<uint32_t>(a.reg_lane[rb + j]);
                    keys[b + j] = (static_cast<unsigned long long>(pop) << 56) |
                                  (static_cast<unsigned long long>(__float_as_uint(a.reg_e[rb + j])) << 24) | local;
                }
            }
        }
        for (int64_t i = S + tid; i < n2; i += kSortThreads) keys[i] = ~0ull;
        __syncthreads();
        block_sort<L>(keys, n2, s_k, in_lds);
        if (multi) {
            const int hs = static_cast<int>(k % a.n_hist);
            unsigned long long* hk = a.hist_key + static_cast<int64_t>(hs) * a.keys_cap;
            for (int i = tid; i < S; i += kSortThreads) hk[i] = keys[i];
            if (tid <= a.n_pops) {
                int32_t* ho = a.hist_off + static_cast<int64_t>(hs) * (a.n_pops + 1);
                ho[tid] = tid < a.n_pops ? a.group_prefix[s_pops[tid].group0] : S;
            }
            if (tid == 0) a.hist_win[hs] = k;
        }
        if (a.rec_node != nullptr) {
            const int rc0 = *a.rec_count;
            const bool rec_ok = static_cast<int64_t>(rc0) + S <= a.rec_cap;
            for (int i = tid; rec_ok && i < S; i += kSortThreads) {
                const unsigned long long key = keys[i];
                const int pop = static_cast<int>(key >> 56);
                a.rec_node[rc0 + i] = s_pops[pop].first + static_cast<int32_t>(key & 0xFFFFFFull);
                a.rec_e[rc0 + i] = __uint_as_float(static_cast<uint32_t>((key >> 24) & 0xFFFFFFFFull));
            }
            __syncthreads();
            if (tid == 0) {
                if (rec_ok) *a.rec_count = rc0 + S;
                else atomicOr(&a.flags[N2H_EV_FLAG_RECORD], 1);
                a.rec_wcount[k % a.rec_windows] = S;
            }
        }
        __syncthreads();
    }
    if (tid == 0) atomicAdd(&a.counters[1], static_cast<unsigned long long>(S));
    if (trow) trow[2] = wall_clock64();
    // 3. one stream per distinct source set, in (arrival, connection, neuron) order: keys
    //    arrival bits << 32 | connection << 28 | source slot.  Up to kBinCap spikes: a counting sort into
    //    kBins arrival-time bins (monotone in the arrival), then each bin's few keys sorted by one thread;
    //    more spikes, or a crowded bin: the bitonic network.
    const float inv_bw = static_cast<float>(kBins) / a.W;
    for (int si = 0; si < a.n_streams; ++si) {
        const EvStreamDev& St = a.streams[si];
        if (tid < a.n_pops) s_c[tid] = -1;
        if (tid == 0) s_n = 0;
        for (int b = tid; b < kBins; b += kSortThreads) s_bin[b] = 0;
        __syncthreads();
        if (tid < St.n_src) s_c[St.src[tid]] = tid;
        // source j's spikes: of this window (lag 1, from the regions or the history), or of window
        // k + 1 - lag from the history; hs_j its history slot (-1: none yet), jb_j its first key there
        int ns = 0;
        int hs_j[N2H_EV_MAX_SOURCES], jb_j[N2H_EV_MAX_SOURCES], nb_j[N2H_EV_MAX_SOURCES + 1];
        for (int j = 0; j < St.n_src; ++j) {
            const EvPopDev& Q = s_pops[St.src[j]];
            nb_j[j] = ns;
            if (!multi) {
                ns += a.group_prefix[Q.group0 + Q.n_groups] - a.group_prefix[Q.group0];
                continue;
            }
            const int64_t w = k + 1 - St.lag[j];
            const int hs = static_cast<int>(((w % a.n_hist) + a.n_hist) % a.n_hist);
            hs_j[j] = (w >= 0 && a.hist_win[hs] == w) ? hs : -1;
            if (hs_j[j] >= 0) {
                const int32_t* ho = a.hist_off + static_cast<int64_t>(hs) * (a.n_pops + 1);
                jb_j[j] = ho[St.src[j]];
                ns += ho[St.src[j] + 1] - ho[St.src[j]];
            }
        }
        nb_j[St.n_src] = ns;
        const bool binned = ns <= kBinCap;
        int64_t n2 = pow2_at_least(ns > 0 ? ns : 1);
        if (n2 < kSmallSort) n2 = kSmallSort;
        const bool in_lds = binned || n2 <= L;
        unsigned long long* keys = in_lds ? s_k : a.keys;
        const float a0 = window_time(k + 1, a.W);  // (the earliest arrival of the next window)
        __syncthreads();
        if (!multi && lds_prefix) {  // one thread per spike of the stream's sources
            int nb_r[N2H_EV_MAX_SOURCES];  // (a register copy: the source search below indexes it statically)
#pragma unroll
            for (int q = 0; q < N2H_EV_MAX_SOURCES; ++q) nb_r[q] = q < St.n_src ? nb_j[q] : INT_MAX;
            for (int i = tid; i < ns; i += kSortThreads) {
                int j = 0;
#pragma unroll
                for (int q = 1; q < N2H_EV_MAX_SOURCES; ++q) j = (i >= nb_r[q]) ? q : j;
                const int pop = St.src[j];
                const EvPopDev& Q = s_pops[pop];
                int nbj = nb_r[0];
#pragma unroll
                for (int q = 1; q < N2H_EV_MAX_SOURCES; ++q) nbj = (q == j) ? nb_r[q] : nbj;
                const int ip = s_gp[Q.group0] + (i - nbj);  // its entry in the window's list
                const int g = group_of_entry(ip, Q.group0, Q.group0 + Q.n_groups);
                const int64_t rb = Q.region0 + static_cast<int64_t>(g - Q.group0) * Q.cap + (ip - s_gp[g]);
                const float arr = a.reg_e[rb] + St.delay[j];
                const uint32_t slot = static_cast<uint32_t>(g) * kWave + static_cast<uint32_t>(a.reg_lane[rb]);
                keys[i] = (static_cast<unsigned long long>(__float_as_uint(arr)) << 32) |
                          (static_cast<unsigned long long>(j) << 28) | slot;
                if (binned) atomicAdd(&s_bin[arrival_bin(arr, a0, inv_bw)], 1);
            }
        } else if (!multi) {
            for (int g = tid; g < G; g += kSortThreads) {
                const int pop = a.group_info[g] >> 16;
                const int cix = s_c[pop];
                const int c = cix < 0 ? 0 : a.reg_count[g];
                if (c == 0) continue;
                const EvPopDev& P = s_pops[pop];
                const int64_t rb = P.region0 + static_cast<int64_t>(g - P.group0) * P.cap;
                const int p0 = atomicAdd(&s_n, c);
                const float dl = St.delay[cix];
                for (int j = 0; j < c && p0 + j < ns; ++j) {
                    const float arr = a.reg_e[rb + j] + dl;
                    const uint32_t slot = static_cast<uint32_t>(g) * kWave + static_cast<uint32_t>(a.reg_lane[rb + j]);
                    keys[p0 + j] = (static_cast<unsigned long long>(__float_as_uint(arr)) << 32) |
                                   (static_cast<unsigned long long>(cix) << 28) | slot;
                    if (binned) atomicAdd(&s_bin[arrival_bin(arr, a0, inv_bw)], 1);
                }
            }
        } else {
            for (int i = tid; i < ns; i += kSortThreads) {
                int j = 0;
                while (i >= nb_j[j + 1]) ++j;
                const unsigned long long ka =
                    a.hist_key[static_cast<int64_t>(hs_j[j]) * a.keys_cap + jb_j[j] + (i - nb_j[j])];
                const EvPopDev& Q = s_pops[static_cast<int>(ka >> 56)];
                const uint32_t local = static_cast<uint32_t>(ka & 0xFFFFFFull);
                const float arr = __uint_as_float(static_cast<uint32_t>((ka >> 24) & 0xFFFFFFFFull)) + St.delay[j];
                const uint32_t slot = (static_cast<uint32_t>(Q.group0) + local / kWave) * kWave + (local & (kWave - 1));
                keys[i] = (static_cast<unsigned long long>(__float_as_uint(arr)) << 32) |
                          (static_cast<unsigned long long>(j) << 28) | slot;
                if (binned) atomicAdd(&s_bin[arrival_bin(arr, a0, inv_bw)], 1);
            }
        }
        __syncthreads();
        if (trow && si == 0) trow[5] = wall_clock64();
        const unsigned long long* out = keys;
        bool crowded = false;
        if (binned) {
            bool big = false;
            for (int b = tid; b < kBins; b += kSortThreads) big = big || s_bin[b] > kBinRankMax;
            crowded = __syncthreads_or(big) != 0;
        }
        if (binned && !crowded) {
            unsigned long long* el2 = s_k + kBinCap;
            int cnt2[kBins / kSortThreads];
            int sum = 0;
#pragma unroll
            for (int j = 0; j < kBins / kSortThreads; ++j) {
                cnt2[j] = s_bin[tid * (kBins / kSortThreads) + j];
                sum += cnt2[j];
            }
            int total;
            int off = block_excl_scan(sum, s_w, &total);
#pragma unroll
            for (int j = 0; j < kBins / kSortThreads; ++j) {
                const int b = tid * (kBins / kSortThreads) + j;
                s_bin[b] = off;
                s_fill[b] = 0;
                off += cnt2[j];
            }
            if (tid == 0) s_bin[kBins] = total;
            __syncthreads();
            for (int i = tid; i < ns; i += kSortThreads) {
                const unsigned long long key = keys[i];
                const int b = arrival_bin(__uint_as_float(static_cast<uint32_t>(key >> 32)), a0, inv_bw);
                el2[s_bin[b] + atomicAdd(&s_fill[b], 1)] = key;
            }
            __syncthreads();
            // each key's place: its bin's start + the keys of its bin below it (keys are distinct; a bin's
            // keys are read by all its threads at once, LDS broadcasts).  Spikes at one exact time (a
            // neuron fires at the arrival of the input that crosses its threshold) crowd one bin.
            static_assert(kBinCap <= 2 * kSortThreads, "two keys a thread");
            const int i0 = tid, i1 = tid + kSortThreads;
            const unsigned long long x0 = i0 < ns ? keys[i0] : 0ull;
            const unsigned long long x1 = i1 < ns ? keys[i1] : 0ull;
            __syncthreads();  // every key read before the placement overwrites keys
            if (trow && si == 0) trow[6] = wall_clock64();
            auto place = [&](unsigned long long x) {
                const int b = arrival_bin(__uint_as_float(static_cast<uint32_t>(x >> 32)), a0, inv_bw);
                const int hi = s_bin[b + 1];
                int r = s_bin[b];
                for (int q = r; q < hi; ++q) r += (el2[q] < x) ? 1 : 0;
                keys[r] = x;
            };
            if (i0 < ns) place(x0);
            if (i1 < ns) place(x1);
            __syncthreads();
            out = keys;
        } else {
            for (int64_t i = ns + tid; i < n2; i += kSortThreads) keys[i] = ~0ull;
            __syncthreads();
            block_sort<L>(keys, n2, s_k, in_lds && n2 <= L);
        }
        if (trow && si == 0) trow[3] = wall_clock64();
        for (int i = tid; i < ns; i += kSortThreads) {
            const unsigned long long kb = out[i];
            a.stream_slot[St.base + i] = static_cast<int32_t>(kb & 0x0FFFFFFFull);
            a.stream_a[St.base + i] = __uint_as_float(static_cast<uint32_t>(kb >> 32));
        }
        if (tid == 0) a.stream_count[si] = ns;
        __syncthreads();
    }
    if (trow) trow[4] = wall_clock64();
    if (tid == 0) *a.window = k + 1;
    if (trow) trow[7] = wall_clock64();
}

__global__ void ev_set_kernel(int64_t* window, int64_t k, int32_t* stream_count, int n_streams, int64_t* hist_win,
                              int n_hist, int clear) {
    *window = k;
    if (clear) {
        for (int i = 0; i < n_streams; ++i) stream_count[i] = 0;
        for (int i = 0; hist_win && i < n_hist; ++i) hist_win[i] = -1;
    }
}

__global__ void ev_math_kernel(const float* x, float* ex, const float* u, float* lg, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        ex[i] = ev_exp(x[i]);
        lg[i] = ev_log(u[i]);
    }
}

}  // namespace

struct n2h_ev_plan {
    EvArgs args;
    int32_t n_neurons = 0;
    int32_t sort_lds = 0;         // keys sorted in LDS (8192 or 4096)
    std::vector<EvPopDev> pops;
    std::vector<EvStreamDev> streams;
    EvPopDev* pops_dev = nullptr;
    EvStreamDev* streams_dev = nullptr;
    int32_t* reg_lane = nullptr;
    float* reg_e = nullptr;
    int32_t* reg_count = nullptr;
    int32_t* stream_count = nullptr;
    int32_t* stream_slot = nullptr;
    float* stream_a = nullptr;
    unsigned long long* keys = nullptr;
    int32_t* group_prefix = nullptr;
    int64_t* window = nullptr;
    int32_t* flags = nullptr;
    unsigned long long* counters = nullptr;
    int32_t* log_conn = nullptr;
    unsigned long long* hist_key = nullptr;
    int32_t* hist_off = nullptr;
    int64_t* hist_win = nullptr;
    int64_t host_window = 0;
    bool bound = false;
};

struct n2h_ev_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    n2h_ev_plan* plan;
    int32_t n_windows;
};

namespace {

void free_plan(n2h_ev_plan* p) {
    if (!p) return;
    void* bufs[] = {p->pops_dev, p->streams_dev, p->reg_lane, p->reg_e, p->reg_count, p->stream_count, p->stream_slot,
                    p->stream_a, p->keys, p->group_prefix, p->window, p->flags, p->counters, p->log_conn,
                    p->hist_key, p->hist_off, p->hist_win};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

template <typename T>
int dev_alloc(T** out, int64_t count) {
    const size_t bytes = static_cast<size_t>(count > 0 ? count : 1) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(out), bytes);
    if (e != hipSuccess) return fail(N2H_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipMemset(*out, 0, bytes);
    if (e != hipSuccess) return fail(N2H_EHIP, "hipMemset: %s", hipGetErrorString(e));
    return N2H_OK;
}

// Enqueue n windows.  ev: null, or 3 events per window recorded before the deliver kernel, between the two
// kernels and after the sort kernel (timing).
int launch_windows(n2h_ev_plan* p, int32_t n, hipStream_t stream, hipEvent_t* ev = nullptr) {
    const int k = p->args.xcd;
    const int64_t span = static_cast<int64_t>(k) * kXcdRun;
    const int64_t blocks = k > 1 ? (p->args.n_groups + span - 1) / span * span : p->args.n_groups;
    const dim3 grid(static_cast<unsigned>(blocks)), block(kThreads);
    const size_t lds = static_cast<size_t>(p->sort_lds) * sizeof(unsigned long long);
    for (int32_t i = 0; i < n; ++i) {
        if (ev) EV_HIP(hipEventRecord(ev[3 * i], stream));
        if (p->args.log_rec) hipLaunchKernelGGL(ev_deliver_kernel<true>, grid, block, 0, stream, p->args);
        else hipLaunchKernelGGL(ev_deliver_kernel<false>, grid, block, 0, stream, p->args);
        if (ev) EV_HIP(hipEventRecord(ev[3 * i + 1], stream));
        if (p->sort_lds == 8192)
            hipLaunchKernelGGL(ev_sort_kernel<8192>, dim3(1), dim3(kSortThreads), lds, stream, p->args);
        else
            hipLaunchKernelGGL(ev_sort_kernel<4096>, dim3(1), dim3(kSortThreads), lds, stream, p->args);
        if (ev) EV_HIP(hipEventRecord(ev[3 * i + 2], stream));
    }
    EV_HIP(hipGetLastError());
    return N2H_OK;
}

}  // namespace

extern "C" {

int n2h_ev_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return N2H_EINVAL;
    snprintf(buf, len, "%s", g_err);
    return N2H_OK;
}

int n2h_ev_create(const n2h_ev_desc* d, n2h_ev_plan** out) {
    if (!d || !out) return fail(N2H_EINVAL, "null argument");
    *out = nullptr;
    if (d->n_neurons < 0 || d->n_pops < 1 || d->n_pops > N2H_MAX_POPS || d->n_groups < 1 || d->n_streams < 0)
        return fail(N2H_EINVAL, "bad sizes (neurons %d, populations %d, groups %d, streams %d)", d->n_neurons,
                    d->n_pops, d->n_groups, d->n_streams);
    if (!d->pop_first || !d->pops || (d->n_streams && !d->streams) || !d->group_first || !d->group_info ||
        !d->seg_off || (d->n_links && !d->links))
        return fail(N2H_EINVAL, "a required pointer is null");
    if (!(d->window > 0.0f)) return fail(N2H_EINVAL, "the window must be positive");
    if (d->pop_first[0] != 0 || d->pop_first[d->n_pops] != d->n_neurons)
        return fail(N2H_EINVAL, "pop_first must run from 0 to n_neurons");
    n2h_ev_plan* p = new (std::nothrow) n2h_ev_plan();
    if (!p) return fail(N2H_ENOMEM, "plan");
    int rc = N2H_OK;
    std::vector<int32_t> info(static_cast<size_t>(d->n_groups)), first(static_cast<size_t>(d->n_groups));
    p->n_neurons = d->n_neurons;
    p->pops.resize(static_cast<size_t>(d->n_pops));
    if (hipMemcpy(info.data(), d->group_info, info.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(first.data(), d->group_first, first.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        free_plan(p);
        return fail(N2H_EHIP, "reading the group layout");
    }
    for (int32_t q = 0; q < d->n_pops; ++q) {
        const n2h_ev_pop& s = d->pops[q];
        EvPopDev& t = p->pops[static_cast<size_t>(q)];
        memset(&t, 0, sizeof(t));
        if (s.kind != N2H_EV_NEURON && s.kind != N2H_EV_POISSON && s.kind != N2H_EV_TRAIN) {
            free_plan(p);
            return fail(N2H_EINVAL, "population %d: unknown kind %d", q, s.kind);
        }
        if (s.region_cap < 1 || s.region_cap > (1 << 20)) {
            free_plan(p);
            return fail(N2H_EINVAL, "population %d: region_cap %d out of [1, 2^20]", q, s.region_cap);
        }
        if (s.kind == N2H_EV_NEURON && (s.stream < -1 || s.stream >= d->n_streams)) {
            free_plan(p);
            return fail(N2H_EINVAL, "population %d: stream %d out of range", q, s.stream);
        }
        if (s.kind == N2H_EV_TRAIN && (!d->train_ptr || !d->train_t)) {
            free_plan(p);
            return fail(N2H_EINVAL, "population %d: spike trains need train_ptr and train_t", q);
        }
        if (d->pop_first[q + 1] - d->pop_first[q] >= (1 << 24)) {
            free_plan(p);
            return fail(N2H_EINVAL, "population %d: at most 2^24 - 1 neurons", q);
        }
        t.kind = s.kind;
        t.cap = s.region_cap;
        t.tau_inv = s.tau_inv;
        t.theta = s.theta;
        t.v_reset = s.v_reset;
        t.t_ref = s.t_ref;
        t.rate_inv = s.rate_inv;
        t.stream = s.kind == N2H_EV_NEURON ? s.stream : -1;
        t.first = d->pop_first[q];
        t.n = d->pop_first[q + 1] - d->pop_first[q];
        t.group0 = -1;
    }
    // every population's groups: contiguous, in population order, each within its population
    for (int32_t g = 0; g < d->n_groups; ++g) {
        const int q = info[static_cast<size_t>(g)] >> 16, c = info[static_cast<size_t>(g)] & 0xFFFF;
        if (q < 0 || q >= d->n_pops || c < 1 || c > kWave) {
            free_plan(p);
            return fail(N2H_EINVAL, "group %d: bad group_info", g);
        }
        EvPopDev& t = p->pops[static_cast<size_t>(q)];
        if (t.group0 < 0) t.group0 = g;
        else if (t.group0 + t.n_groups != g) {
            free_plan(p);
            return fail(N2H_EINVAL, "population %d: groups not contiguous", q);
        }
        if (first[static_cast<size_t>(g)] != t.first + kWave * t.n_groups) {
            free_plan(p);
            return fail(N2H_EINVAL, "group %d: group_first must be 64 neurons apart within a population", g);
        }
        ++t.n_groups;
    }
    int64_t regions = 0;
    for (auto& t : p->pops) {
        if (t.group0 < 0) t.group0 = 0;
        t.region0 = regions;
        regions += static_cast<int64_t>(t.n_groups) * t.cap;
    }
    int64_t stream_total = 0;
    int32_t max_lag = 1;
    p->streams.resize(static_cast<size_t>(d->n_streams));
    for (int32_t si = 0; si < d->n_streams; ++si) {
        const n2h_ev_stream& s = d->streams[si];
        EvStreamDev& t = p->streams[static_cast<size_t>(si)];
        memset(&t, 0, sizeof(t));
        if (s.n_sources < 1 || s.n_sources > N2H_EV_MAX_SOURCES) {
            free_plan(p);
            return fail(N2H_EINVAL, "stream %d: 1 to %d source populations", si, N2H_EV_MAX_SOURCES);
        }
        t.n_src = s.n_sources;
        t.base = stream_total;
        for (int j = 0; j < s.n_sources; ++j) {
            const int q = s.source_pop[j];
            const double lagd = static_cast<double>(s.lag[j]) * static_cast<double>(d->window);
            if (s.lag[j] < 1 || s.lag[j] > N2H_EV_MAX_LAG || !(s.delay[j] > 0.0f) ||
                std::fabs(lagd - static_cast<double>(s.delay[j])) > 1e-5 * static_cast<double>(s.delay[j])) {
                free_plan(p);
                return fail(N2H_EINVAL, "stream %d source %d: delay %g must be lag %d (1..%d) windows of %g", si, j,
                            static_cast<double>(s.delay[j]), s.lag[j], N2H_EV_MAX_LAG,
                            static_cast<double>(d->window));
            }
            t.delay[j] = s.delay[j];
            t.lag[j] = s.lag[j];
            if (s.lag[j] > max_lag) max_lag = s.lag[j];
            for (int j2 = 0; j2 < j; ++j2)
                if (s.source_pop[j2] == q) {
                    free_plan(p);
                    return fail(N2H_EINVAL, "stream %d: population %d twice", si, q);
                }
            if (q < 0 || q >= d->n_pops) {
                free_plan(p);
                return fail(N2H_EINVAL, "stream %d: source population %d out of range", si, q);
            }
            t.src[j] = q;
            stream_total += static_cast<int64_t>(p->pops[static_cast<size_t>(q)].n_groups) * p->pops[static_cast<size_t>(q)].cap;
        }
    }
    int64_t keys_cap = 1;
    while (keys_cap < regions) keys_cap <<= 1;
    if (keys_cap >= (1ll << 31)) {
        free_plan(p);
        return fail(N2H_EINVAL, "%lld region entries: at most 2^30 per window", static_cast<long long>(regions));
    }
    int dev = 0, lds_max = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    p->sort_lds = lds_max >= 8192 * 8 + 40 * 1024 ? 8192 : 4096;  // (+ the sort kernel's static LDS, ~36 KB)
    if (p->sort_lds == 8192 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(ev_sort_kernel<8192>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 8192 * 8) != hipSuccess)
        p->sort_lds = 4096;
    (void)hipGetLastError();
    if ((rc = dev_alloc(&p->pops_dev, d->n_pops)) || (rc = dev_alloc(&p->streams_dev, d->n_streams)) ||
        (rc = dev_alloc(&p->reg_lane, regions)) || (rc = dev_alloc(&p->reg_e, regions)) ||
        (rc = dev_alloc(&p->reg_count, d->n_groups)) || (rc = dev_alloc(&p->stream_count, d->n_streams)) ||
        (rc = dev_alloc(&p->stream_slot, stream_total)) || (rc = dev_alloc(&p->stream_a, stream_total)) ||
        (rc = dev_alloc(&p->keys, keys_cap)) ||
        (rc = dev_alloc(&p->group_prefix, d->n_groups + 1)) || (rc = dev_alloc(&p->window, 1)) ||
        (rc = dev_alloc(&p->flags, N2H_EV_FLAGS)) || (rc = dev_alloc(&p->counters, 2)) ||
        (max_lag > 1 && ((rc = dev_alloc(&p->hist_key, keys_cap * max_lag)) ||
                         (rc = dev_alloc(&p->hist_off, static_cast<int64_t>(max_lag) * (d->n_pops + 1))) ||
                         (rc = dev_alloc(&p->hist_win, max_lag))))) {
        free_plan(p);
        return rc;
    }
    if (hipMemcpy(p->pops_dev, p->pops.data(), p->pops.size() * sizeof(EvPopDev), hipMemcpyHostToDevice) != hipSuccess ||
        (d->n_streams && hipMemcpy(p->streams_dev, p->streams.data(), p->streams.size() * sizeof(EvStreamDev),
                                   hipMemcpyHostToDevice) != hipSuccess)) {
        free_plan(p);
        return fail(N2H_EHIP, "copying the parameters");
    }
    EvArgs& a = p->args;
    memset(&a, 0, sizeof(a));
    a.n_groups = d->n_groups;
    a.n_pops = d->n_pops;
    a.n_streams = d->n_streams;
    a.rec_windows = 1;
    a.group_first = d->group_first;
    a.group_info = d->group_info;
    a.seg_off = d->seg_off;
    a.links = reinterpret_cast<const uint2*>(d->links);
    a.train_ptr = d->train_ptr;
    a.train_t = d->train_t;
    a.pops = p->pops_dev;
    a.streams = p->streams_dev;
    a.reg_lane = p->reg_lane;
    a.reg_e = p->reg_e;
    a.reg_count = p->reg_count;
    a.stream_count = p->stream_count;
    a.stream_slot = p->stream_slot;
    a.stream_a = p->stream_a;
    a.keys = p->keys;
    a.keys_cap = keys_cap;
    a.group_prefix = p->group_prefix;
    a.window = p->window;
    a.flags = p->flags;
    a.counters = p->counters;
    a.n_hist = max_lag;
    a.hist_key = p->hist_key;
    a.hist_off = p->hist_off;
    a.hist_win = p->hist_win;
    if (max_lag > 1) {
        std::vector<int64_t> none(static_cast<size_t>(max_lag), -1);
        if (hipMemcpy(p->hist_win, none.data(), none.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess) {
            free_plan(p);
            return fail(N2H_EHIP, "clearing the window history");
        }
    }
    a.W = d->window;
    a.key0 = static_cast<uint32_t>(d->seed);
    a.key1 = static_cast<uint32_t>(d->seed >> 32);
    a.instance = d->instance;
    *out = p;
    return N2H_OK;
}

int n2h_ev_destroy(n2h_ev_plan* p) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    (void)hipDeviceSynchronize();
    free_plan(p);
    return N2H_OK;
}

int n2h_ev_bind_state(n2h_ev_plan* p, float* v, float* last_time, float* t_ref_end, float* next_spike) {
    if (!p || !v || !last_time || !t_ref_end || !next_spike) return fail(N2H_EINVAL, "null argument");
    p->args.v = v;
    p->args.last = last_time;
    p->args.tre = t_ref_end;
    p->args.nxt = next_spike;
    p->bound = true;
    return N2H_OK;
}

int n2h_ev_bind_record(n2h_ev_plan* p, int32_t* node, float* time, int32_t* count, int64_t capacity,
                       int32_t* window_counts, int32_t n_window_slots) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (!node) {
        p->args.rec_node = nullptr;
        p->args.rec_e = nullptr;
        p->args.rec_count = nullptr;
        p->args.rec_wcount = nullptr;
        p->args.rec_cap = 0;
        p->args.rec_windows = 1;
        return N2H_OK;
    }
    if (!time || !count || !window_counts || capacity < 0 || n_window_slots < 1)
        return fail(N2H_EINVAL, "record: time, count and window_counts are required, capacity >= 0, slots >= 1");
    p->args.rec_node = node;
    p->args.rec_e = time;
    p->args.rec_count = count;
    p->args.rec_wcount = window_counts;
    p->args.rec_cap = capacity;
    p->args.rec_windows = n_window_slots;
    return N2H_OK;
}

int n2h_ev_set_window(n2h_ev_plan* p, int64_t window, int32_t clear_inputs, void* stream) {
    if (!p || window < 0) return fail(N2H_EINVAL, "bad argument");
    hipLaunchKernelGGL(ev_set_kernel, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), p->window, window,
                       p->stream_count, p->args.n_streams, p->hist_win, p->args.n_hist, clear_inputs ? 1 : 0);
    EV_HIP(hipGetLastError());
    p->host_window = window;
    return N2H_OK;
}

int n2h_ev_get_window(n2h_ev_plan* p, int64_t* window) {
    if (!p || !window) return fail(N2H_EINVAL, "null argument");
    EV_HIP(hipMemcpy(window, p->window, sizeof(int64_t), hipMemcpyDeviceToHost));
    return N2H_OK;
}

int n2h_ev_run(n2h_ev_plan* p, int32_t n_windows, void* stream) {
    if (!p || n_windows < 0) return fail(N2H_EINVAL, "bad argument");
    if (!p->bound) return fail(N2H_ESTATE, "state not bound (n2h_ev_bind_state)");
    int rc = launch_windows(p, n_windows, static_cast<hipStream_t>(stream));
    if (rc) return rc;
    p->host_window += n_windows;
    return N2H_OK;
}

int n2h_ev_graph_create(n2h_ev_plan* p, int32_t n_windows, void* stream, n2h_ev_graph** out) {
    (void)stream;
    if (!p || n_windows < 1 || !out) return fail(N2H_EINVAL, "bad argument");
    if (!p->bound) return fail(N2H_ESTATE, "state not bound (n2h_ev_bind_state)");
    *out = nullptr;
    hipStream_t cap;
    EV_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cap);
        return fail(N2H_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    int rc = launch_windows(p, n_windows, cap);
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
    n2h_ev_graph* gr = new (std::nothrow) n2h_ev_graph{graph, exec, p, n_windows};
    if (!gr) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        return fail(N2H_ENOMEM, "graph");
    }
    *out = gr;
    return N2H_OK;
}

int n2h_ev_graph_launch(n2h_ev_graph* g, void* stream) {
    if (!g) return fail(N2H_EINVAL, "null graph");
    EV_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    g->plan->host_window += g->n_windows;
    return N2H_OK;
}

int n2h_ev_graph_destroy(n2h_ev_graph* g) {
    if (!g) return N2H_OK;
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
    return N2H_OK;
}

int n2h_ev_status(n2h_ev_plan* p, int32_t* flags, int32_t n) {
    if (!p || !flags || n < 1) return fail(N2H_EINVAL, "bad argument");
    int32_t f[N2H_EV_FLAGS];
    EV_HIP(hipMemcpy(f, p->flags, sizeof(f), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) flags[i] = i < N2H_EV_FLAGS ? f[i] : 0;
    return N2H_OK;
}

int n2h_ev_bind_log(n2h_ev_plan* p, const int32_t* conn_index, int32_t* records, float* arrival, int32_t* count,
                    int64_t capacity) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (!records) {
        p->args.log_rec = nullptr;
        p->args.log_a = nullptr;
        p->args.log_count = nullptr;
        p->args.log_cap = 0;
        return N2H_OK;
    }
    if (!conn_index || !arrival || !count || capacity < 0)
        return fail(N2H_EINVAL, "log: conn_index, arrival and count are required, capacity >= 0");
    const int n2 = p->args.n_pops * p->args.n_pops;
    if (!p->log_conn) {
        const int rc = dev_alloc(&p->log_conn, n2);
        if (rc) return rc;
    }
    EV_HIP(hipMemcpy(p->log_conn, conn_index, static_cast<size_t>(n2) * sizeof(int32_t), hipMemcpyHostToDevice));
    p->args.log_conn = p->log_conn;
    p->args.log_rec = records;
    p->args.log_a = arrival;
    p->args.log_count = count;
    p->args.log_cap = capacity;
    return N2H_OK;
}

int n2h_ev_set_tuning(n2h_ev_plan* p, const n2h_tuning* t) {
    if (!p || !t) return fail(N2H_EINVAL, "null argument");
    if (t->xcd < 0 || t->xcd > 64) return fail(N2H_EINVAL, "xcd %d out of [0, 64]", t->xcd);
    p->args.xcd = t->xcd;
    return N2H_OK;
}

int n2h_ev_set_trace(n2h_ev_plan* p, uint64_t* buffer) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    p->args.trace = reinterpret_cast<unsigned long long*>(buffer);
    return N2H_OK;
}

int n2h_ev_counters(n2h_ev_plan* p, int64_t* counts, int32_t n) {
    if (!p || !counts || n < 1) return fail(N2H_EINVAL, "bad argument");
    unsigned long long c[2];
    EV_HIP(hipMemcpy(c, p->counters, sizeof(c), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) counts[i] = i < 2 ? static_cast<int64_t>(c[i]) : 0;
    return N2H_OK;
}

int n2h_ev_run_timed(n2h_ev_plan* p, int32_t n_windows, void* stream, double* deliver_ms, double* sort_ms) {
    if (!p || n_windows < 1 || !deliver_ms || !sort_ms) return fail(N2H_EINVAL, "bad argument");
    if (!p->bound) return fail(N2H_ESTATE, "state not bound (n2h_ev_bind_state)");
    std::vector<hipEvent_t> ev(static_cast<size_t>(3 * n_windows), nullptr);
    int rc = N2H_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = fail(N2H_EHIP, "hipEventCreate");
    if (!rc) rc = launch_windows(p, n_windows, static_cast<hipStream_t>(stream), ev.data());
    if (!rc && hipEventSynchronize(ev.back()) != hipSuccess) rc = fail(N2H_EHIP, "hipEventSynchronize");
    double d = 0.0, s = 0.0;
    for (int32_t i = 0; !rc && i < n_windows; ++i) {
        float a = 0.0f, b = 0.0f;
        if (hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 1]) != hipSuccess ||
            hipEventElapsedTime(&b, ev[3 * i + 1], ev[3 * i + 2]) != hipSuccess)
            rc = fail(N2H_EHIP, "hipEventElapsedTime");
        d += a;
        s += b;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    p->host_window += n_windows;
    *deliver_ms = d;
    *sort_ms = s;
    return N2H_OK;
}

int n2h_ev_math(const float* x, float* exp_x, const float* u, float* log_u, int64_t n, void* stream) {
    if (n < 0 || (n && (!x || !exp_x || !u || !log_u))) return fail(N2H_EINVAL, "bad argument");
    if (n == 0) return N2H_OK;
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(ev_math_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), x, exp_x, u,
                       log_u, n);
    EV_HIP(hipGetLastError());
    return N2H_OK;
}

}  // extern "C"
