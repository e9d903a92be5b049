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
//   emission        fused.py:2611-2644 (_emit_spikes): here the step's spikes are one ballot word per
//                   group, so the spike order is the neuron order without atomics.
//
// Summation order (identical in the CPU oracle): for target i and arrival step n, for each emission step s
// ascending, p = +0.0f, p += w over the links into i of sources spiking at s with s + d = n in ascending
// canonical link index; then ring[n % D][i] = ring[n % D][i] + p.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdarg>
#include <new>

#include "naer2_hip.h"

namespace {

constexpr int kWave = 64;
constexpr int kCapS = 512;   // spiking sources staged in LDS per pass
constexpr int kCapL = 1024;  // links staged in LDS per pass

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
    uint32_t key0, key1;
    const int32_t* group_first;
    const int32_t* group_info;
    const int64_t* seg_off;
    const uint2* links;
    const n2h_pop_params* pops;
    float* v;
    float* isyn;
    int32_t* ref;
    float* ring;
    unsigned long long* bits;
    const int64_t* step_ctr;
};

// Philox4x32-10 (Salmon et al., SC'11), word 0 of the output block.
__device__ __forceinline__ uint32_t philox_word0(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                 uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c0;
}

// Inclusive prefix sum over the 64 lanes of the wavefront.
__device__ __forceinline__ int wave_incl_scan(int x, int lane) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

template <int MODEL>
__global__ __launch_bounds__(kWave) void step_kernel(StepArgs a, int32_t k) {
    __shared__ int32_t sp[kCapS];           // spiking sources of this pass, ascending
    __shared__ int64_t seg_lo[kCapS];       // first link of each source's segment into this group
    __shared__ int32_t seg_start[kCapS + 1];  // exclusive scan of the segment lengths
    __shared__ uint2 stage[kCapL];          // links of this pass, in contract order
    extern __shared__ float acc[];          // [d_slots][64] partial sums of this emission step

    const int g = blockIdx.x;
    const int lane = threadIdx.x;
    const int NG = a.n_groups;
    const int64_t N = a.n_neurons;
    const int64_t D = a.d_slots;
    const int64_t s = *a.step_ctr + k;

    const int first = a.group_first[g];
    const int info = a.group_info[g];
    const int count = info & 0xffff;
    const int pop = info >> 16;
    const bool live = lane < count;
    const int nid = first + lane;
    const int64_t cur = (s % D) * N + nid;

    // state and the due ring cell: nothing below writes them before they are used
    float V = 0.0f, I = 0.0f, x = 0.0f;
    int32_t R = 0;
    if (live) {
        V = a.v[nid];
        R = a.ref[nid];
        if (MODEL == N2H_MODEL_LIF_EXP) I = a.isyn[nid];
        x = a.ring[cur];
    }

    // ---- pull intake + delivery of the spikes of step s-1 ----
    const int64_t prev_slot = ((s - 1) % a.bit_slots + a.bit_slots) % a.bit_slots;
    const unsigned long long* prev = a.bits + prev_slot * NG;
    const int W = (NG + kWave - 1) / kWave;
    const int w_lo = lane * W;
    const int w_hi = min(NG, w_lo + W);
    int cnt = 0;
    for (int w = w_lo; w < w_hi; ++w) cnt += __popcll(prev[w]);
    const int incl = wave_incl_scan(cnt, lane);
    const int excl = incl - cnt;
    const int total = __shfl(incl, kWave - 1, kWave);

    unsigned long long t0 = 0ull, t1 = 0ull;  // delays 1..64 and 65..128 touched by this lane's target
    for (int r0 = 0; r0 < total; r0 += kCapS) {
        const int r1 = min(total, r0 + kCapS);
        if (excl < r1 && excl + cnt > r0) {
            int rank = excl;
            for (int w = w_lo; w < w_hi && rank < r1; ++w) {
                unsigned long long m = prev[w];
                const int pc = __popcll(m);
                if (rank + pc <= r0) {
                    rank += pc;
                    continue;
                }
                const int base = a.group_first[w];
                while (m) {
                    const int b = __ffsll(static_cast<long long>(m)) - 1;
                    if (rank >= r0 && rank < r1) sp[rank - r0] = base + b;
                    ++rank;
                    m &= m - 1ull;
                }
            }
        }
        __syncthreads();
        const int n_sp = r1 - r0;
        int carry = 0;
        for (int i0 = 0; i0 < n_sp; i0 += kWave) {
            const int i = i0 + lane;
            int n = 0;
            if (i < n_sp) {
                const int64_t idx = static_cast<int64_t>(sp[i]) * NG + g;
                const int64_t lo = a.seg_off[idx];
                n = static_cast<int>(a.seg_off[idx + 1] - lo);
                seg_lo[i] = lo;
            }
            const int inc = wave_incl_scan(n, lane);
            if (i < n_sp) seg_start[i] = carry + inc - n;
            carry += __shfl(inc, kWave - 1, kWave);
        }
        if (lane == 0) seg_start[n_sp] = carry;
        __syncthreads();
        const int L = carry;
        for (int e0 = 0; e0 < L; e0 += kCapL) {
            const int e1 = min(L, e0 + kCapL);
            for (int e = e0 + lane; e < e1; e += kWave) {
                int lo_i = 0, hi_i = n_sp;  // seg_start[lo_i] <= e < seg_start[hi_i]
                while (hi_i - lo_i > 1) {
                    const int mid = (lo_i + hi_i) >> 1;
                    if (seg_start[mid] <= e) lo_i = mid;
                    else hi_i = mid;
                }
                stage[e - e0] = a.links[seg_lo[lo_i] + (e - seg_start[lo_i])];
            }
            __syncthreads();
            const int nl = e1 - e0;
            for (int e = 0; e < nl; ++e) {
                const uint2 ln = stage[e];
                if (static_cast<int>(ln.x & 0xffffu) == lane) {
                    const int b = static_cast<int>(ln.x >> 16) - 1;
                    const float wv = __uint_as_float(ln.y);
                    bool seen;
                    if (b < 64) {
                        seen = (t0 >> b) & 1ull;
                        t0 |= 1ull << b;
                    } else {
                        seen = (t1 >> (b - 64)) & 1ull;
                        t1 |= 1ull << (b - 64);
                    }
                    float* c = &acc[b * kWave + lane];
                    *c = seen ? *c + wv : 0.0f + wv;
                }
            }
            __syncthreads();
        }
    }

    // ---- ring: add each touched cell's partial sum; the delay-1 cell is the one due now ----
    if (live) {
        unsigned long long m = t0 & ~1ull;
        while (m) {
            const int b = __ffsll(static_cast<long long>(m)) - 1;
            m &= m - 1ull;
            const int64_t idx = ((s + b) % D) * N + nid;  // arrival s - 1 + (b + 1)
            a.ring[idx] = a.ring[idx] + acc[b * kWave + lane];
        }
        m = t1;
        while (m) {
            const int b = __ffsll(static_cast<long long>(m)) - 1 + 64;
            m &= m - 1ull;
            const int64_t idx = ((s + b) % D) * N + nid;
            a.ring[idx] = a.ring[idx] + acc[b * kWave + lane];
        }
        if (t0 & 1ull) x = x + acc[lane];
        a.ring[cur] = 0.0f;
    }

    // ---- background, LIF update, threshold ----
    const n2h_pop_params& P = a.pops[pop];
    bool fired = false;
    if (live) {
        const uint32_t u = philox_word0(static_cast<uint32_t>(nid), static_cast<uint32_t>(s),
                                        static_cast<uint32_t>(static_cast<uint64_t>(s) >> 32), a.instance,
                                        a.key0, a.key1);
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
                if (V >= P.v_th) {
                    fired = true;
                    V = P.v_reset;
                    R = P.ref_steps;
                }
            }
        } else {
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
            }
            I = P.p_i * I;
            if (V > P.v_th) {
                fired = true;
                V = P.v_reset;
                R = P.ref_steps;
            }
        }
        a.v[nid] = V;
        a.ref[nid] = R;
        if (MODEL == N2H_MODEL_LIF_EXP) a.isyn[nid] = I;
    }
    const unsigned long long mask = __ballot(fired);
    if (lane == 0) a.bits[(s % a.bit_slots) * NG + g] = mask;
}

__global__ void advance_kernel(int64_t* ctr, int32_t n) { *ctr += n; }
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
    int64_t bit_slots;
    size_t lds_bytes;
    int device;
};

struct n2h_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    n2h_plan* plan;
    int32_t n_steps;
};

namespace {

int launch_steps(n2h_plan* p, int32_t n_steps, hipStream_t stream) {
    StepArgs a;
    a.n_neurons = p->desc.n_neurons;
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
    a.step_ctr = p->step_ctr;
    const dim3 grid(static_cast<unsigned>(a.n_groups)), block(kWave);
    for (int32_t k = 0; k < n_steps; ++k) {
        if (p->desc.model == N2H_MODEL_LIF_DELTA)
            hipLaunchKernelGGL(step_kernel<N2H_MODEL_LIF_DELTA>, grid, block, p->lds_bytes, stream, a, k);
        else
            hipLaunchKernelGGL(step_kernel<N2H_MODEL_LIF_EXP>, grid, block, p->lds_bytes, stream, a, k);
    }
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, stream, p->step_ctr, n_steps);
    N2H_HIP(hipGetLastError());
    return N2H_OK;
}

int check_ready(n2h_plan* p) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (!p->v || !p->ref || (p->desc.model == N2H_MODEL_LIF_EXP && !p->isyn))
        return fail(N2H_ESTATE, "state not bound (n2h_bind_state)");
    if (!p->ring || !p->bits) return fail(N2H_ESTATE, "ring not bound (n2h_bind_ring)");
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
    }
    if (tables->n_groups <= 0) return fail(N2H_EINVAL, "n_groups must be positive");
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
    p->lds_bytes = sizeof(float) * kWave * static_cast<size_t>(desc->d_slots);
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

int n2h_bind_ring(n2h_plan* p, float* ring, uint64_t* spike_bits, int64_t bit_slots) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (!ring || !spike_bits) return fail(N2H_EINVAL, "null ring pointer");
    if (bit_slots < 2) return fail(N2H_EINVAL, "bit_slots must be at least 2");
    p->ring = ring;
    p->bits = reinterpret_cast<unsigned long long*>(spike_bits);
    p->bit_slots = bit_slots;
    return N2H_OK;
}

int n2h_set_step(n2h_plan* p, int64_t step, void* stream) {
    if (!p) return fail(N2H_EINVAL, "null plan");
    if (step < 0) return fail(N2H_EINVAL, "negative step");
    hipLaunchKernelGGL(set_kernel, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), p->step_ctr, step);
    N2H_HIP(hipGetLastError());
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
    return launch_steps(p, n_steps, static_cast<hipStream_t>(stream));
}

int n2h_graph_create(n2h_plan* p, int32_t n_steps, void* stream, n2h_graph** out) {
    (void)stream;
    if (!out) return fail(N2H_EINVAL, "null out");
    *out = nullptr;
    int rc = check_ready(p);
    if (rc) return rc;
    if (n_steps <= 0) return fail(N2H_EINVAL, "n_steps must be positive");
    hipStream_t cap;
    N2H_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cap);
        return fail(N2H_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    rc = launch_steps(p, n_steps, cap);
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
    n2h_graph* gr = new (std::nothrow) n2h_graph{graph, exec, p, n_steps};
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
    N2H_HIP(hipGraphLaunch(gr->exec, static_cast<hipStream_t>(stream)));
    return N2H_OK;
}

int n2h_graph_destroy(n2h_graph* gr) {
    if (!gr) return N2H_OK;
    (void)hipGraphExecDestroy(gr->exec);
    (void)hipGraphDestroy(gr->graph);
    delete gr;
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
