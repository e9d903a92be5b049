// rlif.hip -- forward and surrogate-gradient adjoint (BPTT) of a dense recurrent LIF layer with
// exponential current synapses, subtractive reset and a leaky readout, on MI355X (gfx950).
//
// Model, t = 0 .. T-1 (all state zero before t = 0):
//   I_t = alpha I_{t-1} + (u_t + W_rec s_{t-1})        u_t = W_in x_t, computed by the caller (a plain GEMM)
//   V_t = (beta V_{t-1} + I_t) - theta s_{t-1}         subtractive reset, one step after the spike
//   s_t = H(V_t - theta)                                mode 0; mode 1 (tests): s_t = f(V_t - theta)
//   J_t = alpha_o J_{t-1} + z_t,  z_t = W_out s_t;   R_t = beta_o R_{t-1} + J_t
// Surrogate derivative (fast sigmoid, Zenke & Ganguli 2018): ds/dV = gamma / (1 + slope |V - theta|)^2,
// which is exactly f'(V - theta) for f(x) = x / (1 + slope |x|) (gamma = 1), so mode 1 is a smooth model
// whose true gradient the adjoint computes.
// Adjoint, t = T-1 .. 0 (gV_T = gI_T = 0):
//   gs_t = W_out^T gJ_t + W_rec^T gI_{t+1} - theta gV_{t+1}
//   gV_t = gs_t ds/dV(V_t) + beta gV_{t+1};   gI_t = gV_t + alpha gI_{t+1}
//
// Precision.  The surrogate's relative sensitivity to V reaches 2 slope (50) at the threshold, so the
// state is carried in compensated fp32: I and V are unevaluated sums hi + lo of two floats (double-float
// arithmetic with error-free TwoSum / TwoProd transformations; Dekker 1971, Knuth TAOCP 4.2.2), the
// recurrent sum is a compensated (Kahan-Babuska style) sum, and the decays are passed as hi + lo pairs.
// That carries about 44 bits of the potential with fp32 arithmetic only; spikes are decided, and the
// surrogate evaluated, on x = (V_hi - theta) + V_lo.
//
// Kernels:
//   step_kernel<false>   one forward step: compacts the nonzero entries of s_{t-1}[b] (ascending) and sums
//                        the matching rows of W_rec^T in that order (event-driven recurrent delivery), then
//                        the LIF update; every K steps the state is checkpointed (4 floats a unit).
//   step_kernel<true>    the same step in the backward pass's recompute, from a checkpoint and the stored
//                        spikes (not re-decided); writes the surrogate derivative h_t of every unit.
//   reverse_kernel       one reverse step: the adjoint of the recurrent delivery m = gI_{t+1} W_rec on the
//                        matrix cores (v_mfma_f32_16x16x4f32, 16x16 output tiles, K split over the four
//                        waves of a workgroup and reduced through LDS in a fixed order), fused with the
//                        elementwise LIF adjoint above.  Deterministic: the same result for any checkpoint
//                        interval and any launch.
// There is no counterpart in the reference: autograd and BPTT do not exist in naer2.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "naer2_hip.h"
#include "twofloat.h"

namespace {

constexpr int kTile = 256;     // hidden units per workgroup of the step kernels
constexpr int kMM = 16;        // output tile edge of the reverse kernel
constexpr int kMaxDevices = 64;

thread_local char r_err[512] = "";

int rfail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(r_err, sizeof(r_err), fmt, ap);
    va_end(ap);
    return code;
}

#define R_HIP(call)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return rfail(N2H_EHIP, "%s: %s", #call, hipGetErrorString(e_));   \
    } while (0)

using n2h_twofloat::df;
using n2h_twofloat::df_add;
using n2h_twofloat::df_mul;
using n2h_twofloat::two_prod;
using n2h_twofloat::two_sum;

__device__ __forceinline__ float spike_fn(float x, int mode, float slope) {
    if (mode == 0) return x > 0.0f ? 1.0f : 0.0f;
    return x / (1.0f + slope * fabsf(x));
}

__device__ __forceinline__ float surrogate(float x, float slope, float gamma) {
    const float d = 1.0f + slope * fabsf(x);
    return gamma / (d * d);
}

// Compact the nonzero entries of s_prev[b] (ascending) into act_idx / act_val; returns their count.
__device__ __forceinline__ int compact_row(const float* __restrict__ s_row, int H, int* act_idx, float* act_val,
                                           int* wave_base) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int n_act = 0;
    for (int base = 0; base < H; base += kTile) {
        const int j = base + threadIdx.x;
        const float sv = j < H ? s_row[j] : 0.0f;
        const bool nz = sv != 0.0f;
        const unsigned long long m = __ballot(nz);
        if (lane == 0) wave_base[wave] = __popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int w = 0; w < kTile / 64; ++w) {
                const int c = wave_base[w];
                wave_base[w] = acc;
                acc += c;
            }
            wave_base[kTile / 64] = acc;
        }
        __syncthreads();
        if (nz) {
            const int pos = n_act + wave_base[wave] + __popcll(m & ((1ull << lane) - 1ull));
            act_idx[pos] = j;
            act_val[pos] = sv;
        }
        n_act += wave_base[kTile / 64];
        __syncthreads();
    }
    return n_act;
}

// One step for batch row b and hidden units [tile*256, tile*256 + 256).  state: [4][B][H] floats
// (I hi, I lo, V hi, V lo).  Forward (kRecompute false): decides s_t into out, checkpoints the state before
// the step into ckpt when given.  Recompute: the forward's spikes s_{t-1} come in, h_t = ds/dV goes out.
template <bool kRecompute>
__global__ __launch_bounds__(kTile) void step_kernel(n2h_rlif p, const float* __restrict__ u_t,
                                                     const float* __restrict__ w_rec_t, float* __restrict__ state,
                                                     const float* __restrict__ s_prev, float* __restrict__ out,
                                                     float* __restrict__ ckpt) {
    extern __shared__ unsigned char smem[];
    int* act_idx = reinterpret_cast<int*>(smem);                    // [H]
    float* act_val = reinterpret_cast<float*>(smem) + p.n_hidden;  // [H]
    __shared__ int wave_base[kTile / 64 + 1];
    const int H = p.n_hidden;
    const int b = blockIdx.y;
    const int h = blockIdx.x * kTile + threadIdx.x;
    const int64_t row = static_cast<int64_t>(b) * H;
    const int64_t bh = static_cast<int64_t>(p.batch) * H;
    const int n_act = s_prev ? compact_row(s_prev + row, H, act_idx, act_val, wave_base) : 0;
    if (h >= H) return;
    // compensated recurrent sum, in ascending presynaptic order
    float rs = 0.0f, rc = 0.0f;
    if (p.mode == 0) {
        for (int q = 0; q < n_act; ++q) {
            const df t = two_sum(rs, w_rec_t[static_cast<int64_t>(act_idx[q]) * H + h]);
            rs = t.hi;
            rc = rc + t.lo;
        }
    } else {
        for (int q = 0; q < n_act; ++q) {
            const df pr = two_prod(w_rec_t[static_cast<int64_t>(act_idx[q]) * H + h], act_val[q]);
            const df t = two_sum(rs, pr.hi);
            rs = t.hi;
            rc = rc + (t.lo + pr.lo);
        }
    }
    const int64_t k = row + h;
    df I{state[k], state[bh + k]}, V{state[2 * bh + k], state[3 * bh + k]};
    if (!kRecompute && ckpt) {  // the segment's starting state
        ckpt[k] = I.hi;
        ckpt[bh + k] = I.lo;
        ckpt[2 * bh + k] = V.hi;
        ckpt[3 * bh + k] = V.lo;
    }
    const float sp = s_prev ? s_prev[k] : 0.0f;
    const df a = df_add(two_sum(u_t[k], rs), df{rc, 0.0f});
    I = df_add(df_mul(df{p.alpha, p.alpha_lo}, I), a);
    V = df_add(df_mul(df{p.beta, p.beta_lo}, V), I);
    V = df_add(V, two_prod(-p.theta, sp));
    state[k] = I.hi;
    state[bh + k] = I.lo;
    state[2 * bh + k] = V.hi;
    state[3 * bh + k] = V.lo;
    const float x = (V.hi - p.theta) + V.lo;
    out[k] = kRecompute ? surrogate(x, p.slope, p.gamma) : spike_fn(x, p.mode, p.slope);
}

// One reverse step, for the 16x16 output tile (rows b0.., columns j0..) of m = gI_{t+1} W_rec and the
// elementwise adjoint over it.  gi_next = gI_{t+1} [B][H] (null at t = T-1: m = 0, gV_{t+1} = gI_{t+1} = 0);
// h_t the surrogate derivatives; gso_t = W_out^T gJ_t; gv [B][H] holds gV_{t+1} on entry and gV_t on
// return; gi_out = gI_t.  epilogue false: writes m itself to gi_out (the delivery adjoint alone).
// The product reads W_rec through its transpose w_rec_t [H][H] (m[b][j] = sum_k gI[b][k] w_rec_t[j][k]), so
// that both operands are rows: H is a multiple of 16, split into chunks of 16 k; wave w takes chunks
// [w nc / 4, (w + 1) nc / 4) of the nc = H / 16.  Lane l (row i = l % 16, quarter g = l / 16) loads the
// float4 k = 16 c + 4 g .. + 3 of gI row b0 + i and of w_rec_t row j0 + i; MFMA (c, q) then takes k =
// 16 c + 4 g + q in lane quarter g, as v_mfma_f32_16x16x4f32 wants (A[i][k] from lane i + 16 k', B[k][j]
// from lane j + 16 k', D[4 (l / 16) + r][l % 16] in accumulator component r).  Eight chunks are loaded
// before their 32 MFMAs.  The four waves' partial tiles are added in wave order through LDS.
template <bool kEpilogue>
__global__ __launch_bounds__(256) void reverse_kernel(n2h_rlif p, const float* __restrict__ gi_next,
                                                      const float* __restrict__ w_rec_t, const float* __restrict__ h_t,
                                                      const float* __restrict__ gso_t, float* __restrict__ gv,
                                                      float* __restrict__ gi_out) {
    __shared__ float part[4][kMM][kMM + 1];
    const int H = p.n_hidden, B = p.batch;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j0 = blockIdx.x * kMM, b0 = blockIdx.y * kMM;
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (gi_next) {
        const int i = lane & 15, g = lane >> 4;
        const int nc = H >> 4;
        const int c0 = nc * wave / 4, c1 = nc * (wave + 1) / 4;
        // rows past the batch are clamped: they only feed output rows that are not written
        const int brow = b0 + i < B ? b0 + i : B - 1;
        const f4* arow = reinterpret_cast<const f4*>(gi_next + static_cast<int64_t>(brow) * H) + g;
        const f4* brw = reinterpret_cast<const f4*>(w_rec_t + static_cast<int64_t>(j0 + i) * H) + g;
        int c = c0;
        for (; c + 8 <= c1; c += 8) {
            f4 a[8], bv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                a[u] = arow[4 * (c + u)];
                bv[u] = brw[4 * (c + u)];
            }
            __builtin_amdgcn_sched_barrier(0);  // all sixteen loads in flight before the first MFMA waits
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][0], bv[u][0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][1], bv[u][1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][2], bv[u][2], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][3], bv[u][3], acc, 0, 0, 0);
            }
        }
        for (; c < c1; ++c) {
            const f4 a = arow[4 * c], bv = brw[4 * c];
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], bv[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], bv[1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], bv[2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], bv[3], acc, 0, 0, 0);
        }
    }
    {
        const int col = lane & 15, r0 = 4 * (lane >> 4);
        part[wave][r0 + 0][col] = acc[0];
        part[wave][r0 + 1][col] = acc[1];
        part[wave][r0 + 2][col] = acc[2];
        part[wave][r0 + 3][col] = acc[3];
    }
    __syncthreads();
    const int ii = threadIdx.x >> 4, jj = threadIdx.x & 15;
    const int b = b0 + ii, j = j0 + jj;
    if (b >= B) return;
    const float m = ((part[0][ii][jj] + part[1][ii][jj]) + part[2][ii][jj]) + part[3][ii][jj];
    const int64_t k = static_cast<int64_t>(b) * H + j;
    if (!kEpilogue) {
        gi_out[k] = m;
        return;
    }
    const float gv1 = gi_next ? gv[k] : 0.0f;
    const float gi1 = gi_next ? gi_next[k] : 0.0f;
    const float gs = (gso_t[k] + m) - p.theta * gv1;
    const float gV = gs * h_t[k] + p.beta * gv1;
    const float gI = gV + p.alpha * gi1;
    gv[k] = gV;
    gi_out[k] = gI;
}

// Readout: J_t = alpha J_{t-1} + z_t, R_t = beta R_{t-1} + J_t; one thread per (b, o).
__global__ void readout_kernel(int32_t T, int64_t bo, float alpha, float beta, const float* __restrict__ z,
                               float* __restrict__ r) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= bo) return;
    float J = 0.0f, R = 0.0f;
    for (int t = 0; t < T; ++t) {
        J = alpha * J + z[t * bo + k];
        R = beta * R + J;
        r[t * bo + k] = R;
    }
}

// Adjoint of the readout: gR'_t = gR_t + beta gR'_{t+1}; gJ_t = gR'_t + alpha gJ_{t+1}.
__global__ void readout_adj_kernel(int32_t T, int64_t bo, float alpha, float beta, const float* __restrict__ gr,
                                   float* __restrict__ gj) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= bo) return;
    float gR = 0.0f, gJ = 0.0f;
    for (int t = T - 1; t >= 0; --t) {
        gR = gr[t * bo + k] + beta * gR;
        gJ = gR + alpha * gJ;
        gj[t * bo + k] = gJ;
    }
}

__global__ void zero_kernel(uint32_t* __restrict__ p, int64_t n) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0u;
}

__global__ void copy_kernel(const float* __restrict__ src, float* __restrict__ dst, int64_t n) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

// Clear n 32-bit words with a kernel rather than hipMemsetAsync, so that a caller may capture these calls
// into a graph.
int zero(void* ptr, int64_t n_words, hipStream_t st) {
    hipLaunchKernelGGL(zero_kernel, dim3((n_words + 255) / 256), dim3(256), 0, st, static_cast<uint32_t*>(ptr), n_words);
    R_HIP(hipGetLastError());
    return N2H_OK;
}

int check(const n2h_rlif* p) {
    if (!p) return rfail(N2H_EINVAL, "null layer");
    if (p->batch <= 0 || p->n_hidden <= 0 || p->n_steps <= 0) return rfail(N2H_EINVAL, "sizes must be positive");
    if (p->mode != 0 && p->mode != 1) return rfail(N2H_EINVAL, "mode must be 0 or 1");
    if (p->checkpoint <= 0) return rfail(N2H_EINVAL, "checkpoint interval must be positive");
    if (p->batch > 65535) return rfail(N2H_EINVAL, "batch %d above 65535", p->batch);
    if (p->n_hidden > 8192) return rfail(N2H_EINVAL, "n_hidden %d above 8192", p->n_hidden);
    return N2H_OK;
}

size_t step_lds(const n2h_rlif* p) { return sizeof(float) * 2 * static_cast<size_t>(p->n_hidden); }

dim3 step_grid(const n2h_rlif* p) { return dim3((p->n_hidden + kTile - 1) / kTile, p->batch); }

dim3 reverse_grid(const n2h_rlif* p) { return dim3((p->n_hidden + kMM - 1) / kMM, (p->batch + kMM - 1) / kMM); }

int64_t work_floats(const n2h_rlif* p) {
    const int32_t K = p->checkpoint < p->n_steps ? p->checkpoint : p->n_steps;
    return (5 + static_cast<int64_t>(K)) * p->batch * p->n_hidden;
}

// The backward pass over the hidden layer with `work` as its scratch ((5 + K) B H floats: the recomputed
// state [4][B][H], gV [B][H], the segment's surrogate derivatives [K][B][H]).
int backward(const n2h_rlif* p, const float* u, const float* w_rec_t, const float* spikes,
             const float* ckpt, const float* gso, float* g_i_all, float* work, hipStream_t st) {
    const int H = p->n_hidden, B = p->batch, T = p->n_steps, K = p->checkpoint;
    const int64_t bh = static_cast<int64_t>(B) * H;
    float* state = work;
    float* gv = work + 4 * bh;
    float* hseg = work + 5 * bh;
    int rc = zero(gv, bh, st);
    if (rc) return rc;
    const int n_seg = (T + K - 1) / K;
    for (int sgm = n_seg - 1; sgm >= 0; --sgm) {
        const int t0 = sgm * K, t1 = (t0 + K < T) ? t0 + K : T;
        // recompute the surrogate derivatives over the segment from its checkpoint and the stored spikes
        hipLaunchKernelGGL(copy_kernel, dim3((4 * bh + 255) / 256), dim3(256), 0, st, ckpt + sgm * 4 * bh, state, 4 * bh);
        for (int t = t0; t < t1; ++t)
            hipLaunchKernelGGL(step_kernel<true>, step_grid(p), dim3(kTile), step_lds(p), st, *p, u + t * bh, w_rec_t,
                               state, t ? spikes + (t - 1) * bh : nullptr, hseg + (t - t0) * bh, nullptr);
        for (int t = t1 - 1; t >= t0; --t)
            hipLaunchKernelGGL(reverse_kernel<true>, reverse_grid(p), dim3(256), 0, st, *p,
                               t < T - 1 ? g_i_all + (t + 1) * bh : nullptr, w_rec_t, hseg + (t - t0) * bh, gso + t * bh,
                               gv, g_i_all + t * bh);
        R_HIP(hipGetLastError());
    }
    return N2H_OK;
}

int check_backward_args(const n2h_rlif* p, const float* u, const float* w_rec_t,
                        const float* spikes, const float* ckpt, const float* gso, const float* g_i_all) {
    int rc = check(p);
    if (rc) return rc;
    if (!u || !w_rec_t || !spikes || !ckpt || !gso || !g_i_all) return rfail(N2H_EINVAL, "null buffer");
    if (step_lds(p) > 64 * 1024) return rfail(N2H_EINVAL, "n_hidden %d too large", p->n_hidden);
    if (p->n_hidden % kMM) return rfail(N2H_EINVAL, "n_hidden %d is not a multiple of %d", p->n_hidden, kMM);
    return N2H_OK;
}

// The scratch of n2h_rlif_backward_noalloc: one array per GPU (each GPU the library runs on holds its own
// copy until the process ends), the size of the largest work area that entry point takes.
constexpr int64_t kScratchFloats = N2H_RLIF_SCRATCH_FLOATS;
__device__ __align__(256) float scratch[kScratchFloats];

std::mutex scratch_mu;
float* scratch_dev[kMaxDevices] = {};

// This GPU's address of `scratch`, looked up on the first call with the device current and kept per device.
// Host code passes only this address to kernels: the array's own name is a host-side stand-in.
int scratch_address(float** out) {
    int dev = 0;
    R_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return rfail(N2H_EINVAL, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return rfail(N2H_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        scratch_dev[dev] = static_cast<float*>(ptr);
    }
    *out = scratch_dev[dev];
    return N2H_OK;
}

}  // namespace

extern "C" {

int n2h_rlif_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return N2H_EINVAL;
    strncpy(buf, r_err, len - 1);
    buf[len - 1] = '\0';
    return N2H_OK;
}

int n2h_rlif_forward(const n2h_rlif* p, const float* u, const float* w_rec_t, float* state, float* spikes,
                     float* ckpt, void* stream) {
    int rc = check(p);
    if (rc) return rc;
    if (!u || !w_rec_t || !state || !spikes || !ckpt) return rfail(N2H_EINVAL, "null buffer");
    if (step_lds(p) > 64 * 1024) return rfail(N2H_EINVAL, "n_hidden %d too large", p->n_hidden);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t bh = static_cast<int64_t>(p->batch) * p->n_hidden;
    rc = zero(state, 4 * bh, st);
    if (rc) return rc;
    for (int32_t t = 0; t < p->n_steps; ++t) {
        float* ck = (t % p->checkpoint == 0) ? ckpt + (t / p->checkpoint) * 4 * bh : nullptr;
        hipLaunchKernelGGL(step_kernel<false>, step_grid(p), dim3(kTile), step_lds(p), st, *p, u + t * bh, w_rec_t,
                           state, t ? spikes + (t - 1) * bh : nullptr, spikes + t * bh, ck);
    }
    R_HIP(hipGetLastError());
    return N2H_OK;
}

int n2h_rlif_readout(int32_t n_steps, int64_t bo, float alpha, float beta, const float* z, float* r, void* stream) {
    if (n_steps <= 0 || bo <= 0 || !z || !r) return rfail(N2H_EINVAL, "bad readout arguments");
    hipLaunchKernelGGL(readout_kernel, dim3((bo + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), n_steps,
                       bo, alpha, beta, z, r);
    R_HIP(hipGetLastError());
    return N2H_OK;
}

int n2h_rlif_readout_adjoint(int32_t n_steps, int64_t bo, float alpha, float beta, const float* g_r, float* g_j,
                             void* stream) {
    if (n_steps <= 0 || bo <= 0 || !g_r || !g_j) return rfail(N2H_EINVAL, "bad readout arguments");
    hipLaunchKernelGGL(readout_adj_kernel, dim3((bo + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       n_steps, bo, alpha, beta, g_r, g_j);
    R_HIP(hipGetLastError());
    return N2H_OK;
}

int n2h_rlif_delivery_adjoint(const n2h_rlif* p, const float* g_i, const float* w_rec_t, float* m, void* stream) {
    int rc = check(p);
    if (rc) return rc;
    if (!g_i || !w_rec_t || !m) return rfail(N2H_EINVAL, "null buffer");
    if (p->n_hidden % kMM) return rfail(N2H_EINVAL, "n_hidden %d is not a multiple of %d", p->n_hidden, kMM);
    hipLaunchKernelGGL(reverse_kernel<false>, reverse_grid(p), dim3(256), 0, static_cast<hipStream_t>(stream), *p, g_i,
                       w_rec_t, nullptr, nullptr, nullptr, m);
    R_HIP(hipGetLastError());
    return N2H_OK;
}

int64_t n2h_rlif_work_floats(const n2h_rlif* p) {
    if (check(p)) return -1;
    return work_floats(p);
}

int n2h_rlif_backward(const n2h_rlif* p, const float* u, const float* w_rec_t, const float* spikes,
                      const float* ckpt, const float* gso, float* g_i_all, float* work, void* stream) {
    int rc = check_backward_args(p, u, w_rec_t, spikes, ckpt, gso, g_i_all);
    if (rc) return rc;
    if (!work) return rfail(N2H_EINVAL, "null work buffer");
    return backward(p, u, w_rec_t, spikes, ckpt, gso, g_i_all, work, static_cast<hipStream_t>(stream));
}

int n2h_rlif_backward_noalloc(const n2h_rlif* p, const float* u, const float* w_rec_t, const float* spikes,
                              const float* ckpt, const float* gso, float* g_i_all, void* stream) {
    int rc = check_backward_args(p, u, w_rec_t, spikes, ckpt, gso, g_i_all);
    if (rc) return rc;
    const int64_t need = work_floats(p);
    if (need > kScratchFloats)
        return rfail(N2H_EINVAL, "work area of %lld floats ((5 + min(K, T)) * batch * n_hidden) above the %lld of "
                     "the library's scratch", static_cast<long long>(need), static_cast<long long>(kScratchFloats));
    float* work = nullptr;
    rc = scratch_address(&work);
    if (rc) return rc;
    return backward(p, u, w_rec_t, spikes, ckpt, gso, g_i_all, work, static_cast<hipStream_t>(stream));
}

}  // extern "C"
