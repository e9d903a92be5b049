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
// The recurrent delivery W_rec s_{t-1} is event-driven: a workgroup compacts the nonzero entries of
// s_{t-1}[b] (ascending) and each lane sums the matching rows of W_rec^T in that order.  The dense
// adjoint product W_rec^T gI_{t+1} is a plain GEMM (rocBLAS) per step, followed by the fused elementwise
// adjoint.  The forward carries a double-precision shadow of (I, V) beside its fp32 state (which decides the
// spikes); every K steps the shadow is checkpointed, and the backward pass recomputes the shadow's V inside
// each segment from its checkpoint and the stored spikes -- the same trajectory whatever K -- for the
// surrogate derivative.
// There is no counterpart in the reference: autograd and BPTT do not exist in naer2.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>

#include "naer2_hip.h"

namespace {

constexpr int kTile = 256;  // hidden units per workgroup

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

// The double-precision shadow of one step (identical expression in the forward and the recompute).
__device__ __forceinline__ void shadow_update(const n2h_rlif& p, double rec, float u, float sp, double& I, double& V) {
    I = static_cast<double>(p.alpha) * I + (static_cast<double>(u) + rec);
    V = (static_cast<double>(p.beta) * V + I) - static_cast<double>(p.theta) * static_cast<double>(sp);
}

// One forward step for batch row b and hidden units [tile*256, tile*256 + 256): the fp32 state decides the
// spikes; a double shadow of (I, V) runs alongside and is what the checkpoints keep, so that the
// backward's recompute reproduces one continuous double trajectory whatever the checkpoint interval.
__global__ __launch_bounds__(kTile) void fwd_step_kernel(n2h_rlif p, const float* __restrict__ u_t,
                                                         const float* __restrict__ w_rec_t,
                                                         float* __restrict__ i_state, float* __restrict__ v_state,
                                                         double* __restrict__ shadow, const float* __restrict__ s_prev,
                                                         float* __restrict__ s_out, double* __restrict__ ckpt) {
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
    float rec = 0.0f;
    double rec64 = 0.0;
    for (int q = 0; q < n_act; ++q) {
        const float w = w_rec_t[static_cast<int64_t>(act_idx[q]) * H + h];
        rec = rec + w * act_val[q];
        rec64 = rec64 + static_cast<double>(w) * static_cast<double>(act_val[q]);
    }
    const int64_t k = row + h;
    double I64 = shadow[k], V64 = shadow[bh + k];
    if (ckpt) {  // the segment's starting state
        ckpt[k] = I64;
        ckpt[bh + k] = V64;
    }
    const float sp = s_prev ? s_prev[k] : 0.0f;
    const float a = u_t[k] + rec;
    const float I = p.alpha * i_state[k] + a;
    const float V = (p.beta * v_state[k] + I) - p.theta * sp;
    i_state[k] = I;
    v_state[k] = V;
    s_out[k] = spike_fn(V - p.theta, p.mode, p.slope);
    shadow_update(p, rec64, u_t[k], sp, I64, V64);
    shadow[k] = I64;
    shadow[bh + k] = V64;
}

// Recompute of one step for the backward pass: the double shadow of fwd_step_kernel, with the forward's
// spikes (not re-decided).  The surrogate derivative's relative sensitivity to V reaches 2 slope (50) at
// the threshold, so it is evaluated on this trajectory rather than on the fp32 one.
__global__ __launch_bounds__(kTile) void recompute_step_kernel(n2h_rlif p, const float* __restrict__ u_t,
                                                               const float* __restrict__ w_rec_t,
                                                               double* __restrict__ shadow,
                                                               const float* __restrict__ s_prev,
                                                               double* __restrict__ v_out) {
    extern __shared__ unsigned char smem[];
    int* act_idx = reinterpret_cast<int*>(smem);
    float* act_val = reinterpret_cast<float*>(smem) + p.n_hidden;
    __shared__ int wave_base[kTile / 64 + 1];
    const int H = p.n_hidden;
    const int b = blockIdx.y;
    const int h = blockIdx.x * kTile + threadIdx.x;
    const int64_t row = static_cast<int64_t>(b) * H;
    const int64_t bh = static_cast<int64_t>(p.batch) * H;
    const int n_act = s_prev ? compact_row(s_prev + row, H, act_idx, act_val, wave_base) : 0;
    if (h >= H) return;
    double rec64 = 0.0;
    for (int q = 0; q < n_act; ++q)
        rec64 = rec64 + static_cast<double>(w_rec_t[static_cast<int64_t>(act_idx[q]) * H + h]) *
                            static_cast<double>(act_val[q]);
    const int64_t k = row + h;
    double I64 = shadow[k], V64 = shadow[bh + k];
    shadow_update(p, rec64, u_t[k], s_prev ? s_prev[k] : 0.0f, I64, V64);
    shadow[k] = I64;
    shadow[bh + k] = V64;
    v_out[k] = V64;
}

__global__ void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

// Elementwise adjoint of one step: gs = gso + m - theta gV_{t+1}; gV = gs s'(V) + beta gV_{t+1};
// gI = gV + alpha gI_{t+1}.  gv / gi hold step t+1's values on entry and step t's on return.
__global__ void adj_step_kernel(n2h_rlif p, int64_t n, const double* __restrict__ v_t,
                                const float* __restrict__ gso_t, const float* __restrict__ m_t,
                                float* __restrict__ gv, float* __restrict__ gi, float* __restrict__ gi_out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gv1 = gv[k], gi1 = gi[k];
    const float gs = (gso_t[k] + (m_t ? m_t[k] : 0.0f)) - p.theta * gv1;
    const double x = v_t[k] - static_cast<double>(p.theta);
    const double d = 1.0 + static_cast<double>(p.slope) * fabs(x);
    const float sg = static_cast<float>(static_cast<double>(p.gamma) / (d * d));
    const float gV = gs * sg + p.beta * gv1;
    const float gI = gV + p.alpha * gi1;
    gv[k] = gV;
    gi[k] = gI;
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
    return N2H_OK;
}

size_t fwd_lds(const n2h_rlif* p) { return sizeof(float) * 2 * static_cast<size_t>(p->n_hidden); }

dim3 step_grid(const n2h_rlif* p) { return dim3((p->n_hidden + kTile - 1) / kTile, p->batch); }

}  // namespace

extern "C" {

int n2h_rlif_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return N2H_EINVAL;
    strncpy(buf, r_err, len - 1);
    buf[len - 1] = '\0';
    return N2H_OK;
}

int n2h_rlif_forward(const n2h_rlif* p, const float* u, const float* w_rec_t, float* i_state, float* v_state,
                     double* shadow, float* spikes, double* ckpt, void* stream) {
    int rc = check(p);
    if (rc) return rc;
    if (!u || !w_rec_t || !i_state || !v_state || !shadow || !spikes || !ckpt) return rfail(N2H_EINVAL, "null buffer");
    if (fwd_lds(p) > 64 * 1024) return rfail(N2H_EINVAL, "n_hidden %d too large", p->n_hidden);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t bh = static_cast<int64_t>(p->batch) * p->n_hidden;
    rc = zero(i_state, bh, st);
    if (!rc) rc = zero(v_state, bh, st);
    if (!rc) rc = zero(shadow, 4 * bh, st);
    if (rc) return rc;
    for (int32_t t = 0; t < p->n_steps; ++t) {
        double* ck = (t % p->checkpoint == 0) ? ckpt + (t / p->checkpoint) * 2 * bh : nullptr;
        hipLaunchKernelGGL(fwd_step_kernel, step_grid(p), dim3(kTile), fwd_lds(p), st, *p, u + t * bh, w_rec_t, i_state,
                           v_state, shadow, t ? spikes + (t - 1) * bh : nullptr, spikes + t * bh, ck);
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

// Backward through the hidden layer.  gso = W_out^T gJ for every step ([T][B][H]); writes gI_t for every step
// into g_i_all.  work: (7 + 2 K) * B * H floats: I, V (double), gV, gI, M (float), segment V ([K][B][H] double).
int n2h_rlif_backward(const n2h_rlif* p, const float* u, const float* w_rec, const float* w_rec_t, const float* spikes,
                      const double* ckpt, const float* gso, float* g_i_all, float* work, void* stream) {
    int rc = check(p);
    if (rc) return rc;
    if (!u || !w_rec || !w_rec_t || !spikes || !ckpt || !gso || !g_i_all || !work)
        return rfail(N2H_EINVAL, "null buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int H = p->n_hidden, B = p->batch, T = p->n_steps, K = p->checkpoint;
    const int64_t bh = static_cast<int64_t>(B) * H;
    double* shadow = reinterpret_cast<double*>(work);  // [2][B][H]
    float* gv = work + 4 * bh;
    float* gi = work + 5 * bh;
    float* m = work + 6 * bh;
    double* vseg = reinterpret_cast<double*>(work + 7 * bh);
    rocblas_handle handle;
    if (rocblas_create_handle(&handle) != rocblas_status_success) return rfail(N2H_EHIP, "rocblas_create_handle");
    rocblas_set_stream(handle, st);
    const float f_one = 1.0f, f_zero = 0.0f;
    rc = zero(gv, bh, st);
    if (!rc) rc = zero(gi, bh, st);
    if (rc) {
        rocblas_destroy_handle(handle);
        return rc;
    }
    hipError_t e = hipSuccess;
    const int n_seg = (T + K - 1) / K;
    for (int sgm = n_seg - 1; sgm >= 0 && rc == N2H_OK; --sgm) {
        const int t0 = sgm * K, t1 = (t0 + K < T) ? t0 + K : T;
        // recompute V over the segment from its checkpoint and the stored spikes
        hipLaunchKernelGGL(copy_kernel, dim3((2 * bh + 255) / 256), dim3(256), 0, st, ckpt + sgm * 2 * bh, shadow, 2 * bh);
        for (int t = t0; t < t1; ++t)
            hipLaunchKernelGGL(recompute_step_kernel, step_grid(p), dim3(kTile), fwd_lds(p), st, *p, u + t * bh, w_rec_t,
                               shadow, t ? spikes + (t - 1) * bh : nullptr, vseg + (t - t0) * bh);
        e = hipGetLastError();
        if (e != hipSuccess) {
            rc = rfail(N2H_EHIP, "recompute: %s", hipGetErrorString(e));
            break;
        }
        for (int t = t1 - 1; t >= t0 && rc == N2H_OK; --t) {
            const float* mt = nullptr;
            if (t < T - 1) {  // m = gI_{t+1} W_rec (row-major), as column-major m^T = W_rec^T gI^T
                if (rocblas_sgemm(handle, rocblas_operation_none, rocblas_operation_none, H, B, H, &f_one, w_rec, H, gi,
                                  H, &f_zero, m, H) != rocblas_status_success) {
                    rc = rfail(N2H_EHIP, "rocblas_sgemm");
                    break;
                }
                mt = m;
            }
            hipLaunchKernelGGL(adj_step_kernel, dim3((bh + 255) / 256), dim3(256), 0, st, *p, bh, vseg + (t - t0) * bh,
                               gso + t * bh, mt, gv, gi, g_i_all + t * bh);
            e = hipGetLastError();
            if (e != hipSuccess) rc = rfail(N2H_EHIP, "adj_step_kernel: %s", hipGetErrorString(e));
        }
    }
    rocblas_destroy_handle(handle);
    return rc;
}

}  // extern "C"
