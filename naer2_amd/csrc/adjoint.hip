// Surrogate-gradient adjoint of the exponential-synapse LIF step (the LIF_EXP model of naer2_hip.hip) for a
// per-population rate loss, batched over networks: one launch per reverse step.
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
// One wavefront per source neuron gathers its out-links (lanes stride the list; per-lane sums per target
// population, then a fixed butterfly): every sum has a fixed order, so results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "naer2_hip.h"

namespace {

constexpr int kWave = 64;
constexpr int kAdjPops = 8;      // target-population bins per lane
constexpr int kAdjNets = 16;     // networks in one launch
constexpr int kWavesPerBlock = 4;

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
    int32_t ring_pos;  // step mod ring_slots: the slot this step's gx is written to
    n2h_surrogate sur;
};

__global__ __launch_bounds__(kWave * kWavesPerBlock) void adj_step_kernel(AdjBatch b) {
    const int blk = blockIdx.x;
    int q = 0;
    while (q + 1 < b.n_nets && b.first_block[q + 1] <= blk) ++q;
    const n2h_adj_net& a = b.net[q];
    const int lane = threadIdx.x & (kWave - 1);
    const int j = (blk - b.first_block[q]) * kWavesPerBlock + static_cast<int>(threadIdx.x >> 6);
    if (j >= a.n_neurons) return;
    const int n = a.n_neurons;
    const int R = a.ring_slots;

    int pj = 0;
    while (pj + 1 < a.n_pops && a.pop_first[pj + 1] <= j) ++pj;
    const float* C = a.pop_coef + pj * 8;
    const float P_V = C[0], P_I = C[1], P_VI = C[2], theta = C[4], dv = C[5], c_loss = C[6];
    const float vp = a.vtrace[j];
    const bool s = a.spikes[j] != 0;
    const float u = fabsf(vp - theta) / dv;  // +inf on refractory steps
    const bool near = u <= b.sur.cutoff;     // the surrogate is nonzero
    float bins[kAdjPops];
#pragma unroll
    for (int p = 0; p < kAdjPops; ++p) bins[p] = 0.0f;
    const int64_t e0 = a.out_ptr[j], e1 = (near || s) ? a.out_ptr[j + 1] : e0;
    for (int64_t e = e0 + lane; e < e1; e += kWave) {
        const uint2 l = reinterpret_cast<const uint2*>(a.out_links)[e];
        const uint32_t tgt = l.x & 0x1ffffu;
        const int d = static_cast<int>((l.x >> 17) & 127u) + 1;
        const uint32_t tp = l.x >> 24;
        int slot = b.ring_pos + d;
        slot -= (slot >= R) ? R : 0;
        const float g = a.gx_ring[static_cast<int64_t>(slot) * n + tgt];
        const float c = __uint_as_float(l.y) * g;
#pragma unroll
        for (int p = 0; p < kAdjPops; ++p) bins[p] += (tp == static_cast<uint32_t>(p)) ? c : 0.0f;
    }
    // fixed butterfly: every lane ends with the same sums, in the same order on every run (a neuron that
    // gathered nothing has all-zero sums: no exchange)
    if (e1 > e0) {
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
            for (int p = 0; p < kAdjPops; ++p) bins[p] += __shfl_xor(bins[p], off, kWave);
        }
    }
    if (lane != 0) return;

    float coup = 0.0f;
    for (int p = 0; p < a.n_pops; ++p) coup += a.alpha[p * a.n_pops + pj] * bins[p];
    if (s) {
        double* gs = a.g_sum + static_cast<int64_t>(j) * a.n_pops;
        for (int p = 0; p < a.n_pops; ++p) gs[p] += static_cast<double>(bins[p]);
    }
    const float gv = a.g_v[j], gi = a.g_i[j];
    float gip, gv_new;
    if (isfinite(vp)) {
        const float gsum = c_loss + coup;
        const float den = 1.0f + b.sur.kappa * u;
        const float h = near ? (b.sur.gamma * 0.5f * b.sur.kappa / dv) / (den * den) : 0.0f;
        const float gvpre = (s ? 0.0f : gv) + h * gsum;
        gip = P_VI * gvpre + P_I * gi;
        gv_new = P_V * gvpre;
    } else {
        gip = P_I * gi;
        gv_new = 0.0f;
    }
    a.g_v[j] = gv_new;
    a.g_i[j] = gip;
    a.gx_ring[static_cast<int64_t>(b.ring_pos) * n + j] = gip;
}

}  // namespace

extern "C" {

int n2h_adj_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return N2H_EINVAL;
    snprintf(buf, len, "%s", g_err);
    return N2H_OK;
}

int n2h_adj_step(const n2h_adj_net* nets, int32_t n_nets, int64_t step, const n2h_surrogate* sur, void* stream) {
    if (!nets || n_nets < 1 || n_nets > kAdjNets) return fail(N2H_EINVAL, "1 to %d networks per launch", kAdjNets);
    if (step < 0) return fail(N2H_EINVAL, "negative step");
    if (!sur || !(sur->kappa > 0.0f) || !(sur->cutoff >= 0.0f) || !(sur->gamma >= 0.0f))
        return fail(N2H_EINVAL, "surrogate: kappa > 0, cutoff >= 0, gamma >= 0");
    AdjBatch b;
    memset(&b, 0, sizeof(b));
    b.n_nets = n_nets;
    b.sur = *sur;
    int64_t blocks = 0;
    int ring = -1;
    for (int i = 0; i < n_nets; ++i) {
        const n2h_adj_net& a = nets[i];
        if (a.n_neurons <= 0 || a.n_neurons > (1 << 17)) return fail(N2H_EINVAL, "net %d: n_neurons outside [1, 2^17]", i);
        if (a.n_pops < 1 || a.n_pops > kAdjPops) return fail(N2H_EINVAL, "net %d: 1 to %d populations", i, kAdjPops);
        if (a.ring_slots < 2 || a.ring_slots > N2H_MAX_SLOTS + 1) return fail(N2H_EINVAL, "net %d: ring_slots", i);
        if (!a.pop_first || !a.out_ptr || !a.out_links || !a.pop_coef || !a.alpha || !a.vtrace || !a.spikes ||
            !a.gx_ring || !a.g_v || !a.g_i || !a.g_sum)
            return fail(N2H_EINVAL, "net %d: null pointer", i);
        if (ring >= 0 && ring != a.ring_slots) return fail(N2H_EINVAL, "every network needs the same ring_slots");
        ring = a.ring_slots;
        b.net[i] = a;
        b.first_block[i] = static_cast<int32_t>(blocks);
        blocks += (a.n_neurons + kWavesPerBlock - 1) / kWavesPerBlock;
    }
    b.first_block[n_nets] = static_cast<int32_t>(blocks);
    b.ring_pos = static_cast<int32_t>(step % ring);
    hipLaunchKernelGGL(adj_step_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kWave * kWavesPerBlock), 0,
                       static_cast<hipStream_t>(stream), b);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(N2H_EHIP, "adj_step_kernel: %s", hipGetErrorString(e));
    return N2H_OK;
}

}  // extern "C"
