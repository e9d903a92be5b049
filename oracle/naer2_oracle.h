/*
 * naer2_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded CPU restatement of the time-driven spiking-network step that
 * libnaer2_hip.so runs on the GPU.  It is the checker of the HIP path: only the tests, the smoke check
 * and the CPU-baseline leg of the benchmark load it, and the product never calls it.
 *
 * Followed reference code (MatteoSchmider/naer2):
 *   - tools/brunel_reference.py:34-48   the time-driven step: t=(n+1)dt, input = ring[n % D] + J*Poisson,
 *                                        clear, refractory gate, V = V*decay + input, fire at V >= theta,
 *                                        reset, deliver to the ring D steps later
 *   - test/networks/PD14/neuron.py:37-60 and test/fused_reference.py:76-107
 *                                        the exp-synapse LIF: deliver (current += w) -> update with the exact
 *                                        propagator -> current *= P_I -> threshold V > V_th, reset
 *   - naer2/transpiler/templates/fused.py:2308-2604 (_deliver) and :1626-1965 (_pull)
 *                                        delivery of a source's links into the target's ring slots
 *   - naer2/ring.py:132-245              the ring of arrival slots (here: fp32 input accumulators)
 *
 * Deviations from the reference, kept identical on the GPU:
 *   - refractoriness counts whole steps (ref_steps) instead of comparing float times;
 *   - propagators are fp32 constants for the fixed dt (no exp in the loop);
 *   - every delay is a whole number of steps in [1, d_slots];
 *   - summation order (the reference's GPU order is not deterministic): for target i and arrival step n,
 *     for each emission step s ascending: p = +0.0f; p += w_l over the links l -> i of sources spiking at
 *     s with s + d_l = n, in ascending canonical link index; then ring[n % D][i] = ring[n % D][i] + p
 *     (only for cells that received at least one link).
 *   - background: count = #{j : u >= thr[j]}, u = word 0 of Philox4x32-10(ctr = {neuron, step lo,
 *     step hi, instance}, key = {seed lo, seed hi}); input = x + j_ext * (float)count.
 *
 * Multi-rate (naer2/population.py:191-204, 382-410; naer2/network.py:80-152): a population whose timestep is
 * stride * dt (its parameters computed for that interval) updates at the steps s with (s + 1) % stride == 0
 * only; x is then its window's ring slots summed oldest first, ((r[s-k+1] + r[s-k+2]) + ... + r[s-1]) + r[s],
 * each slot cleared; its other steps leave the ring and its state alone and emit no spike.  d_slots must be
 * at least the largest delay + the largest stride - 1.  (or_mt_run takes stride-1 networks only.)
 */
#ifndef NAER2_ORACLE_H
#define NAER2_ORACLE_H
#include <stdint.h>

#define OR_POISSON_K 32
#define OR_MAX_STRIDE 64
#define OR_LIF_DELTA 1
#define OR_LIF_EXP 2

/* Same byte layout as the product's per-population parameter block (fp32 constants + inversion table). */
typedef struct or_pop {
    int32_t ref_steps;
    int32_t stride;   /* steps per update (0 or 1: every step) */
    float decay, p_i, p_vi, v_rest, v_th, v_reset, j_ext, reserved_f;
    uint32_t poisson_thr[OR_POISSON_K];
} or_pop;

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int or_poisson_count(uint32_t u, const uint32_t* thr);

/* Runs n_steps steps from step0.  Topology: canonical CSR by source (src_ptr [N+1], tgt/w/delay [E]).
 * prev_spk [N]: the spikes of step0 - 1 on entry, of the last step on return.  raster [n_steps][N] or NULL.
 * accum_mode 0: p += w (the product's); 1: p = p*0.5f + w and ring = ring*0.5f + p (an order hash, to pin
 * the order).  events: links delivered (out).  Returns 0, or -1 on a bad argument. */
int or_run(int model, int32_t n_neurons, int32_t n_pops, const int32_t* pop_first, const or_pop* pops,
           int32_t d_slots, uint64_t seed, uint32_t instance,
           const int64_t* src_ptr, const int32_t* tgt, const float* w, const int32_t* delay,
           float* v, float* isyn, int32_t* ref, float* ring, uint8_t* prev_spk,
           int64_t step0, int32_t n_steps, uint8_t* raster, int32_t accum_mode, int64_t* events);


/* The exp-LIF step with per-link edge kinds and ports (the reference's open decisions D-6 / D-7 and edge
 * forward, TODO.md:179-204, test/test_fused_generated_gpu.py:361-398, as libnaer2_hip.so's n2h_bind_edges
 * defines them): per canonical link, kind 0 static delivers w, 1 gain delivers w * attr2, 2 plastic first does
 * w += attr2 (stored: w is updated in place) then delivers the new w; port 0 adds to the current's ring, port 1
 * to ring_v [d_slots][N], whose due input is added to V after the propagation, V = (V_rest + P_V (V - V_rest))
 * + P_VI I, then V = V + x_v (not on refractory steps; the slot is cleared either way).  Per port the
 * summation order is the plain step's.  Single-rate networks only. */
int or_run_edges(int32_t n_neurons, int32_t n_pops, const int32_t* pop_first, const or_pop* pops, int32_t d_slots,
                 uint64_t seed, uint32_t instance, const int64_t* src_ptr, const int32_t* tgt, float* w,
                 const int32_t* delay, const uint8_t* kind, const uint8_t* port, const float* attr2, float* ring_v,
                 float* v, float* isyn, int32_t* ref, float* ring, uint8_t* prev_spk, int64_t step0, int32_t n_steps,
                 uint8_t* raster, int64_t* events);

/* The same step, OpenMP-parallel over contiguous target ranges: identical results (every per-target sum
 * sees the same links in the same order).  or_mt_create splits the canonical CSR by target range. */
typedef struct or_mt or_mt;
or_mt* or_mt_create(int32_t n_neurons, const int64_t* src_ptr, const int32_t* tgt, const float* w,
                    const int32_t* delay, int32_t n_threads);
void or_mt_destroy(or_mt* m);
int or_mt_run(or_mt* m, int model, int32_t n_pops, const int32_t* pop_first, const or_pop* pops, int32_t d_slots,
              uint64_t seed, uint32_t instance, float* v, float* isyn, int32_t* ref, float* ring, uint8_t* prev_spk,
              int64_t step0, int32_t n_steps, uint8_t* raster, int64_t* events);

#endif
