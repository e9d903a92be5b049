/*
 * naer2_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU restatement that checks libnaer2_hip.so.
 * Built with -ffp-contract=off so every fp32 operation rounds once, as in the HIP kernels.
 * The algorithm, the followed reference lines and the deviations are listed in naer2_oracle.h.
 */
#include "naer2_oracle.h"

#include <omp.h>

#include <stdlib.h>
#include <string.h>

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3", SC'11):
 * ten rounds of two 32x32->64 multiplies, the key bumped by the Weyl constants between rounds. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

int or_poisson_count(uint32_t u, const uint32_t* thr) {
    int k = 0;
    for (int j = 0; j < OR_POISSON_K; ++j) k += (u >= thr[j]) ? 1 : 0;
    return k;
}

/* 1 if some population updates every stride > 1 steps; -1 if a stride is out of range. */
static int multirate(int32_t n_pops, const or_pop* pops, int32_t d_slots) {
    int any = 0;
    for (int32_t p = 0; p < n_pops; ++p) {
        if (pops[p].stride < 0 || pops[p].stride > OR_MAX_STRIDE || pops[p].stride > d_slots) return -1;
        any |= pops[p].stride > 1;
    }
    return any;
}

static int pop_of(int32_t i, int32_t n_pops, const int32_t* pop_first) {
    int p = 0;
    while (p + 1 < n_pops && i >= pop_first[p + 1]) ++p;
    return p;
}

/* The step with optional edges (or_run_edges): kind/port/attr2 NULL is the plain step.  The cells of the
 * potential port follow those of the current port in acc / touched (cell + N * D). */
static int run_core(int model, int32_t n_neurons, int32_t n_pops, const int32_t* pop_first, const or_pop* pops,
                    int32_t d_slots, uint64_t seed, uint32_t instance,
                    const int64_t* src_ptr, const int32_t* tgt, float* w, const int32_t* delay,
                    const uint8_t* kind, const uint8_t* port, const float* attr2, float* ring_v,
                    float* v, float* isyn, int32_t* ref, float* ring, uint8_t* prev_spk,
                    int64_t step0, int32_t n_steps, uint8_t* raster, int32_t accum_mode, int64_t* events) {
    if (n_neurons < 0 || n_pops < 1 || d_slots < 1 || n_steps < 0) return -1;
    if (model != OR_LIF_DELTA && model != OR_LIF_EXP) return -1;
    if (model == OR_LIF_EXP && !isyn) return -1;
    const int mr = multirate(n_pops, pops, d_slots);
    if (mr < 0 || (mr && accum_mode != 0)) return -1;
    const int edges = kind != NULL;
    if (edges && (!port || !ring_v || model != OR_LIF_EXP || mr || accum_mode != 0)) return -1;
    const int64_t N = n_neurons, D = d_slots;
    const int64_t n_cells = (edges ? 2 : 1) * N * D;
    /* per-(port, slot, target) partial sums of one emission step, with a touched list */
    float* acc = (float*)malloc(sizeof(float) * (size_t)(n_cells + 1));
    uint8_t* touched = (uint8_t*)calloc((size_t)(n_cells + 1), 1);
    int64_t* list = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_cells + 1));
    int32_t* pop = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N + 1));
    if (!acc || !touched || !list || !pop) {
        free(acc); free(touched); free(list); free(pop);
        return -1;
    }
    for (int64_t i = 0; i < N; ++i) pop[i] = pop_of((int32_t)i, n_pops, pop_first);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t n_ev = 0;
    int rc = 0;

    for (int32_t k = 0; k < n_steps; ++k) {
        const int64_t s = step0 + k;
        /* 1. delivery of the spikes of step s-1: sources ascending, links in canonical order */
        int64_t n_list = 0;
        for (int64_t src = 0; src < N; ++src) {
            if (!prev_spk[src]) continue;
            for (int64_t l = src_ptr[src]; l < src_ptr[src + 1]; ++l) {
                const int32_t d = delay[l];
                if (d < 1 || d > d_slots) { rc = -1; goto done; }
                int64_t cell = ((s - 1 + d) % D) * N + tgt[l];
                float val = w[l];
                if (edges) {
                    /* the edge forward before receive (TODO.md D-6/D-7; test_fused_generated_gpu.py:361-398):
                     * gain w * g; plastic w += f, stored, then the new w; the port picks the cells */
                    if (kind[l] == 1) {
                        val = w[l] * attr2[l];
                    } else if (kind[l] == 2) {
                        w[l] = w[l] + attr2[l];
                        val = w[l];
                    } else if (kind[l] != 0) {
                        rc = -1;
                        goto done;
                    }
                    if (port[l] == 1) cell += N * D;
                }
                if (!touched[cell]) {
                    touched[cell] = 1;
                    list[n_list++] = cell;
                    acc[cell] = 0.0f + val;
                } else if (accum_mode == 1) {
                    acc[cell] = acc[cell] * 0.5f + val;
                } else {
                    acc[cell] = acc[cell] + val;
                }
                ++n_ev;
            }
        }
        for (int64_t j = 0; j < n_list; ++j) {
            const int64_t cell = list[j];
            if (cell >= N * D) {
                ring_v[cell - N * D] = ring_v[cell - N * D] + acc[cell];
            } else {
                ring[cell] = accum_mode == 1 ? ring[cell] * 0.5f + acc[cell] : ring[cell] + acc[cell];
            }
            touched[cell] = 0;
        }
        /* 2-4. read and clear the due slot, background, LIF update, threshold */
        const int64_t slot = (s % D) * N;
        for (int64_t i = 0; i < N; ++i) {
            const or_pop* P = &pops[pop[i]];
            /* multi-rate (population.py:382-410): stride k > 1 updates at (s + 1) % k == 0 only, over its
             * window's slots oldest first, ((r[s-k+1] + r[s-k+2]) + ... + r[s-1]) + r[s]; other steps leave
             * the ring and the state as they are and emit nothing */
            const int64_t k_i = P->stride > 1 ? P->stride : 1;
            if ((s + 1) % k_i != 0) {
                prev_spk[i] = 0;
                if (raster) raster[(int64_t)k * N + i] = 0;
                continue;
            }
            float x = ring[slot + i];
            ring[slot + i] = 0.0f;
            float xv = 0.0f; /* the potential port's due input (edge plans) */
            if (edges) {
                xv = ring_v[slot + i];
                ring_v[slot + i] = 0.0f;
            }
            if (k_i > 1) {
                int64_t q = (s - k_i + 1) % D;
                float xw = ring[q * N + i];
                ring[q * N + i] = 0.0f;
                for (int64_t j = 2; j < k_i; ++j) {
                    q = (q + 1) % D;
                    xw = xw + ring[q * N + i];
                    ring[q * N + i] = 0.0f;
                }
                x = xw + x;
            }
            const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)s, (uint32_t)((uint64_t)s >> 32), instance};
            uint32_t r[4];
            or_philox4x32_10(ctr, key, r);
            const int cnt = or_poisson_count(r[0], P->poisson_thr);
            const float in = x + P->j_ext * (float)cnt;
            uint8_t fired = 0;
            if (model == OR_LIF_DELTA) {
                if (ref[i] > 0) {
                    ref[i] -= 1;
                } else {
                    const float vd = v[i] * P->decay;
                    v[i] = vd + in;
                    if (v[i] >= P->v_th) {
                        fired = 1;
                        v[i] = P->v_reset;
                        ref[i] = P->ref_steps;
                    }
                }
            } else {
                /* receive (current += input), then update(t_end) as test/fused_reference.py:86-99 */
                const float I = isyn[i] + in;
                if (ref[i] > 0) {
                    v[i] = P->v_reset;
                    ref[i] -= 1;
                } else {
                    const float dv = v[i] - P->v_rest;
                    const float pv = P->decay * dv;
                    const float vr = P->v_rest + pv;
                    const float pvi = P->p_vi * I;
                    v[i] = vr + pvi;
                    if (edges) v[i] = v[i] + xv; /* the potential port: a jump after the propagation */
                }
                isyn[i] = P->p_i * I;
                /* send(): threshold, reset, refractory (test/fused_reference.py:100-103) */
                if (v[i] > P->v_th) {
                    fired = 1;
                    v[i] = P->v_reset;
                    ref[i] = P->ref_steps;
                }
            }
            prev_spk[i] = fired;
            if (raster) raster[(int64_t)k * N + i] = fired;
        }
    }
done:
    if (events) *events = n_ev;
    free(acc); free(touched); free(list); free(pop);
    return rc;
}

int or_run(int model, int32_t n_neurons, int32_t n_pops, const int32_t* pop_first, const or_pop* pops,
           int32_t d_slots, uint64_t seed, uint32_t instance,
           const int64_t* src_ptr, const int32_t* tgt, const float* w, const int32_t* delay,
           float* v, float* isyn, int32_t* ref, float* ring, uint8_t* prev_spk,
           int64_t step0, int32_t n_steps, uint8_t* raster, int32_t accum_mode, int64_t* events) {
    return run_core(model, n_neurons, n_pops, pop_first, pops, d_slots, seed, instance, src_ptr, tgt, (float*)w, delay,
                    NULL, NULL, NULL, NULL, v, isyn, ref, ring, prev_spk, step0, n_steps, raster, accum_mode, events);
}

int or_run_edges(int32_t n_neurons, int32_t n_pops, const int32_t* pop_first, const or_pop* pops, int32_t d_slots,
                 uint64_t seed, uint32_t instance, const int64_t* src_ptr, const int32_t* tgt, float* w,
                 const int32_t* delay, const uint8_t* kind, const uint8_t* port, const float* attr2, float* ring_v,
                 float* v, float* isyn, int32_t* ref, float* ring, uint8_t* prev_spk, int64_t step0, int32_t n_steps,
                 uint8_t* raster, int64_t* events) {
    if (!kind || !port || !ring_v) return -1;
    return run_core(OR_LIF_EXP, n_neurons, n_pops, pop_first, pops, d_slots, seed, instance, src_ptr, tgt, w, delay,
                    kind, port, attr2, ring_v, v, isyn, ref, ring, prev_spk, step0, n_steps, raster, 0, events);
}

/* ---- the same step, OpenMP-parallel over target ranges (the CPU baseline at all host cores) ----
 * Thread t owns the targets [lo_t, hi_t): its own source-major list of the links into them (canonical
 * order kept per source), its partial sums, its ring cells and its neurons' update.  Every per-target sum
 * therefore sees the same links in the same order as or_run, and the results are identical. */
struct or_mt {
    int32_t n_neurons, n_threads;
    int64_t* lo;        /* [T + 1] target range bounds */
    int64_t** ptr;      /* per thread [N + 1] */
    int32_t** tgt;
    float** w;
    int32_t** d;
};

or_mt* or_mt_create(int32_t n_neurons, const int64_t* src_ptr, const int32_t* tgt, const float* w,
                    const int32_t* delay, int32_t n_threads) {
    if (n_neurons < 1 || n_threads < 1) return NULL;
    if (n_threads > n_neurons) n_threads = n_neurons;
    or_mt* m = (or_mt*)calloc(1, sizeof(or_mt));
    if (!m) return NULL;
    const int64_t N = n_neurons, T = n_threads;
    m->n_neurons = n_neurons;
    m->n_threads = n_threads;
    m->lo = (int64_t*)malloc(sizeof(int64_t) * (size_t)(T + 1));
    m->ptr = (int64_t**)calloc((size_t)T, sizeof(int64_t*));
    m->tgt = (int32_t**)calloc((size_t)T, sizeof(int32_t*));
    m->w = (float**)calloc((size_t)T, sizeof(float*));
    m->d = (int32_t**)calloc((size_t)T, sizeof(int32_t*));
    if (!m->lo || !m->ptr || !m->tgt || !m->w || !m->d) {
        or_mt_destroy(m);
        return NULL;
    }
    for (int64_t t = 0; t <= T; ++t) m->lo[t] = N * t / T;
    int ok = 1;
#pragma omp parallel for num_threads(n_threads) schedule(static, 1)
    for (int64_t t = 0; t < T; ++t) {
        const int64_t lo = m->lo[t], hi = m->lo[t + 1];
        int64_t cnt = 0;
        for (int64_t l = 0; l < src_ptr[N]; ++l) cnt += (tgt[l] >= lo && tgt[l] < hi);
        int64_t* p = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N + 1));
        int32_t* tg = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cnt + 1));
        float* ww = (float*)malloc(sizeof(float) * (size_t)(cnt + 1));
        int32_t* dd = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cnt + 1));
        m->ptr[t] = p;
        m->tgt[t] = tg;
        m->w[t] = ww;
        m->d[t] = dd;
        if (!p || !tg || !ww || !dd) {
#pragma omp atomic write
            ok = 0;
            continue;
        }
        int64_t k = 0;
        for (int64_t s = 0; s < N; ++s) {
            p[s] = k;
            for (int64_t l = src_ptr[s]; l < src_ptr[s + 1]; ++l) {
                if (tgt[l] >= lo && tgt[l] < hi) {
                    tg[k] = tgt[l];
                    ww[k] = w[l];
                    dd[k] = delay[l];
                    ++k;
                }
            }
        }
        p[N] = k;
    }
    if (!ok) {
        or_mt_destroy(m);
        return NULL;
    }
    return m;
}

void or_mt_destroy(or_mt* m) {
    if (!m) return;
    for (int32_t t = 0; t < m->n_threads; ++t) {
        if (m->ptr) free(m->ptr[t]);
        if (m->tgt) free(m->tgt[t]);
        if (m->w) free(m->w[t]);
        if (m->d) free(m->d[t]);
    }
    free(m->ptr);
    free(m->tgt);
    free(m->w);
    free(m->d);
    free(m->lo);
    free(m);
}

int or_mt_run(or_mt* m, int model, int32_t n_pops, const int32_t* pop_first, const or_pop* pops, int32_t d_slots,
              uint64_t seed, uint32_t instance, float* v, float* isyn, int32_t* ref, float* ring, uint8_t* prev_spk,
              int64_t step0, int32_t n_steps, uint8_t* raster, int64_t* events) {
    if (!m || n_pops < 1 || d_slots < 1 || n_steps < 0) return -1;
    if (multirate(n_pops, pops, d_slots) != 0) return -1; /* every population at stride 1 */
    if (model != OR_LIF_DELTA && model != OR_LIF_EXP) return -1;
    if (model == OR_LIF_EXP && !isyn) return -1;
    const int64_t N = m->n_neurons, D = d_slots;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t n_ev = 0;
    int rc = 0;
#pragma omp parallel num_threads(m->n_threads) reduction(+ : n_ev)
    {
        const int t = omp_get_thread_num();
        const int64_t lo = m->lo[t], hi = m->lo[t + 1], nt = hi - lo;
        const int64_t* p = m->ptr[t];
        const int32_t* tg = m->tgt[t];
        const float* ww = m->w[t];
        const int32_t* dd = m->d[t];
        float* acc = (float*)malloc(sizeof(float) * (size_t)(nt * D + 1));
        uint8_t* touched = (uint8_t*)calloc((size_t)(nt * D + 1), 1);
        int64_t* list = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nt * D + 1));
        int32_t* pop = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nt + 1));
        int mine_ok = acc && touched && list && pop;
        if (mine_ok)
            for (int64_t i = lo; i < hi; ++i) pop[i - lo] = pop_of((int32_t)i, n_pops, pop_first);
        for (int32_t k = 0; k < n_steps; ++k) {
            const int64_t s = step0 + k;
            int64_t n_list = 0;
            if (mine_ok) {
                for (int64_t src = 0; src < N; ++src) {
                    if (!prev_spk[src]) continue;
                    for (int64_t l = p[src]; l < p[src + 1]; ++l) {
                        const int32_t d = dd[l];
                        if (d < 1 || d > d_slots) {
                            mine_ok = 0;
                            break;
                        }
                        const int64_t cell = ((s - 1 + d) % D) * nt + (tg[l] - lo);
                        if (!touched[cell]) {
                            touched[cell] = 1;
                            list[n_list++] = cell;
                            acc[cell] = 0.0f + ww[l];
                        } else {
                            acc[cell] = acc[cell] + ww[l];
                        }
                        ++n_ev;
                    }
                }
                for (int64_t j = 0; j < n_list; ++j) {
                    const int64_t cell = list[j];
                    const int64_t slot = cell / nt, i = lo + cell % nt;
                    ring[slot * N + i] = ring[slot * N + i] + acc[cell];
                    touched[cell] = 0;
                }
            }
#pragma omp barrier
            if (mine_ok) {
                const int64_t slot = (s % D) * N;
                for (int64_t i = lo; i < hi; ++i) {
                    const or_pop* P = &pops[pop[i - lo]];
                    const float x = ring[slot + i];
                    ring[slot + i] = 0.0f;
                    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)s, (uint32_t)((uint64_t)s >> 32), instance};
                    uint32_t r[4];
                    or_philox4x32_10(ctr, key, r);
                    const int cnt = or_poisson_count(r[0], P->poisson_thr);
                    const float in = x + P->j_ext * (float)cnt;
                    uint8_t fired = 0;
                    if (model == OR_LIF_DELTA) {
                        if (ref[i] > 0) {
                            ref[i] -= 1;
                        } else {
                            const float vd = v[i] * P->decay;
                            v[i] = vd + in;
                            if (v[i] >= P->v_th) {
                                fired = 1;
                                v[i] = P->v_reset;
                                ref[i] = P->ref_steps;
                            }
                        }
                    } else {
                        const float I = isyn[i] + in;
                        if (ref[i] > 0) {
                            v[i] = P->v_reset;
                            ref[i] -= 1;
                        } else {
                            const float dv = v[i] - P->v_rest;
                            const float pv = P->decay * dv;
                            const float vr = P->v_rest + pv;
                            const float pvi = P->p_vi * I;
                            v[i] = vr + pvi;
                        }
                        isyn[i] = P->p_i * I;
                        if (v[i] > P->v_th) {
                            fired = 1;
                            v[i] = P->v_reset;
                            ref[i] = P->ref_steps;
                        }
                    }
                    prev_spk[i] = fired;
                    if (raster) raster[(int64_t)k * N + i] = fired;
                }
            }
#pragma omp barrier
        }
        if (!mine_ok) {
#pragma omp atomic write
            rc = -1;
        }
        free(acc);
        free(touched);
        free(list);
        free(pop);
    }
    if (events) *events = n_ev;
    return rc;
}
