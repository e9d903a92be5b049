/* event_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU checker of the event-driven path (event_oracle.h). */
#include "event_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "naer2_oracle.h"

/* ---------------------------------------------------------------- fp32 exp / log of the contract */

static float bits_f(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static uint32_t f_bits(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}

/* exp(x) for x <= 88: Cody-Waite reduction x = n ln2 + r (ln 2 split so that n * LN2_HI is exact), the degree-7
 * Taylor polynomial of exp(r) by Estrin's scheme (pairs, then r^2, then r^4: a short dependency chain), scaled
 * by 2^n; 0 below -87. */
float or_ev_exp(float x) {
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
    return ldexpf(q03 + r4 * q47, (int)n);
}

/* log(u) for finite u > 0: u = 2^e m with m in [sqrt(1/2), sqrt(2)), log(m) = t P(t), t = m - 1 (exact), P a
 * degree-10 least-squares fit of log1p(t)/t on 4097 Chebyshev nodes of [sqrt(1/2) - 1, sqrt(2) - 1]
 * (coefficients rounded to fp32; relative error below 2^-22 over the interval). */
float or_ev_log(float u) {
    const uint32_t b = f_bits(u);
    int e = (int)((b >> 23) & 0xFF) - 127;
    float m = bits_f((b & 0x7FFFFFu) | 0x3F800000u);
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
    const float fe = (float)e;
    float r = t * p;
    r = r + fe * 1.42860677e-06f;
    return r + fe * 0.693145751953125f;
}

float or_ev_interval(uint32_t node, int64_t window, uint32_t m, uint64_t seed, uint32_t instance, float rate_inv) {
    const uint32_t ctr[4] = {node, (uint32_t)window, m, instance};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    or_philox4x32_10(ctr, key, r);
    const float u = ((float)(r[0] >> 9) + 0.5f) * 1.1920928955078125e-07f;
    return (-or_ev_log(u)) * rate_inv;
}

/* ---------------------------------------------------------------- one run */

typedef struct {
    float a;
    int32_t c;
    float e;
    int32_t node;
} entry;

static int entry_cmp(const void* x, const void* y) {
    const entry* p = (const entry*)x;
    const entry* q = (const entry*)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->c != q->c) return p->c < q->c ? -1 : 1;
    return (p->node > q->node) - (p->node < q->node);
}

typedef struct {
    int32_t pop;
    float e;
    int32_t node;
} spike;

static int spike_cmp(const void* x, const void* y) {
    const spike* p = (const spike*)x;
    const spike* q = (const spike*)y;
    if (p->pop != q->pop) return p->pop < q->pop ? -1 : 1;
    if (p->e != q->e) return p->e < q->e ? -1 : 1;
    return (p->node > q->node) - (p->node < q->node);
}

typedef struct {
    spike* s;
    int64_t n, cap;
} spikes;

static int push(spikes* l, int32_t pop, float e, int32_t node) {
    if (l->n == l->cap) {
        const int64_t cap = l->cap ? 2 * l->cap : 1024;
        spike* s = (spike*)realloc(l->s, (size_t)cap * sizeof(spike));
        if (!s) return -2;
        l->s = s;
        l->cap = cap;
    }
    l->s[l->n].pop = pop;
    l->s[l->n].e = e;
    l->s[l->n].node = node;
    l->n++;
    return 0;
}

static int32_t pop_of(int32_t n_pops, const int32_t* pop_first, int32_t node) {
    int32_t p = 0;
    while (p + 1 < n_pops && node >= pop_first[p + 1]) ++p;
    return p;
}

/* The threshold of neuron i at time t (send after the inputs at one time). */
static int check(int32_t i, int32_t pop, float t, const or_ev_pop* P, float* v, float* last, float* tre,
                 spikes* out) {
    if (v[i] >= P->theta) {
        v[i] = P->v_reset;
        tre[i] = t + P->t_ref;
        last[i] = tre[i];
        return push(out, pop, t, i);
    }
    return 0;
}

int or_ev_run(int32_t n, int32_t n_pops, const int32_t* pop_first, const or_ev_pop* pops, int32_t n_conns,
              const int32_t* conn_src, const int32_t* conn_tgt, const float* conn_delay, const int32_t* conn_lag,
              float window, uint64_t seed, uint32_t instance,
              const int64_t* src_ptr, const int32_t* tgt, const float* w, const int64_t* train_ptr,
              const float* train_t, float* v, float* last, float* tre, float* nxt, int64_t* prev_n,
              int32_t* prev_node, float* prev_e, int64_t* prev_win, int64_t prev_cap, int64_t k0, int32_t n_windows,
              int32_t* rec_node,
              float* rec_e, int64_t rec_cap, int64_t* rec_n, int32_t* rec_wcount, int64_t* events, int32_t* log_rec,
              float* log_a, int64_t log_cap, int64_t* log_n) {
    if (n < 0 || n_pops < 1 || n_windows < 0 || !(window > 0.0f)) return -1;
    int32_t M = 1;  /* windows of spikes kept: the largest lag */
    for (int32_t c = 0; c < n_conns; ++c) {
        if (conn_lag[c] < 1) return -1;
        if (conn_lag[c] > M) M = conn_lag[c];
    }
    int rc = 0;
    spikes cur = {0, 0, 0};
    spikes* hist = (spikes*)calloc((size_t)M, sizeof(spikes));
    int64_t* hist_w = (int64_t*)malloc((size_t)M * sizeof(int64_t));
    entry* ent = NULL;
    int64_t ent_cap = 0;
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    int64_t* fill = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    float* in_a = NULL;
    float* in_w = NULL;
    int32_t* in_c = NULL;
    int32_t* in_s = NULL;
    int64_t in_cap = 0;
    int32_t* pop_at = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
    if (!cnt || !fill || !pop_at || !hist || !hist_w) {
        rc = -2;
        goto done;
    }
    for (int32_t i = 0; i < n; ++i) pop_at[i] = pop_of(n_pops, pop_first, i);
    for (int32_t j = 0; j < M; ++j) hist_w[j] = -1;
    for (int64_t j = 0; j < *prev_n; ++j) {
        const int64_t pw = prev_win[j];
        if (pw < k0 - M || pw >= k0) continue;
        const int32_t sl = (int32_t)(pw % M);
        if (hist_w[sl] != pw) {
            hist[sl].n = 0;
            hist_w[sl] = pw;
        }
        if ((rc = push(&hist[sl], pop_at[prev_node[j]], prev_e[j], prev_node[j]))) goto done;
    }
    if (events) *events = 0;

    for (int32_t kw = 0; kw < n_windows; ++kw) {
        const int64_t k = k0 + kw;
        const float t_start = (float)((double)k * (double)window);
        const float t_end = (float)((double)(k + 1) * (double)window);
        cur.n = 0;
        /* 1. deliveries of the previous window's spikes, target population by target population */
        for (int32_t T = 0; T < n_pops; ++T) {
            if (pops[T].kind != OR_EV_NEURON) continue;
            int64_t ne = 0;
            for (int32_t c = 0; c < n_conns; ++c) {
                if (conn_tgt[c] != T) continue;
                const int64_t from = k - conn_lag[c];  /* the emission window whose spikes arrive now */
                if (from < 0 || hist_w[from % M] != from) continue;
                const spikes* prev = &hist[from % M];
                for (int64_t j = 0; j < prev->n; ++j) {
                    if (prev->s[j].pop != conn_src[c]) continue;
                    if (ne == ent_cap) {
                        ent_cap = ent_cap ? 2 * ent_cap : 1024;
                        entry* x = (entry*)realloc(ent, (size_t)ent_cap * sizeof(entry));
                        if (!x) {
                            rc = -2;
                            goto done;
                        }
                        ent = x;
                    }
                    ent[ne].a = prev->s[j].e + conn_delay[c];
                    ent[ne].c = c;
                    ent[ne].e = prev->s[j].e;
                    ent[ne].node = prev->s[j].node;
                    ++ne;
                }
            }
            if (ne == 0) continue;
            qsort(ent, (size_t)ne, sizeof(entry), entry_cmp);
            const int32_t lo = pop_first[T], hi = pop_first[T + 1];
            /* each target's inputs in entry order, then link order: count, place, walk */
            for (int32_t i = lo; i <= hi; ++i) cnt[i] = 0;
            int64_t total = 0;
            for (int64_t q = 0; q < ne; ++q)
                for (int64_t l = src_ptr[ent[q].node]; l < src_ptr[ent[q].node + 1]; ++l)
                    if (tgt[l] >= lo && tgt[l] < hi) {
                        cnt[tgt[l]]++;
                        ++total;
                    }
            if (total > in_cap) {
                float* xa = (float*)realloc(in_a, (size_t)total * sizeof(float));
                if (!xa) {
                    rc = -2;
                    goto done;
                }
                in_a = xa;
                float* xw = (float*)realloc(in_w, (size_t)total * sizeof(float));
                if (!xw) {
                    rc = -2;
                    goto done;
                }
                in_w = xw;
                if (log_rec) {
                    int32_t* xc = (int32_t*)realloc(in_c, (size_t)total * sizeof(int32_t));
                    if (!xc) {
                        rc = -2;
                        goto done;
                    }
                    in_c = xc;
                    int32_t* xs = (int32_t*)realloc(in_s, (size_t)total * sizeof(int32_t));
                    if (!xs) {
                        rc = -2;
                        goto done;
                    }
                    in_s = xs;
                }
                in_cap = total;
            }
            int64_t acc = 0;
            for (int32_t i = lo; i < hi; ++i) {
                fill[i] = acc;
                acc += cnt[i];
            }
            for (int64_t q = 0; q < ne; ++q)
                for (int64_t l = src_ptr[ent[q].node]; l < src_ptr[ent[q].node + 1]; ++l)
                    if (tgt[l] >= lo && tgt[l] < hi) {
                        in_a[fill[tgt[l]]] = ent[q].a;
                        in_w[fill[tgt[l]]] = w[l];
                        if (log_rec) {
                            in_c[fill[tgt[l]]] = ent[q].c;
                            in_s[fill[tgt[l]]] = ent[q].node;
                        }
                        fill[tgt[l]]++;
                    }
            if (events) *events += total;
            const or_ev_pop* P = &pops[T];
            int64_t at = 0;
            for (int32_t i = lo; i < hi; ++i) {
                float aprev = 0.0f;
                int pend = 0;
                for (int64_t q = 0; q < cnt[i]; ++q, ++at) {
                    const float a = in_a[at];
                    if (log_rec) {
                        if (*log_n < log_cap) {
                            int32_t* r = log_rec + 4 * *log_n;
                            r[0] = (int32_t)k;
                            r[1] = in_c[at];
                            r[2] = in_s[at] - pop_first[pop_at[in_s[at]]];
                            r[3] = i - lo;
                            log_a[*log_n] = a;
                        }
                        ++*log_n;
                    }
                    if (pend && a != aprev)
                        if ((rc = check(i, T, aprev, P, v, last, tre, &cur))) goto done;
                    if (a >= tre[i]) {
                        v[i] = v[i] * or_ev_exp((last[i] - a) * P->tau_inv) + in_w[at];
                        last[i] = a;
                    }
                    aprev = a;
                    pend = 1;
                }
                if (pend)
                    if ((rc = check(i, T, aprev, P, v, last, tre, &cur))) goto done;
            }
        }
        /* 2. the sources' spikes of this window */
        for (int32_t p = 0; p < n_pops; ++p) {
            if (pops[p].kind == OR_EV_POISSON) {
                for (int32_t i = pop_first[p]; i < pop_first[p + 1]; ++i) {
                    float next = nxt[i];
                    uint32_t m = 0;
                    if (next < 0.0f) next = or_ev_interval((uint32_t)i, k, m++, seed, instance, pops[p].rate_inv);
                    while (next < t_end) {
                        if ((rc = push(&cur, p, next, i))) goto done;
                        next = next + or_ev_interval((uint32_t)i, k, m++, seed, instance, pops[p].rate_inv);
                    }
                    nxt[i] = next;
                }
            } else if (pops[p].kind == OR_EV_TRAIN) {
                if (!train_ptr) {
                    rc = -1;
                    goto done;
                }
                for (int32_t i = pop_first[p]; i < pop_first[p + 1]; ++i)
                    for (int64_t j = train_ptr[i]; j < train_ptr[i + 1]; ++j)
                        if (train_t[j] >= t_start && train_t[j] < t_end)
                            if ((rc = push(&cur, p, train_t[j], i))) goto done;
            }
        }
        /* 3. the window's spikes in (pop, e, node) order: recorded, and delivered by the next window */
        if (cur.n > 1) qsort(cur.s, (size_t)cur.n, sizeof(spike), spike_cmp);
        if (rec_wcount) rec_wcount[kw] = (int32_t)cur.n;
        if (rec_node && rec_e) {
            if (*rec_n + cur.n > rec_cap) {
                rc = -2;
                goto done;
            }
            for (int64_t j = 0; j < cur.n; ++j) {
                rec_node[*rec_n + j] = cur.s[j].node;
                rec_e[*rec_n + j] = cur.s[j].e;
            }
            *rec_n += cur.n;
        }
        const int32_t sl = (int32_t)(k % M);
        spikes t = hist[sl];
        hist[sl] = cur;
        hist_w[sl] = k;
        cur = t;
    }
    {
        /* the spikes of the last M windows, oldest first */
        int64_t np = 0;
        const int64_t k_end = k0 + n_windows;
        for (int64_t pw = k_end - M; pw < k_end; ++pw) {
            if (pw < 0 || hist_w[pw % M] != pw) continue;
            const spikes* h = &hist[pw % M];
            if (np + h->n > prev_cap) {
                rc = -2;
                goto done;
            }
            for (int64_t j = 0; j < h->n; ++j, ++np) {
                prev_node[np] = h->s[j].node;
                prev_e[np] = h->s[j].e;
                prev_win[np] = pw;
            }
        }
        *prev_n = np;
    }
done:
    if (hist)
        for (int32_t j = 0; j < M; ++j) free(hist[j].s);
    free(hist);
    free(hist_w);
    free(cur.s);
    free(ent);
    free(cnt);
    free(fill);
    free(in_a);
    free(in_w);
    free(in_c);
    free(in_s);
    free(pop_at);
    return rc;
}
