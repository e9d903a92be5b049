/*
 * event_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded CPU restatement of the event-driven, exact-time spiking network that
 * libnaer2_hip.so runs on the GPU (n2h_ev_*).  It is the checker of the HIP path: only the tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it; the product never calls it.
 *
 * Followed reference code (MatteoSchmider/naer2):
 *   test/networks/Brunel/neuron.py:6-45   the neuron: receive decays V over the time since its last exact
 *                                         change and adds the weight, unless refractory (time < t_ref_end);
 *                                         send, after the inputs at one exact time, fires at V >= theta,
 *                                         resets to V_r, t_ref_end = last_time = time + t_ref
 *   test/networks/Brunel/poisson.py:7-27  the source: exponential intervals from 0, every spike before the
 *                                         run's end emitted, next_spike kept
 *   naer2/transpiler/templates/fused.py:2084-2301 (_pull_queues), 1300-1370 (_merge_segments),
 *                 1421-1526 (_walk_batch) the delivery order: a source run's spikes in (e, node) order,
 *                                         connections merged by arrival e + D_c (ties: the lower connection
 *                                         first), each spike's links in link order; per target the inputs
 *                                         at one time received one by one, then send once
 *   BRUNEL.md / TODO.md:148-171 (D-1..D-3) exact arrival e + D; spikes inside delivery at the arrival time
 *
 * The contract (identical on the GPU):
 *   - Every connection c has one delay D_c, a whole multiple lag_c >= 1 of the window W (the smallest delay).
 *     Window k emits spikes; window k + lag_c delivers them through c.  t_end(k) = (float)((double)(k + 1) * W),
 *     t_start(k) = (float)((double)k * W).
 *   - Per target neuron, the inputs of a window in ascending (a, c, source, link): a = e + D_c (fp32), c the
 *     connection index, link the canonical link index.  (The reference orders a connection's spikes of one
 *     arrival time by emission time, then source; the two differ only when two emission times round to the
 *     same fp32 arrival.)  For each input: if its a
 *     differs from the previous input's, the threshold is checked at the previous a first; then, if
 *     a >= t_ref_end: V = V * ev_exp((last - a) * tau_inv) + w, last = a.  After the window's last input, the
 *     threshold is checked at its a.  Check at t: if V >= theta: spike at t, V = V_r,
 *     t_ref_end = last = t + t_ref.
 *   - Poisson source i in window k: next < 0: next = interval(m = 0); while next < t_end(k): spike at next,
 *     next = next + interval(m + 1), ...; interval = (-ev_log(u)) * rate_inv, u = ((x >> 9) + 0.5) 2^-23,
 *     x = word 0 of Philox4x32-10(ctr = {node, (uint32)k, m, instance}, key = {seed lo, seed hi}).
 *   - Spike-train source (tests): its listed times t with t_start(k) <= t < t_end(k).
 *   - ev_exp / ev_log: the fp32 functions below (exactly these operations, -ffp-contract=off).
 */
#ifndef EVENT_ORACLE_H
#define EVENT_ORACLE_H
#include <stdint.h>

#define OR_EV_NEURON 0
#define OR_EV_POISSON 1
#define OR_EV_TRAIN 2

typedef struct or_ev_pop {
    int32_t kind;
    int32_t reserved;
    float tau_inv;   /* neuron: 1 / tau (fp32) */
    float theta;
    float v_reset;
    float t_ref;
    float rate_inv;  /* Poisson: 1 / rate, ms */
    float reserved_f;
} or_ev_pop;

float or_ev_exp(float x);
float or_ev_log(float u);
float or_ev_interval(uint32_t node, int64_t window, uint32_t m, uint64_t seed, uint32_t instance, float rate_inv);

/* Runs n_windows windows from k0.  Connections: conn_src/conn_tgt population of each (one connection per
 * (source, target) population pair), in creation order, conn_delay (ms, fp32) and conn_lag (= delay / window).  Topology: canonical CSR by source over all
 * connections.  Trains: train_ptr [n + 1] / train_t (sorted per source), or NULL.
 * State [n]: v, last, tre (t_ref_end), nxt (Poisson next spike, -1 before the first).
 * prev: the spikes of the M = max lag windows before k0 on entry, of the last M windows on return (each
 *   window's in (pop, e, node) order, oldest window first): *prev_n entries of prev_node / prev_e / prev_win
 *   (the emission window), capacity prev_cap.
 * rec_node / rec_e (capacity rec_cap, *rec_n advanced) and rec_wcount [n_windows]: every window's spikes in
 * (pop, e, node) order, or NULL.  events: inputs delivered (out).  log_rec / log_a (capacity log_cap, *log_n
 * advanced, or NULL): every delivered input as {window, connection, source, target} (indices in their
 * populations) and its arrival, per target in delivery order, targets ascending.  Returns 0, -1 on a bad
 * argument, -2 when memory or a capacity runs out. */
int or_ev_run(int32_t n_neurons, int32_t n_pops, const int32_t* pop_first, const or_ev_pop* pops,
              int32_t n_conns, const int32_t* conn_src, const int32_t* conn_tgt, const float* conn_delay,
              const int32_t* conn_lag, float window, uint64_t seed,
              uint32_t instance, const int64_t* src_ptr, const int32_t* tgt, const float* w,
              const int64_t* train_ptr, const float* train_t, float* v, float* last, float* tre, float* nxt,
              int64_t* prev_n, int32_t* prev_node, float* prev_e, int64_t* prev_win, int64_t prev_cap, int64_t k0,
              int32_t n_windows,
              int32_t* rec_node, float* rec_e, int64_t rec_cap, int64_t* rec_n, int32_t* rec_wcount,
              int64_t* events, int32_t* log_rec, float* log_a, int64_t log_cap, int64_t* log_n);

#endif
