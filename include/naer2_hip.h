/*
 * naer2_hip.h -- the C-ABI of libnaer2_hip.so, the MI355X (gfx950) execution backend of naer2's
 * time-driven spiking-network step.
 *
 * One call of n2h_step advances a network by whole steps of dt.  Each step is ONE fused kernel launch:
 *   1. pull intake + delivery of the previous step's spikes into the delayed-input ring
 *      (each target wavefront gathers the group segments of the spiking sources, in source order),
 *   2. the ring slot that is due is read and cleared,
 *   3. the LIF update (delta synapse, or exponential current synapse with the exact propagator)
 *      with an in-kernel Poisson background drawn from Philox4x32-10,
 *   4. the threshold, and the spikes written as one 64-bit word per target group (the step's bitmap).
 *
 * What each entry point replaces in the reference (file:line in MatteoSchmider/naer2):
 *   n2h_plan_create / n2h_plan_destroy  naer2/runtime.py:213-323 (PopulationRuntime.__init__ / connect:
 *                                       state, ring, plan, history of one population) and
 *                                       naer2/transpiler/templates/fused.py:121-172 (FusedPlan)
 *   n2h_bind_state                      naer2/runtime.py:490-504 (_launch_args: one pointer per dynamic attr)
 *   n2h_bind_ring                       naer2/ring.py:132-245 (Ring) and naer2/history.py:76-130
 *                                       (EmissionHistory: the spikes of a run)
 *   n2h_step                            naer2/runtime.py:506-526 (_launch) and naer2/fusion.py:137-141
 *                                       (FusedRound.launch: every population of a round in one grid)
 *   n2h_graph_create / _launch / _destroy
 *                                       naer2/graph.py:35-136 (WindowGraph: a captured window of rounds)
 *   n2h_set_step / n2h_get_step         naer2/schedule.py:30-36 (write_entry: the run's round and times)
 *   n2h_status                          plan metadata.  The reference's check (naer2/runtime.py:724-735) judges
 *                                       flags its kernels raise on ring, history or log overflow; the
 *                                       time-driven step here has nothing that can overflow (its spike lists
 *                                       hold every neuron, its ring every delay), so it raises none
 *   n2h_last_error                      (error text of the last failed call; the reference raises RuntimeError)
 *
 * Conventions.  Every function returns an int status: N2H_OK (0) or a negative N2H_E* code; the text of
 * the last error is copied out by n2h_last_error.  No exception crosses the ABI.  Device buffers are owned
 * by the caller (PyTorch-ROCm tensors) and passed as plain pointers; the library allocates only its small
 * per-plan parameter block and step counter.  Calls that take a stream only enqueue work on it and never
 * synchronize, so they can be captured into a hipGraph.  A plan is not safe for concurrent calls from two
 * host threads.
 */
#ifndef NAER2_HIP_H
#define NAER2_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define N2H_ABI_VERSION 1

/* status codes */
#define N2H_OK 0
#define N2H_EINVAL (-1)   /* a bad argument (null pointer, size, layout) */
#define N2H_EHIP (-2)     /* a HIP runtime call failed */
#define N2H_ESTATE (-3)   /* the plan is missing a binding */
#define N2H_ENOMEM (-4)   /* a host or device allocation failed */

/* model families (one per plan) */
#define N2H_MODEL_LIF_DELTA 1 /* delta-synapse LIF: V = V*decay + input (tools/brunel_reference.py:34-48) */
#define N2H_MODEL_LIF_EXP 2   /* exp-synapse LIF, exact propagator (test/networks/PD14/neuron.py:37-60) */

#define N2H_GROUP 64        /* neurons per target group: one wavefront */
#define N2H_POISSON_K 32    /* entries of the Poisson inversion table */
#define N2H_MAX_POPS 64
#define N2H_MAX_SLOTS 128   /* ring slots: the largest delay in steps + the largest stride - 1 */
#define N2H_MAX_STRIDE 64   /* steps per update of a population with its own, coarser timestep */

/* Parameters of one population (fp32 constants precomputed on the host for its update interval).
 *
 * Multi-rate (naer2/population.py:191-204 _own_timestep, 382-410 _execute_update_kernel; naer2/network.py:80-152):
 * a population whose timestep is stride * dt runs windows of stride steps.  It updates at the steps s with
 * (s + 1) % stride == 0 only, once, with every parameter below computed for stride * dt (propagators, the
 * Poisson mean, ref_steps counted in its own updates: test/networks/PD14/neuron.py:37-49 integrates over
 * time - last_time).  Its input is the sum of the ring slots of its window, ((r[s-stride+1] + r[s-stride+2])
 * + ... + r[s-1]) + r[s]; it fires at its update steps only, spikes delivered as every other with their
 * delay from the step of the update.  At its other steps the step kernel only takes its intake into the
 * ring.  d_slots must be at least the largest delay + the largest stride - 1, so that no arrival lands in
 * a slot of a window before it is read. */
typedef struct n2h_pop_params {
    int32_t ref_steps;   /* updates a neuron stays refractory after a spike */
    int32_t stride;      /* steps per update: 0 or 1 every step, else at most N2H_MAX_STRIDE */
    float decay;         /* LIF_DELTA: exp(-dt/tau_m); LIF_EXP: P_V = exp(-dt/tau_m) */
    float p_i;           /* LIF_EXP: P_I = exp(-dt/tau_syn) */
    float p_vi;          /* LIF_EXP: R tau_syn/(tau_m - tau_syn) (P_V - P_I) */
    float v_rest;        /* LIF_EXP: resting potential */
    float v_th;          /* threshold: LIF_DELTA fires at V >= v_th, LIF_EXP at V > v_th */
    float v_reset;       /* potential after a spike (and during refractoriness for LIF_EXP) */
    float j_ext;         /* weight of one background Poisson event */
    float reserved_f;
    /* Poisson background: the count is the number of j with u >= poisson_thr[j], u a 32-bit Philox word;
     * poisson_thr[j] = min(floor(2^32 CDF(j)), 2^32-1), non-decreasing (all 0xFFFFFFFF for no drive). */
    uint32_t poisson_thr[N2H_POISSON_K];
} n2h_pop_params;

/* The network: neurons numbered 0..n_neurons-1, population p owning [pop_first[p], pop_first[p+1]). */
typedef struct n2h_model_desc {
    int32_t model;                 /* N2H_MODEL_* */
    int32_t n_neurons;
    int32_t n_pops;
    int32_t d_slots;               /* ring slots: the largest delay in steps (every delay in [1, d_slots]),
                                    * + the largest population stride - 1 */
    const int32_t* pop_first;      /* host [n_pops + 1] */
    const n2h_pop_params* pops;    /* host [n_pops] */
    uint64_t seed;                 /* Philox key */
    uint32_t instance;             /* Philox counter word 3 (batched instances draw distinct streams) */
    uint32_t reserved;
} n2h_model_desc;

/* Group-segment synapse tables, on the device.  Target group g covers neurons
 * [group_first[g], group_first[g] + (group_info[g] & 0xffff)) of population group_info[g] >> 16;
 * groups never straddle populations.  A neuron's source slot is 64 * (its group) + (its index in the
 * group).  The links of source slot p into group g are
 * links[seg_off[p * n_groups + g] .. seg_off[p * n_groups + g + 1]), in canonical link order, each packed
 * as {x = delay_steps << 16 | target - group_first[g], y = fp32 weight bits}. */
typedef struct n2h_tables {
    int32_t n_groups;
    int32_t uniform_delay;         /* delay_steps of every link when all are equal, else 0: the step then
                                    * keeps one partial sum per target instead of one per delay */
    const int32_t* group_first;    /* device [n_groups] */
    const int32_t* group_info;     /* device [n_groups] */
    const int64_t* seg_off;        /* device [64 * n_groups * n_groups + 1] */
    const uint32_t* links;         /* device [2 * n_links] */
    int64_t n_links;
} n2h_tables;

/* Launch mapping of a batch (chosen by naer2_amd/autotune.py). */
typedef struct n2h_tuning {
    int32_t order;    /* 0: plans in the given order; 1: plans with the most links first */
    int32_t xcd;      /* 0 or 1: block b runs tile b.  k > 1: the blocks the hardware deals to each of k XCDs (block
                         b to XCD b mod k) run consecutive tiles, neighbouring groups that share cache lines: */
    int32_t stage;    /* 0: the lean batched stage (128 sources and 384 links a pass, 4 row cells in flight; less
                         LDS and fewer registers a group, more groups a CU); 1: the single-plan stage (256, 512,
                         8), for single-model batches of plans of at most 1,024 groups (otherwise 0 is used) */
    int32_t run;      /* with xcd = k > 1: 0, each XCD one contiguous chunk of ceil(T / k) tiles (block b runs tile
                         (b % k) * ceil(T / k) + b / k); r in [1, 4096], runs of r consecutive tiles dealt to the
                         k XCDs in turn (block b runs tile ((b / k / r) * k + b % k) * r + (b / k) % r), so the
                         XCDs' shares interleave over the whole tile order.  Ignored by n2h_ev_set_tuning. */
} n2h_tuning;

typedef struct n2h_plan n2h_plan;
typedef struct n2h_graph n2h_graph;
typedef struct n2h_batch n2h_batch;

int n2h_version(void);
int n2h_last_error(char* buf, size_t len);

/* Validates desc and tables, copies the population parameters to the device (current device). */
int n2h_plan_create(const n2h_model_desc* desc, const n2h_tables* tables, n2h_plan** out);
int n2h_plan_destroy(n2h_plan* plan);

/* State, fp32/int32 [n_neurons] device tensors.  i_syn is required for LIF_EXP and ignored for LIF_DELTA. */
int n2h_bind_state(n2h_plan* plan, float* v, float* i_syn, int32_t* ref);

/* ring: fp32 [d_slots][n_neurons], the input arriving at step n in slot n % d_slots.
 * spike_bits: uint64 [bit_slots][n_groups], step n's spikes in slot n % bit_slots, bit b of word g =
 * neuron group_first[g] + b (bit_slots >= 2).
 * spike_list: int32 [4][1 + 64 * n_groups], step n's spikes in slot n % 4: a count, then the source slots
 * 64 * g + b of its spikes in any order.
 * Before a step n0: bit slot and list slot of n0 - 1 hold the spikes of step n0 - 1 (zero before the first
 * step) and list slot n0 % 4 has count 0. */
int n2h_bind_ring(n2h_plan* plan, float* ring, uint64_t* spike_bits, int64_t bit_slots, int32_t* spike_list);

/* The device step counter: the next step n2h_step computes. */
int n2h_set_step(n2h_plan* plan, int64_t step, void* stream);
int n2h_get_step(n2h_plan* plan, int64_t* step);   /* synchronizes with the device */

/* Enqueue n_steps fused step kernels on stream (a hipStream_t; NULL = the null stream) and advance the
 * counter by n_steps. */
int n2h_step(n2h_plan* plan, int32_t n_steps, void* stream);

/* A captured hipGraph of n_steps steps (each replay continues from the device counter).  A graph replays
 * only at steps with the phase (step mod 4) it was captured at: windows of a multiple of 4 steps keep it. */
int n2h_graph_create(n2h_plan* plan, int32_t n_steps, void* stream, n2h_graph** out);
int n2h_graph_launch(n2h_graph* graph, void* stream);
int n2h_graph_destroy(n2h_graph* graph);

/* A batch: the groups of up to 256 independent plans (at the same step) in one launch per step.  Plans of
 * both models may share a batch (a heterogeneous round: each group runs its own network's model).
 * Replaces naer2/fusion.py:31-141 (FusedRound over several runtimes).  The plans' bound buffers must stay
 * in place while the batch exists; a plan stepped on its own leaves the batch unusable (N2H_ESTATE). */
int n2h_batch_create(n2h_plan* const* plans, int32_t n_plans, const n2h_tuning* tuning, n2h_batch** out);
int n2h_batch_destroy(n2h_batch* batch);
int n2h_batch_step(n2h_batch* batch, int32_t n_steps, void* stream);
/* Bring a batch in line with its plans again (no counterpart in the reference: FusedRound is rebuilt there):
 * the step they are at, all equal, after n2h_set_step moved them (a restored checkpoint), and every plan's
 * launch arguments as n2h_batch_create copied them, so that bindings changed since (n2h_set_vtrace, the ring
 * or the state moved) take effect.  Enqueued on `stream` as one small kernel per 8 plans; allocates nothing.
 * A plan bound to a delivery log or to edges since is refused, as at creation.  Captured graphs of the
 * batch replay with the new arguments but keep the step phase they were captured at. */
int n2h_batch_sync(n2h_batch* batch, void* stream);
int n2h_batch_graph_create(n2h_batch* batch, int32_t n_steps, void* stream, n2h_graph** out);

/* Dense recurrent LIF layer for surrogate-gradient BPTT (no counterpart in the reference, which has no
 * autograd: a surrogate-gradient adjoint beside the plain path).  Model and adjoint: naer2_amd/csrc/rlif.hip.
 * All arithmetic is fp32; the state is compensated (each of I and V an unevaluated sum hi + lo of two floats).
 * Buffers are fp32, row-major: u [T][B][H] (= x W_in^T), w_rec [H][H], w_rec_t = its transpose,
 * spikes [T][B][H]; state [4][B][H] = (I hi, I lo, V hi, V lo) after the last step;
 * checkpoints [ceil(T/K)][4][B][H], the state before steps 0, K, 2K, ... */
typedef struct n2h_rlif {
    int32_t batch, n_hidden, n_steps, checkpoint;  /* checkpoint: K steps between saved states */
    int32_t mode;                                   /* 0: s = H(V - theta); 1: s = f(V - theta), smooth (tests) */
    float alpha, beta, theta, slope, gamma;         /* synaptic / membrane decay, threshold, surrogate slope, scale */
    float alpha_lo, beta_lo;                        /* the decays' remainders: alpha = alpha + alpha_lo exactly
                                                       enough for the compensated state (0 for fp32 decays) */
} n2h_rlif;
int n2h_rlif_last_error(char* buf, size_t len);
int n2h_rlif_forward(const n2h_rlif* layer, const float* u, const float* w_rec_t, float* state, float* spikes,
                     float* checkpoints, void* stream);
int n2h_rlif_readout(int32_t n_steps, int64_t bo, float alpha, float beta, const float* z, float* r, void* stream);
int n2h_rlif_readout_adjoint(int32_t n_steps, int64_t bo, float alpha, float beta, const float* g_r, float* g_j,
                             void* stream);
/* The adjoint of the recurrent delivery alone: m [B][H] = g_i [B][H] w_rec [H][H], read through
 * w_rec_t = w_rec^T (hand-written, on the matrix cores; the product every reverse step of n2h_rlif_backward
 * computes).  n_hidden must be a multiple of 16. */
int n2h_rlif_delivery_adjoint(const n2h_rlif* layer, const float* g_i, const float* w_rec_t, float* m, void* stream);
/* Floats of n2h_rlif_backward's work area: (5 + min(K, T)) * B * H (-1: bad layer). */
int64_t n2h_rlif_work_floats(const n2h_rlif* layer);
/* gso = W_out^T gJ per step [T][B][H]; writes gI_t for every step to g_i_all [T][B][H];
 * work: n2h_rlif_work_floats(layer) floats of device memory.  n_hidden must be a multiple of 16. */
int n2h_rlif_backward(const n2h_rlif* layer, const float* u, const float* w_rec_t, const float* spikes,
                      const float* checkpoints, const float* gso, float* g_i_all, float* work, void* stream);
/* n2h_rlif_backward without a work area and without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  The work area is a static device array of N2H_RLIF_SCRATCH_FLOATS floats (16 MiB),
 * one per GPU, held by each GPU the library runs on until the process ends; layers whose
 * n2h_rlif_work_floats exceeds it are refused (N2H_EINVAL).  The call allocates nothing and keeps no state
 * between calls beyond the array's contents (every value it reads there it first writes).  All calls on one
 * GPU share that array: calls that overlap on one GPU (two streams, two host threads) race. */
#define N2H_RLIF_SCRATCH_FLOATS (1 << 22)
int n2h_rlif_backward_noalloc(const n2h_rlif* layer, const float* u, const float* w_rec_t, const float* spikes,
                              const float* checkpoints, const float* gso, float* g_i_all, void* stream);

/* Row intake (the target-major pull): each target's in-links in contract order (ascending canonical link
 * index).  Per group a block of 16-byte cells [row_depth[g]][64]: cell q of lane l (cell index
 * row_base[g] + 64 q + l, i.e. int32 words 4 * that index on) holds the lane's entries per*q .. per*q+per-1.
 *   N2H_ROWS_WIDE    (per = 2): {source slot << 7 | (delay_steps - 1), fp32 weight bits}
 *   N2H_ROWS_COMPACT (per = 4): source slot << 15 | (delay_steps - 1) << 8 | weight index into wtab[256]
 *                               (at most 256 distinct weights)
 *   N2H_ROWS_SLOT16  (per = 8): the 16-bit source slot alone, entry 2k in the low half of word k; every link
 *                               has tables->uniform_delay, and the weight is wtab[0] for source slots below
 *                               split, wtab[1] from it on, split = the bits of wtab[2] as uint32 (networks
 *                               whose weights are set by two source populations, Brunel's; 64 n_groups
 *                               <= 65535)
 * Entries past a target's in-links name the sentinel source slot 64 * n_groups (never spiking).
 * With tables->uniform_delay set, the walk is branch-free.
 * Steps that follow at least dense_min_spikes spikes walk every in-link against the previous step's spike
 * bitmap (held in LDS: n_groups <= 1535) instead of looking up the spiking sources' segments; both intakes
 * give bit-identical results.  rows = NULL: off.
 * Replaces: the reference's pull, fused.py:1626-1965 (here target-major, as receives.py names it). */
#define N2H_ROWS_WIDE 0
#define N2H_ROWS_COMPACT 1
#define N2H_ROWS_SLOT16 2
int n2h_bind_rows(n2h_plan* plan, const uint32_t* rows, const int64_t* row_base, const int32_t* row_depth,
                  int32_t format, const float* wtab, int32_t dense_min_spikes);

/* Surrogate-gradient adjoint of the LIF_EXP step for a per-population rate loss (no counterpart in the
 * reference, which is forward-only: a surrogate-gradient adjoint beside the plain path).
 * One n2h_adj_step call runs reverse step `step` for up to 16 networks in one launch.  Per network: its
 * out-links (canonical CSR by source) split by delay into a short list (delay < long_delay steps) and a long
 * list (delay >= long_delay), the long list further split by target range (range k: targets t with
 * t * long_parts / n == k, or with range_pop the caller's population-pure ranges), each run kept in canonical
 * order and packed {target | (delay_steps - 1) << 17 |
 * target population << 24, base weight bits}; alpha[target pop][source pop] the weight scales the forward
 * ran with (its weights are alpha * base); pop_coef[p] = {P_V, P_I, P_VI, V_reset, V_th, V_th - V_reset,
 * dL/ds of one spike of population p, 0}; this step's distances to the threshold (n2h_set_vtrace) and
 * spikes; the adjoint state gV, gI [n]; the ring gx [n][ring_slots + long_delay - 1] of the input adjoints of
 * the following steps (zero at the start; ring_slots > the largest delay; step s in slot s mod ring_slots,
 * the first long_delay - 1 slots mirrored past the end), and g_sum [n][n_pops][2], where every spiking
 * source j adds sum over its links into population a of w gx_{step+d} to the compensated pair (hi, lo) of
 * g_sum[j][a] (dL/dalpha[a][pop j] = sum over j in pop j of hi + lo).  The long links' sums come from long_bins [n][8], which n2h_adj_window filled for the
 * window holding this step (a neuron reads its row only where it is within the surrogate's support or
 * spiked).  The reset and the refractory gate carry no gradient. */
typedef struct n2h_adj_net {
    int32_t n_neurons;           /* <= 2^17 */
    int32_t n_pops;              /* <= 8 */
    int32_t ring_slots;
    int32_t long_delay;          /* 1 .. 8: the window length and the long-list delay bound */
    int32_t long_parts;          /* 1 .. 64: target ranges of the long list (a multiple of 8: one launch for all) */
    int32_t reserved;
    const int32_t* pop_first;    /* device [n_pops + 1] */
    const int64_t* out_ptr;      /* device [n + 1], short list */
    const uint32_t* out_links;   /* device [2 * n_short] */
    const int64_t* long_ptr;     /* device [n * long_parts + 1], long list by (source, target range) */
    const uint32_t* long_links;  /* device [2 * n_long] */
    const float* alpha;          /* device [n_pops][n_pops] */
    const float* pop_coef;       /* device [n_pops][8] */
    const float* vtrace;         /* device [n] (n2h_adj_window: [n_win][n] from the window's first step) */
    const uint8_t* spikes;       /* device [n] (n2h_adj_window: [n_win][n]) */
    float* gx_ring;              /* device [n][ring_slots + long_delay - 1] */
    float* gx_recent;            /* device [long_delay][n]: the short ring, step s in row s mod long_delay */
    float* long_bins;            /* device [n][8] of this step (n2h_adj_window: [n_win][n][8]) */
    float* g_v;                  /* device [n] */
    float* g_i;                  /* device [n] */
    float* g_sum;                /* device [n][n_pops][2]: a compensated sum, hi + lo */
    const int32_t* range_pop;    /* NULL, or device [long_parts] with long_parts a multiple of 8: the long
                                    list's ranges are then population-pure (range k holds targets of population
                                    range_pop[k] only; any split of each population's targets into ranges) */
} n2h_adj_net;
/* The surrogate: ds/dV = gamma (kappa/2) / dv / (1 + kappa |u|)^2 for |u| = |V - V_th| / dv <= cutoff, else 0
 * (dv = V_th - V_reset); neurons beyond the cutoff that did not spike skip their out-link gather. */
typedef struct n2h_surrogate {
    float kappa;
    float cutoff;
    float gamma;
    float reserved;
} n2h_surrogate;
int n2h_adj_last_error(char* buf, size_t len);
/* win_step = step - the first step of its window (n2h_adj_window); work as for n2h_adj_window, unchanged
 * since that call. */
int n2h_adj_step(const n2h_adj_net* nets, int32_t n_nets, int64_t step, int32_t win_step, const n2h_surrogate* sur,
                 int32_t* work, int64_t work_words, void* stream);
/* The long links of a window of reverse steps [step0, step0 + n_win), n_win <= long_delay, run before that
 * window's n2h_adj_step calls (after step0 + n_win's): for every source j and window step w where j is within
 * the surrogate's support or spiked, long_bins[w][j][a] = sum over j's long links into population a of
 * w_base gx_{step0+w+d}.  Every such input adjoint lies past the window (d >= long_delay >= n_win), so one
 * pass over a source's long links serves the whole window, and a link's W input adjoints are adjacent in
 * the ring.  The sums do not depend on where windows start.  work: device int32 scratch of
 * n2h_adj_work_words words, where the call lists the neurons that gather (the window's and each step's) for
 * the gather kernels. */
int n2h_adj_window(const n2h_adj_net* nets, int32_t n_nets, int64_t step0, int32_t n_win, const n2h_surrogate* sur,
                   int32_t* work, int64_t work_words, void* stream);
/* The int32 words of the work buffer of n2h_adj_window / n2h_adj_step for networks of these sizes (-1: bad
 * arguments).  With long_parts a multiple of 8 the buffer also holds the range sums of every possible window
 * entry (n * long_parts * 8 * long_delay rounded up to a power of 2, in floats): n2h_adj_window then walks all
 * ranges in one launch, the XCDs taking ranges in turn, and adds their sums in range order. */
int64_t n2h_adj_work_words(const int32_t* n_neurons, int32_t n_nets, int32_t long_delay, int32_t long_parts);

/* Potential trace for the adjoint (no counterpart in the reference, which is forward-only): with a buffer,
 * every step s in [step0, step0 + n_steps) writes each neuron's distance to the threshold as the threshold
 * sees it (V before reset minus V_th) to buffer[(s - step0) * n_neurons + i], or +inf on a refractory step.
 * v_lo, i_lo (LIF_EXP; both or neither, fp32 [n_neurons] each): while bound, every step also tracks what fp32
 * rounding dropped from V and I -- each operation of the update with its exact error (TwoSum / TwoProd), so
 * that V + v_lo and I + i_lo follow exact arithmetic on the same inputs -- and the trace is then the
 * compensated distance (V - V_th) + v_lo.  V, I, the spikes and every other result are unchanged (bit for bit).
 * buffer may be NULL with v_lo set (tracking without a trace).  Captured graphs and batches keep the setting
 * they were made with. */
int n2h_set_vtrace(n2h_plan* plan, float* buffer, int64_t step0, int32_t n_steps, float* v_lo, float* i_lo);

/* The delivery log (naer2/runtime.py:747-752, PopulationRuntime.deliveries: (slot, e, connection, source,
 * target) in log order).  While bound, every link a step delivers appends a record
 * records[i][0..5) = {arrival step, emission step, connection, source neuron, target neuron} at
 * i = (*count)++; past `capacity` the count still grows and nothing is written.  A step delivers the spikes of
 * the step before it, every delay at once (the pull intake), so the records of one step are its emission
 * step's links; each target's records appear in contract order (ascending canonical link index), targets
 * interleaved.  link_conn: device int16 [n_links], each link's connection (creation order), parallel to
 * tables->links; NULL: -1.  Logging steps run a variant without the row intake (results are identical); a
 * logging plan (at most 2^32 - 1 links) cannot join a batch.  records NULL: off.  Graphs captured before the call keep the setting
 * they were captured with. */
int n2h_bind_log(n2h_plan* plan, const int16_t* link_conn, int32_t* records, int32_t* count, int64_t capacity);

/* Edge kinds and ports per connection (the reference's open decisions D-6, per-connection receive code, and D-7,
 * per-link attribute blocks, with edge forward: TODO.md:179-204, test/test_fused_generated_gpu.py:361-398).
 * Every link's packed word x carries its edge kind in bits 8-9 and its port in bit 10
 * (x = delay << 16 | port << 10 | kind << 8 | local target), and edge_attr [n_links] (parallel to
 * tables->links; NULL when every link is static) its second attribute:
 *   kind 0  static: delivers its weight w
 *   kind 1  gain: delivers w * edge_attr
 *   kind 2  plastic: w += edge_attr first, stored back into the link table (the update persists), then
 *           delivers the new w (edge forward before receive)
 *   port 0  current: adds to I with the other current inputs (the summation contract per port)
 *   port 1  potential: ring_v [d_slots][n_neurons], read and cleared like the ring; at an update that is not
 *           refractory, V = (V_rest + P_V (V - V_rest)) + P_VI I, then V = V + the potential input
 * v_slots: at least the largest delay (steps) of a potential-port link, at most d_slots; the step keeps that
 * many extra LDS partial-sum cells per group (4 * 64 * v_slots bytes), and a potential-port link of a longer
 * delay is not delivered.  ring_v and v_slots > 0 go together (0 and NULL: no potential-port links).
 * LIF_EXP plans of at most 2^32 - 1 links only, every delay cell kept (tables->uniform_delay 0), every
 * population updating every step; such a plan steps without the row intake (same results), takes no
 * compensated trace and cannot join a batch.  edge_attr and ring_v both NULL: off (kinds and ports in x must
 * then be 0). */
int n2h_bind_edges(n2h_plan* plan, const float* edge_attr, float* ring_v, int32_t v_slots);

/* Status rows (the per-group status of naer2/transpiler/templates/fused.py:865-883, which the reference
 * writes for its runs: blocks held, most logged, lost bits, done): with rows, device int32 [n_groups][4], every
 * step writes row g = {the step, spikes group g emitted, its intake path (0: no spikes came in, 1: sorted
 * spike list, 2: spike bitmap scan, 3: row intake), spiking sources it took in}.  Nothing here can be lost or
 * held, so the rows carry the step's account instead.  NULL: off.  Graphs and batches keep the setting they
 * were made with. */
int n2h_set_status(n2h_plan* plan, int32_t* rows);

/* In-kernel phase trace (replaces naer2/trace.py's in-kernel stamps, compiled in by TRACE): with a buffer,
 * lane 0 of every group stamps the 100 MHz wall clock at 8 points of each step it runs -- entry, spike list
 * read, sources ranked, segment bounds, links staged, walked, updated, exit -- into
 * buffer[group][8] (the last step's stamps; points a step does not reach keep their old value).
 * NULL turns it off.  Graphs captured before the call keep the setting they were captured with. */
int n2h_set_trace(n2h_plan* plan, uint64_t* buffer);

/* Launch mapping of the plan's own step kernels (replaces the reference's per-device launch choice,
 * naer2/runtime.py:88-95 _warp_cap, and its autotuning backlog TODO.md:107-109): tuning->xcd and ->run as in
 * n2h_tuning, tiles being the plan's target groups.  With k > 1 the blocks the hardware deals to one XCD run
 * consecutive groups, which read neighbouring entries of a spiking source's offset row and neighbouring
 * link segments through one L2.  The default is k = 8 (MI355X's XCDs), run 0.  Results do not change.
 * Graphs captured before the call keep the mapping they were captured with; batches take theirs from
 * n2h_batch_create. */
int n2h_set_tuning(n2h_plan* plan, const n2h_tuning* tuning);

/* info[0] = n_groups, info[1] = d_slots, info[2] = dynamic LDS bytes per workgroup, info[3] = model. */
int n2h_status(n2h_plan* plan, int32_t* info, int32_t n_info);

/* ------------------------------------------------------------------------------------------------------------
 * The event-driven, exact-time network (Brunel's model A as the reference runs it, test/networks/Brunel):
 * delta-synapse LIF neurons whose potential changes only at an input's exact arrival time e + D_c and which
 * fire inside delivery at that time; Poisson sources with exact spike times; one delay D_c per connection,
 * each a whole multiple lag_c of the window (the smallest delay).  Window k emits spikes; window k + lag_c
 * delivers them through connection c.  Per target, a window's inputs run in ascending (arrival, connection,
 * source, link); inputs at one arrival
 * time are received one by one, then the threshold is checked once (the reference's send after every
 * delivery).  Receive: if a >= t_ref_end: V = V * exp((last - a) / tau) + w, last = a.  Threshold at t:
 * V >= theta: spike at t, V = V_r, t_ref_end = last = t + t_ref.  exp and log are this library's own fp32
 * functions (identical in the CPU oracle).  Poisson source i, window k: next < 0: next = interval(0); while
 * next < t_end(k): spike at next, next += interval(m); interval(m) = -log(u) / rate, u from word 0 of
 * Philox4x32-10(ctr = {i, k, m, instance}, key = seed), t_end(k) = (float)((double)(k + 1) * window).
 *
 * What each entry point replaces in the reference:
 *   n2h_ev_create / n2h_ev_destroy   naer2/runtime.py:213-323 (PopulationRuntime of queue populations) with
 *                                    naer2/ring.py:132-245 (the queues) and naer2/history.py:76-130 (region
 *                                    histories, EmissionHistory(regions=...))
 *   n2h_ev_run                       naer2/network.py:80-152 (Network.step over D): per window the fused
 *                                    kernel's queue pull (fused.py:2084-2301, _pull_queues), the merged
 *                                    delivery (fused.py:2428-2472 with _merge_segments 1300-1370 and
 *                                    _walk_batch 1421-1526), spikes inside delivery (_emit_delivered
 *                                    2667-2702) and the sorted regions (_sort_region 2715-2758)
 *   n2h_ev_graph_*                   naer2/graph.py:35-136 (captured windows)
 *   n2h_ev_bind_record               naer2/history.py:119-124 (EmissionHistory.spikes: a run's (e, node))
 *   n2h_ev_status                    naer2/runtime.py:724-735 (check: flags raised by the kernel)
 * ------------------------------------------------------------------------------------------------------------ */
#define N2H_EV_NEURON 0      /* delta-synapse LIF with exact times (test/networks/Brunel/neuron.py:6-45) */
#define N2H_EV_POISSON 1     /* Poisson source with exact times (test/networks/Brunel/poisson.py:7-27) */
#define N2H_EV_TRAIN 2       /* listed spike times per source (the reference tests' commanded sources) */
#define N2H_EV_MAX_SOURCES 8 /* source populations of one input stream */
#define N2H_EV_MAX_LAG 64    /* a delay of at most 64 windows */
#define N2H_EV_FLAGS 4
#define N2H_EV_FLAG_REGION 0   /* a group emitted more spikes in a window than its region holds (lost) */
#define N2H_EV_FLAG_SEGMENT 1  /* a source had more than 2048 links into one group (the excess lost) */
#define N2H_EV_FLAG_RECORD 2   /* the record buffer was full (a window not recorded) */
#define N2H_EV_FLAG_SORT 3     /* internal: more spikes than the sort holds */

typedef struct n2h_ev_pop {
    int32_t kind;        /* N2H_EV_* */
    int32_t region_cap;  /* spikes one group of the population can emit per window (<= 2^20) */
    float tau_inv;       /* neuron: 1 / tau, 1/ms */
    float theta;         /* neuron: threshold, fires at V >= theta */
    float v_reset;       /* neuron: V_r */
    float t_ref;         /* neuron: refractory period, ms */
    float rate_inv;      /* Poisson: 1 / rate, ms (inf: silent) */
    int32_t stream;      /* neuron: index of its input stream, -1: no inputs */
} n2h_ev_pop;

/* An input stream: the spikes of its source populations, listed in the order of the target populations'
 * incoming connections (connection index order), merged by arrival.  Target populations whose incoming
 * connections come from the same populations with the same delays in the same order share one stream.
 * Source j's connection has delay[j] = lag[j] windows: window k delivers its spikes of window k - lag[j]. */
typedef struct n2h_ev_stream {
    int32_t n_sources;
    int32_t source_pop[N2H_EV_MAX_SOURCES];
    float delay[N2H_EV_MAX_SOURCES];    /* ms (fp32), a whole multiple of the window */
    int32_t lag[N2H_EV_MAX_SOURCES];    /* delay / window, 1 .. N2H_EV_MAX_LAG */
} n2h_ev_stream;

/* Neurons 0..n-1, population p owning [pop_first[p], pop_first[p+1]); groups of <= 64 neurons of one
 * population, a population's groups contiguous and 64 neurons apart.  seg_off / links: the group segments of
 * n2h_tables (x & 0xffff = target lane); at most one connection per (source, target) population pair, so a
 * segment holds the links of one connection, in its link order.  train_ptr [n + 1] / train_t: sorted spike
 * times of N2H_EV_TRAIN sources (else NULL). */
typedef struct n2h_ev_desc {
    int32_t n_neurons;
    int32_t n_pops;
    int32_t n_groups;
    int32_t n_streams;
    const int32_t* pop_first;        /* host [n_pops + 1] */
    const n2h_ev_pop* pops;          /* host [n_pops] */
    const n2h_ev_stream* streams;    /* host [n_streams] */
    float window;                    /* ms: the smallest delay (every delay a whole multiple of it) */
    uint32_t instance;               /* Philox counter word 3 */
    uint64_t seed;                   /* Philox key */
    const int32_t* group_first;      /* device [n_groups] */
    const int32_t* group_info;       /* device [n_groups]: count | pop << 16 */
    const int64_t* seg_off;          /* device [64 * n_groups * n_groups + 1] */
    const uint32_t* links;           /* device [2 * n_links] */
    int64_t n_links;
    const int64_t* train_ptr;        /* device [n + 1] or NULL */
    const float* train_t;            /* device or NULL */
} n2h_ev_desc;

typedef struct n2h_ev_plan n2h_ev_plan;
typedef struct n2h_ev_graph n2h_ev_graph;

int n2h_ev_last_error(char* buf, size_t len);
/* Validates the description and allocates the window buffers (regions, input streams, sort scratch) on the
 * current device; the window counter starts at 0 with empty input streams. */
int n2h_ev_create(const n2h_ev_desc* desc, n2h_ev_plan** out);
int n2h_ev_destroy(n2h_ev_plan* plan);
/* State, fp32 [n] device tensors: V, the time of V's last exact change, the refractory end, and (Poisson
 * sources) the next spike time (-1 before the first draw). */
int n2h_ev_bind_state(n2h_ev_plan* plan, float* v, float* last_time, float* t_ref_end, float* next_spike);
/* Every window's spikes in (population, time, neuron) order appended at node/time[*count...] (device), and
 * the window's spike count at window_counts[k % n_window_slots]; a window that does not fit is not recorded
 * (flag).  node NULL: off. */
int n2h_ev_bind_record(n2h_ev_plan* plan, int32_t* node, float* time, int32_t* count, int64_t capacity,
                       int32_t* window_counts, int32_t n_window_slots);
/* The next window to run; clear_inputs: empty the input streams (a fresh start). */
int n2h_ev_set_window(n2h_ev_plan* plan, int64_t window, int32_t clear_inputs, void* stream);
int n2h_ev_get_window(n2h_ev_plan* plan, int64_t* window);   /* synchronizes */
/* Enqueue n_windows windows (two kernels each) on stream; never synchronizes. */
int n2h_ev_run(n2h_ev_plan* plan, int32_t n_windows, void* stream);
int n2h_ev_graph_create(n2h_ev_plan* plan, int32_t n_windows, void* stream, n2h_ev_graph** out);
int n2h_ev_graph_launch(n2h_ev_graph* graph, void* stream);
int n2h_ev_graph_destroy(n2h_ev_graph* graph);
/* flags[0 .. n) = the sticky N2H_EV_FLAG_* words (synchronizes). */
int n2h_ev_status(n2h_ev_plan* plan, int32_t* flags, int32_t n);
/* The delivery log (naer2/runtime.py:747-752, PopulationRuntime.deliveries: (slot, e, connection, source,
 * target) in log order): while bound, every delivered input appends records[i] = {window, connection index,
 * source (index in its population), target (index in its population)} and arrival[i] = its arrival time at
 * i = (*count)++ -- per target in delivery order, targets interleaved; past `capacity` the count still grows
 * and nothing is written.  conn_index: host [n_pops][n_pops], the connection from population s into t at
 * [t * n_pops + s].  A separate kernel variant runs while logging; records NULL: off.  Graphs captured before
 * the call keep the setting they were captured with. */
int n2h_ev_bind_log(n2h_ev_plan* plan, const int32_t* conn_index, int32_t* records, float* arrival, int32_t* count,
                    int64_t capacity);
/* Launch mapping of the deliver kernel: tuning->xcd = 0: block b runs group b; k > 1: the blocks dealt to each
 * of k XCDs run runs of 16 consecutive groups (neighbouring groups share their segments' cache lines in one
 * L2).  Results do not change.  Graphs captured before the call keep the mapping they were captured with. */
int n2h_ev_set_tuning(n2h_ev_plan* plan, const n2h_tuning* tuning);
/* In-kernel phase trace of the windows that follow (the reference's naer2/trace.py stamps): with a buffer
 * uint64 [n_groups + 1][8], every neuron group g writes row g for the last window it ran -- its start, the
 * 100 MHz ticks it spent on the input-stream entries, link staging, inbox places, inbox fill and the walk
 * (summed over its rounds), its rounds, and its end -- and the sort kernel writes row n_groups: start, group
 * prefix done, record done, first stream sorted, streams written (three times), end.  NULL: off.  Graphs
 * captured before the call keep the setting they were captured with. */
int n2h_ev_set_trace(n2h_ev_plan* plan, uint64_t* buffer);
/* counts[0] = inputs delivered (links walked), counts[1] = spikes emitted, since the plan was made
 * (synchronizes). */
int n2h_ev_counters(n2h_ev_plan* plan, int64_t* counts, int32_t n);
/* n2h_ev_run with a HIP event around each kernel: the summed durations of the deliver and the sort kernels
 * (synchronizes; not capturable).  The reference's in-kernel phase trace (naer2/trace.py) is the nearest
 * counterpart. */
int n2h_ev_run_timed(n2h_ev_plan* plan, int32_t n_windows, void* stream, double* deliver_ms, double* sort_ms);
/* The library's fp32 exp and log over n inputs (no counterpart in the reference: a check of the functions
 * that the event path's results depend on). */
int n2h_ev_math(const float* x, float* exp_x, const float* u, float* log_u, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NAER2_HIP_H */
