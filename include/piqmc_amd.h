/*
 * piqmc_amd.h -- C ABI of libpiqmc_amd.so, the MI355X (gfx950) implementation
 * of piqmc's path-integral quantum-annealing sweep.
 *
 * The reference (hadsed/pathintegral-qmc) exposes its sweep only as Cython
 * cpdef functions inside CPython extension modules; there is no C symbol to
 * bind.  The entry points below are what a Cython/ctypes shim for those
 * functions binds instead.  Each cites the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns a status (PIQMC_OK or one of the codes below);
 *    piqmc_last_error() returns a message for the calling thread.
 *  - "host" pointers are ordinary CPU memory, "dev" pointers are HIP device
 *    memory of the current device; `stream` is a hipStream_t (NULL = default).
 *  - Spins are +1/-1.  Device spin arrays are int8.
 *  - Neighbour tables are tools.GenerateNeighbors arrays: float64
 *    [nspins][maxnb][2] holding (neighbour index, coupling); an entry whose
 *    index equals its row is a local field (piqmc/qmc.pyx:107-109).
 */
#ifndef PIQMC_AMD_H
#define PIQMC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIQMC_OK 0
#define PIQMC_EINVAL 1    /* bad shape / value      -> ValueError        */
#define PIQMC_EINDEX 2    /* neighbour row overflow -> IndexError        */
#define PIQMC_EZERODIV 3  /* slices*temp == 0 or temp == 0 -> ZeroDivisionError */
#define PIQMC_EHIP 4      /* HIP runtime / device error -> RuntimeError  */
#define PIQMC_ENOMEM 5    /* allocation failure     -> MemoryError       */

const char *piqmc_last_error(void);
int piqmc_version(void);
/* Device time of the main sweep kernel of the last replay / parallel call made
 * from this host thread (HIP events on the call's stream; waits for it).
 * No reference counterpart: the reference has no device work to time. */
int piqmc_last_kernel_ms(float *ms);

/* ------------------------------------------------------------------ */
/* Problem handle: a neighbour table resident in HBM plus its graph   */
/* coloring (for the parallel sweep).                                 */
/* Replaces the `nbs` argument conversion of QuantumAnneal            */
/* (piqmc/qmc.pyx:36, memoryview unpack piqmc/qmc.c:2655).            */
/* ------------------------------------------------------------------ */
typedef struct piqmc_problem piqmc_problem;

int piqmc_problem_create(int nspins, int maxnb, const double *nbs_host,
                         piqmc_problem **out);
/* A problem from a dense coupling matrix J (element (a,b) at J[a*s0 + b*s1],
 * float64; diagonal = fields; only the upper triangle is read), the argument
 * of qmc.QuantumAnneal_dense / sa.Anneal_dense (piqmc/qmc.pyx:148,
 * piqmc/sa.pyx:129).  Replays on such a problem follow the *_dense functions:
 * couplings in double precision, and sa replay accepts ediff > 0
 * (piqmc/sa.pyx:181) where sa.Anneal accepts ediff >= 0. */
int piqmc_problem_create_dense(int nspins, const double *J_host, int64_t s0, int64_t s1,
                               piqmc_problem **out);
int piqmc_problem_destroy(piqmc_problem *pb);
/* ncolors: spatial colors found; pm_j: 1 if every coupling is +-1 between
 * distinct spins (bit-sliced path), 2 if in addition spins carry fields with
 * 2^-27 <= |h| <= 2^20 (bit-sliced path with per-spin thresholds), else 0;
 * zmax: max couplings per spin. */
int piqmc_problem_info(const piqmc_problem *pb, int *nspins, int *maxnb,
                       int *ncolors, int *pm_j, int *zmax);
/* 1 if every row's local field sums exactly in double whatever the order
 * (its couplings, as floats, are multiples of 2^L and the sum of their
 * magnitudes is below 2^E with E - L <= 52), else 0.  The lane-per-slice
 * sweep then splits each field over several lanes.  No counterpart in the
 * reference: it reports the problem analysis the parallel sweep acts on. */
int piqmc_problem_exact_rows(const piqmc_problem *pb, int *exact);
/* Replace the spatial coloring (color_host[nspins], values 0..ncolors-1).
 * Rejected unless no two coupled spins share a color. */
int piqmc_problem_set_coloring(piqmc_problem *pb, int ncolors, const int32_t *color_host);
/* The coloring in use (color_host[nspins]). */
int piqmc_problem_get_coloring(const piqmc_problem *pb, int32_t *color_host);
/* Parallel-sweep kernel: 0 automatic (dense when color classes average fewer
 * than 8 spins), 1 one parity word per thread, 2 one replica per wavefront
 * with a lane per slice.  Every choice gives identical results. */
int piqmc_problem_set_sweep_path(piqmc_problem *pb, int path);
/* Layout of the lane-per-slice (dense) sweep, the tile the autotuner picks
 * for dense problems: 0 automatic; 1 one replica per wavefront, a lane per
 * slice; 2 each slice's field split over 64/(P/2) lanes; 3 one replica per
 * workgroup of four wavefronts (the field split over them); 4 / 5 that
 * workgroup pipelined, one wavefront deciding a spin while three / seven sum
 * the next one's field; 6 one replica per wavefront with every spin's local
 * field in every slice of the swept parity kept in LDS and updated on flips
 * (rows that list each neighbour once with the same coupling both ways).
 * 2-6 need rows that sum exactly in any order (piqmc_problem_exact_rows); a
 * sweep whose variant does not apply fails with PIQMC_EINVAL.  Every variant
 * gives identical results.  No reference counterpart. */
int piqmc_problem_set_dense_variant(piqmc_problem *pb, int variant);

/* ------------------------------------------------------------------ */
/* Random streams of the reference, for replay.                       */
/* ------------------------------------------------------------------ */
/* glibc srand(seed) image of the TYPE_3 state used by libc rand()
 * (piqmc/qmc.pyx:22,132): state[0..30] words, state[31] = front index. */
int piqmc_rand_seed(uint32_t seed, uint32_t *state32_host);
/* Advance a TYPE_3 state by n draws on the host. */
int piqmc_rand_advance(uint32_t *state32_host, uint64_t n);
/* numpy RandomState(seed) permutation streams as QuantumAnneal draws them
 * (piqmc/qmc.pyx:90-91,136): stream r, permutation g, position t lands at
 * perms[(g*N + t)*R + r].  Stream r starts from RandomState(seeds[r]). */
int piqmc_perm_streams(const uint32_t *seeds_host, int R, int N, int G,
                       uint16_t *perms_host, int nthreads);

/* ------------------------------------------------------------------ */
/* Replay sweeps (bit-exact with the reference).                      */
/* ------------------------------------------------------------------ */
/* qmc.QuantumAnneal (piqmc/qmc.pyx:30-136) for R independent replicas.
 *  sched_host [S] transverse fields; M sweeps per field; P slices; temp.
 *  confs_dev  int8 [R][P][N] in/out.
 *  perms_dev  uint16 [S*M][N][R]: the permutation used by sweep g.
 *  rand_dev   uint32 [R][32] TYPE_3 states in/out (see piqmc_rand_seed).
 *  draws_dev  uint64 [R] out: rand() calls made.
 *  ties_dev   uint32 [R] out: acceptance tests where the device exp() and a
 *             correctly rounded exp could disagree (|exp - u| <= 2^-50 exp);
 *             zero means the run is bit-exact with the libm reference. */
int piqmc_qa_replay(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                    const double *sched_host, int8_t *confs_dev, const uint16_t *perms_dev,
                    uint32_t *rand_dev, uint64_t *draws_dev, uint32_t *ties_dev,
                    void *stream);
/* sa.Anneal (piqmc/sa.pyx:50-120) for R replicas; svec_dev int8 [R][N]. */
int piqmc_sa_replay(const piqmc_problem *pb, int R, int S, int M, const double *sched_host,
                    int8_t *svec_dev, const uint16_t *perms_dev, uint32_t *rand_dev,
                    uint64_t *draws_dev, uint32_t *ties_dev, void *stream);

/* Host-memory drop-ins: same arguments and in-place semantics as the Cython
 * functions, with the permutation stream pre-drawn by the caller from its
 * RandomState and the libc rand() state passed explicitly.
 *  qmc.QuantumAnneal: confs element (i,k) at confs[i*cs0 + k*cs1] (float64,
 *  +-1); perms int64 [S*M][N]; rand_state32 in/out; *draws out. */
int piqmc_quantum_anneal(const double *sched, int S, int mcsteps, int slices, float temp,
                         int nspins, double *confs, int64_t cs0, int64_t cs1,
                         const double *nbs, int maxnb, const int64_t *perms,
                         uint32_t *rand_state32, uint64_t *draws, uint32_t *ties);
/* sa.Anneal: svec element i at svec[i*ss]. */
int piqmc_anneal(const double *sched, int S, int mcsteps, double *svec, int64_t ss, int nspins,
                 const double *nbs, int maxnb, const int64_t *perms, uint32_t *rand_state32,
                 uint64_t *draws, uint32_t *ties);
/* qmc.QuantumAnneal_dense (piqmc/qmc.pyx:141-242): as piqmc_quantum_anneal
 * with the dense matrix J (element (a,b) at J[a*js0 + b*js1]) in place of
 * the neighbour table. */
int piqmc_quantum_anneal_dense(const double *sched, int S, int mcsteps, int slices, float temp,
                               int nspins, double *confs, int64_t cs0, int64_t cs1,
                               const double *J, int64_t js0, int64_t js1, const int64_t *perms,
                               uint32_t *rand_state32, uint64_t *draws, uint32_t *ties);
/* sa.Anneal_dense (piqmc/sa.pyx:126-187). */
int piqmc_anneal_dense(const double *sched, int S, int mcsteps, double *svec, int64_t ss,
                       int nspins, const double *J, int64_t js0, int64_t js1,
                       const int64_t *perms, uint32_t *rand_state32, uint64_t *draws,
                       uint32_t *ties);

/* sa.Anneal_multispin (piqmc/sa.pyx:282-405) for R packed sets of 64
 * configurations: bits_dev uint64 [R][N] in/out (configuration k at bit 63-k,
 * bit 1 = spin -1), G sweeps at temperatures temps_host[G] (float, one per
 * sweep), perms_dev uint16 [G][R][N], U_dev float64 [G][N][R][64]: the
 * uniform configuration k of set r compares against at attempt t of sweep g
 * (the reference's rng.rand(64) stream, piqmc/sa.pyx:331,398).  ties_dev
 * uint32 [R] is incremented by the near ties of the exp comparison. */
int piqmc_sa_multispin(const piqmc_problem *pb, int R, int G, const float *temps_host,
                       uint64_t *bits_dev, const uint16_t *perms_dev, const double *U_dev,
                       uint32_t *ties_dev, void *stream);

/* ------------------------------------------------------------------ */
/* Parallel sweep ("fixed" semantics: periodic Trotter neighbours,     */
/* per-spin reset) -- graph-colored in space, parity-alternated in     */
/* imaginary time, Philox4x32-10 draws.  The parallel counterpart of   */
/* qmc.QuantumAnneal_parallel (piqmc/qmc.pyx:247-357).                 */
/*  confs_dev int8 [R][P][N] in/out (P even).                          */
/*  etrace_dev float64 [R][S] (or NULL): mean over slices of the       */
/*  classical energy after each field step.                            */
/*  block: threads per workgroup (0 = default), rpw: replicas per      */
/*  workgroup (0 = default); every choice gives identical results.     */
/* ------------------------------------------------------------------ */
int piqmc_qa_colored(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                     const double *sched_host, int8_t *confs_dev, uint64_t seed,
                     uint32_t replica_base, uint32_t sweep_base, double *etrace_dev,
                     int block, int rpw, void *stream);
/* piqmc_qa_colored that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as piqmc_qa_colored; the per-call buffers
 * (J_perp and temperature tables, the +-1 threshold tables, the packed state
 * [R][2][N] words) come from a fixed 64 MiB device region per GPU, held until
 * the process ends.  A call needing more (e.g. more than ~8000 replicas of
 * 1024 spins at P <= 64) fails with PIQMC_EINVAL.  All calls on one GPU share
 * that region: calls that overlap on one GPU (other streams or threads) race. */
int piqmc_qa_colored_noalloc(const piqmc_problem *pb, int R, int P, int S, int M, float temp,
                             const double *sched_host, int8_t *confs_dev, uint64_t seed,
                             uint32_t replica_base, uint32_t sweep_base, double *etrace_dev,
                             int block, int rpw, void *stream);
/* Classical annealing with the parallel sweep (the preanneal of
 * examples/spinglass32.py:121, sa.Anneal, on the GPU): C independent chains
 * spins_dev int8 [C][N] (C even) in place, temperature sched_host[s] in field
 * step s, M sweeps per step.  Fixed semantics and Philox draws as
 * piqmc_qa_colored: chain c is slice c % P' of replica replica_base + c / P'
 * (P' the largest even divisor of C up to 64) with zero Trotter coupling. */
int piqmc_sa_colored(const piqmc_problem *pb, int C, int S, int M, const double *sched_host,
                     int8_t *spins_dev, uint64_t seed, uint32_t replica_base, uint32_t sweep_base,
                     int block, int rpw, void *stream);

/* ------------------------------------------------------------------ */
/* Batched heterogeneous parallel sweeps (no reference counterpart: the */
/* reference runs one replica per call).  Jobs of any mix of problems,  */
/* slice counts, schedule lengths and replica counts run in ONE         */
/* persistent launch that takes work items from a device counter;       */
/* each job's result equals its own piqmc_qa_colored call.              */
/* ------------------------------------------------------------------ */
typedef struct piqmc_batch piqmc_batch;
int piqmc_batch_create(piqmc_batch **out);
int piqmc_batch_destroy(piqmc_batch *b);
/* Same arguments and meaning as piqmc_qa_colored (sched is copied). */
int piqmc_batch_add(piqmc_batch *b, const piqmc_problem *pb, int R, int P, int S, int M,
                    float temp, const double *sched_host, int8_t *confs_dev, uint64_t seed,
                    uint32_t replica_base, uint32_t sweep_base, double *etrace_dev);
/* One launch over every job; block: threads per workgroup (0 = 256), rpw:
 * replicas per work item for word-per-thread jobs (0 = automatic). */
int piqmc_batch_run(piqmc_batch *b, int block, int rpw, void *stream);
/* piqmc_batch_run in two parts.  prepare: tables, packed-state buffers and
 * the sorted work queue on the device (owned by the batch until the next
 * prepare or destroy).  launch: pack every job's spins, the persistent
 * work-queue kernel, unpack -- kernel launches only, so a stream capture of
 * it gives a HIP graph that re-runs the whole batch (each replay continues
 * from the spins the previous one left). */
int piqmc_batch_prepare(piqmc_batch *b, int block, int rpw, void *stream);
int piqmc_batch_launch(piqmc_batch *b, void *stream);
int piqmc_batch_info(const piqmc_batch *b, int *njobs, int *nitems, int *grid);

/* ------------------------------------------------------------------ */
/* Energies.                                                           */
/* ------------------------------------------------------------------ */
/* Classical energy of each of B int8 configurations [B][N] from the
 * neighbour table (pair entries weighted 1/2, field entries 1). */
int piqmc_slice_energies(const piqmc_problem *pb, int B, const int8_t *spins_dev,
                         double *E_dev, void *stream);
/* sa.ClassicalIsingEnergy (piqmc/sa.pyx:25-44): E = -s.(J_off s) - sum(diag(J) s)
 * for B float64 configurations [B][N] and a COO coupling list (host). */
int piqmc_classical_energy(int nspins, int64_t nnz, const int64_t *rows, const int64_t *cols,
                           const double *vals, int B, const double *spins, double *E);

/* ------------------------------------------------------------------ */
/* Learned slice proposal (no reference counterpart: the reference has */
/* no learned component).  A masked autoregressive network (MADE)      */
/* N -> H1 -> H2 -> N over one slice; BF16 MFMA kernels, FP32 accum.    */
/* Slices are int8 [B][N] device arrays.                               */
/* ------------------------------------------------------------------ */
typedef struct piqmc_made piqmc_made;
int piqmc_made_create(int nspins, int h1, int h2, uint64_t seed, piqmc_made **out);
int piqmc_made_destroy(piqmc_made *m);
/* fp32 parameters / last gradients to host: w0 [h1][N], b0 [h1], w1 [h2][h1],
 * b1 [h2], w2 [N][h2], b2 [N] (any pointer may be NULL). */
int piqmc_made_get_params(const piqmc_made *m, float *w0, float *b0, float *w1, float *b1,
                          float *w2, float *b2);
int piqmc_made_get_grads(const piqmc_made *m, float *w0, float *b0, float *w1, float *b1,
                         float *w2, float *b2);
/* log q of B slices -> logq_dev float64 [B]. */
int piqmc_made_logprob(piqmc_made *m, int B, const int8_t *s_dev, double *logq_dev, void *stream);
/* Gradient of the mean negative log-likelihood over B slices; an Adam step
 * (beta 0.9/0.999, eps 1e-8) when lr > 0; *mean_nll (host, may be NULL). */
int piqmc_made_train_step(piqmc_made *m, int B, const int8_t *s_dev, float lr, double *mean_nll,
                          void *stream);
/* One epoch of piqmc_made_train_step over n slices data_dev int8 [n][N] in
 * the order perm_dev int32 [n]: floor(n / batch) Adam steps of `batch` rows,
 * launched back to back on the stream with one host synchronisation at the
 * end, for *mean_nll (host, may be NULL), the epoch's mean NLL. */
int piqmc_made_train_epoch(piqmc_made *m, int n, const int8_t *data_dev, const int32_t *perm_dev, int batch,
                           float lr, double *mean_nll, void *stream);
/* B autoregressive draws -> int8 [B][N] and their log q (float64 [B]).
 * Draw d of row b uses philox(key=seed, ctr={round, b, d, 0x5A4D0000})[0].
 * The N-step launch chain runs as a captured HIP graph (replayed per call)
 * unless piqmc_made_set_graph(m, 0) selects eager launches. */
int piqmc_made_sample(piqmc_made *m, int B, uint64_t seed, uint32_t round, int8_t *s_out_dev,
                      double *logq_dev, void *stream);
int piqmc_made_set_graph(piqmc_made *m, int use_graph);
/* Independence Metropolis-Hastings slice move for R replicas of confs_dev
 * int8 [R][P][N]: replica r proposes new_dev[r] (log q in logq_new_dev[r])
 * for slice k = slice_dev[r]; accepted with min(1, exp(-dS/T) q(old)/q(new)),
 * dS the fixed-semantics action change (classical energy of the slice and
 * its Trotter bonds to k-1, k+1 at this field).  accepted_dev int32 [R]. */
int piqmc_made_mh(piqmc_made *m, const piqmc_problem *pb, int R, int P, int8_t *confs_dev,
                  const int8_t *new_dev, const double *logq_new_dev, const int32_t *slice_dev,
                  float jperp, float temp, uint64_t seed, uint32_t round, int32_t *accepted_dev,
                  void *stream);
/* The fused autoregressive sampler: one kernel pass per slice instead of N
 * forward passes (running pre-activations, each unit added once when it
 * completes).  It defines q_beta(s) = prod_d sigmoid(+-beta l_d(s_<d)), a
 * normalized density for every beta > 0 (beta = 1: the network; beta > 1:
 * sharpened towards its modes).  Rows [0, Bs): draws into s_out_dev int8
 * [Bs][N]; rows [Bs, Bs + Bf): the given slices forced_dev int8 [Bf][N];
 * logq_dev float64 [Bs + Bf] gets log q_beta of every row.  A draw and the
 * evaluation of the drawn slice give the same log q bit for bit.  Draw d of
 * row b uses philox(key=seed, ctr={round, b, d/4, 0x4D415200})[d % 4].
 * Supported shapes: 2 <= N <= 65 with H <= 4(N-1), 66 <= N <= 129 with
 * H <= 2(N-1), 194 <= N <= 257 with H <= 3(N-1), 1666 <= N <= 1729 with
 * H < N (else PIQMC_EINVAL). */
int piqmc_made_ar(piqmc_made *m, int Bs, int Bf, const int8_t *forced_dev, float beta, uint64_t seed,
                  uint32_t round, int8_t *s_out_dev, double *logq_dev, void *stream);
/* World-line Metropolis-Hastings move for R replicas of confs_dev int8
 * [R][P][N]: y ~ q_beta, f = y * s^0 (s^0 the replica's slice 0), and every
 * slice s^k becomes s^k * f, which keeps every Trotter bond; accepted with
 * min(1, exp(-sum_k [E(s^k f) - E(s^k)] / T) q_beta(s^0) / q_beta(y)).  The
 * stationary law of the sweep is unchanged.  accepted_dev int32 [R]. */
int piqmc_made_worldline_mh(piqmc_made *m, const piqmc_problem *pb, int R, int P, int8_t *confs_dev,
                            float beta, float temp, uint64_t seed, uint32_t round, int32_t *accepted_dev,
                            void *stream);
/* piqmc_made_mh with the proposal drawn by the fused sampler: replica r's
 * slice k = slice_dev[r] is replaced by y ~ q_beta, accepted with
 * min(1, exp(-dS/T) q_beta(old) / q_beta(y)) (dS as in piqmc_made_mh); both
 * densities come from the fused sampler's arithmetic. */
int piqmc_made_slice_mh(piqmc_made *m, const piqmc_problem *pb, int R, int P, int8_t *confs_dev,
                        const int32_t *slice_dev, float beta, float jperp, float temp, uint64_t seed,
                        uint32_t round, int32_t *accepted_dev, void *stream);
/* C[M][N] = A[M][K] . B[N][K]^T, bf16 (uint16 bit patterns) in, fp32 out. */
int piqmc_gemm_bf16_nt(const uint16_t *A, const uint16_t *B, int M, int N, int K, float *C,
                       void *stream);

/* ------------------------------------------------------------------ */
/* tools.GenerateNeighbors (piqmc/tools.pyx:28-96).                    */
/*  (rows, cols, vals) list the DOK keys in insertion order.           */
/*  order 0: CPython 2.7 dict iteration order (the reference's);       */
/*  order 1: insertion order; order 2: row-major sorted.               */
/*  nbs_out float64 [nspins][maxnb][2], zero-filled then written.      */
/*  PIQMC_EINDEX when a spin has more than maxnb entries.              */
/* ------------------------------------------------------------------ */
int piqmc_generate_neighbors(int nspins, int maxnb, int64_t nnz, const int64_t *rows,
                             const int64_t *cols, const double *vals, int order,
                             double *nbs_out);

#ifdef __cplusplus
}
#endif
#endif
