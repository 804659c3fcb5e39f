/*
 * piqmc_oracle.c -- CPU restatement of piqmc's annealing sweeps.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product is
 * compared against; it is built into oracle/lib/liborc.so and loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing in
 * the product (pathintegral-qmc_amd/) links or calls it.
 *
 * Every function cites the reference (hadsed/pathintegral-qmc) lines it
 * restates.  Arithmetic types follow the Cython-generated C
 * (piqmc/qmc.c:1752-1766): ediff/jval/jperp/temp are float, confs/nbs/sched
 * are double, every "+=" is evaluated in double and rounded to float on store,
 * exp() is libm double, rand() is glibc random() (TYPE_3).
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_rand_*        : libc srand()/rand() through ctypes, many seeds.
 *   orc_mt_*, shuffle : numpy.random.RandomState.permutation, chained.
 *   orc_philox4x32    : Random123 known-answer vectors (Salmon et al. SC'11).
 *   orc_qa_compat     : hand traces on N=1..4: defects D1/D2, a field row, and
 *                       non-dyadic couplings (0.1/0.2/-0.3) where the float32
 *                       running ediff takes a different decision than a double
 *                       sum or an uncast J would; Boixo outcome test
 *                       (testing/test_boixo.py:90-114).
 *   orc_jperp         : values fixed by the mathematics (tanh = e^-2 gives PT,
 *                       small- and large-Gamma limits) and a float-P*T case.
 *   orc_sa_compat     : hand traces, Boixo SA outcome test (test_boixo.py:72-88).
 *   orc_qa_dense,     : equal to orc_qa_compat / orc_sa_compat (and so to their
 *   orc_sa_dense        pins) on matrices whose entries are float-exact, given
 *                       the neighbour table of the same matrix; a hand trace
 *                       where double-precision couplings round differently.
 *   orc_qa_fixed_seq  : exact enumeration of the stationary law on tiny
 *                       systems (N=2,P=2) at fixed Gamma.
 *   orc_qa_colored    : the same stationarity check; brute-force equality
 *                       with orc_qa_fixed_seq's per-attempt arithmetic.
 *   orc_energy_*      : ClassicalIsingEnergy known answers
 *                       (testing/test_boixo.py:60-70), instance ground states.
 *   orc_sa_multispin  : each of the 64 packed configurations equals an
 *                       independent +-1 Metropolis chain fed its own uniforms
 *                       (tests/test_oracle_multispin.py); hand trace.
 *   orc_qa_parallel   : hand traces at nthreads=1, where it is deterministic
 *                       (ascending spins, per-spin double ediffs, D1, libc
 *                       rand() only on the exp branch); with more threads it
 *                       races like the reference and serves only as a timing twin.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../pathintegral-qmc_amd/csrc/piqmc_fexp64.h"

#define ORC_RAND_MAX 2147483647

/* ------------------------------------------------------------------ */
/* glibc random() TYPE_3: additive feedback r[i] = r[i-31] + r[i-3].   */
/* State layout: st[0..30] the 31 words, st[31] = index of "fptr";     */
/* "rptr" trails it by 3 (mod 31).  glibc stdlib/random_r.c semantics. */
/* ------------------------------------------------------------------ */
uint32_t orc_rand_next(uint32_t *st)
{
    uint32_t f = st[31];
    uint32_t r = (f + 28u) % 31u;
    uint32_t v = st[f] + st[r];
    st[f] = v;
    st[31] = (f + 1u) % 31u;
    return v >> 1;
}

void orc_rand_seed(uint32_t seed, uint32_t *st)
{
    int32_t *w = (int32_t *)st;
    if (seed == 0) seed = 1;
    w[0] = (int32_t)seed;
    int32_t word = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
        long hi = word / 127773;
        long lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        w[i] = word;
    }
    st[31] = 3;
    for (int i = 0; i < 310; ++i) (void)orc_rand_next(st);
}

/* ------------------------------------------------------------------ */
/* numpy legacy RandomState: MT19937 + Fisher-Yates shuffle with      */
/* masked-rejection random_interval.  State: st[0..623] words,        */
/* st[624] = position.  Used at piqmc/qmc.pyx:90-91,136.              */
/* ------------------------------------------------------------------ */
void orc_mt_seed(uint32_t seed, uint32_t *st)
{
    st[0] = seed;
    for (uint32_t i = 1; i < 624; ++i)
        st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + i;
    st[624] = 624;
}

static void orc_mt_twist(uint32_t *mt)
{
    const uint32_t UP = 0x80000000u, LO = 0x7fffffffu, A = 0x9908b0dfu;
    int i;
    uint32_t y;
    for (i = 0; i < 624 - 397; ++i) {
        y = (mt[i] & UP) | (mt[i + 1] & LO);
        mt[i] = mt[i + 397] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
    }
    for (; i < 623; ++i) {
        y = (mt[i] & UP) | (mt[i + 1] & LO);
        mt[i] = mt[i + 397 - 624] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
    }
    y = (mt[623] & UP) | (mt[0] & LO);
    mt[623] = mt[396] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
}

uint32_t orc_mt_next32(uint32_t *st)
{
    if (st[624] >= 624) {
        orc_mt_twist(st);
        st[624] = 0;
    }
    uint32_t y = st[st[624]++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static uint64_t orc_interval(uint32_t *st, uint64_t max)
{
    if (max == 0) return 0;
    uint64_t mask = max;
    mask |= mask >> 1; mask |= mask >> 2; mask |= mask >> 4;
    mask |= mask >> 8; mask |= mask >> 16; mask |= mask >> 32;
    uint64_t v;
    do { v = orc_mt_next32(st) & mask; } while (v > max);
    return v;
}

/* RandomState.shuffle of a 1-D int64 array (legacy _shuffle_raw). */
void orc_mt_shuffle_i64(uint32_t *st, int64_t *a, int64_t n)
{
    for (int64_t i = n - 1; i >= 1; --i) {
        int64_t j = (int64_t)orc_interval(st, (uint64_t)i);
        int64_t t = a[i]; a[i] = a[j]; a[j] = t;
    }
}

/* G permutations exactly as piqmc draws them: perms[0] =
 * rng.permutation(range(N)), perms[g] = rng.permutation(perms[g-1]).
 * st is advanced in place (caller may seed it or bring a used state). */
void orc_perm_stream(uint32_t *st, int N, int G, int64_t *perms)
{
    for (int i = 0; i < N; ++i) perms[i] = i;
    orc_mt_shuffle_i64(st, perms, N);
    for (int g = 1; g < G; ++g) {
        memcpy(perms + (size_t)g * N, perms + (size_t)(g - 1) * N, sizeof(int64_t) * N);
        orc_mt_shuffle_i64(st, perms + (size_t)g * N, N);
    }
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                  */
/* ------------------------------------------------------------------ */
void orc_philox4x32(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* J_perp for one schedule entry, as piqmc/qmc.pyx:95 / qmc.c:2055-2075:
 * ((-0.5*P)*T) * log(tanh(G / (float)(P*T))) evaluated in double, stored float. */
float orc_jperp(double gamma, int P, float temp)
{
    float pt = (float)P * temp;
    return (float)(((-0.5 * P) * temp) * log(tanh(gamma / pt)));
}

/* ------------------------------------------------------------------ */
/* qmc.QuantumAnneal, piqmc/qmc.pyx:30-136 (C body qmc.c:1752-2470).   */
/* confs element (i,k) at confs[i*cs0 + k*cs1]; nbs[N][Z][2] dense.    */
/* perms[g*N + t] is the t-th spin of sweep g (g = step*M + sweep).    */
/* Reproduces D1 (tidx never assigned -> slices P-1 and 1, :83,115-117) */
/* and D2 (ediff reset once per slice, :134-135).                      */
/* ------------------------------------------------------------------ */
int orc_qa_compat(const double *sched, int S, int M, int P, float temp, int N,
                  double *confs, long cs0, long cs1, const double *nbs, int Z,
                  const int64_t *perms, uint32_t *rs, uint64_t *draws)
{
    if (P < 2 || N < 1 || Z < 1) return -1;
    uint64_t nd = 0;
    for (int s = 0; s < S; ++s) {
        float jperp = orc_jperp(sched[s], P, temp);
        for (int m = 0; m < M; ++m) {
            const int64_t *perm = perms + ((size_t)s * M + m) * N;
            for (int k = 0; k < P; ++k) {
                float ediff = 0.0f;
                const int tleft = P - 1, tright = 1;   /* D1 */
                for (int t = 0; t < N; ++t) {
                    int sidx = (int)perm[t];
                    double ci = confs[sidx * cs0 + k * cs1];
                    for (int si = 0; si < Z; ++si) {
                        const double *row = nbs + ((size_t)sidx * Z + si) * 2;
                        int spinidx = (int)row[0];
                        float jval = (float)row[1];
                        if (spinidx == sidx)
                            ediff = (float)((double)ediff + (-2.0 * ci) * (double)jval);
                        else
                            ediff = (float)((double)ediff + (-2.0 * ci) *
                                            ((double)jval * confs[spinidx * cs0 + k * cs1]));
                    }
                    ediff = (float)((double)ediff + (-2.0 * ci) *
                                    ((double)jperp * confs[sidx * cs0 + tleft * cs1]));
                    ediff = (float)((double)ediff + (-2.0 * ci) *
                                    ((double)jperp * confs[sidx * cs0 + tright * cs1]));
                    if ((double)ediff > 0.0) {
                        confs[sidx * cs0 + k * cs1] *= -1.0;
                    } else {
                        uint32_t r = orc_rand_next(rs);
                        ++nd;
                        if (exp((double)(ediff / temp)) > (double)r / (double)ORC_RAND_MAX)
                            confs[sidx * cs0 + k * cs1] *= -1.0;
                    }
                }
                /* D2: no reset inside the spin loop */
            }
        }
    }
    if (draws) *draws = nd;
    return 0;
}

/* ------------------------------------------------------------------ */
/* sa.Anneal, piqmc/sa.pyx:50-120 (cdivision; per-spin reset :119;     */
/* accept ediff >= 0 :114; temp = (float)sched[t] :96).                */
/* ------------------------------------------------------------------ */
int orc_sa_compat(const double *sched, int S, int M, double *svec, long ss, int N,
                  const double *nbs, int Z, const int64_t *perms, uint32_t *rs,
                  uint64_t *draws)
{
    if (N < 1 || Z < 1) return -1;
    uint64_t nd = 0;
    for (int s = 0; s < S; ++s) {
        float temp = (float)sched[s];
        for (int m = 0; m < M; ++m) {
            const int64_t *perm = perms + ((size_t)s * M + m) * N;
            for (int t = 0; t < N; ++t) {
                int sidx = (int)perm[t];
                float ediff = 0.0f;
                double ci = svec[sidx * ss];
                for (int si = 0; si < Z; ++si) {
                    const double *row = nbs + ((size_t)sidx * Z + si) * 2;
                    int spinidx = (int)row[0];
                    float jval = (float)row[1];
                    if (spinidx == sidx)
                        ediff = (float)((double)ediff + (-2.0 * ci) * (double)jval);
                    else
                        ediff = (float)((double)ediff + (-2.0 * ci) *
                                        ((double)jval * svec[spinidx * ss]));
                }
                if ((double)ediff >= 0.0) {
                    svec[sidx * ss] *= -1.0;
                } else {
                    uint32_t r = orc_rand_next(rs);
                    ++nd;
                    if (exp((double)(ediff / temp)) > (double)r / (double)ORC_RAND_MAX)
                        svec[sidx * ss] *= -1.0;
                }
            }
        }
    }
    if (draws) *draws = nd;
    return 0;
}

/* ------------------------------------------------------------------ */
/* qmc.QuantumAnneal_dense, piqmc/qmc.pyx:197-242 (qmc.c:2957-3390):   */
/* every si in 0..N-1 in order, J read as double from the upper        */
/* triangle (J[sidx,si] if sidx<si else J[si,sidx], diagonal = field), */
/* each term added in double and stored to the float ediff; D1 and D2  */
/* as QuantumAnneal.  J element (a,b) at J[a*js0 + b*js1].             */
/* ------------------------------------------------------------------ */
int orc_qa_dense(const double *sched, int S, int M, int P, float temp, int N,
                 double *confs, long cs0, long cs1, const double *J, long js0, long js1,
                 const int64_t *perms, uint32_t *rs, uint64_t *draws)
{
    if (P < 2 || N < 1) return -1;
    uint64_t nd = 0;
    for (int s = 0; s < S; ++s) {
        float jperp = orc_jperp(sched[s], P, temp);
        for (int m = 0; m < M; ++m) {
            const int64_t *perm = perms + ((size_t)s * M + m) * N;
            for (int k = 0; k < P; ++k) {
                float ediff = 0.0f;
                const int tleft = P - 1, tright = 1;   /* D1: tidx stays 0 */
                for (int t = 0; t < N; ++t) {
                    int sidx = (int)perm[t];
                    double ci = confs[sidx * cs0 + k * cs1];
                    for (int si = 0; si < N; ++si) {
                        if (si == sidx)
                            ediff = (float)((double)ediff + (-2.0 * ci) * J[sidx * js0 + si * js1]);
                        else if (sidx < si)
                            ediff = (float)((double)ediff + (-2.0 * ci) *
                                            (J[sidx * js0 + si * js1] * confs[si * cs0 + k * cs1]));
                        else
                            ediff = (float)((double)ediff + (-2.0 * ci) *
                                            (J[si * js0 + sidx * js1] * confs[si * cs0 + k * cs1]));
                    }
                    ediff = (float)((double)ediff + (-2.0 * ci) *
                                    ((double)jperp * confs[sidx * cs0 + tleft * cs1]));
                    ediff = (float)((double)ediff + (-2.0 * ci) *
                                    ((double)jperp * confs[sidx * cs0 + tright * cs1]));
                    if ((double)ediff > 0.0) {
                        confs[sidx * cs0 + k * cs1] *= -1.0;
                    } else {
                        uint32_t r = orc_rand_next(rs);
                        ++nd;
                        if (exp((double)(ediff / temp)) > (double)r / (double)ORC_RAND_MAX)
                            confs[sidx * cs0 + k * cs1] *= -1.0;
                    }
                }
            }
        }
    }
    if (draws) *draws = nd;
    return 0;
}

/* ------------------------------------------------------------------ */
/* sa.Anneal_dense, piqmc/sa.pyx:160-187 (sa.c:3287-3490): per-spin    */
/* reset, accept ediff > 0 (:181), temp = (float)sched[t] (:163).      */
/* ------------------------------------------------------------------ */
int orc_sa_dense(const double *sched, int S, int M, double *svec, long ss, int N,
                 const double *J, long js0, long js1, const int64_t *perms, uint32_t *rs,
                 uint64_t *draws)
{
    if (N < 1) return -1;
    uint64_t nd = 0;
    for (int s = 0; s < S; ++s) {
        float temp = (float)sched[s];
        for (int m = 0; m < M; ++m) {
            const int64_t *perm = perms + ((size_t)s * M + m) * N;
            for (int t = 0; t < N; ++t) {
                int sidx = (int)perm[t];
                float ediff = 0.0f;
                double ci = svec[sidx * ss];
                for (int si = 0; si < N; ++si) {
                    if (si == sidx)
                        ediff = (float)((double)ediff + (-2.0 * ci) * J[sidx * js0 + si * js1]);
                    else if (sidx < si)
                        ediff = (float)((double)ediff + (-2.0 * ci) *
                                        (J[sidx * js0 + si * js1] * svec[si * ss]));
                    else
                        ediff = (float)((double)ediff + (-2.0 * ci) *
                                        (J[si * js0 + sidx * js1] * svec[si * ss]));
                }
                if ((double)ediff > 0.0) {
                    svec[sidx * ss] *= -1.0;
                } else {
                    uint32_t r = orc_rand_next(rs);
                    ++nd;
                    if (exp((double)(ediff / temp)) > (double)r / (double)ORC_RAND_MAX)
                        svec[sidx * ss] *= -1.0;
                }
            }
        }
    }
    if (draws) *draws = nd;
    return 0;
}

/* ------------------------------------------------------------------ */
/* sa.Anneal_multispin, piqmc/sa.pyx:282-405: 64 configurations packed */
/* in one uint64 per spin, configuration k at bit 63-k, bit 1 = spin   */
/* -1 (:339-345).  Per attempt: ediffs[k] (double, reset per attempt   */
/* :395-396) -= 2*jval where the configuration's bit of svec[sidx]     */
/* (field) or of svec[sidx]^svec[spinidx] (coupling) is 0, += where it */
/* is 1 (:355-379), jval = (float)nbs; flip configuration k iff        */
/* exp(ediffs[k]/temp) > U[k] (:382-393), temp = (float)sched (:349).  */
/* The uniforms come from the caller: U[(g*N + t)*64 + k] is what      */
/* configuration k compares against at attempt t of sweep g (the rng   */
/* stream order is reproduced by the caller).  perms[g*N + t].         */
/* ------------------------------------------------------------------ */
int orc_sa_multispin(const float *temps, int G, int N, uint64_t *svec, const double *nbs,
                     int Z, const int64_t *perms, const double *U)
{
    if (N < 1 || Z < 1) return -1;
    double ed[64];
    for (int g = 0; g < G; ++g) {
        const double temp = (double)temps[g];
        for (int t = 0; t < N; ++t) {
            const int sidx = (int)perms[(size_t)g * N + t];
            for (int k = 0; k < 64; ++k) ed[k] = 0.0;
            for (int si = 0; si < Z; ++si) {
                const double *row = nbs + ((size_t)sidx * Z + si) * 2;
                const int spinidx = (int)row[0];
                const float jval = (float)row[1];
                const uint64_t w = spinidx == sidx ? svec[sidx] : (svec[sidx] ^ svec[spinidx]);
                for (int k = 0; k < 64; ++k) {
                    if (!((w >> (63 - k)) & 1u)) ed[k] -= 2.0 * (double)jval;
                    else ed[k] += 2.0 * (double)jval;
                }
            }
            uint64_t fm = 0;
            for (int k = 0; k < 64; ++k) {
                fm <<= 1;
                if (exp(ed[k] / temp) > U[((size_t)g * N + t) * 64 + k]) fm ^= 1u;
            }
            svec[sidx] ^= fm;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* qmc.QuantumAnneal_parallel, piqmc/qmc.pyx:247-357: OpenMP guided    */
/* prange over spins, float64 ediffs[] (reset after each slice, :357), */
/* no permutation, D1 kept, libc rand() shared by all threads.  Timing */
/* twin for the CPU baseline; nondeterministic like the original.      */
/* ------------------------------------------------------------------ */
int orc_qa_parallel(const double *sched, int S, int M, int P, float temp, int N,
                    double *confs, long cs0, long cs1, const double *nbs, int Z,
                    int nthreads)
{
    if (P < 2 || N < 1 || Z < 1) return -1;
    double *ediffs = (double *)calloc((size_t)N, sizeof(double));
    if (!ediffs) return -2;
    for (int s = 0; s < S; ++s) {
        float jperp = orc_jperp(sched[s], P, temp);
        for (int m = 0; m < M; ++m) {
            for (int k = 0; k < P; ++k) {
#pragma omp parallel for schedule(guided) num_threads(nthreads)
                for (int sidx = 0; sidx < N; ++sidx) {
                    double ci = confs[sidx * cs0 + k * cs1];
                    for (int si = 0; si < Z; ++si) {
                        const double *row = nbs + ((size_t)sidx * Z + si) * 2;
                        int spinidx = (int)row[0];
                        float jval = (float)row[1];
                        if (spinidx == sidx)
                            ediffs[sidx] += -2.0 * ci * jval;
                        else
                            ediffs[sidx] += -2.0 * ci * (jval * confs[spinidx * cs0 + k * cs1]);
                    }
                    ediffs[sidx] += -2.0 * ci * (jperp * confs[sidx * cs0 + (P - 1) * cs1]);
                    ediffs[sidx] += -2.0 * ci * (jperp * confs[sidx * cs0 + 1 * cs1]);
                    if (ediffs[sidx] > 0.0)
                        confs[sidx * cs0 + k * cs1] *= -1.0;
                    else if (exp(ediffs[sidx] / temp) > rand() / (double)ORC_RAND_MAX)
                        confs[sidx * cs0 + k * cs1] *= -1.0;
                }
                memset(ediffs, 0, sizeof(double) * (size_t)N);
            }
        }
    }
    free(ediffs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* "Fixed" semantics (D1 and D2 corrected): periodic Trotter neighbours */
/* k-1, k+1; per-spin reset; exact double local field                  */
/*   ediff = -2 s_i^k ( sum_rows[j==i ? J : J s_j^k] + Jp (s^{k-1}+s^{k+1}) ) */
/* accepted iff ediff > 0 or piqmc_fexp64(ediff/T) > u.                */
/* spins int8 [P][N] (+1/-1), slice-major.                             */
/* ------------------------------------------------------------------ */
static double orc_fixed_ediff(const int8_t *sp, int P, int N, int k, int i,
                              const double *nbs, int Z, float jperp)
{
    const int8_t *row_k = sp + (size_t)k * N;
    double si = (double)row_k[i];
    double h = 0.0;
    for (int z = 0; z < Z; ++z) {
        const double *e = nbs + ((size_t)i * Z + z) * 2;
        int j = (int)e[0];
        double J = (double)(float)e[1];
        if (j == i) h += J;
        else h += J * (double)row_k[j];
    }
    int kl = (k == 0) ? P - 1 : k - 1;
    int kr = (k == P - 1) ? 0 : k + 1;
    double ts = (double)sp[(size_t)kl * N + i] + (double)sp[(size_t)kr * N + i];
    return -2.0 * si * (h + (double)jperp * ts);
}

/* Classical energy of every slice from the neighbour table; pair terms
 * carry 1/2 because GenerateNeighbors lists each bond in both rows. */
void orc_energy_nbs(const int8_t *sp, int P, int N, const double *nbs, int Z, double *E)
{
    for (int k = 0; k < P; ++k) {
        const int8_t *row_k = sp + (size_t)k * N;
        double e = 0.0;
        for (int i = 0; i < N; ++i) {
            for (int z = 0; z < Z; ++z) {
                const double *en = nbs + ((size_t)i * Z + z) * 2;
                int j = (int)en[0];
                double J = (double)(float)en[1];
                if (j == i) e -= J * row_k[i];
                else e -= 0.5 * J * row_k[i] * row_k[j];
            }
        }
        E[k] = e;
    }
}

static double orc_mean_energy(const int8_t *sp, int P, int N, const double *nbs, int Z)
{
    double tot = 0.0;
    double *E = (double *)malloc(sizeof(double) * (size_t)P);
    orc_energy_nbs(sp, P, N, nbs, Z, E);
    for (int k = 0; k < P; ++k) tot += E[k];
    free(E);
    return tot / (double)P;
}

int orc_qa_fixed_seq(const double *sched, int S, int M, int P, float temp, int N,
                     int8_t *sp, const double *nbs, int Z, const int64_t *perms,
                     uint32_t *rs, uint64_t *draws, double *etrace)
{
    if (P < 2 || N < 1 || Z < 1) return -1;
    uint64_t nd = 0;
    for (int s = 0; s < S; ++s) {
        float jperp = orc_jperp(sched[s], P, temp);
        for (int m = 0; m < M; ++m) {
            const int64_t *perm = perms + ((size_t)s * M + m) * N;
            for (int k = 0; k < P; ++k) {
                for (int t = 0; t < N; ++t) {
                    int i = (int)perm[t];
                    double ed = orc_fixed_ediff(sp, P, N, k, i, nbs, Z, jperp);
                    int flip;
                    if (ed > 0.0) {
                        flip = 1;
                    } else {
                        uint32_t r = orc_rand_next(rs);
                        ++nd;
                        flip = piqmc_fexp64(ed / (double)temp) > (double)r / (double)ORC_RAND_MAX;
                    }
                    if (flip) sp[(size_t)k * N + i] = (int8_t)-sp[(size_t)k * N + i];
                }
            }
        }
        if (etrace) etrace[s] = orc_mean_energy(sp, P, N, nbs, Z);
    }
    if (draws) *draws = nd;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Colored (parallel) fixed-semantics sweep: phases (tau, c) for tau in */
/* {0,1} (slice parity) and c over the spatial colors; a phase updates  */
/* every spin i of color c in every slice k = tau (mod 2).              */
/* Draws are bit-sliced per (sweep g, spin i, replica r, parity tau):   */
/* with w = 32 when P/2 <= 32 (else 64), random words R_0..R_31 of w    */
/* bits come from philox4x32_10(key = seed, ctr = {g, i, r, tau<<24 | q}) */
/*   w = 32: R_b = out[b&3] with q = b>>2                               */
/*   w = 64: R_b = out[2(b&1)] | out[2(b&1)+1]<<32 with q = b>>1        */
/* and the 32-bit uniform of slice k (bit m = k>>1) is                  */
/*   u = sum_b ((R_b >> m) & 1) << (31 - b).                            */
/* The attempt is accepted iff u < thr of piqmc_fixed_class.            */
/* ------------------------------------------------------------------ */
static void orc_sliced_words(uint32_t g, uint32_t i, uint32_t rep, uint32_t tau, int wide,
                             const uint32_t *key, uint64_t *R)
{
    for (int q = 0; q < (wide ? 16 : 8); ++q) {
        uint32_t ctr[4] = {g, i, rep, (tau << 24) | (uint32_t)q}, out[4];
        orc_philox4x32(ctr, key, out);
        if (wide) {
            R[2 * q] = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
            R[2 * q + 1] = (uint64_t)out[2] | ((uint64_t)out[3] << 32);
        } else {
            for (int t = 0; t < 4; ++t) R[4 * q + t] = out[t];
        }
    }
}
/* stats (optional, uint64[4]): parity words swept, words with at least one
 * stochastic attempt, stochastic attempts, and the Philox blocks an
 * MSB-first comparator with per-word early exit consumes (a bit is decided
 * at the first uniform bit that differs from its threshold's; a block
 * supplies 32 / words-per-call uniform bits of every slice of the word). */
static int colored_core(const float *jps, const float *temps, int S, int M, int P, int N,
                        int8_t *sp, const double *nbs, int Z, const int32_t *color,
                        int ncolors, uint64_t seed, uint32_t replica, uint32_t sweep_base,
                        double *etrace, uint64_t *stats)
{
    if (P < 2 || (P & 1) || N < 1 || Z < 1 || ncolors < 1) return -1;
    for (int i = 0; i < N; ++i) {
        if (color[i] < 0 || color[i] >= ncolors) return -2;
        for (int z = 0; z < Z; ++z) {
            const double *e = nbs + ((size_t)i * Z + z) * 2;
            int j = (int)e[0];
            if (j < 0 || j >= N) return -3;
            if (j != i && (float)e[1] != 0.0f && color[j] == color[i]) return -2;
        }
    }
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int s = 0; s < S; ++s) {
        const float jperp = jps[s], temp = temps[s];
        for (int m = 0; m < M; ++m) {
            uint32_t g = sweep_base + (uint32_t)(s * M + m);
            for (int tau = 0; tau < 2; ++tau) {
                for (int c = 0; c < ncolors; ++c) {
                    for (int i = 0; i < N; ++i) {
                        if (color[i] != c) continue;
                        uint64_t R[32];
                        int have_R = 0, nstoch = 0, need = 0;
                        for (int k = tau; k < P; k += 2) {
                            double ed = orc_fixed_ediff(sp, P, N, k, i, nbs, Z, jperp);
                            unsigned int thr = 0;
                            int cls = piqmc_fixed_class(ed, (double)temp, &thr);
                            int flip = cls == 1;
                            if (cls == 2) {
                                if (!have_R) {
                                    orc_sliced_words(g, (uint32_t)i, replica, (uint32_t)tau, P / 2 > 32,
                                                     key, R);
                                    have_R = 1;
                                }
                                int mb = k >> 1;
                                uint32_t u = 0;
                                for (int b = 0; b < 32; ++b)
                                    u |= (uint32_t)((R[b] >> mb) & 1u) << (31 - b);
                                flip = u < thr;
                                if (stats) {
                                    /* uniform bits examined: up to the first that differs */
                                    const uint32_t d = u ^ (uint32_t)thr;
                                    int used = 32;
                                    for (int b = 0; b < 32; ++b)
                                        if ((d >> (31 - b)) & 1u) { used = b + 1; break; }
                                    const int per = P / 2 > 32 ? 2 : 4;   /* uniform bits per block */
                                    const int blocks = (used + per - 1) / per;
                                    if (blocks > need) need = blocks;
                                    ++nstoch;
                                }
                            }
                            if (flip) sp[(size_t)k * N + i] = (int8_t)-sp[(size_t)k * N + i];
                        }
                        if (stats) {
                            stats[0] += 1;
                            stats[1] += nstoch > 0;
                            stats[2] += (uint64_t)nstoch;
                            stats[3] += (uint64_t)need;
                        }
                    }
                }
            }
        }
        if (etrace) etrace[s] = orc_mean_energy(sp, P, N, nbs, Z);
    }
    return 0;
}

int orc_qa_colored(const double *sched, int S, int M, int P, float temp, int N,
                   int8_t *sp, const double *nbs, int Z, const int32_t *color,
                   int ncolors, uint64_t seed, uint32_t replica, uint32_t sweep_base,
                   double *etrace)
{
    if (P < 2) return -1;
    float *jps = (float *)malloc(sizeof(float) * (S > 0 ? 2 * S : 1));
    if (!jps) return -4;
    for (int s = 0; s < S; ++s) {
        jps[s] = orc_jperp(sched[s], P, temp);
        jps[S + s] = temp;
    }
    int rc = colored_core(jps, jps + S, S, M, P, N, sp, nbs, Z, color, ncolors, seed, replica,
                          sweep_base, etrace, NULL);
    free(jps);
    return rc;
}

/* orc_qa_colored that also counts its work (see colored_core): stats[0..3]
 * = words, stochastic words, stochastic attempts, Philox blocks. */
int orc_qa_colored_stats(const double *sched, int S, int M, int P, float temp, int N,
                         int8_t *sp, const double *nbs, int Z, const int32_t *color,
                         int ncolors, uint64_t seed, uint32_t replica, uint64_t *stats)
{
    if (P < 2) return -1;
    float *jps = (float *)malloc(sizeof(float) * (S > 0 ? 2 * S : 1));
    if (!jps) return -4;
    for (int s = 0; s < S; ++s) {
        jps[s] = orc_jperp(sched[s], P, temp);
        jps[S + s] = temp;
    }
    for (int t = 0; t < 4; ++t) stats[t] = 0;
    int rc = colored_core(jps, jps + S, S, M, P, N, sp, nbs, Z, color, ncolors, seed, replica, 0,
                          NULL, stats);
    free(jps);
    return rc;
}

/* Classical annealing of C chains with the parallel sweep: chain c is slice
 * c % P of replica rbase + c / P (P the largest even divisor of C up to 64),
 * with zero Trotter coupling and temperature (float)sched[s] in step s. */
int orc_sa_colored(const double *sched, int S, int M, int C, int N, int8_t *spins,
                   const double *nbs, int Z, const int32_t *color, int ncolors,
                   uint64_t seed, uint32_t rbase, uint32_t sweep_base)
{
    if (C < 2 || (C & 1)) return -1;
    int P = 64;
    while (C % P) P -= 2;
    float *jps = (float *)malloc(sizeof(float) * (S > 0 ? 2 * S : 1));
    if (!jps) return -4;
    for (int s = 0; s < S; ++s) {
        jps[s] = 0.0f;
        jps[S + s] = (float)sched[s];
    }
    int rc = 0;
    for (int q = 0; q < C / P && !rc; ++q)
        rc = colored_core(jps, jps + S, S, M, P, N, spins + (size_t)q * P * N, nbs, Z, color,
                          ncolors, seed, rbase + (uint32_t)q, sweep_base, NULL, NULL);
    free(jps);
    return rc;
}

/* sa.ClassicalIsingEnergy, piqmc/sa.pyx:41-44:  -s.(J_off s) - sum(diag(J) s)
 * from a COO list (duplicates add, as todense() sums them). */
double orc_energy_coo(const double *spins, int N, long nnz, const int64_t *ri,
                      const int64_t *ci, const double *v)
{
    (void)N;
    double off = 0.0, diag = 0.0;
    for (long t = 0; t < nnz; ++t) {
        int64_t i = ri[t], j = ci[t];
        if (i == j) diag += v[t] * spins[i];
        else off += spins[i] * v[t] * spins[j];
    }
    return -off - diag;
}

/* The fixed-semantics acceptance exponential (shared header), exported so the
 * tests can compare it with libm exp. */
double orc_fexp64(double x) { return piqmc_fexp64(x); }
