// Host-side services of libpiqmc_amd: error state, the reference's random
// streams (glibc TYPE_3, numpy legacy MT19937 permutations), the problem
// handle (neighbour table in HBM + graph coloring) and GenerateNeighbors.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <utility>
#include <unordered_map>

#include <climits>

#include <tuple>
#include <cstring>
#include "piqmc_internal.h"

namespace piqmc {
static thread_local std::string g_err;
int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char *what)
{
    return set_error(PIQMC_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace piqmc

using piqmc::set_error;

extern "C" const char *piqmc_last_error(void) { return piqmc::g_err.c_str(); }
extern "C" int piqmc_version(void) { return 1; }

namespace piqmc {
struct KernelTiming {
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending = false;
    ~KernelTiming()
    {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};
static thread_local KernelTiming g_timing;
void timing_begin(hipStream_t st)
{
    for (auto &e : g_timing.ev)
        if (!e && hipEventCreate(&e) != hipSuccess) e = nullptr;
    if (g_timing.ev[0]) (void)hipEventRecord(g_timing.ev[0], st);
    g_timing.pending = false;
}
void timing_end(hipStream_t st)
{
    if (g_timing.ev[1] && hipEventRecord(g_timing.ev[1], st) == hipSuccess) g_timing.pending = true;
}
}  // namespace piqmc

extern "C" int piqmc_last_kernel_ms(float *ms)
{
    if (!ms) return set_error(PIQMC_EINVAL, "null output");
    if (!piqmc::g_timing.pending) return set_error(PIQMC_EINVAL, "no timed kernel on this thread");
    PIQMC_HIP(hipEventSynchronize(piqmc::g_timing.ev[1]));
    PIQMC_HIP(hipEventElapsedTime(ms, piqmc::g_timing.ev[0], piqmc::g_timing.ev[1]));
    return PIQMC_OK;
}

float host_jperp(double gamma, int P, float temp)
{
    float pt = (float)P * temp;
    return (float)(((-0.5 * P) * temp) * log(tanh(gamma / pt)));
}

// ------------------------------------------------------------------ TYPE_3
// glibc random_r.c: r[f] += r[f-3 mod 31]; output >> 1.
static inline uint32_t t3_next(uint32_t *st)
{
    uint32_t f = st[31];
    uint32_t r = f >= 3 ? f - 3 : f + 28;
    uint32_t v = st[f] + st[r];
    st[f] = v;
    st[31] = f == 30 ? 0 : f + 1;
    return v >> 1;
}

extern "C" int piqmc_rand_seed(uint32_t seed, uint32_t *st)
{
    if (!st) return set_error(PIQMC_EINVAL, "null state");
    if (seed == 0) seed = 1;
    int32_t *w = reinterpret_cast<int32_t *>(st);
    w[0] = (int32_t)seed;
    int32_t word = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {  // Schrage's 16807 LCG, truncating division
        long hi = word / 127773;
        long lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        w[i] = word;
    }
    st[31] = 3;
    for (int i = 0; i < 310; ++i) t3_next(st);
    return PIQMC_OK;
}

extern "C" int piqmc_rand_advance(uint32_t *st, uint64_t n)
{
    if (!st || st[31] > 30) return set_error(PIQMC_EINVAL, "bad TYPE_3 state");
    for (uint64_t i = 0; i < n; ++i) t3_next(st);
    return PIQMC_OK;
}

// ------------------------------------------------------------------ MT19937
namespace {
struct MT {
    uint32_t k[624];
    int pos;
    explicit MT(uint32_t seed)
    {
        k[0] = seed;
        for (uint32_t i = 1; i < 624; ++i) k[i] = 1812433253u * (k[i - 1] ^ (k[i - 1] >> 30)) + i;
        pos = 624;
    }
    void twist()
    {
        const uint32_t UP = 0x80000000u, LO = 0x7fffffffu, A = 0x9908b0dfu;
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (k[i] & UP) | (k[(i + 1) % 624] & LO);
            k[i] = k[(i + 397) % 624] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
        }
        pos = 0;
    }
    uint32_t next()
    {
        if (pos >= 624) twist();
        uint32_t y = k[pos++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    // numpy legacy random_interval(max) for max < 2^32
    uint32_t interval(uint32_t max)
    {
        if (max == 0) return 0;
        uint32_t mask = max;
        mask |= mask >> 1; mask |= mask >> 2; mask |= mask >> 4; mask |= mask >> 8; mask |= mask >> 16;
        uint32_t v;
        do { v = next() & mask; } while (v > max);
        return v;
    }
};
}  // namespace

extern "C" int piqmc_perm_streams(const uint32_t *seeds, int R, int N, int G, uint16_t *perms,
                                  int nthreads)
{
    if (!seeds || !perms || R < 1 || N < 1 || G < 1)
        return set_error(PIQMC_EINVAL, "piqmc_perm_streams: bad arguments");
    if (N > 65536) return set_error(PIQMC_EINVAL, "permutation streams hold N <= 65536");
    if (nthreads < 1) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::min(nthreads, R);
    auto work = [&](int t) {
        std::vector<uint32_t> a(N);
        for (int r = t; r < R; r += nthreads) {
            MT mt(seeds[r]);
            for (int i = 0; i < N; ++i) a[i] = (uint32_t)i;
            for (int g = 0; g < G; ++g) {  // perm g+1 = shuffle(copy of perm g)
                for (int i = N - 1; i >= 1; --i) std::swap(a[i], a[mt.interval((uint32_t)i)]);
                uint16_t *out = perms + (size_t)g * N * R + r;
                for (int i = 0; i < N; ++i) out[(size_t)i * R] = (uint16_t)a[i];
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    return PIQMC_OK;
}

// ------------------------------------------------------------------ problem
static int greedy_color(const piqmc_problem &p, std::vector<int> &color, int &nc)
{
    const int N = p.N, Z = p.Z;
    // Bipartite check by BFS first (2 colors); otherwise greedy (Welsh-Powell).
    color.assign(N, -1);
    bool bip = true;
    std::vector<int> q;
    for (int s = 0; s < N && bip; ++s) {
        if (color[s] >= 0) continue;
        color[s] = 0;
        q.assign(1, s);
        for (size_t h = 0; h < q.size() && bip; ++h) {
            int i = q[h];
            for (int z = 0; z < Z; ++z) {
                int j = p.idx[(size_t)i * Z + z];
                if (j == i || p.J[(size_t)i * Z + z] == 0.0f) continue;
                if (color[j] < 0) {
                    color[j] = 1 - color[i];
                    q.push_back(j);
                } else if (color[j] == color[i]) {
                    bip = false;
                    break;
                }
            }
        }
    }
    // Coupling lists can be one-directional; verify on both directions.
    if (bip) {
        for (int i = 0; i < N && bip; ++i)
            for (int z = 0; z < Z; ++z) {
                int j = p.idx[(size_t)i * Z + z];
                if (j != i && p.J[(size_t)i * Z + z] != 0.0f && color[j] == color[i]) bip = false;
            }
    }
    if (bip) {
        nc = 1;
        for (int i = 0; i < N; ++i) nc = std::max(nc, color[i] + 1);
        return 0;
    }
    std::vector<std::vector<int>> adj(N);
    for (int i = 0; i < N; ++i)
        for (int z = 0; z < Z; ++z) {
            int j = p.idx[(size_t)i * Z + z];
            if (j == i || p.J[(size_t)i * Z + z] == 0.0f) continue;
            adj[i].push_back(j);
            adj[j].push_back(i);
        }
    std::vector<int> ord(N);
    for (int i = 0; i < N; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return adj[a].size() > adj[b].size(); });
    color.assign(N, -1);
    nc = 0;
    std::vector<char> used;
    for (int i : ord) {
        used.assign(adj[i].size() + 2, 0);
        for (int j : adj[i])
            if (color[j] >= 0 && color[j] < (int)used.size()) used[color[j]] = 1;
        int c = 0;
        while (used[c]) ++c;
        color[i] = c;
        nc = std::max(nc, c + 1);
    }
    return 0;
}

static int upload_coloring(piqmc_problem &p)
{
    const int N = p.N;
    p.color_off.assign(p.ncolors + 1, 0);
    for (int i = 0; i < N; ++i) p.color_off[p.color[i] + 1]++;
    for (int c = 0; c < p.ncolors; ++c) p.color_off[c + 1] += p.color_off[c];
    std::vector<int32_t> order(N);
    std::vector<int> fill(p.color_off.begin(), p.color_off.end() - 1);
    for (int i = 0; i < N; ++i) order[fill[p.color[i]]++] = i;
    if (p.d_order) (void)hipFree(p.d_order);
    if (p.d_cnb) (void)hipFree(p.d_cnb);
    if (p.d_cinfo) (void)hipFree(p.d_cinfo);
    if (p.d_hq) (void)hipFree(p.d_hq);
    if (p.d_prevj) (void)hipFree(p.d_prevj);
    p.d_order = nullptr; p.d_cnb = nullptr; p.d_cinfo = nullptr; p.d_hq = nullptr; p.d_prevj = nullptr;
    PIQMC_HIP(hipMalloc(&p.d_order, sizeof(int32_t) * N));
    PIQMC_HIP(hipMemcpy(p.d_order, order.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
    // coupling of spin order[q] to its predecessor order[q-1], summed over
    // the row's entries for that neighbour (0 for q = 0)
    std::vector<double> prevj(N, 0.0);
    for (int q = 1; q < N; ++q) {
        const int i = order[q], pj = order[q - 1];
        for (int e = 0; e < p.Z; ++e)
            if (p.idx[(size_t)i * p.Z + e] == pj) prevj[q] += (double)p.J[(size_t)i * p.Z + e];
    }
    PIQMC_HIP(hipMalloc(&p.d_prevj, sizeof(double) * N));
    PIQMC_HIP(hipMemcpy(p.d_prevj, prevj.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    if (p.pmj || p.pmjf) {
        const int zm = std::max(p.zmax, 1);
        std::vector<int32_t> cnb((size_t)zm * N, 0);
        std::vector<uint32_t> cinfo(N, 0);
        std::vector<float> hq(N, 0.0f);
        for (int q = 0; q < N; ++q) {
            int i = order[q], z = 0;
            uint32_t neg = 0;
            for (int e = 0; e < p.Z; ++e) {
                float Jv = p.J[(size_t)i * p.Z + e];
                if (Jv == 0.0f) continue;
                if (p.idx[(size_t)i * p.Z + e] == i) {  // the field (only with pmjf)
                    hq[q] = Jv;
                    continue;
                }
                cnb[(size_t)z * N + q] = p.idx[(size_t)i * p.Z + e];
                if (Jv < 0.0f) neg |= 1u << z;
                ++z;
            }
            cinfo[q] = neg | ((uint32_t)z << 8);
        }
        if (p.pmjf) {
            PIQMC_HIP(hipMalloc(&p.d_hq, sizeof(float) * N));
            PIQMC_HIP(hipMemcpy(p.d_hq, hq.data(), sizeof(float) * N, hipMemcpyHostToDevice));
        }
        PIQMC_HIP(hipMalloc(&p.d_cnb, sizeof(int32_t) * cnb.size()));
        PIQMC_HIP(hipMemcpy(p.d_cnb, cnb.data(), sizeof(int32_t) * cnb.size(), hipMemcpyHostToDevice));
        PIQMC_HIP(hipMalloc(&p.d_cinfo, sizeof(uint32_t) * N));
        PIQMC_HIP(hipMemcpy(p.d_cinfo, cinfo.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
    }
    return PIQMC_OK;
}

extern "C" int piqmc_problem_create(int N, int Z, const double *nbs, piqmc_problem **out)
{
    if (!out) return set_error(PIQMC_EINVAL, "null output handle");
    *out = nullptr;
    if (N < 1 || Z < 1 || !nbs) return set_error(PIQMC_EINVAL, "nspins and maxnb must be >= 1");
    auto *p = new piqmc_problem();
    p->N = N;
    p->Z = Z;
    p->idx.resize((size_t)N * Z);
    p->J.resize((size_t)N * Z);
    p->pmj = 1;
    p->pmjf = 1;
    int nfield = 0;
    p->zmax = 0;
    for (int i = 0; i < N; ++i) {
        int z = 0, fields = 0;
        for (int e = 0; e < Z; ++e) {
            double jd = nbs[((size_t)i * Z + e) * 2];
            if (!(jd > -1.0) || !(jd < (double)N)) {  // (int) truncation must land in [0, N)
                delete p;
                return set_error(PIQMC_EINVAL, "neighbour index out of range in row " + std::to_string(i));
            }
            int j = (int)jd;
            float Jv = (float)nbs[((size_t)i * Z + e) * 2 + 1];
            if (!isfinite(Jv)) {
                delete p;
                return set_error(PIQMC_EINVAL, "non-finite coupling in row " + std::to_string(i));
            }
            p->idx[(size_t)i * Z + e] = j;
            p->J[(size_t)i * Z + e] = Jv;
            if (Jv != 0.0f) {
                if (j == i) {
                    p->pmj = 0;
                    ++fields;
                    ++nfield;
                    const float ah = fabsf(Jv);
                    if (fields > 1 || ah < 0x1p-27f || ah > 0x1p20f) p->pmjf = 0;
                } else if (Jv != 1.0f && Jv != -1.0f) {
                    p->pmj = 0;
                    p->pmjf = 0;
                } else {
                    ++z;
                }
            }
        }
        p->zmax = std::max(p->zmax, z);
    }
    p->exact_rows = 1;
    for (int i = 0; i < N && p->exact_rows; ++i) {
        double bound = 0.0;
        int lsb = INT_MAX;
        for (int e = 0; e < Z; ++e) {
            const float Jv = p->J[(size_t)i * Z + e];
            if (Jv == 0.0f) continue;
            bound += fabs((double)Jv);
            int ex = 0;
            (void)frexpf(Jv, &ex);                  // |Jv| = m 2^ex, 1/2 <= m < 1
            lsb = std::min(lsb, std::max(ex - 24, -149));  // float: 24-bit significand
        }
        if (lsb == INT_MAX) continue;               // empty row
        int eb = 0;
        (void)frexp(bound, &eb);                    // every partial sum is below 2^eb
        if (eb - lsb > 52) p->exact_rows = 0;
    }
    {
        // fields_ok: every nonzero coupling (i, j != i, J) is the row's only
        // entry for j, and (j, i, J) is in row j
        std::vector<std::tuple<int, int, uint32_t>> ent;
        for (int i = 0; i < N; ++i)
            for (int e = 0; e < Z; ++e) {
                const float Jv = p->J[(size_t)i * Z + e];
                const int j = p->idx[(size_t)i * Z + e];
                if (Jv != 0.0f && j != i) {
                    uint32_t b;
                    std::memcpy(&b, &Jv, 4);
                    ent.emplace_back(i, j, b);
                }
            }
        std::sort(ent.begin(), ent.end());
        bool ok = true;
        for (size_t t = 1; t < ent.size() && ok; ++t)
            if (std::get<0>(ent[t]) == std::get<0>(ent[t - 1]) && std::get<1>(ent[t]) == std::get<1>(ent[t - 1]))
                ok = false;
        for (size_t t = 0; t < ent.size() && ok; ++t)
            ok = std::binary_search(ent.begin(), ent.end(),
                                    std::make_tuple(std::get<1>(ent[t]), std::get<0>(ent[t]), std::get<2>(ent[t])));
        p->fields_ok = ok ? 1 : 0;
    }
    if (p->zmax > 7) p->pmj = p->pmjf = 0;  // bit-sliced counter holds 0..7
    if (p->pmj || nfield == 0) p->pmjf = 0;
    p->zuni = p->zmax;
    for (int i = 0; i < N && p->zuni >= 0; ++i) {
        int z = 0;
        for (int e = 0; e < Z; ++e)
            if (p->J[(size_t)i * Z + e] != 0.0f) ++z;
        if (z != p->zmax) p->zuni = -1;
    }
    p->zuni_nf = -2;
    for (int i = 0; i < N && p->zuni_nf != -1; ++i) {
        int z = 0;
        for (int e = 0; e < Z; ++e)
            if (p->J[(size_t)i * Z + e] != 0.0f && p->idx[(size_t)i * Z + e] != i) ++z;
        if (p->zuni_nf == -2) p->zuni_nf = z;
        else if (z != p->zuni_nf) p->zuni_nf = -1;
    }
    if (p->zuni_nf < 0) p->zuni_nf = -1;
    std::vector<int2> rows((size_t)N * Z);
    for (size_t t = 0; t < rows.size(); ++t) {
        rows[t].x = p->idx[t];
        float f = p->J[t];
        memcpy(&rows[t].y, &f, 4);
    }
    hipError_t e = hipMalloc(&p->d_rows, sizeof(int2) * rows.size());
    if (e == hipSuccess) e = hipMemcpy(p->d_rows, rows.data(), sizeof(int2) * rows.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        piqmc_problem_destroy(p);
        return piqmc::hip_fail(e, "problem upload");
    }
    greedy_color(*p, p->color, p->ncolors);
    int rc = upload_coloring(*p);
    if (rc) {
        piqmc_problem_destroy(p);
        return rc;
    }
    *out = p;
    return PIQMC_OK;
}

// Rows of a dense coupling matrix as the *_dense functions read it
// (piqmc/qmc.pyx:206-219, piqmc/sa.pyx:169-179): spin i meets every si in
// ascending order, J[i,i] as its field, J[min(i,si), max(i,si)] otherwise (the
// lower triangle is never read).  A zero coupling adds +-0 to the running
// float ediff, which changes neither its value nor any later comparison, so
// zero entries are dropped and the surviving terms keep their order.
extern "C" int piqmc_problem_create_dense(int N, const double *Jm, int64_t s0, int64_t s1,
                                          piqmc_problem **out)
{
    if (!out) return set_error(PIQMC_EINVAL, "null output handle");
    *out = nullptr;
    if (N < 1 || !Jm) return set_error(PIQMC_EINVAL, "nspins must be >= 1 and J non-null");
    std::vector<std::vector<std::pair<int, double>>> rows(N);
    int Z = 1;
    for (int i = 0; i < N; ++i) {
        for (int si = 0; si < N; ++si) {
            const int a = std::min(i, si), b = std::max(i, si);
            const double v = Jm[a * s0 + b * s1];
            if (!std::isfinite(v) || !std::isfinite((float)v))
                return set_error(PIQMC_EINVAL, "non-finite coupling J[" + std::to_string(a) + "," +
                                                   std::to_string(b) + "]");
            if (v != 0.0) rows[i].push_back({si, v});
        }
        Z = std::max(Z, (int)rows[i].size());
    }
    std::vector<double> nbs((size_t)N * Z * 2, 0.0), jd((size_t)N * Z, 0.0);
    for (int i = 0; i < N; ++i)
        for (size_t e = 0; e < rows[i].size(); ++e) {
            nbs[((size_t)i * Z + e) * 2] = rows[i][e].first;
            nbs[((size_t)i * Z + e) * 2 + 1] = rows[i][e].second;
            jd[(size_t)i * Z + e] = rows[i][e].second;
        }
    piqmc_problem *p = nullptr;
    int rc = piqmc_problem_create(N, Z, nbs.data(), &p);
    if (rc) return rc;
    p->dense = 1;
    hipError_t e = hipMalloc(&p->d_jd, sizeof(double) * jd.size());
    if (e == hipSuccess) e = hipMemcpy(p->d_jd, jd.data(), sizeof(double) * jd.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        piqmc_problem_destroy(p);
        return piqmc::hip_fail(e, "problem upload");
    }
    *out = p;
    return PIQMC_OK;
}

extern "C" int piqmc_problem_destroy(piqmc_problem *p)
{
    if (!p) return PIQMC_OK;
    if (p->d_rows) (void)hipFree(p->d_rows);
    if (p->d_jd) (void)hipFree(p->d_jd);
    if (p->d_order) (void)hipFree(p->d_order);
    if (p->d_cnb) (void)hipFree(p->d_cnb);
    if (p->d_cinfo) (void)hipFree(p->d_cinfo);
    if (p->d_hq) (void)hipFree(p->d_hq);
    if (p->d_prevj) (void)hipFree(p->d_prevj);
    delete p;
    return PIQMC_OK;
}

extern "C" int piqmc_problem_info(const piqmc_problem *p, int *n, int *z, int *nc, int *pmj, int *zmax)
{
    if (!p) return set_error(PIQMC_EINVAL, "null problem");
    if (n) *n = p->N;
    if (z) *z = p->Z;
    if (nc) *nc = p->ncolors;
    if (pmj) *pmj = p->pmj ? 1 : (p->pmjf ? 2 : 0);
    if (zmax) *zmax = p->zmax;
    return PIQMC_OK;
}

extern "C" int piqmc_problem_exact_rows(const piqmc_problem *p, int *exact)
{
    if (!p || !exact) return set_error(PIQMC_EINVAL, "null argument");
    *exact = p->exact_rows;
    return PIQMC_OK;
}

extern "C" int piqmc_problem_set_coloring(piqmc_problem *p, int nc, const int32_t *color)
{
    if (!p || !color || nc < 1) return set_error(PIQMC_EINVAL, "bad coloring arguments");
    for (int i = 0; i < p->N; ++i) {
        if (color[i] < 0 || color[i] >= nc) return set_error(PIQMC_EINVAL, "color out of range");
        for (int e = 0; e < p->Z; ++e) {
            int j = p->idx[(size_t)i * p->Z + e];
            if (j != i && p->J[(size_t)i * p->Z + e] != 0.0f && color[j] == color[i])
                return set_error(PIQMC_EINVAL, "coloring is not proper: spins " + std::to_string(i) +
                                                   " and " + std::to_string(j) + " share a color");
        }
    }
    p->ncolors = nc;
    p->color.assign(color, color + p->N);
    return upload_coloring(*p);
}

extern "C" int piqmc_problem_set_sweep_path(piqmc_problem *p, int path)
{
    if (!p || path < 0 || path > 2) return set_error(PIQMC_EINVAL, "sweep path must be 0 (auto), 1 or 2");
    p->sweep_path = path;
    return PIQMC_OK;
}

extern "C" int piqmc_problem_set_dense_variant(piqmc_problem *p, int variant)
{
    if (!p || variant < 0 || variant > 6) return set_error(PIQMC_EINVAL, "dense variant must be 0 (auto) .. 6");
    p->dense_variant = variant;
    return PIQMC_OK;
}

extern "C" int piqmc_problem_get_coloring(const piqmc_problem *p, int32_t *color)
{
    if (!p || !color) return set_error(PIQMC_EINVAL, "bad arguments");
    for (int i = 0; i < p->N; ++i) color[i] = p->color[i];
    return PIQMC_OK;
}

// ------------------------------------------------------------------ GenerateNeighbors
// CPython 2.7 dict iteration order over (i, j) tuple keys: tuplehash
// (Objects/tupleobject.c) and lookdict probing / dictresize growth
// (Objects/dictobject.c).  That is the order in which piqmc/tools.pyx:84
// visits J.todok().iterkeys().
namespace {
inline int64_t py27_int_hash(int64_t v) { return v == -1 ? -2 : v; }
inline int64_t py27_tuple_hash(int64_t a, int64_t b)
{
    uint64_t x = 0x345678u, mult = 1000003u;
    int64_t items[2] = {a, b};
    int64_t len = 2;
    for (int t = 0; t < 2; ++t) {
        --len;
        x = (x ^ (uint64_t)py27_int_hash(items[t])) * mult;
        mult += (uint64_t)(82520 + len + len);
    }
    x += 97531u;
    int64_t h = (int64_t)x;
    return h == -1 ? -2 : h;
}
struct Slot {
    int64_t h;
    int64_t a, b;
    bool used;
};
size_t py27_probe(const std::vector<Slot> &tab, size_t mask, int64_t h, int64_t a, int64_t b)
{
    size_t i = (size_t)h, perturb = (size_t)h, s = i & mask;
    while (tab[s].used && !(tab[s].a == a && tab[s].b == b)) {
        i = (i << 2) + i + perturb + 1;
        s = i & mask;
        perturb >>= 5;
    }
    return s;
}
std::vector<std::pair<int64_t, int64_t>> py27_order(const std::vector<std::pair<int64_t, int64_t>> &keys)
{
    size_t mask = 7, used = 0, fill = 0;
    std::vector<Slot> tab(8, Slot{0, 0, 0, false});
    for (auto &k : keys) {
        int64_t h = py27_tuple_hash(k.first, k.second);
        size_t s = py27_probe(tab, mask, h, k.first, k.second);
        if (tab[s].used) continue;
        tab[s] = Slot{h, k.first, k.second, true};
        ++used;
        ++fill;
        if (fill * 3 >= (mask + 1) * 2) {
            size_t minused = (used > 50000 ? 2 : 4) * used, size = 8;
            while (size <= minused) size <<= 1;
            std::vector<Slot> old;
            old.swap(tab);
            tab.assign(size, Slot{0, 0, 0, false});
            mask = size - 1;
            for (auto &e : old)
                if (e.used) tab[py27_probe(tab, mask, e.h, e.a, e.b)] = e;
            fill = used;
        }
    }
    std::vector<std::pair<int64_t, int64_t>> out;
    out.reserve(used);
    for (auto &e : tab)
        if (e.used) out.emplace_back(e.a, e.b);
    return out;
}
struct PairHash {
    size_t operator()(const std::pair<int64_t, int64_t> &p) const
    {
        return std::hash<int64_t>()(p.first * 1000003 ^ p.second);
    }
};
}  // namespace

extern "C" int piqmc_generate_neighbors(int N, int maxnb, int64_t nnz, const int64_t *rows,
                                        const int64_t *cols, const double *vals, int order,
                                        double *nbs)
{
    if (N < 0 || maxnb < 0 || nnz < 0 || !nbs || (nnz && (!rows || !cols || !vals)))
        return set_error(PIQMC_EINVAL, "GenerateNeighbors: bad arguments");
    memset(nbs, 0, sizeof(double) * 2 * (size_t)N * maxnb);
    // Last value wins per key (dict assignment); position is the first insertion.
    std::unordered_map<std::pair<int64_t, int64_t>, double, PairHash> val;
    std::vector<std::pair<int64_t, int64_t>> keys;
    keys.reserve(nnz);
    for (int64_t t = 0; t < nnz; ++t) {
        auto k = std::make_pair(rows[t], cols[t]);
        if (!val.count(k)) keys.push_back(k);
        val[k] = vals[t];
    }
    std::vector<std::pair<int64_t, int64_t>> it;
    if (order == 0) it = py27_order(keys);
    else if (order == 1) it = keys;
    else if (order == 2) {
        it = keys;
        std::sort(it.begin(), it.end());
    } else return set_error(PIQMC_EINVAL, "GenerateNeighbors: unknown order");
    // One pass over the key order fills every row in the order the reference's
    // per-spin scans (tools.pyx:81-92) would.
    std::vector<int> fillc(N, 0);
    auto put = [&](int64_t row, int64_t other, double v) -> int {
        if (row < 0 || row >= N) return 0;  // the reference scans ispin in [0, N) only
        if (fillc[row] >= maxnb)
            return set_error(PIQMC_EINDEX, "Out of bounds on buffer access (axis 1): spin " +
                                               std::to_string(row) + " has more than maxnb entries");
        double *e = nbs + ((size_t)row * maxnb + fillc[row]) * 2;
        e[0] = (double)other;
        e[1] = v;
        fillc[row]++;
        return 0;
    };
    for (auto &k : it) {
        double v = val[k];
        int rc = put(k.first, k.second, v);
        if (rc) return rc;
        if (k.second != k.first) {
            rc = put(k.second, k.first, v);
            if (rc) return rc;
        }
    }
    return PIQMC_OK;
}
