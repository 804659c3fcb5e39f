// State of a MADE network (learned slice proposal), shared by the HIP
// translation units that implement it.
#pragma once

#include "piqmc_internal.h"

// ------------------------------------------------------------------ MADE state
struct piqmc_made {
    int N = 0, H1 = 0, H2 = 0, cap = 0;  // cap = max batch rows of the buffers
    int layers_in[3], layers_out[3];
    float *w[3] = {}, *b[3] = {}, *mw[3] = {}, *vw[3] = {}, *mb[3] = {}, *vb[3] = {};
    uint16_t *wb[3] = {}, *wt[3] = {};
    uint8_t *mask[3] = {};
    float *gw[3] = {}, *gb[3] = {};
    // activations [cap][*]
    uint16_t *x = nullptr, *h1 = nullptr, *h2 = nullptr;
    float *logits = nullptr;
    uint16_t *dlog = nullptr, *dh2 = nullptr, *dh1 = nullptr;
    uint16_t *tA = nullptr, *tB = nullptr;  // transposes for weight gradients
    double *nll = nullptr;
    int8_t *s = nullptr;
    int step = 0;
    // sampling graph (captured once per batch size on a private stream)
    uint32_t *rng = nullptr;  // {key0, key1, round}
    hipStream_t gs = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gB = 0;
    int use_graph = 1;
    // epoch training: the gathered batch rows and the running NLL sum
    int8_t *epoch_rows = nullptr;
    int epoch_cap = 0;
    double *nll_acc = nullptr;
    // fused autoregressive sampler: the weights in its lane layout, rebuilt
    // from the bf16 weights whenever they change (ar_dirty)
    int ar_dirty = 1;
    int ar_ds = 0, ar_jm = 0;
    uint16_t *ar_w0t = nullptr, *ar_w1t = nullptr, *ar_w2t = nullptr;
    float *ar_b0l = nullptr, *ar_b1l = nullptr, *ar_b2l = nullptr;
    size_t ar_cap = 0;     // rows of the buffers below
    int8_t *ar_rows = nullptr;   // [2 cap][N] scratch rows (world-line move)
    double *ar_lq = nullptr;     // [2 cap]
};

namespace piqmc {
// Frees the fused sampler's buffers (called by piqmc_made_destroy).
void made_ar_release(piqmc_made *m);
// The fused sampler on stream st: rows [0, Bs) drawn into s_out, rows
// [Bs, Bs + Bf) evaluated from forced, log q_beta of every row into logq.
int made_ar_launch(piqmc_made *m, int Bs, int Bf, const int8_t *forced, float beta, uint64_t seed, uint32_t round,
                   int8_t *s_out, double *logq, hipStream_t st);
// The network's move buffers for R replicas: rows [2R][N], log q [2R].
int made_ar_scratch(piqmc_made *m, int R, int8_t **rows, double **lq);
}  // namespace piqmc
