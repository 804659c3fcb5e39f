/*
 * piqmc_fexp64.h -- the acceptance exponential of the "fixed" (colored /
 * parallel) semantics, shared bit-for-bit between device and host code.
 *
 * The replay (compat) mode uses the platform exp() exactly like the Cython
 * reference (piqmc/qmc.pyx:132 calls libm exp).  The fixed semantics instead
 * define their acceptance probability with this function so that a GPU run
 * and a CPU run with the same counter-based random stream take identical
 * decisions: every operation below is a single IEEE-754 binary64 operation
 * (fma, rint, ldexp, +, *), which the x86-64 host and gfx950 both round the
 * same way.
 *
 * Method (Cody & Waite argument reduction + Taylor polynomial):
 *   k = rint(x / ln 2),  r = x - k*ln2 (two-part ln2, fused),  |r| <= 0.3466
 *   exp(r) = sum_{n=0}^{13} r^n / n!   (truncation error < 5e-18 relative)
 *   exp(x) = 2^k * exp(r)
 * Accuracy is about 1 ulp over the range used for acceptance (x <= 0).
 *
 * Acceptance threshold.  A uniform u = (k + 1/2) * 2^-32 with k a 32-bit
 * counter-based draw is accepted iff p > u; p*2^32 is exact and so is
 * p*2^32 - 1/2 whenever it is >= 0, hence
 *     accept  <=>  k < piqmc_accept_y(p)   (compared in double).
 */
#ifndef PIQMC_FEXP64_H
#define PIQMC_FEXP64_H

#if defined(__HIPCC__)
#define PIQMC_HD static inline __host__ __device__
#else
#define PIQMC_HD static inline
#include <math.h>
#endif

/* ln2 split so that k*PIQMC_LN2_HI is exact for |k| < 2^11. */
#define PIQMC_LN2_HI 6.93147180369123816490e-01
#define PIQMC_LN2_LO 1.90821492927058770002e-10
#define PIQMC_INV_LN2 1.44269504088896338700e+00

PIQMC_HD double piqmc_fexp64(double x)
{
    if (x != x) return x;                 /* NaN propagates (never accepted) */
    if (x > 709.0) return __builtin_inf();  /* acceptance only uses x <= 0 */
    if (x < -746.0) return 0.0;
    double kd = rint(x * PIQMC_INV_LN2);
    double r = fma(-kd, PIQMC_LN2_HI, x);
    r = fma(-kd, PIQMC_LN2_LO, r);
    double p = 1.0 / 6227020800.0;              /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);           /* 1/12! */
    p = fma(p, r, 1.0 / 39916800.0);            /* 1/11! */
    p = fma(p, r, 1.0 / 3628800.0);             /* 1/10! */
    p = fma(p, r, 1.0 / 362880.0);              /* 1/9!  */
    p = fma(p, r, 1.0 / 40320.0);               /* 1/8!  */
    p = fma(p, r, 1.0 / 5040.0);                /* 1/7!  */
    p = fma(p, r, 1.0 / 720.0);                 /* 1/6!  */
    p = fma(p, r, 1.0 / 120.0);                 /* 1/5!  */
    p = fma(p, r, 1.0 / 24.0);                  /* 1/4!  */
    p = fma(p, r, 1.0 / 6.0);                   /* 1/3!  */
    p = fma(p, r, 0.5);                         /* 1/2!  */
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)kd);
}

/* y = p*2^32 - 1/2; the draw k (uint32) accepts iff (double)k < y. */
PIQMC_HD double piqmc_accept_y(double p)
{
    return p * 4294967296.0 - 0.5;
}

/*
 * Fixed-semantics decision class for a flip whose (reference-sign) energy
 * change is ediff = -Delta E:
 *   returns 1 -> always accepted, 0 -> never accepted,
 *   2 -> accepted iff draw < *thr   (1 <= *thr <= 2^32-1).
 */
PIQMC_HD int piqmc_fixed_class(double ediff, double temp, unsigned int *thr)
{
    if (ediff > 0.0) return 1;
    double y = piqmc_accept_y(piqmc_fexp64(ediff / temp));
    if (!(y > 0.0)) return 0;
    if (y > 4294967295.0) return 1;
    *thr = (unsigned int)ceil(y);
    return 2;
}

#endif
