/*
 * posit32_oracle.c -- scalar posit32 (es=2) arithmetic for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see posit32_oracle.h).
 *
 * add / sub / mul restate, step for step, the SoftPosit routines that the reference
 * carries as libRpack/kernel_entry.cl:
 *   - regime decoding by scanning the run of equal bits        (kernel_entry.cl:112-126)
 *   - the 64-bit fraction frames (hidden bit at 62 for add,
 *     at 60/61 for the product)                                 (:129-130, :452, :470)
 *   - alignment by a right shift that drops bits               (:151-154, :288-298)
 *   - the far-operand shortcut of the magnitude subtraction    (:292-296)
 *   - normalisation by 4-bit then 1-bit shifts                  (:302-316)
 *   - the shared rounding tail: saturation for regime > 30,
 *     fraction window ">> (reg+2)", the special cases for
 *     regime 29 and 30 and the bit-string round-to-even        (:167-213, :487-533)
 * The tail is kept exactly as the reference has it, including two properties that
 * differ from a correctly rounded result when the regime is 29 or 30 long:
 *   (1) bits of the 64-bit frame below (reg+2) are not part of the sticky bit, and
 *   (2) when the guard bit is set, the sticky bit is recomputed from the fraction
 *       window only, so for regime 30 the low exponent bit is ignored.
 * For every result whose regime is at most 28 long the routines are correctly
 * rounded (round to nearest, ties to the even bit pattern).
 *
 * div, sqrt and double->posit are not in the reference tree (they come from the
 * SoftPosit library the reference links).  They are written here as correctly rounded
 * operations: exact integer quotient / square root, then round-to-nearest-even on the
 * posit bit string, saturating to minpos/maxpos.  The tests pin them against exact
 * rational arithmetic.
 */
#include "posit32_oracle.h"

#include <math.h>
#include <string.h>

#define NAR ORC_P32_NAR

/* ------------------------------------------------------------------------- */
/* SoftPosit-style helpers                                                   */
/* ------------------------------------------------------------------------- */

/* Scan the regime of a non-negative pattern (sign bit clear) one bit at a time.
 * Returns k; *after gets the bits that follow the terminating regime bit, with the
 * two exponent bits at 30..29 and the fraction below (bit 31 is clear). */
static int sp_scan_regime(uint32_t ui, uint32_t *after)
{
    uint32_t t = ui << 2;
    int k;
    if (ui & 0x40000000u) {
        k = 0;
        while (t & 0x80000000u) {
            k++;
            t <<= 1;
        }
    } else {
        k = -1;
        while (!(t & 0x80000000u)) {
            k--;
            t <<= 1;
        }
        t &= 0x7FFFFFFFu;
    }
    *after = t;
    return k;
}

/* Fraction with its hidden bit, in the 31-bit layout SoftPosit uses (hidden at 30). */
static uint32_t sp_frac31(uint32_t after) { return ((after << 1) | 0x40000000u) & 0x7FFFFFFFu; }

/* The rounding tail shared by add, sub and mul.  `f` holds the magnitude with the
 * hidden bit at position 62.  Returns the positive pattern. */
static uint32_t sp_round_tail(int k, int e, uint64_t f)
{
    int reg;
    int reg_ones;
    if (k < 0) {
        reg = -k;
        reg_ones = 0;
    } else {
        reg = k + 1;
        reg_ones = 1;
    }
    if (reg > 30)
        return reg_ones ? 0x7FFFFFFFu : 0x1u;

    uint32_t regime = reg_ones ? (0x7FFFFFFFu - (0x7FFFFFFFu >> reg)) : (0x40000000u >> reg);
    uint64_t w = (f & 0x3FFFFFFFFFFFFFFFull) >> (reg + 2);
    uint32_t frac = (uint32_t)(w >> 32);
    uint32_t ex = (uint32_t)e;
    int guard = 0;
    int sticky = 0;
    if (reg <= 28) {
        guard = (int)((w >> 31) & 1u);
        ex <<= (28 - reg);
    } else {
        if (reg == 30) {
            guard = (int)((ex >> 1) & 1u);
            sticky = (int)(ex & 1u);
            ex = 0;
        } else { /* reg == 29 */
            guard = (int)(ex & 1u);
            ex >>= 1;
        }
        if (frac > 0) {
            frac = 0;
            sticky = 1;
        }
    }
    uint32_t ui = regime + ex + frac;
    if (guard) {
        /* the reference recomputes the sticky bit from the window here */
        sticky = (w & 0x7FFFFFFFull) != 0;
        ui += (ui & 1u) | (uint32_t)sticky;
    }
    (void)sticky;
    return ui;
}

/* |a| + |b| for operands of equal sign (kernel_entry.cl:88-217). */
static uint32_t sp_add_mags(uint32_t ua, uint32_t ub)
{
    int neg = (int)(ua >> 31);
    if (neg) {
        ua = 0u - ua;
        ub = 0u - ub;
    }
    if ((int32_t)ua < (int32_t)ub) {
        uint32_t t = ua;
        ua = ub;
        ub = t;
    }
    uint32_t ra, rb;
    int ka = sp_scan_regime(ua, &ra);
    int kb = sp_scan_regime(ub, &rb);
    int ea = (int)(ra >> 29);
    int eb = (int)(rb >> 29);
    uint64_t fa = (uint64_t)sp_frac31(ra) << 32;
    uint64_t fb = (uint64_t)sp_frac31(rb) << 32;
    int shift = (ka - kb) * 4 + ea - eb;
    fb = (shift > 63) ? 0 : (fb >> shift);
    fa += fb;
    if (fa & 0x8000000000000000ull) {
        ea++;
        if (ea > 3) {
            ka++;
            ea &= 3;
        }
        fa >>= 1;
    }
    uint32_t ui = sp_round_tail(ka, ea, fa);
    return neg ? (0u - ui) : ui;
}

/* |a| - |b| for operands of opposite sign (kernel_entry.cl:219-368). */
static uint32_t sp_sub_mags(uint32_t ua, uint32_t ub)
{
    int neg = (int)(ua >> 31);
    if (neg)
        ua = 0u - ua;
    else
        ub = 0u - ub;
    if (ua == ub)
        return 0;
    if ((int32_t)ua < (int32_t)ub) {
        uint32_t t = ua;
        ua = ub;
        ub = t;
        neg = !neg;
    }
    uint32_t ra, rb;
    int ka = sp_scan_regime(ua, &ra);
    int kb = sp_scan_regime(ub, &rb);
    int ea = (int)(ra >> 29);
    int eb = (int)(rb >> 29);
    uint64_t fa = (uint64_t)sp_frac31(ra) << 32;
    uint64_t fb = (uint64_t)sp_frac31(rb) << 32;
    int shift = (ka - kb) * 4 + ea - eb;
    if (shift > 63)
        return neg ? (0u - ua) : ua;
    fb >>= shift;
    fa -= fb;
    while ((fa >> 59) == 0) {
        ka--;
        fa <<= 4;
    }
    while (!(fa & 0x4000000000000000ull)) {
        if (ea == 0) {
            ka--;
            ea = 3;
        } else {
            ea--;
        }
        fa <<= 1;
    }
    uint32_t ui = sp_round_tail(ka, ea, fa);
    return neg ? (0u - ui) : ui;
}

uint32_t orc_p32_add(uint32_t a, uint32_t b)
{
    if (a == 0 || b == 0)
        return a | b;
    if (a == NAR || b == NAR)
        return NAR;
    if ((a ^ b) >> 31)
        return sp_sub_mags(a, b);
    return sp_add_mags(a, b);
}

uint32_t orc_p32_neg(uint32_t a) { return 0u - a; }

/* SoftPosit's p32_sub: a + (-b); negation of a posit is exact. */
uint32_t orc_p32_sub(uint32_t a, uint32_t b)
{
    if (a == NAR || b == NAR)
        return NAR;
    return orc_p32_add(a, 0u - b);
}

/* kernel_entry.cl:399-538 */
uint32_t orc_p32_mul(uint32_t a, uint32_t b)
{
    if (a == NAR || b == NAR)
        return NAR;
    if (a == 0 || b == 0)
        return 0;
    int sa = (int)(a >> 31), sb = (int)(b >> 31);
    int neg = sa ^ sb;
    if (sa)
        a = 0u - a;
    if (sb)
        b = 0u - b;
    uint32_t ra, rb;
    int k = sp_scan_regime(a, &ra) + sp_scan_regime(b, &rb);
    int e = (int)(ra >> 29) + (int)(rb >> 29);
    uint64_t p = (uint64_t)sp_frac31(ra) * (uint64_t)sp_frac31(rb);
    if (e > 3) {
        k++;
        e &= 3;
    }
    if (p >> 61) {
        e++;
        if (e > 3) {
            k++;
            e &= 3;
        }
        p >>= 1;
    }
    /* hidden bit now at 60; the shared tail expects it at 62 */
    uint32_t ui = sp_round_tail(k, e, p << 2);
    return neg ? (0u - ui) : ui;
}

/* ------------------------------------------------------------------------- */
/* Correctly rounded encoder (div, sqrt, double conversion)                   */
/* ------------------------------------------------------------------------- */

typedef unsigned __int128 u128;

/* Round sign * m * 2^(scale - hb) to posit32, where bit hb (29 <= hb <= 63) is the
 * leading one of m and `sticky` says whether nonzero bits were dropped below m.  Round
 * to nearest, ties to the even pattern, on the posit bit string; saturates to minpos /
 * maxpos and never returns 0 or NaR for a nonzero finite value. */
static uint32_t enc_round(int neg, int scale, uint64_t m, int hb, int sticky)
{
    /* k = floor(scale / 4), e = scale - 4k */
    int k = (scale >= 0) ? scale / 4 : -((-scale + 3) / 4);
    int e = scale - 4 * k;
    int reg = (k < 0) ? -k : k + 1;
    uint32_t ui;
    if (reg > 30) {
        ui = (k >= 0) ? 0x7FFFFFFFu : 0x1u;
    } else {
        /* bit string after the sign: regime (reg+1 bits), exponent (2), fraction (hb
         * bits).  Its length reg+3+hb is always above 31 because hb >= 29. */
        u128 regime = (k >= 0) ? (((u128)1 << reg) - 1) << 1 : (u128)1;
        u128 frac = (u128)(m & ((1ull << hb) - 1));
        int len = reg + 1 + 2 + hb;
        u128 s = (regime << (2 + hb)) | ((u128)(uint32_t)e << hb) | frac;
        int drop = len - 31;
        uint32_t kept = (uint32_t)(s >> drop);
        int guard = (int)((s >> (drop - 1)) & 1u);
        int rest = ((s & (((u128)1 << (drop - 1)) - 1)) != 0) || sticky;
        if (guard && (rest || (kept & 1u)))
            kept++;
        ui = kept;
    }
    return neg ? (0u - ui) : ui;
}

/* Exact decode of a finite nonzero pattern: value = (-1)^neg * m * 2^(scale - 27),
 * m has its leading one at bit 27. */
static void dec_exact(uint32_t a, int *neg, int *scale, uint32_t *m)
{
    *neg = (int)(a >> 31);
    if (*neg)
        a = 0u - a;
    uint32_t rest;
    int k = sp_scan_regime(a, &rest);
    int e = (int)(rest >> 29);
    *scale = 4 * k + e;
    *m = (1u << 27) | ((rest & 0x1FFFFFFFu) >> 2);
}

uint32_t orc_p32_div(uint32_t a, uint32_t b)
{
    if (a == NAR || b == NAR || b == 0)
        return NAR;
    if (a == 0)
        return 0;
    int na, nb_, sa, sb;
    uint32_t ma, mb;
    dec_exact(a, &na, &sa, &ma);
    dec_exact(b, &nb_, &sb, &mb);
    /* q = (ma << 36) / mb lies in (2^35, 2^37) */
    uint64_t num = (uint64_t)ma << 36;
    uint64_t q = num / mb;
    uint64_t r = num % mb;
    int hb = (q >> 36) ? 36 : 35;
    /* value = q * 2^(sa - sb - 36) = q * 2^(scale - hb) */
    int scale = sa - sb - 36 + hb;
    return enc_round(na ^ nb_, scale, q, hb, r != 0);
}

static u128 isqrt128(u128 x, int *exact)
{
    u128 res = 0;
    u128 bit = (u128)1 << 126;
    u128 v = x;
    while (bit > v)
        bit >>= 2;
    while (bit != 0) {
        if (v >= res + bit) {
            v -= res + bit;
            res = (res >> 1) + bit;
        } else {
            res >>= 1;
        }
        bit >>= 2;
    }
    *exact = (v == 0);
    return res;
}

uint32_t orc_p32_sqrt(uint32_t a)
{
    if (a == 0)
        return 0;
    if (a & 0x80000000u)
        return NAR; /* NaR and negative numbers */
    int neg, s;
    uint32_t m;
    dec_exact(a, &neg, &s, &m);
    /* value = m * 2^(s-27); make the exponent even */
    int ex = s - 27;
    uint64_t mm = m;
    if (ex & 1) {
        mm <<= 1;
        ex -= 1;
    }
    /* sqrt(mm * 2^96) = R * ... with R about 2^62 */
    int exact;
    u128 r = isqrt128((u128)mm << 96, &exact);
    uint64_t r64 = (uint64_t)r;
    int hb = 63 - __builtin_clzll(r64);
    /* value = r64 * 2^(ex/2 - 48) = r64 * 2^(scale - hb) */
    int scale = ex / 2 - 48 + hb;
    return enc_round(0, scale, r64, hb, !exact);
}

uint32_t orc_p32_from_double(double x)
{
    if (x == 0.0)
        return 0;
    if (isnan(x) || isinf(x))
        return NAR;
    uint64_t bits;
    memcpy(&bits, &x, sizeof bits);
    int neg = (int)(bits >> 63);
    int bexp = (int)((bits >> 52) & 0x7FF);
    uint64_t frac = bits & ((1ull << 52) - 1);
    if (bexp == 0) {
        /* subnormal doubles are far below minpos (2^-120) */
        return neg ? (0u - 1u) : 1u;
    }
    uint64_t m = frac | (1ull << 52);
    return enc_round(neg, bexp - 1023, m, 52, 0);
}

double orc_p32_to_double(uint32_t a)
{
    if (a == 0)
        return 0.0;
    if (a == NAR)
        return NAN;
    int neg, s;
    uint32_t m;
    dec_exact(a, &neg, &s, &m);
    double v = ldexp((double)m, s - 27);
    return neg ? -v : v;
}

int orc_p32_lt(uint32_t a, uint32_t b) { return (int32_t)a < (int32_t)b; }

/* util.cpp:12-17: negate when below zero (signed pattern order) */
uint32_t orc_p32_abs(uint32_t a) { return ((int32_t)a < 0) ? (0u - a) : a; }

/* ------------------------------------------------------------------------- */
/* bulk forms                                                                */
/* ------------------------------------------------------------------------- */

void orc_p32_add_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
        out[i] = orc_p32_add(a[i], b[i]);
}

void orc_p32_mul_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
        out[i] = orc_p32_mul(a[i], b[i]);
}

void orc_p32_sub_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
        out[i] = orc_p32_sub(a[i], b[i]);
}

void orc_p32_div_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
        out[i] = orc_p32_div(a[i], b[i]);
}

void orc_p32_sqrt_v(const uint32_t *a, uint32_t *out, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
        out[i] = orc_p32_sqrt(a[i]);
}

void orc_p32_from_double_v(const double *x, uint32_t *out, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
        out[i] = orc_p32_from_double(x[i]);
}

void orc_p32_to_double_v(const uint32_t *a, double *out, size_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++)
        out[i] = orc_p32_to_double(a[i]);
}
