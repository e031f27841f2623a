/*
 * pm_oracle.c -- CPU restatement of cvt's PatchMatch stereo hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function cites the reference code it restates (paths relative to
 * the cvt source tree).  Build: oracle/Makefile (gcc -O2 -fopenmp
 * -ffp-contract=off).  See pm_oracle.h for layouts and the arithmetic
 * contract.
 *
 * Restatement choices where the reference is implementation-defined
 * (OpenCL built-ins, FP_CONTRACT, evaluation order):
 *  - exp/sin/cos/acospi: own Cody-Waite reduction + Taylor polynomials;
 *  - nd_state_transform (a x + b y) + c: fmaf(a, x, b * y) + c, the
 *    contraction FP_CONTRACT (on by default in OpenCL C) permits;
 *  - dot() of the support weight: fmaf(z, wz, fmaf(y, wy, x * wx));
 *  - KLT J = J + g * p as one fmaf per component;
 *  - linear sampling: texel t0 + frac*(t1 - t0) as one fmaf, coordinate
 *    u = x' (the sampler's +0.5/-0.5 half-texel shift cancels);
 *  - x / 0.05f as x * 20.0f (OpenCL permits 2.5 ulp division);
 *  - census += w * cost as one fmaf;
 *  - (int2)(RNG_float, RNG_float): x draw first, then y;
 *  - view-propagation scatter: the two nearest surfaces (largest
 *    disparity) win, instead of the reference's atomic race
 *    (pmhstereo.cl:1057-1067).
 */
#include "pm_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z, w; } f4;
typedef struct { float x, y, z; } f3;

#define NORMAL_Z_MIN 1e-4f        /* pmhstereo.cl:50 */
#define NORMALCOMPMAX 0.95f       /* pmhstereo.cl:32 */
#define DEPTHREFINEMUL 2.0f       /* pmhstereo.cl:30 */
#define NORMALREFINEMUL 0.05f     /* pmhstereo.cl:31 */
#define NUMRNDSAMPLE 4            /* pmhstereo.cl:34 */
#define NUMRNDSAMPLERADIUS 5.0f   /* pmhstereo.cl:35 */
#define VIEWSAMPLES 2             /* pmhstereo.cl:46 */
#define KLT_PASSES 6              /* pmhstereo.cl:1036-1038 */
#define SLOT_EMPTY 0xFFFFFFFFu

/* ======================================================================
 * Scalar math.  Own implementations (published methods: Cody & Waite
 * argument reduction, Taylor series), so that the HIP engine can evaluate
 * the identical operation sequence.
 * ==================================================================== */

/* e^x: k = rint(x log2 e), r = x - k ln2 (two-part ln2), e^r by a degree-7
 * Taylor polynomial, result r * 2^k.  x < -87 returns 0 (no subnormals). */
float pmo_expf(float x)
{
    if (x != x) return x;
    if (x < -87.0f) return 0.0f;
    if (x > 88.5f) return INFINITY;
    float k = rintf(x * 0x1.715476p+0f);
    float r = fmaf(-k, 0x1.62e400p-1f, x);
    r = fmaf(-k, 0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;          /* 1/7! */
    p = fmaf(p, r, 0x1.6c16c2p-10f);    /* 1/6! */
    p = fmaf(p, r, 0x1.111112p-7f);     /* 1/5! */
    p = fmaf(p, r, 0x1.555556p-5f);     /* 1/4! */
    p = fmaf(p, r, 0x1.555556p-3f);     /* 1/3! */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}

/* sin and cos: k = rint(x 2/pi), r = x - k pi/2 (three-part pi/2),
 * Taylor polynomials on |r| <= pi/4, quadrant selection.  Valid for
 * |x| < 2^20 (the engine only needs [0, 2 pi]). */
void pmo_sincosf(float x, float* s, float* c)
{
    float k = rintf(x * 0x1.45f306p-1f);
    float r = fmaf(-k, 0x1.920000p+0f, x);
    r = fmaf(-k, 0x1.fb4000p-12f, r);
    r = fmaf(-k, 0x1.4442d2p-24f, r);
    float r2 = r * r;
    float ps = -0x1.ae6456p-26f;          /* -1/11! */
    ps = fmaf(ps, r2, 0x1.71de3ap-19f);   /* 1/9! */
    ps = fmaf(ps, r2, -0x1.a01a02p-13f);  /* -1/7! */
    ps = fmaf(ps, r2, 0x1.111112p-7f);    /* 1/5! */
    ps = fmaf(ps, r2, -0x1.555556p-3f);   /* -1/3! */
    float sr = fmaf(r * r2, ps, r);
    float pc = 0x1.1eed8ep-29f;           /* 1/12! */
    pc = fmaf(pc, r2, -0x1.27e4fcp-22f);  /* -1/10! */
    pc = fmaf(pc, r2, 0x1.a01a02p-16f);   /* 1/8! */
    pc = fmaf(pc, r2, -0x1.6c16c2p-10f);  /* -1/6! */
    pc = fmaf(pc, r2, 0x1.555556p-5f);    /* 1/4! */
    pc = fmaf(pc, r2, -0.5f);
    float cr = fmaf(pc, r2, 1.0f);
    int q = ((int)k) & 3;
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

/* asin on |t| <= 0.5: Taylor series through t^27. */
static float asin_small(float t)
{
    float t2 = t * t;
    float p = 0x1.782ddap-8f;
    p = fmaf(p, t2, 0x1.a6863ep-8f);
    p = fmaf(p, t2, 0x1.df3bd4p-8f);
    p = fmaf(p, t2, 0x1.12ef3cp-7f);
    p = fmaf(p, t2, 0x1.3fde50p-7f);
    p = fmaf(p, t2, 0x1.7a8788p-7f);
    p = fmaf(p, t2, 0x1.c9999ap-7f);
    p = fmaf(p, t2, 0x1.1c4ec4p-6f);
    p = fmaf(p, t2, 0x1.6e8ba2p-6f);
    p = fmaf(p, t2, 0x1.f1c71cp-6f);
    p = fmaf(p, t2, 0x1.6db6dcp-5f);
    p = fmaf(p, t2, 0x1.333334p-4f);
    p = fmaf(p, t2, 0x1.555556p-3f);
    return fmaf(t * t2, p, t);
}

/* acos(x)/pi on [-1, 1]: acos = pi/2 - asin(x) for |x| <= 1/2, otherwise
 * acos(|x|) = 2 asin(sqrt((1-|x|)/2)) and acos(-x) = pi - acos(x). */
float pmo_acospif(float x)
{
    if (x != x) return x;
    float ax = fabsf(x);
    float a;
    if (ax <= 0.5f) {
        a = (0x1.921fb6p+0f - asin_small(x)) + -0x1.777a5cp-25f;
    } else {
        float z = (1.0f - ax) * 0.5f;
        float h = 2.0f * asin_small(sqrtf(z));
        a = (x > 0.0f) ? h : (0x1.921fb6p+1f - h) + -0x1.777a5cp-24f;
    }
    return a * 0x1.45f306p-2f;
}

/* ======================================================================
 * MWC64X (cvt/cl/kernel/MWC64X.cl, third-party: D. B. Thomas 2011,
 * vendored) and RNG_init / RNG_float (cvt/cl/kernel/RNG.cl:30-35,54-59).
 * SeedStreams(0,0) puts every work-item at distance 0 of one global
 * stream (MWC64X.cl:94-97); Skip(d) multiplies the LCG-equivalent value
 * v = x*A + c by A^d mod M (MWC64X.cl:78-84).  Draw k of the stream is
 * therefore decode(BASE * A^k mod M).
 * ==================================================================== */
#define MWC_A 4294883355ULL                 /* MWC64X.cl:106 */
#define MWC_M 18446383549859758079ULL       /* MWC64X.cl:107 */
#define MWC_BASE 4077358422479273989ULL     /* MWC64X.cl:94 */

#define MWC_R 360523849793537ULL            /* 2^64 - M */

/* a*b mod M by folding the high word with 2^64 = R (mod M) */
static uint64_t mulmod_m(uint64_t a, uint64_t b)
{
    unsigned __int128 p = (unsigned __int128)a * b;
    while (p >> 64) {
        uint64_t hi = (uint64_t)(p >> 64), lo = (uint64_t)p;
        p = (unsigned __int128)hi * MWC_R + lo;
    }
    uint64_t r = (uint64_t)p;
    if (r >= MWC_M) r -= MWC_M;
    return r;
}

/* A^(2^i) mod M, i = 0..63 */
static uint64_t g_apow[64];

__attribute__((constructor)) static void init_apow(void)
{
    uint64_t sq = MWC_A;
    for (int i = 0; i < 64; i++) {
        g_apow[i] = sq;
        sq = mulmod_m(sq, sq);
    }
}

static uint64_t powmod_a(uint64_t e)
{
    uint64_t acc = 1;
    for (int i = 0; e; i++, e >>= 1)
        if (e & 1) acc = mulmod_m(acc, g_apow[i]);
    return acc;
}

typedef struct { uint32_t x, c; } mwc_t;

static mwc_t mwc_at(uint64_t k)
{
    uint64_t v = mulmod_m(MWC_BASE, powmod_a(k));
    mwc_t s;
    s.x = (uint32_t)(v / MWC_A);
    s.c = (uint32_t)(v % MWC_A);
    return s;
}

/* MWC64X_NextFloat (RNG.cl:30-35) with MWC64X_Step (MWC64X.cl:109-120) */
static float mwc_next(mwc_t* s)
{
    uint32_t res = s->x ^ s->c;
    uint32_t X = s->x, C = s->c;
    uint32_t Xn = (uint32_t)MWC_A * X + C;
    uint32_t carry = (uint32_t)(Xn < C);
    uint32_t Cn = (uint32_t)(((uint64_t)(uint32_t)MWC_A * X) >> 32) + carry;
    s->x = Xn;
    s->c = Cn;
    return 0x1.0p-32f * (float)res;
}

void pmo_mwc_state(uint64_t k, uint32_t* x, uint32_t* c)
{
    mwc_t s = mwc_at(k);
    *x = s.x;
    *c = s.c;
}

void pmo_draws(uint64_t k, int n, float* out)
{
    for (int i = 0; i < n; i++) {
        mwc_t s = mwc_at(k + (uint64_t)i);
        out[i] = mwc_next(&s);
    }
}

void pmo_draws_stepped(uint64_t k, int n, float* out)
{
    mwc_t s = mwc_at(k);
    for (int i = 0; i < n; i++) out[i] = mwc_next(&s);
}

/* First draw of a propagate at pixel p, iteration iter.  The call is
 * RNG_init(&rng, p + iter, <numdraws text>) (pmhstereo.cl:942) and the
 * macro expands to MWC64X_Skip(x, (linpos) * numdraws) (RNG.cl:54-57)
 * with the argument text unparenthesised:
 *   (p + iter) * (((2*0+1)*(2*0+1) - 1) + 2) * 3 + 2 * 4
 * = 6 (p + iter) + 8   (PROPSIZE 0, NUMRNDTRIES 2, NUMRNDSAMPLE 4,
 * pmhstereo.cl:28,32-33).  Neighbouring pixels' 14-draw windows overlap. */
uint64_t pmo_prop_draw_start(uint64_t p, int iter)
{
    return (p + (uint64_t)iter) * 6u + 8u;
}

/* ======================================================================
 * Plane algebra (pmhstereo.cl:77-224)
 * ==================================================================== */
static inline f4 mk4(float x, float y, float z, float w)
{
    f4 r = {x, y, z, w};
    return r;
}

/* nd_state_viewprop, pmhstereo.cl:174-181 */
static f4 viewprop(f4 n)
{
    f4 r = mk4(1.0f / n.x, -n.y / n.x, -n.z / n.x, 0.0f);
    if (fabsf(n.x) < 1e-8f) r = mk4(0.0f, -n.y, -n.z, 0.0f);
    return r;
}

/* nd_state_transform, pmhstereo.cl:183-186: (a x + b y) + c with a x
 * fused into the first sum */
static inline float transform(f4 s, float x, float y)
{
    return fmaf(s.x, x, s.y * y) + s.z;
}

/* (1,0,0,0) - s */
static inline f4 one_minus(f4 s)
{
    return mk4(1.0f - s.x, 0.0f - s.y, 0.0f - s.z, 0.0f - s.w);
}

/* OpenCL normalize(): v / |v|, zero vector returned unchanged */
static f3 normalize3(f3 v)
{
    float l2 = fmaf(v.z, v.z, fmaf(v.y, v.y, v.x * v.x));
    if (l2 == 0.0f) return v;
    float l = sqrtf(l2);
    f3 r = {v.x / l, v.y / l, v.z / l};
    return r;
}

/* nd_state_to_normal, pmhstereo.cl:201-210 (native_rsqrt as 1/sqrt) */
static f3 to_normal(f4 s0)
{
    f4 s = one_minus(s0);
    f3 n;
    n.z = 1.0f / sqrtf(fmaf(s.x, s.x, s.y * s.y) + 1.0f);
    n.x = -s.x * n.z;
    n.y = -s.y * n.z;
    n.z = sqrtf(fmaf(-n.y, n.y, fmaf(-n.x, n.x, 1.0f)));
    return n;
}

/* plane through (x, y) with disparity d and normal n:
 * (1,0,0,0) - (-nx/nz, -ny/nz, (nx x + ny y)/nz + d, 0) */
static f4 plane_from_nd(f3 n, float x, float y, float d)
{
    f4 t = mk4(-n.x / n.z, -n.y / n.z, fmaf(n.x, x, n.y * y) / n.z + d,
               0.0f);
    return mk4(1.0f - t.x, 0.0f - t.y, 0.0f - t.z, 0.0f - t.w);
}

/* nd_state_from_disparity_normal, pmhstereo.cl:99-106 */
static f4 from_disparity_normal(float x, float y, float d, f3 n)
{
    n.z = fmaxf(n.z, NORMAL_Z_MIN);
    n = normalize3(n);
    return plane_from_nd(n, x, y, d);
}

/* 2*M_PI*u: M_PI is a double constant in OpenCL C, so the product is
 * formed in double and rounded to float (pmhstereo.cl:82,144). */
static inline float azimuth_of(float u)
{
    return (float)(2.0 * 3.14159265358979323846 * (double)u);
}

/* nd_state_init, pmhstereo.cl:77-97 (u0, u1, u2 = three draws in order) */
static f4 state_init(float u0, float u1, float u2, float x, float y, int lr,
                     float dmax)
{
    float z = u0 * dmax;
    float az = azimuth_of(u1);
    float u = u2 * NORMALCOMPMAX;
    float w = sqrtf(fmaf(-u, u, 1.0f));
    float sn, cs;
    pmo_sincosf(az, &sn, &cs);
    f3 n = {w * cs, w * sn, u};
    n.z = fmaxf(n.z, NORMAL_Z_MIN);
    n = normalize3(n);
    f4 r = plane_from_nd(n, x, y, z);
    if (!lr) r = viewprop(r);
    return r;
}

/* nd_state_refine, pmhstereo.cl:129-172.  For the right view (!lr) the
 * reference returns viewprop of the unperturbed converted state
 * (pmhstereo.cl:169-170); reproduced. */
static f4 state_refine(float u0, float u1, float u2, f4 s0, float x, float y,
                       float dmax, int lr)
{
    f4 st = s0;
    if (!lr) st = viewprop(st);
    st = one_minus(st);
    float z = fmaf(st.x, x, st.y * y) + st.z;
    f3 n;
    n.z = 1.0f / sqrtf(fmaf(st.x, st.x, st.y * st.y) + 1.0f);
    n.x = -st.x * n.z;
    n.y = -st.y * n.z;
    z = fmaf((u0 - 0.5f) * 2.0f, DEPTHREFINEMUL, z);
    z = fminf(fmaxf(z, 0.0f), dmax);
    float az = azimuth_of(u1);
    float u = u2 * (1.0f - NORMALREFINEMUL);
    float w = sqrtf(fmaf(-u, u, 1.0f));
    float sn, cs;
    pmo_sincosf(az, &sn, &cs);
    f3 n2 = {w * cs, w * sn, u};
    n.x = (n.x + n2.x) * 0.5f;
    n.y = (n.y + n2.y) * 0.5f;
    n.z = (n.z + n2.z) * 0.5f;
    n.z = fmaxf(n.z, NORMAL_Z_MIN);
    n = normalize3(n);
    f4 r = plane_from_nd(n, x, y, z);
    if (!lr) return viewprop(st);
    return r;
}

/* nd_state_to_ref_normal_depth, pmhstereo.cl:108-127 */
static f4 to_ref_normal_depth(f4 s0, float x, float y, int lr)
{
    f4 st = s0;
    float sidesign = lr ? -1.0f : 1.0f;
    if (!lr) st = viewprop(s0);
    st = one_minus(st);
    f4 r;
    r.w = (transform(s0, x, y) - x) * sidesign;
    r.z = 1.0f / sqrtf(fmaf(st.x, st.x, st.y * st.y) + 1.0f);
    r.x = -st.x * r.z;
    r.y = -st.y * r.z;
    if (!(isfinite(r.x) && isfinite(r.y) && isfinite(r.z)))
        return mk4(0.0f, 0.0f, 1.0f, 0.0f);
    return r;
}

/* smoothDistance, pmhstereo.cl:846-871: returns dist(s) to smooth */
static float smooth_dist(f4 s, f4 smooth, float x, float y, float dmax,
                         int lr)
{
    f4 a = to_ref_normal_depth(s, x, y, lr);
    float id2 = 1.0f / (dmax * dmax);
    float den = id2 + 0.3f;
    f4 diag = mk4(0.1f / den, 0.1f / den, 0.1f / den, id2 / den);
    f4 d = mk4(a.x - smooth.x, a.y - smooth.y, a.z - smooth.z,
               a.w - smooth.w);
    return fmaf(d.w, diag.w * d.w,
                fmaf(d.z, diag.z * d.z,
                     fmaf(d.y, diag.y * d.y, d.x * (diag.x * d.x))));
}

void pmo_viewprop(const float* s, float* out)
{
    f4 r = viewprop(mk4(s[0], s[1], s[2], s[3]));
    memcpy(out, &r, sizeof r);
}

float pmo_transform(const float* s, float x, float y)
{
    return transform(mk4(s[0], s[1], s[2], s[3]), x, y);
}

void pmo_to_normal(const float* s, float* n3)
{
    f3 n = to_normal(mk4(s[0], s[1], s[2], s[3]));
    n3[0] = n.x;
    n3[1] = n.y;
    n3[2] = n.z;
}

void pmo_from_disparity_normal(float x, float y, float d, const float* n3,
                               float* out)
{
    f3 n = {n3[0], n3[1], n3[2]};
    f4 r = from_disparity_normal(x, y, d, n);
    memcpy(out, &r, sizeof r);
}

void pmo_state_init(float u0, float u1, float u2, float x, float y, int lr,
                    float dmax, float* out)
{
    f4 r = state_init(u0, u1, u2, x, y, lr, dmax);
    memcpy(out, &r, sizeof r);
}

void pmo_state_refine(float u0, float u1, float u2, const float* s, float x,
                      float y, float dmax, int lr, float* out)
{
    f4 r = state_refine(u0, u1, u2, mk4(s[0], s[1], s[2], s[3]), x, y, dmax,
                        lr);
    memcpy(out, &r, sizeof r);
}

void pmo_to_ref_normal_depth(const float* s, float x, float y, int lr,
                             float* out)
{
    f4 r = to_ref_normal_depth(mk4(s[0], s[1], s[2], s[3]), x, y, lr);
    memcpy(out, &r, sizeof r);
}

/* ======================================================================
 * Image precompute
 * ==================================================================== */

/* the built-in dot(v, (0.2126, 0.7152, 0.0722)) of the support weight as
 * a chain of fused multiply-adds */
static inline float gray3(float x, float y, float z)
{
    return fmaf(z, 0.0722f, fmaf(y, 0.7152f, x * 0.2126f));
}

/* support weight exp(-pow(|dot(d, gray)|, 1) * w), pmhstereo.cl:701 */
static inline float support_weight(float dx, float dy, float dz, float cw)
{
    return pmo_expf(-fabsf(gray3(dx, dy, dz)) * cw);
}

static inline void fetch_rgba(const float* img, int channels, int stride,
                              int x, int y, float v[4])
{
    const float* p = img + (size_t)y * stride;
    if (channels == 1) {
        v[0] = v[1] = v[2] = v[3] = p[x];      /* CL_INTENSITY */
    } else {
        v[0] = p[4 * x + 0];
        v[1] = p[4 * x + 1];
        v[2] = p[4 * x + 2];
        v[3] = p[4 * x + 3];
    }
}

/* pmhstereo_gradxy's gray value dot(rgba, (0.2126, 0.7152, 0.0722, 0))
 * with the CLK_ADDRESS_CLAMP zero border (pmhstereo.cl:1551,1571) */
static inline float gray_at(const float* img, int W, int H, int channels,
                            int stride, int x, int y)
{
    if (x < 0 || x >= W || y < 0 || y >= H) return 0.0f;
    float v[4];
    fetch_rgba(img, channels, stride, x, y, v);
    return fmaf(v[3], 0.0f, fmaf(v[2], 0.0722f,
                                 fmaf(v[1], 0.7152f, v[0] * 0.2126f)));
}

/* pmhstereo_bilateral_weight_to_alpha (pmhstereo.cl:1462-1498) and
 * pmhstereo_gradxy (pmhstereo.cl:1549-1602) into one texel record. */
void pmo_prepare(const float* img, int W, int H, int channels, int stride,
                 int radius, float cw, float* tex)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            float c[4];
            fetch_rgba(img, channels, stride, x, y, c);
            float wsum = 0.0f;
            for (int dy = -radius; dy <= radius; dy++) {
                int qy = y + dy;
                if (qy < 0 || qy >= H) continue;
                for (int dx = -radius; dx <= radius; dx++) {
                    int qx = x + dx;
                    if (qx < 0 || qx >= W) continue;
                    float q[4];
                    fetch_rgba(img, channels, stride, qx, qy, q);
                    wsum += support_weight(q[0] - c[0], q[1] - c[1],
                                           q[2] - c[2], cw);
                }
            }
            float g = gray_at(img, W, H, channels, stride, x, y);
            float gxp = gray_at(img, W, H, channels, stride, x + 1, y);
            float gxm = gray_at(img, W, H, channels, stride, x - 1, y);
            float gyp = gray_at(img, W, H, channels, stride, x, y + 1);
            float gym = gray_at(img, W, H, channels, stride, x, y - 1);
            float* t = tex + ((size_t)y * W + x) * 8;
            t[0] = c[0];
            t[1] = c[1];
            t[2] = c[2];
            t[3] = wsum;
            t[4] = gxp - g;
            t[5] = gyp - g;
            t[6] = gxp - gxm;
            t[7] = gyp - gym;
        }
    }
}

/* ======================================================================
 * Costs (patch_eval_census, pmhstereo.cl:661-736, and
 * patch_eval_color_grad_weighted, :510-565) and KLT step
 * (patch_census_klt2, pmhstereo.cl:326-410)
 * ==================================================================== */

#define COLORGRADALPHA 0.05f      /* pmhstereo.cl:39 */
#define COLORMAXDIFF 0.05f        /* pmhstereo.cl:40 */
#define GRADMAXDIFF 0.05f         /* pmhstereo.cl:41 */

/* which patch_eval the kernels use, and its support-weight factors */
typedef struct {
    float cw;  /* census: CENSUS_COLOR_WEIGHT (pmhstereo.cl:47), also the
                  bilateral weight sum's and the KLT step's */
    int kind;  /* PMO_COST_CENSUS or PMO_COST_COLOR_GRAD */
    float kw;  /* colour+gradient cost: 16 in the reference (:531) */
} cost_model;

/* CLK_FILTER_LINEAR + CLK_ADDRESS_CLAMP_TO_EDGE on integer row `row` at
 * column coordinate u (texel centres at integers).  Only r,g,b and the
 * four gradient channels are produced (out[0..2], out[4..7]). */
static inline void sample_row(const float* row, int W, float u, float out[8])
{
    u = fmaxf(fminf(u, (float)W), -1.0f);
    float fl = floorf(u);
    float a = u - fl;
    int i0 = (int)fl;
    int i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > W - 1 ? W - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > W - 1 ? W - 1 : i1);
    const float* t0 = row + (size_t)i0 * 8;
    const float* t1 = row + (size_t)i1 * 8;
    for (int k = 0; k < 8; k++) {
        if (k == 3) continue;
        out[k] = fmaf(a, t1[k] - t0[k], t0[k]);
    }
}

static inline uint32_t signmask4(const float* g)
{
    uint32_t m = 0;
    for (int k = 0; k < 4; k++) m |= (uint32_t)(signbit(g[k]) ? 1 : 0) << k;
    return m;
}

static inline int popc4(uint32_t m)
{
    return (int)(m & 1) + (int)((m >> 1) & 1) + (int)((m >> 2) & 1) +
           (int)((m >> 3) & 1);
}

/* patch_eval_color_grad_weighted, pmhstereo.cl:510-565: support weight
 * exp(-|dot(I1(p) - I1(q), gray)| * kw) (16 in the reference), tap cost
 * C = 0.05 sum_3 min(|I1(q) - I2(x')|, 0.05) + 0.95 sum_4 min(|G1(q) -
 * G2(x')|, 0.05), result sum(w C) / sum(w).  No early exit and no
 * invalid-plane test; a tap counts when its own pixel is in the image and
 * x' lies in [0, W).  No qualifying tap gives 0 / 0 = NaN, which the
 * accept rule's <= never takes.  A non-finite x' fails the range test
 * here; the reference samples the image at it, which OpenCL leaves
 * undefined. */
static float cost_cg_at(const float* tex1, const float* tex2, int W, int H,
                        int R, float kw, int x, int y, f4 s)
{
    const float xf = (float)x, yf = (float)y;
    const float* tc = tex1 + ((size_t)y * W + x) * 8;
    float ret = 0.0f, wsum = 0.0f;
    for (int dy = -R; dy <= R; dy++) {
        for (int dx = -R; dx <= R; dx++) {
            float px = xf + (float)dx, py = yf + (float)dy;
            if (px < 0.0f || px >= (float)W || py < 0.0f || py >= (float)H)
                continue;
            float xq = transform(s, px, py);
            if (!(xq >= 0.0f && xq < (float)W)) continue;
            int qx = x + dx, qy = y + dy;
            const float* t1 = tex1 + ((size_t)qy * W + qx) * 8;
            /* dot(valcenter - val1, gray): the negated difference gives
             * the negated dot bit for bit, so |.| is the same */
            float w = support_weight(t1[0] - tc[0], t1[1] - tc[1],
                                     t1[2] - tc[2], kw);
            float v2[8];
            sample_row(tex2 + (size_t)qy * W * 8, W, xq, v2);
            float sc = (fminf(fabsf(t1[0] - v2[0]), COLORMAXDIFF) +
                        fminf(fabsf(t1[1] - v2[1]), COLORMAXDIFF)) +
                       fminf(fabsf(t1[2] - v2[2]), COLORMAXDIFF);
            float sg = ((fminf(fabsf(t1[4] - v2[4]), GRADMAXDIFF) +
                         fminf(fabsf(t1[5] - v2[5]), GRADMAXDIFF)) +
                        fminf(fabsf(t1[6] - v2[6]), GRADMAXDIFF)) +
                       fminf(fabsf(t1[7] - v2[7]), GRADMAXDIFF);
            /* alpha dot(.) + (1 - alpha) dot(.): the left product fused */
            float C = fmaf(COLORGRADALPHA, sc, (1.0f - COLORGRADALPHA) * sg);
            wsum += w;
            ret = fmaf(w, C, ret);
        }
    }
    return ret / wsum;
}

static float cost_census_at(const float* tex1, const float* tex2, int W,
                            int H, int R, float cw, int x, int y, f4 s,
                            float mincost)
{
    const float xf = (float)x, yf = (float)y;
    float xp = transform(s, xf, yf);
    if (!(isfinite(s.x) && isfinite(s.y) && isfinite(s.z)) ||
        to_normal(s).z < NORMAL_Z_MIN || xp < 0.0f || xp >= (float)W)
        return 1e5f;
    const float* tc = tex1 + ((size_t)y * W + x) * 8;
    const float bilw = tc[3];
    float census = 0.0f, wsum = 0.0f;
    for (int dy = -R; dy <= R; dy++) {
        for (int dx = -R; dx <= R; dx++) {
            float px = xf + (float)dx, py = yf + (float)dy;
            float xq = transform(s, px, py);
            if (xq < 0.0f || xq >= (float)W || px < 0.0f || px >= (float)W ||
                py < 0.0f || py >= (float)H)
                continue;
            int qx = x + dx, qy = y + dy;
            const float* t1 = tex1 + ((size_t)qy * W + qx) * 8;
            float w = support_weight(t1[0] - tc[0], t1[1] - tc[1],
                                     t1[2] - tc[2], cw);
            float v2[8];
            sample_row(tex2 + (size_t)qy * W * 8, W, xq, v2);
            float c0 = fminf(fabsf(v2[0] - t1[0]), 0.05f) * 20.0f;
            float c1 = fminf(fabsf(v2[1] - t1[1]), 0.05f) * 20.0f;
            float c2 = fminf(fabsf(v2[2] - t1[2]), 0.05f) * 20.0f;
            float cad = 0.33f * ((c0 + c1) + c2);
            float cost = 0.125f * cad;
            int nd = popc4(signmask4(t1 + 4) ^ signmask4(v2 + 4));
            float cc = 0.25f * (float)nd;
            cost += (1.0f - 0.125f) * cc;
            census = fmaf(w, cost, census);
            wsum += w;
            if (census / bilw >= mincost) return 1.0f;
        }
    }
    if (wsum <= 1.0f) return 1.1f;
    return census / wsum;
}

/* the candidate cost the propagate's patch_eval selects */
static float cost_at(const float* tex1, const float* tex2, int W, int H,
                     int R, const cost_model* m, int x, int y, f4 s,
                     float mincost)
{
    if (m->kind == PMO_COST_COLOR_GRAD)
        return cost_cg_at(tex1, tex2, W, H, R, m->kw, x, y, s);
    return cost_census_at(tex1, tex2, W, H, R, m->cw, x, y, s, mincost);
}

/* one census term of patch_census_klt2 at column u; val1 is the centred
 * colour I1(q) - I1(p) (the reference compares against it,
 * pmhstereo.cl:352,364) */
static inline float klt_term(const float* row2, int W, float u,
                             const float* val1, uint32_t ga, float w1)
{
    float v2[8];
    sample_row(row2, W, u, v2);
    float c0 = fminf(fabsf(v2[0] - val1[0]), 0.05f) * 20.0f;
    float c1 = fminf(fabsf(v2[1] - val1[1]), 0.05f) * 20.0f;
    float c2 = fminf(fabsf(v2[2] - val1[2]), 0.05f) * 20.0f;
    float C = ((w1 * 0.125f) * 0.33f) * ((c0 + c1) + c2);
    int nd = popc4(ga ^ signmask4(v2 + 4));
    /* C += w1 * (1 - alpha) * 0.25 * n with 0.25 a double literal
     * (pmhstereo.cl:370): the double product is exact, one rounding. */
    return fmaf(w1 * (1.0f - 0.125f), 0.25f * (float)nd, C);
}

/* mat3f_lusolve without pivoting, cvt/cl/kernel/Matrix3.cl:286-311 */
static void lusolve3(const float Hm[9], const float b_in[3], float out[3])
{
    float m[9];
    memcpy(m, Hm, sizeof m);
    float b0 = b_in[0], b1 = b_in[1], b2 = b_in[2];
    m[3] /= m[0];
    m[6] /= m[0];
    /* x -= a * b contracts to fma(-a, b, x); the built-in dot of two
     * components is fma(y0, y1, x0 * x1) */
    m[4] = fmaf(-m[3], m[1], m[4]);
    m[5] = fmaf(-m[3], m[2], m[5]);
    m[7] = fmaf(-m[6], m[1], m[7]);
    m[8] = fmaf(-m[6], m[2], m[8]);
    m[7] /= m[4];
    m[8] = fmaf(-m[7], m[5], m[8]);
    b1 = fmaf(-m[3], b0, b1);
    b2 -= fmaf(m[7], b1, m[6] * b0);
    out[2] = b2 / m[8];
    out[1] = fmaf(-m[5], out[2], b1) / m[4];
    out[0] = (b0 - fmaf(m[2], out[2], m[1] * out[1])) / m[0];
}

static f4 klt_at(const float* tex1, const float* tex2, int W, int H, int R,
                 float cw, int x, int y, f4 s)
{
    const float xf = (float)x, yf = (float)y;
    const float* tc = tex1 + ((size_t)y * W + x) * 8;
    float J[3] = {0.0f, 0.0f, 0.0f};
    float Hm[9] = {0};
    for (int dy = -R; dy <= R; dy++) {
        for (int dx = -R; dx <= R; dx++) {
            float px = xf + (float)dx, py = yf + (float)dy;
            if (px < 0.0f || px >= (float)W || py < 0.0f || py >= (float)H)
                continue;
            int qx = x + dx, qy = y + dy;
            const float* t1 = tex1 + ((size_t)qy * W + qx) * 8;
            float val1[3] = {t1[0] - tc[0], t1[1] - tc[1], t1[2] - tc[2]};
            uint32_t ga = signmask4(t1 + 4);
            float w1 = support_weight(val1[0], val1[1], val1[2], cw);
            float xq = transform(s, px, py);
            if (px < 1.0f || px >= (float)(W - 1)) continue;
            const float* row2 = tex2 + (size_t)qy * W * 8;
            float C = klt_term(row2, W, xq, val1, ga, w1);
            float Cxp = klt_term(row2, W, xq + 1.0f, val1, ga, w1);
            float Cxm = klt_term(row2, W, xq - 1.0f, val1, ga, w1);
            float curv = (Cxp + Cxm) - 2.0f * C;
            if (curv > 0.0f) {
                float p3[3] = {px, py, 1.0f};
                float g = (w1 * (Cxm - Cxp)) * 0.5f;
                /* J = J + g * p, the last product fused
                 * (pmhstereo.cl:394) */
                for (int i = 0; i < 3; i++) J[i] = fmaf(g, p3[i], J[i]);
                float h = w1 * curv;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        Hm[3 * i + j] = Hm[3 * i + j] + h * (p3[i] * p3[j]);
            }
        }
    }
    float d[3];
    lusolve3(Hm, J, d);
    return mk4(s.x - d[0], s.y - d[1], s.z - d[2], s.w - 1e5f);
}

float pmo_cost_at(const float* tex1, const float* tex2, int W, int H,
                  int radius, float cw, int kind, float kw, int x, int y,
                  const float* s, float mincost)
{
    cost_model m = {cw, kind, kw};
    return cost_at(tex1, tex2, W, H, radius, &m, x, y,
                   mk4(s[0], s[1], s[2], s[3]), mincost);
}

void pmo_cost_field(const float* tex1, const float* tex2, int W, int H,
                    int radius, float cw, int kind, float kw,
                    const float* planes, float* cost)
{
    cost_model m = {cw, kind, kw};
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* s = planes + ((size_t)y * W + x) * 4;
            cost[(size_t)y * W + x] =
                cost_at(tex1, tex2, W, H, radius, &m, x, y,
                        mk4(s[0], s[1], s[2], s[3]), 1e10f);
        }
}

void pmo_klt_at(const float* tex1, const float* tex2, int W, int H,
                int radius, float cw, int x, int y, const float* s,
                float* out)
{
    f4 r = klt_at(tex1, tex2, W, H, radius, cw, x, y,
                  mk4(s[0], s[1], s[2], s[3]));
    memcpy(out, &r, sizeof r);
}

/* ======================================================================
 * Kernels over a whole view
 * ==================================================================== */

/* pmhstereo_init, pmhstereo.cl:755-773: RNG_init(p, 3) */
void pmo_init(const float* tex1, const float* tex2, int W, int H, int radius,
              float cw, int kind, float kw, float dmax, int lr,
              float* states)
{
    cost_model m = {cw, kind, kw};
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            mwc_t rng = mwc_at((uint64_t)p * 3u);
            float u0 = mwc_next(&rng), u1 = mwc_next(&rng),
                  u2 = mwc_next(&rng);
            f4 s = state_init(u0, u1, u2, (float)x, (float)y, lr, dmax);
            s.w = cost_at(tex1, tex2, W, H, radius, &m, x, y, s, 1e10f);
            memcpy(states + p * 4, &s, sizeof s);
        }
}

/* pmhstereo_init_disparity_normal, pmhstereo.cl:802-828: a prior
 * disparity > 0 seeds the plane (its sign flipped for the right view),
 * elsewhere the random plane of pmhstereo_init. */
void pmo_init_prior(const float* tex1, const float* tex2, int W, int H,
                    int radius, float cw, int kind, float kw, float dmax,
                    int lr, const float* disparity, const float* normals,
                    float* states)
{
    cost_model m = {cw, kind, kw};
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            mwc_t rng = mwc_at((uint64_t)p * 3u);
            float d = disparity[p];
            f4 s;
            if (d > 0.0f) {
                f3 n = {normals[p * 4 + 0], normals[p * 4 + 1],
                        normals[p * 4 + 2]};
                n = normalize3(n);
                s = from_disparity_normal((float)x, (float)y,
                                          lr ? d : -d, n);
            } else {
                float u0 = mwc_next(&rng), u1 = mwc_next(&rng),
                      u2 = mwc_next(&rng);
                s = state_init(u0, u1, u2, (float)x, (float)y, lr, dmax);
            }
            s.w = cost_at(tex1, tex2, W, H, radius, &m, x, y, s, 1e10f);
            memcpy(states + p * 4, &s, sizeof s);
        }
}

void pmo_viewbuf_clear(uint64_t* vbuf, int W, int H)
{
    for (size_t i = 0; i < (size_t)W * H; i++) vbuf[i] = ~(uint64_t)0;
}

/* keep the two smallest keys of a slot (deterministic view scatter) */
static void slot_insert(uint64_t* slot, uint32_t key)
{
    uint32_t k0 = (uint32_t)(*slot & 0xFFFFFFFFu);
    uint32_t k1 = (uint32_t)(*slot >> 32);
    if (key < k0) {
        k1 = k0;
        k0 = key;
    } else if (key < k1) {
        k1 = key;
    } else {
        return;
    }
    *slot = ((uint64_t)k1 << 32) | k0;
}

/* view scatter of a final state, pmhstereo.cl:1060-1067: slot rte(x') of
 * the other view when the integer disparity (int)((x'-x)*sidesign) > 0 */
static void scatter_one(f4 self, int x, int y, int W, int lr,
                        uint64_t* viewout)
{
    float xf = (float)x, yf = (float)y;
    float xs = transform(self, xf, yf);
    float xr = rintf(xs);
    float disp = (xs - xf) * (lr ? -1.0f : 1.0f);
    if (xr >= 0.0f && xr < (float)W && disp >= 1.0f) {
        uint32_t key = (uint32_t)(lr ? W - 1 - x : x);
        slot_insert(&viewout[(size_t)y * W + (int)xr], key);
    }
}

void pmo_scatter(const float* states, int W, int H, int lr, uint64_t* vbuf)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* s = states + ((size_t)y * W + x) * 4;
            f4 st = {s[0], s[1], s[2], s[3]};
            scatter_one(st, x, y, W, lr, vbuf);
        }
}

static inline f4 load4(const float* a, size_t p)
{
    return mk4(a[p * 4], a[p * 4 + 1], a[p * 4 + 2], a[p * 4 + 3]);
}

typedef struct {
    const float *tex1, *tex2;
    int W, H, R;
    const cost_model* m;
    float dmax, theta;
    int lr;
    float xf, yf;
    f4 smooth;
    int x, y;
} prop_ctx;

/* one candidate try, pmhstereo.cl:972-976 (minimalCost :873-877) */
static void try_candidate(const prop_ctx* c, f4* self, f4 cand)
{
    float ds = smooth_dist(*self, c->smooth, c->xf, c->yf, c->dmax, c->lr);
    float dc = smooth_dist(cand, c->smooth, c->xf, c->yf, c->dmax, c->lr);
    float mincost = fmaf(-c->theta, dc, fmaf(c->theta, ds, self->w));
    cand.w = cost_at(c->tex1, c->tex2, c->W, c->H, c->R, c->m, c->x, c->y,
                     cand, mincost);
    if (fmaf(c->theta, dc, cand.w) <= fmaf(c->theta, ds, self->w))
        *self = cand;
}

void pmo_propagate(const float* tex1, const float* tex2, int W, int H,
                   int radius, float cw, int kind, float kw, float dmax,
                   float theta, int lr, int iter, const float* old,
                   const float* smooth, const uint64_t* viewin,
                   const float* src_other, uint64_t* viewout, float* out,
                   int color, const float* fresh)
{
    const cost_model m = {cw, kind, kw};
    /* rows are independent: the scatter never leaves the row, so each
     * row's inserts run in one thread in x order */
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            if (color >= 0 && ((x + y) & 1) != color) continue;
            size_t p = (size_t)y * W + x;
            prop_ctx c = {tex1, tex2, W, H, radius, &m, dmax, theta, lr,
                          (float)x, (float)y, mk4(0, 0, 0, 0), x, y};
            f4 sm = smooth ? load4(smooth, p) : mk4(0, 0, 0, 0);
            sm.w = fminf(fmaxf(sm.w, 0.0f), 1.0f) * dmax;
            c.smooth = sm;
            f4 self = load4(old, p);
            self.w = cost_at(tex1, tex2, W, H, radius, &m, x, y, self, 1e10f);
            mwc_t rng = mwc_at(pmo_prop_draw_start(p, iter));

            /* smooth proposal, pmhstereo.cl:968-976 */
            f3 sn = {sm.x, sm.y, sm.z};
            f4 cand = from_disparity_normal(c.xf, c.yf,
                                            sm.w * (lr ? -1.0f : 1.0f),
                                            normalize3(sn));
            if (!lr) cand = viewprop(cand);
            try_candidate(&c, &self, cand);

            /* random neighbours, pmhstereo.cl:981-991 */
            for (int i = 0; i < NUMRNDSAMPLE; i++) {
                float ux = mwc_next(&rng);
                float uy = mwc_next(&rng);
                int ox = (int)fmaf((ux - 0.5f) * 2.0f, NUMRNDSAMPLERADIUS,
                                   0.5f);
                int oy = (int)fmaf((uy - 0.5f) * 2.0f, NUMRNDSAMPLERADIUS,
                                   0.5f);
                int nx = x + ox, ny = y + oy;
                nx = nx < 0 ? 0 : (nx > W - 1 ? W - 1 : nx);
                ny = ny < 0 ? 0 : (ny > H - 1 ? H - 1 : ny);
                /* red-black pass 1: colour-0 neighbours from pass 0 */
                const float* src =
                    (color == 1 && ((nx + ny) & 1) == 0) ? fresh : old;
                try_candidate(&c, &self, load4(src, (size_t)ny * W + nx));
            }

            /* random plane, pmhstereo.cl:996-1002 */
            {
                float u0 = mwc_next(&rng), u1 = mwc_next(&rng),
                      u2 = mwc_next(&rng);
                try_candidate(&c, &self,
                              state_init(u0, u1, u2, c.xf, c.yf, lr, dmax));
            }

            /* other view, pmhstereo.cl:1007-1016 */
            if (viewin) {
                uint64_t slot = viewin[p];
                uint32_t keys[2] = {(uint32_t)(slot & 0xFFFFFFFFu),
                                    (uint32_t)(slot >> 32)};
                for (int i = 0; i < VIEWSAMPLES; i++) {
                    if (keys[i] == SLOT_EMPTY) break;
                    /* the source is the other view (its lr is !lr) */
                    int xs = lr ? (int)keys[i] : W - 1 - (int)keys[i];
                    f4 src = load4(src_other, (size_t)y * W + xs);
                    try_candidate(&c, &self, viewprop(src));
                }
            }

            /* randomized refinement, pmhstereo.cl:1021-1028 */
            {
                float u0 = mwc_next(&rng), u1 = mwc_next(&rng),
                      u2 = mwc_next(&rng);
                try_candidate(&c, &self,
                              state_refine(u0, u1, u2, self, c.xf, c.yf, dmax,
                                           lr));
            }

            /* KLT step, pmhstereo.cl:1034-1053 */
            {
                f4 k = self;
                for (int it = 0; it < KLT_PASSES; it++)
                    k = klt_at(tex1, tex2, W, H, radius, cw, x, y, k);
                try_candidate(&c, &self, k);
            }

            if (viewout) scatter_one(self, x, y, W, lr, viewout);
            memcpy(out + p * 4, &self, sizeof self);
        }
    }
}

/* pmhstereo_consistency, pmhstereo.cl:1095-1129 */
void pmo_consistency(const float* left, const float* right, int W, int H,
                     float maxdispdiff, float maxdeg, float* out)
{
    const float athr = maxdeg / 180.0f;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            f4 sl = load4(left, p);
            float xf = (float)x, yf = (float)y;
            float x2 = transform(sl, xf, yf);
            f4 val = mk4(0.0f, 0.0f, 0.0f, -1.0f);
            float fl = floorf(x2), ce = ceilf(x2);
            if (ce >= 0.0f && fl < (float)W) {
                int i0 = (int)fl, i1 = (int)ce;
                i0 = i0 < 0 ? 0 : (i0 > W - 1 ? W - 1 : i0);
                i1 = i1 < 0 ? 0 : (i1 > W - 1 ? W - 1 : i1);
                f4 r1 = load4(right, (size_t)y * W + i0);
                f4 r2 = load4(right, (size_t)y * W + i1);
                f3 nl = to_normal(sl);
                f3 n1 = to_normal(viewprop(r1));
                f3 n2 = to_normal(viewprop(r2));
                float nd1 = fmaf(nl.z, n1.z, fmaf(nl.y, n1.y, nl.x * n1.x));
                float nd2 = fmaf(nl.z, n2.z, fmaf(nl.y, n2.y, nl.x * n2.x));
                float dm1 = fabsf(xf - transform(r1, x2, yf));
                float dm2 = fabsf(xf - transform(r2, x2, yf));
                float a1 = pmo_acospif(fminf(fmaxf(nd1, -1.0f), 1.0f));
                float a2 = pmo_acospif(fminf(fmaxf(nd2, -1.0f), 1.0f));
                if (!(dm1 >= maxdispdiff || a1 >= athr) ||
                    !(dm2 >= maxdispdiff || a2 >= athr))
                    val = sl;
            }
            memcpy(out + p * 4, &val, sizeof val);
        }
}

/* pmhstereo_to_disparity, pmhstereo.cl:1232-1257 */
void pmo_to_disparity(const float* states, int W, int H, float scale, int lr,
                      float* disp)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            f4 s = load4(states, p);
            float d = (transform(s, (float)x, (float)y) - (float)x) *
                      (lr ? -1.0f : 1.0f);
            float len = sqrtf(fmaf(s.z, s.z, fmaf(s.y, s.y, s.x * s.x)));
            float v;
            if (len < 1e-4f || d < 0.0f || d > (float)W)
                v = 0.0f;
            else
                v = fmaxf(d, 0.0f);
            disp[p] = v * scale;
        }
}

static inline float len3(f4 s)
{
    return sqrtf(fmaf(s.z, s.z, fmaf(s.y, s.y, s.x * s.x)));
}

/* nearest usable state left and right of x on row y (length(xyz) >= 0.1,
 * missing -> xyz = -1e5), each with w = -(x'(x, y) - x): the row scans of
 * pmhstereo_fill_depthmap / pmhstereo_fill_normalmap */
static void row_fill_pair(const float* in, int W, int x, int y, f4* lft,
                          f4* rgt)
{
    f4 l = mk4(0, 0, 0, 0), r = mk4(0, 0, 0, 0);
    for (int i = x - 1; len3(l) < 1e-1f && i >= 0; i--)
        l = load4(in, (size_t)y * W + i);
    for (int i = x + 1; len3(r) < 1e-1f && i < W; i++)
        r = load4(in, (size_t)y * W + i);
    if (len3(l) < 1e-1f) l.x = l.y = l.z = -1e5f;
    if (len3(r) < 1e-1f) r.x = r.y = r.z = -1e5f;
    float xf = (float)x, yf = (float)y;
    l.w = -(transform(l, xf, yf) - xf);
    r.w = -(transform(r, xf, yf) - xf);
    *lft = l;
    *rgt = r;
}

/* pmhstereo_fill_depthmap, pmhstereo.cl:1187-1230 (the .x channel the
 * GRAY_FLOAT output keeps) */
void pmo_fill_depthmap(const float* states, int W, int H, float scale,
                       float* disp)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            f4 s = load4(states, p);
            float v;
            if (len3(s) < 1e-4f) {
                f4 l, r;
                row_fill_pair(states, W, x, y, &l, &r);
                v = fminf(l.w, r.w);
            } else {
                v = fabsf(transform(s, (float)x, (float)y) - (float)x);
            }
            disp[p] = v * scale;
        }
}

/* pmhstereo_fill_normalmap, pmhstereo.cl:1259-1305 */
void pmo_fill_normalmap(const float* states, int W, int H, float* normals)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            f4 s = load4(states, p);
            f3 n;
            if (len3(s) < 1e-1f) {
                f4 l, r;
                row_fill_pair(states, W, x, y, &l, &r);
                n = (l.w < r.w) ? to_normal(l) : to_normal(r);
            } else {
                n = to_normal(s);
            }
            normals[p * 4 + 0] = n.x;
            normals[p * 4 + 1] = n.y;
            normals[p * 4 + 2] = n.z;
            normals[p * 4 + 3] = 0.0f;
        }
}

/* pmhstereo_disparity_to_normal, pmhstereo.cl:1604-1637: forward
 * differences with a zero border, n = normalize(cross((1,0,dx),(0,1,dy))) */
void pmo_disparity_to_normal(const float* disp, int W, int H, float* normals)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            float d = disp[p];
            float dxp = x + 1 < W ? disp[p + 1] : 0.0f;
            float dyp = y + 1 < H ? disp[p + W] : 0.0f;
            float dx = dxp - d, dy = dyp - d;
            f3 n = {0.0f * dy - dx * 1.0f, dx * 0.0f - 1.0f * dy,
                    1.0f * 1.0f - 0.0f * 0.0f};
            n = normalize3(n);
            normals[p * 4 + 0] = n.x;
            normals[p * 4 + 1] = n.y;
            normals[p * 4 + 2] = n.z;
            normals[p * 4 + 3] = 0.0f;
        }
}

/* ======================================================================
 * PM-Huber smoothing coupling
 * ==================================================================== */

/* PDDiffTensorColor / PDDiffTensorGray (PDDiffTensor.cl:25-116), beta = 1
 * so pow(len, beta) = len.  Gray value: the 4-component dot with
 * (0.2126, 0.7152, 0.0722, 0) for RGBA images, the intensity itself for
 * gray ones; CLK_ADDRESS_CLAMP reads 0 outside the image. */
static float dt_gray(const float* img, int W, int H, int channels,
                     int stride, int x, int y)
{
    if (x < 0 || x >= W || y < 0 || y >= H) return 0.0f;
    if (channels == 1) return img[(size_t)y * stride + x];
    return gray_at(img, W, H, channels, stride, x, y);
}

void pmo_diffusion_tensor(const float* img, int W, int H, int channels,
                          int stride, float alpha, float* dt)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float g = dt_gray(img, W, H, channels, stride, x, y);
            float dx = dt_gray(img, W, H, channels, stride, x + 1, y) - g;
            float dy = dt_gray(img, W, H, channels, stride, x, y + 1) - g;
            float len = sqrtf(fmaf(dx, dx, dy * dy));
            float w = pmo_expf(-alpha * len);
            float* o = dt + ((size_t)y * W + x) * 4;
            if (len <= 0.001f) {
                o[0] = 1.0f;
                o[1] = 0.0f;
                o[2] = 0.0f;
                o[3] = 1.0f;
            } else {
                dx = dx / len;
                dy = dy / len;
                float dxo = dy, dyo = -dx;
                o[0] = fmaf(w * dx, dx, dxo * dxo);
                o[1] = fmaf(w * dx, dy, dxo * dyo);
                o[2] = fmaf(w * dy, dx, dyo * dxo);
                o[3] = fmaf(w * dy, dy, dyo * dyo);
            }
        }
}

/* pmhstereo_normal_depth, pmhstereo.cl:1384-1402 */
void pmo_normal_depth(const float* states, int W, int H, float dispscale,
                      int lr, float* out)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t p = (size_t)y * W + x;
            f4 v = to_ref_normal_depth(load4(states, p), (float)x, (float)y,
                                       lr);
            v.w = v.w * dispscale;
            memcpy(out + p * 4, &v, sizeof v);
        }
}

/* float to int, truncating towards zero, with NaN -> 0 and values
 * outside the int range saturated.  OpenCL leaves those cases undefined;
 * the engine uses the same rule, so that host and device agree. */
static int f2i_trunc(float v)
{
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT_MAX;
    if (v <= -2147483648.0f) return INT_MIN;
    return (int)v;
}

/* pmhstereo_fill_state, pmhstereo.cl:1131-1185: a plane whose cost is
 * negative or whose disparity lies outside [0, W] is replaced by the
 * nearest plane to its left or right whose cost is not negative (the
 * search tests only the cost); of two, the one with the smaller positive
 * disparity at this pixel (<= 0 counts as W, ties to the left).  Then
 * non-finite a, b become 0 and the plane is converted to (normal,
 * disparity * dispscale) as pmhstereo_normal_depth does. */
void pmo_fill_state(const float* states, int W, int H, float dispscale,
                    int lr, float* out)
{
    const float sidesign = lr ? -1.0f : 1.0f;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        const size_t row = (size_t)y * W;
        for (int x = 0; x < W; x++) {
            const float xf = (float)x, yf = (float)y;
            f4 state = load4(states, row + x);
            f4 val = state;
            float disp = (transform(state, xf, yf) - xf) * sidesign;
            if (state.w < 0.0f || disp < 0.0f || disp > (float)W) {
                f4 left = mk4(-1.0f, -1.0f, -1.0f, -1.0f);
                for (int i = x - 1; left.w < 0.0f && i >= 0; i--)
                    left = load4(states, row + i);
                f4 right = mk4(-1.0f, -1.0f, -1.0f, -1.0f);
                for (int i = x + 1; right.w < 0.0f && i < W; i++)
                    right = load4(states, row + i);
                if (left.w >= 0.0f && right.w >= 0.0f) {
                    int dl = f2i_trunc(sidesign * (transform(left, xf, yf) - xf));
                    int dr = f2i_trunc(sidesign * (transform(right, xf, yf) - xf));
                    dl = dl <= 0 ? W : dl;
                    dr = dr <= 0 ? W : dr;
                    val = dl <= dr ? left : right;
                } else {
                    val = right.w < 0.0f ? left : right;
                }
            }
            if (!(isfinite(val.x) && isfinite(val.y))) {
                val.x = 0.0f;
                val.y = 0.0f;
            }
            val = to_ref_normal_depth(val, xf, yf, lr);
            val.w = val.w * dispscale;
            memcpy(out + (row + x) * 4, &val, sizeof val);
        }
    }
}

#define ROF_SIGMA 4.0f  /* PDDROFDiffTensor.cl:36-39 */
#define ROF_TAU 0.05f
#define ROF_ETA 8.0f
#define ROF_THETA 0.5f  /* PDDROFDiffTensor.cl:41 */

/* the tensor-applied projected dual at pixel (i, j) of one step
 * (PDDROFDiffTensor.cl:69-98): u read clamped to the edge, tensor and old
 * dual read as 0 outside the image.  eps >= 0 selects the Huber dual of
 * PDDHuberDiffTensor.cl:64-92 instead: the ascent step divided by
 * 1 + SIGMA eps, and one norm over all four channels (the ROF dual
 * projects xyz jointly and w on its own). */
static void pd_dual(const float* u, const float* p, const float* dt, int W,
                    int H, int i, int j, float eps, float lo[4], float hi[4])
{
    int inside = i >= 0 && i < W && j >= 0 && j < H;
    int ci = i < 0 ? 0 : (i > W - 1 ? W - 1 : i);
    int cj = j < 0 ? 0 : (j > H - 1 ? H - 1 : j);
    int ci1 = i + 1 < 0 ? 0 : (i + 1 > W - 1 ? W - 1 : i + 1);
    int cj1 = j + 1 < 0 ? 0 : (j + 1 > H - 1 ? H - 1 : j + 1);
    const float* u0 = u + ((size_t)cj * W + ci) * 4;
    const float* ux = u + ((size_t)cj * W + ci1) * 4;
    const float* uy = u + ((size_t)cj1 * W + ci) * 4;
    float D[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float plo[4] = {0.0f, 0.0f, 0.0f, 0.0f}, phi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (inside) {
        size_t q = (size_t)j * W + i;
        memcpy(D, dt + q * 4, sizeof D);
        memcpy(plo, p + q * 8, sizeof plo);
        memcpy(phi, p + q * 8 + 4, sizeof phi);
    }
    float n[4];
    const float shrink = fmaf(ROF_SIGMA, eps, 1.0f);
    for (int k = 0; k < 4; k++) {
        float dx = u0[k] - ux[k], dy = u0[k] - uy[k];
        float DTdx = fmaf(D[0], dx, D[2] * dy);
        float DTdy = fmaf(D[1], dx, D[3] * dy);
        plo[k] = fmaf(ROF_SIGMA, DTdx, plo[k]);
        phi[k] = fmaf(ROF_SIGMA, DTdy, phi[k]);
        if (eps >= 0.0f) {
            plo[k] = plo[k] / shrink;
            phi[k] = phi[k] / shrink;
        }
        n[k] = fmaf(plo[k], plo[k], phi[k] * phi[k]);
    }
    float proj[4];
    if (eps >= 0.0f) {
        float nall = fmaxf(1.0f, sqrtf(((n[0] + n[1]) + n[2]) + n[3]));
        proj[0] = proj[1] = proj[2] = proj[3] = nall;
    } else {
        float nxyz = fmaxf(1.0f, sqrtf((n[0] + n[1]) + n[2]));
        proj[0] = proj[1] = proj[2] = nxyz;
        proj[3] = fmaxf(1.0f, sqrtf(n[3]));
    }
    for (int k = 0; k < 4; k++) {
        plo[k] = plo[k] / proj[k];
        phi[k] = phi[k] / proj[k];
        lo[k] = fmaf(D[0], plo[k], D[2] * phi[k]);
        hi[k] = fmaf(D[1], plo[k], D[3] * phi[k]);
    }
}

/* PDDROFDiffTensor, PDDROFDiffTensor.cl:43-125, and with eps >= 0
 * PDDHuberDiffTensor, PDDHuberDiffTensor.cl:38-109 (the same primal) */
static void pd_step(const float* u, const float* p, const float* q,
                    const float* f, const float* dt, const float lambda[4],
                    float eps, int W, int H, float* u_out, float* p_out,
                    float* q_out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t o = (size_t)y * W + x;
            float Plo[4], Phi[4], Llo[4], Lhi[4], Ulo[4], Uhi[4];
            pd_dual(u, p, dt, W, H, x, y, eps, Plo, Phi);
            pd_dual(u, p, dt, W, H, x - 1, y, eps, Llo, Lhi);
            pd_dual(u, p, dt, W, H, x, y - 1, eps, Ulo, Uhi);
            for (int k = 0; k < 4; k++) {
                float div = ((Plo[k] - Llo[k]) + Phi[k]) - Uhi[k];
                float img = f[o * 4 + k], qq = q[o * 4 + k];
                float uo = u[o * 4 + k];
                float un = fmaf(ROF_TAU,
                                fmaf(lambda[k], img, -div) - qq, uo) /
                           fmaf(ROF_TAU, lambda[k], 1.0f);
                un = fmaf(ROF_THETA, un - uo, un);
                /* 1 + 4 ETA / lambda: a quotient, nothing to fuse */
                float qn = fmaf(ROF_ETA, un - img, qq) /
                           (1.0f + 4.0f * ROF_ETA / lambda[k]);
                qn = fmaf(ROF_THETA, qn - qq, qn);
                u_out[o * 4 + k] = un;
                q_out[o * 4 + k] = qn;
                p_out[o * 8 + k] = Plo[k];
                p_out[o * 8 + 4 + k] = Phi[k];
            }
        }
}

void pmo_rof_step(const float* u, const float* p, const float* q,
                  const float* f, const float* dt, const float lambda[4],
                  int W, int H, float* u_out, float* p_out, float* q_out)
{
    pd_step(u, p, q, f, dt, lambda, -1.0f, W, H, u_out, p_out, q_out);
}

void pmo_huber_step(const float* u, const float* p, const float* q,
                    const float* f, const float* dt, const float lambda[4],
                    float eps, int W, int H, float* u_out, float* p_out,
                    float* q_out)
{
    pd_step(u, p, q, f, dt, lambda, eps, W, H, u_out, p_out, q_out);
}

/* denoiseDiffusionTensorROF_PDD, PDOptCL.cpp:1114-1168, and with
 * eps >= 0 denoiseDiffusionTensorHuber_PDD, PDOptCL.cpp:1170-1225 (the
 * same host loop): output = input, duals 0, `iters` steps (one more when
 * the count is odd, so that the result lands in the output image: never
 * for the 40 the callers use) */
static int pd_run(const float* f, const float* dt, const float lambda[4],
                  float eps, int iters, int W, int H, float* u_out)
{
    size_t n = (size_t)W * H;
    if (iters < 1) return -1;
    float* u[2] = {(float*)malloc(n * 16), (float*)malloc(n * 16)};
    float* p[2] = {(float*)calloc(n * 8, 4), (float*)calloc(n * 8, 4)};
    float* q[2] = {(float*)calloc(n * 4, 4), (float*)calloc(n * 4, 4)};
    if (!u[0] || !u[1] || !p[0] || !p[1] || !q[0] || !q[1]) return -1;
    memcpy(u[1], f, n * 16);
    int swap = 0;
    for (int i = 0; i < iters; i++) {
        swap = i & 1;
        pd_step(u[1 - swap], p[1 - swap], q[1 - swap], f, dt, lambda, eps,
                W, H, u[swap], p[swap], q[swap]);
    }
    if (!swap)
        pd_step(u[0], p[0], q[0], f, dt, lambda, eps, W, H, u[1], p[1],
                q[1]);
    memcpy(u_out, u[1], n * 16);
    for (int k = 0; k < 2; k++) {
        free(u[k]);
        free(p[k]);
        free(q[k]);
    }
    return 0;
}

int pmo_rof(const float* f, const float* dt, const float lambda[4],
            int iters, int W, int H, float* u_out)
{
    return pd_run(f, dt, lambda, -1.0f, iters, W, H, u_out);
}

int pmo_huber(const float* f, const float* dt, const float lambda[4],
              float eps, int iters, int W, int H, float* u_out)
{
    if (!(eps >= 0.0f)) return -1;
    return pd_run(f, dt, lambda, eps, iters, W, H, u_out);
}

int pmo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Plain PatchMatch: the disparityMapAll sequence (PMHuberStereo.cpp:
 * 406-581) with theta = 0, cleared smooth field and no PD smoothing,
 * then checkConsistencyLR + convertDisparity (PMHuberStereo.cpp:129-134). */
/* one view's propagate in the run's schedule: the whole view (Jacobi) or
 * the two red-black passes into the same output and scatter buffer */
static void propagate_sched(const float* t1, const float* t2, int W, int H,
                            int radius, float cw, int kind, float kw,
                            float dmax, int lr, int it, const float* old,
                            const uint64_t* viewin, const float* src_other,
                            uint64_t* viewout, float* out, int schedule)
{
    if (schedule == PMO_SCHEDULE_RED_BLACK) {
        for (int color = 0; color < 2; color++)
            pmo_propagate(t1, t2, W, H, radius, cw, kind, kw, dmax, 0.0f, lr,
                          it, old, NULL, viewin, src_other, viewout, out,
                          color, out);
        return;
    }
    pmo_propagate(t1, t2, W, H, radius, cw, kind, kw, dmax, 0.0f, lr, it, old,
                  NULL, viewin, src_other, viewout, out, -1, NULL);
}

int pmo_run(const float* left, const float* right, int W, int H,
            int channels, int stride, int radius, float dmax, int iters,
            int kind, float kw, int schedule, float maxdispdiff,
            float maxdeg, float* stateL, float* stateR, float* disp,
            float* cost_trace)
{
    const float cw = 12.0f;
    size_t n = (size_t)W * H;
    float* texL = (float*)malloc(n * 8 * sizeof(float));
    float* texR = (float*)malloc(n * 8 * sizeof(float));
    float* sL[2] = {(float*)malloc(n * 16), (float*)malloc(n * 16)};
    float* sR[2] = {(float*)malloc(n * 16), (float*)malloc(n * 16)};
    uint64_t* vb1 = (uint64_t*)malloc(n * 8);
    uint64_t* vb2 = (uint64_t*)malloc(n * 8);
    float* chk = (float*)malloc(n * 16);
    if (!texL || !texR || !sL[0] || !sL[1] || !sR[0] || !sR[1] || !vb1 ||
        !vb2 || !chk)
        return -1;
    pmo_prepare(left, W, H, channels, stride, radius, cw, texL);
    pmo_prepare(right, W, H, channels, stride, radius, cw, texR);
    pmo_init(texL, texR, W, H, radius, cw, kind, kw,
             dmax, 1, sL[0]);
    pmo_init(texR, texL, W, H, radius, cw, kind, kw,
             dmax, 0, sR[0]);
    pmo_viewbuf_clear(vb2, W, H);
    int cur = 0;
    for (int it = 0; it < iters; it++) {
        pmo_viewbuf_clear(vb1, W, H);
        propagate_sched(texL, texR, W, H, radius, cw, kind, kw, dmax, 1, it,
                        sL[cur], vb2, sR[cur], vb1, sL[1 - cur], schedule);
        pmo_viewbuf_clear(vb2, W, H);
        propagate_sched(texR, texL, W, H, radius, cw, kind, kw, dmax, 0, it,
                        sR[cur], vb1, sL[1 - cur], vb2, sR[1 - cur],
                        schedule);
        cur = 1 - cur;
        if (cost_trace) {
            for (size_t i = 0; i < n; i++) {
                cost_trace[((size_t)it * 2 + 0) * n + i] = sL[cur][i * 4 + 3];
                cost_trace[((size_t)it * 2 + 1) * n + i] = sR[cur][i * 4 + 3];
            }
        }
    }
    if (stateL) memcpy(stateL, sL[cur], n * 16);
    if (stateR) memcpy(stateR, sR[cur], n * 16);
    if (disp) {
        pmo_consistency(sL[cur], sR[cur], W, H, maxdispdiff, maxdeg, chk);
        pmo_to_disparity(chk, W, H, 1.0f, 1, disp);
    }
    free(texL);
    free(texR);
    free(sL[0]);
    free(sL[1]);
    free(sR[0]);
    free(sR[1]);
    free(vb1);
    free(vb2);
    free(chk);
    return 0;
}

/* PM-Huber: disparityMapAll (PMHuberStereo.cpp:406-581) with its
 * constants (:416-428): alpha 24, beta 1, lambda = (10 theta + 1,
 * 10 theta + 1, 10 theta + 1, 1), theta * 50 in the accept rule, 40 ROF
 * steps; theta starts at theta_min and becomes (i / N)^4 + theta_min after
 * iteration i (:575-576, Math::pow = powf). */
int pmo_run_huber(const float* left, const float* right, int W, int H,
                  int channels, int stride, int radius, float dmax,
                  int iters, float theta_min, float maxdispdiff,
                  float maxdeg, float* stateL, float* stateR, float* disp,
                  float* smoothL, float* smoothR)
{
    const float cw = 12.0f;
    size_t n = (size_t)W * H;
    float* texL = (float*)malloc(n * 8 * sizeof(float));
    float* texR = (float*)malloc(n * 8 * sizeof(float));
    float* sL[2] = {(float*)malloc(n * 16), (float*)malloc(n * 16)};
    float* sR[2] = {(float*)malloc(n * 16), (float*)malloc(n * 16)};
    float* dtL = (float*)malloc(n * 16);
    float* dtR = (float*)malloc(n * 16);
    float* smL = (float*)calloc(n * 4, 4);
    float* smR = (float*)calloc(n * 4, 4);
    float* nd = (float*)malloc(n * 16);
    uint64_t* vb1 = (uint64_t*)malloc(n * 8);
    uint64_t* vb2 = (uint64_t*)malloc(n * 8);
    float* chk = (float*)malloc(n * 16);
    if (!texL || !texR || !sL[0] || !sL[1] || !sR[0] || !sR[1] || !dtL ||
        !dtR || !smL || !smR || !nd || !vb1 || !vb2 || !chk)
        return -1;
    pmo_prepare(left, W, H, channels, stride, radius, cw, texL);
    pmo_prepare(right, W, H, channels, stride, radius, cw, texR);
    pmo_diffusion_tensor(left, W, H, channels, stride, 24.0f, dtL);
    pmo_diffusion_tensor(right, W, H, channels, stride, 24.0f, dtR);
    pmo_init(texL, texR, W, H, radius, cw, PMO_COST_CENSUS,
             16.0f, dmax, 1, sL[0]);
    pmo_init(texR, texL, W, H, radius, cw, PMO_COST_CENSUS,
             16.0f, dmax, 0, sR[0]);
    pmo_viewbuf_clear(vb2, W, H);
    float theta = theta_min;
    int cur = 0;
    for (int it = 0; it < iters; it++) {
        const float lam = theta * 10.0f + 1.0f;
        const float lambda[4] = {lam, lam, lam, 1.0f};
        pmo_viewbuf_clear(vb1, W, H);
        pmo_propagate(texL, texR, W, H, radius, cw, PMO_COST_CENSUS,
                      16.0f, dmax, theta * 50.0f, 1,
                      it, sL[cur], smL, vb2, sR[cur], vb1, sL[1 - cur], -1, NULL);
        pmo_viewbuf_clear(vb2, W, H);
        pmo_propagate(texR, texL, W, H, radius, cw, PMO_COST_CENSUS,
                      16.0f, dmax, theta * 50.0f, 0,
                      it, sR[cur], smR, vb1, sL[1 - cur], vb2, sR[1 - cur], -1, NULL);
        pmo_normal_depth(sL[1 - cur], W, H, 1.0f / dmax, 1, nd);
        if (pmo_rof(nd, dtL, lambda, 40, W, H, smL)) return -1;
        pmo_normal_depth(sR[1 - cur], W, H, 1.0f / dmax, 0, nd);
        if (pmo_rof(nd, dtR, lambda, 40, W, H, smR)) return -1;
        theta = powf((float)it / (float)iters, 4.0f) + theta_min;
        cur = 1 - cur;
    }
    if (stateL) memcpy(stateL, sL[cur], n * 16);
    if (stateR) memcpy(stateR, sR[cur], n * 16);
    if (smoothL) memcpy(smoothL, smL, n * 16);
    if (smoothR) memcpy(smoothR, smR, n * 16);
    if (disp) {
        pmo_consistency(sL[cur], sR[cur], W, H, maxdispdiff, maxdeg, chk);
        pmo_to_disparity(chk, W, H, 1.0f, 1, disp);
    }
    free(texL);
    free(texR);
    free(sL[0]);
    free(sL[1]);
    free(sR[0]);
    free(sR[1]);
    free(dtL);
    free(dtR);
    free(smL);
    free(smR);
    free(nd);
    free(vb1);
    free(vb2);
    free(chk);
    return 0;
}

/* PMHuberStereo::depthMap (PMHuberStereo.cpp:147-347): the PM-Huber
 * coupling with its own constants (:153-165): diffusion tensor alpha 12,
 * beta 1; theta starts at 1e-3, the accept rule uses theta * 10, lambda
 * is theta * 15 + 2 on all four channels; each view's smoothing input is
 * its new field filled and converted by pmhstereo_fill_state (:273, :295)
 * and smoothed by 40 diffusion-tensor Huber steps, eps 1e-4; after
 * iteration i >= 3, theta = ((i - 3) / (N - 4))^3 + 1e-3 (:315-318).  The
 * occlusion maps it computes feed only disabled code and are left out.
 * Then the LR check and disparity conversion at scale 1. */
int pmo_run_depthmap(const float* left, const float* right, int W, int H,
                     int channels, int stride, int radius, float dmax,
                     int iters, float maxdispdiff, float maxdeg,
                     float* stateL, float* stateR, float* disp,
                     float* smoothL, float* smoothR)
{
    const float cw = 12.0f;
    size_t n = (size_t)W * H;
    float* texL = (float*)malloc(n * 8 * sizeof(float));
    float* texR = (float*)malloc(n * 8 * sizeof(float));
    float* sL[2] = {(float*)malloc(n * 16), (float*)malloc(n * 16)};
    float* sR[2] = {(float*)malloc(n * 16), (float*)malloc(n * 16)};
    float* dtL = (float*)malloc(n * 16);
    float* dtR = (float*)malloc(n * 16);
    float* smL = (float*)calloc(n * 4, 4);
    float* smR = (float*)calloc(n * 4, 4);
    float* nd = (float*)malloc(n * 16);
    uint64_t* vb1 = (uint64_t*)malloc(n * 8);
    uint64_t* vb2 = (uint64_t*)malloc(n * 8);
    float* chk = (float*)malloc(n * 16);
    if (!texL || !texR || !sL[0] || !sL[1] || !sR[0] || !sR[1] || !dtL ||
        !dtR || !smL || !smR || !nd || !vb1 || !vb2 || !chk)
        return -1;
    pmo_prepare(left, W, H, channels, stride, radius, cw, texL);
    pmo_prepare(right, W, H, channels, stride, radius, cw, texR);
    pmo_diffusion_tensor(left, W, H, channels, stride, 12.0f, dtL);
    pmo_diffusion_tensor(right, W, H, channels, stride, 12.0f, dtR);
    pmo_init(texL, texR, W, H, radius, cw, PMO_COST_CENSUS,
             16.0f, dmax, 1, sL[0]);
    pmo_init(texR, texL, W, H, radius, cw, PMO_COST_CENSUS,
             16.0f, dmax, 0, sR[0]);
    pmo_viewbuf_clear(vb2, W, H);
    float theta = 1e-3f;
    int cur = 0;
    for (int it = 0; it < iters; it++) {
        const float lam = theta * 15.0f + 2.0f;
        const float lambda[4] = {lam, lam, lam, lam};
        pmo_viewbuf_clear(vb1, W, H);
        pmo_propagate(texL, texR, W, H, radius, cw, PMO_COST_CENSUS,
                      16.0f, dmax, theta * 10.0f, 1,
                      it, sL[cur], smL, vb2, sR[cur], vb1, sL[1 - cur], -1, NULL);
        pmo_viewbuf_clear(vb2, W, H);
        pmo_propagate(texR, texL, W, H, radius, cw, PMO_COST_CENSUS,
                      16.0f, dmax, theta * 10.0f, 0,
                      it, sR[cur], smR, vb1, sL[1 - cur], vb2, sR[1 - cur], -1, NULL);
        pmo_fill_state(sL[1 - cur], W, H, 1.0f / dmax, 1, nd);
        if (pmo_huber(nd, dtL, lambda, 1e-4f, 40, W, H, smL)) return -1;
        pmo_fill_state(sR[1 - cur], W, H, 1.0f / dmax, 0, nd);
        if (pmo_huber(nd, dtR, lambda, 1e-4f, 40, W, H, smR)) return -1;
        if (it >= 3) {
            theta = ((float)it - 3.0f) / ((float)iters - 4.0f);
            theta = theta * theta * theta + 1e-3f;
        }
        cur = 1 - cur;
    }
    if (stateL) memcpy(stateL, sL[cur], n * 16);
    if (stateR) memcpy(stateR, sR[cur], n * 16);
    if (smoothL) memcpy(smoothL, smL, n * 16);
    if (smoothR) memcpy(smoothR, smR, n * 16);
    if (disp) {
        pmo_consistency(sL[cur], sR[cur], W, H, maxdispdiff, maxdeg, chk);
        pmo_to_disparity(chk, W, H, 1.0f, 1, disp);
    }
    free(texL);
    free(texR);
    free(sL[0]);
    free(sL[1]);
    free(sR[0]);
    free(sR[1]);
    free(dtL);
    free(dtR);
    free(smL);
    free(smR);
    free(nd);
    free(vb1);
    free(vb2);
    free(chk);
    return 0;
}
