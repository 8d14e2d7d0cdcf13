/*
 * pmh_oracle.c -- CPU restatement of CVT's slanted-plane PatchMatch stereo
 * path (PMHuberStereo + pmhstereo.cl).  TEST INFRASTRUCTURE ONLY: the checker
 * for the HIP path and the timed CPU baseline.  Nothing in the product loads
 * this file.  Conventions are listed in pmh_oracle.h.
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include "pmh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MWC_A 4294883355U
#define MWC_M 18446383549859758079ULL
#define MWC_BASEID 4077358422479273989ULL

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* MWC64X, restated from RNG.cl:35-156 (David Thomas' MWC64X).              */
/* ------------------------------------------------------------------------ */

/* RNG.cl:35-41 */
static uint64_t mwc_addmod(uint64_t a, uint64_t b)
{
    uint64_t v = a + b;
    if ((v >= MWC_M) || (v < a))
        v = v - MWC_M;
    return v;
}

/* RNG.cl:49-59: double-and-add modular multiply */
uint64_t orc_mwc_mulmod(uint64_t a, uint64_t b)
{
    uint64_t r = 0;
    while (a != 0) {
        if (a & 1)
            r = mwc_addmod(r, b);
        b = mwc_addmod(b, b);
        a = a >> 1;
    }
    return r;
}

/* RNG.cl:66-76: square-and-multiply */
uint64_t orc_mwc_powmod(uint64_t a, uint64_t e)
{
    uint64_t sqr = a, acc = 1;
    while (e != 0) {
        if (e & 1)
            acc = orc_mwc_mulmod(acc, sqr);
        sqr = orc_mwc_mulmod(sqr, sqr);
        e = e >> 1;
    }
    return acc;
}

/* RNG_init (RNG.cl:151-154) = MWC64X_SeedStreams(s, 0, 0) (RNG.cl:86-100,
 * 127-132) followed by MWC64X_Skip(s, linpos * numdraws) (RNG.cl:78-84,
 * 120-125).  With base 0 and gap 0 the seed distance is 0, so the seeded
 * value is MWC_BASEID * A^0 mod M. */
void orc_rng_init(uint64_t linpos, uint32_t numdraws, uint32_t st[2])
{
    uint64_t m = orc_mwc_powmod(MWC_A, 0);
    uint64_t x = orc_mwc_mulmod(MWC_BASEID, m);
    uint32_t sx = (uint32_t)(x / MWC_A), sc = (uint32_t)(x % MWC_A);
    uint64_t dist = linpos * (uint64_t)numdraws;
    m = orc_mwc_powmod(MWC_A, dist);
    x = (uint64_t)sx * (uint64_t)MWC_A + sc;
    x = orc_mwc_mulmod(x, m);
    st[0] = (uint32_t)(x / MWC_A);
    st[1] = (uint32_t)(x % MWC_A);
}

/* Independent second implementation of the same seeding, through a 128-bit
 * product and remainder; used to cross-check orc_rng_init. */
static uint64_t mulmod_u128(uint64_t a, uint64_t b)
{
    return (uint64_t)(((unsigned __int128)a * b) % MWC_M);
}

void orc_rng_init_u128(uint64_t linpos, uint32_t numdraws, uint32_t st[2])
{
    uint64_t e = linpos * (uint64_t)numdraws, p = 1, sq = MWC_A;
    while (e) {
        if (e & 1)
            p = mulmod_u128(p, sq);
        sq = mulmod_u128(sq, sq);
        e >>= 1;
    }
    uint64_t x = mulmod_u128(MWC_BASEID, p);
    st[0] = (uint32_t)(x / MWC_A);
    st[1] = (uint32_t)(x % MWC_A);
}

/* RNG.cl:108-118 */
void orc_rng_step(uint32_t st[2])
{
    uint32_t X = st[0], C = st[1];
    uint32_t Xn = MWC_A * X + C;
    uint32_t carry = (uint32_t)(Xn < C);
    uint32_t Cn = (uint32_t)(((uint64_t)MWC_A * (uint64_t)X) >> 32) + carry;
    st[0] = Xn;
    st[1] = Cn;
}

/* RNG.cl:142-147 */
float orc_rng_float(uint32_t st[2])
{
    uint32_t res = st[0] ^ st[1];
    orc_rng_step(st);
    return 2.3283064365386962890625e-10f * (float)res;
}

void orc_rng_draws(uint64_t linpos, uint32_t numdraws, int count, float* out)
{
    uint32_t st[2];
    orc_rng_init(linpos, numdraws, st);
    for (int i = 0; i < count; i++)
        out[i] = orc_rng_float(st);
}

/* ------------------------------------------------------------------------ */
/* Plane algebra (pmhstereo.cl:50-153).                                     */
/* ------------------------------------------------------------------------ */

static inline float clampf(float v, float lo, float hi)
{
    return fminf(fmaxf(v, lo), hi); /* OpenCL clamp */
}

/* pmhstereo.cl:119-124 */
void orc_state_viewprop(const float s[4], float out[4])
{
    float a = s[0], b = s[1], c = s[2];
    out[0] = 1.0f / a;
    out[1] = -b / a;
    out[2] = -c / a;
    out[3] = 0.0f;
}

/* pmhstereo.cl:126-129 */
float orc_state_transform(const float s[4], float cx, float cy)
{
    return fmaf(s[0], cx, fmaf(s[1], cy, s[2]));
}

/* pmhstereo.cl:50-68; u = the three draws (depth, nx, ny) in draw order */
void orc_state_init(const float u[3], float cx, float cy, int lr, float normmul,
                    float depthmax, float out[4])
{
    float z = u[0] * depthmax;
    float nx = ((u[1] - 0.5f) * normmul) * 0.95f;
    float ny = ((u[2] - 0.5f) * normmul) * 0.95f;
    nx = clampf(nx, -0.95f, 0.95f);
    ny = clampf(ny, -0.95f, 0.95f);
    float nfactor = fmaxf(sqrtf(nx * nx + ny * ny) + 0.001f, 1.0f);
    nx = nx / nfactor;
    ny = ny / nfactor;
    float nz = sqrtf((1.0f - nx * nx) - ny * ny);
    float r[4];
    r[0] = 1.0f - (-nx / nz);
    r[1] = 0.0f - (-ny / nz);
    r[2] = 0.0f - ((nx * cx + ny * cy) / nz + z);
    r[3] = 0.0f;
    if (!lr)
        orc_state_viewprop(r, out);
    else
        memcpy(out, r, sizeof(r));
}

static inline int finite3(const float s[4])
{
    return isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2]);
}

/* pmhstereo.cl:70-86 */
void orc_state_to_ref_normal_depth(const float s[4], float cx, float cy, int lr, float out[4])
{
    float t[4];
    if (!lr)
        orc_state_viewprop(s, t);
    else
        memcpy(t, s, sizeof(t));
    t[0] = 1.0f - t[0];
    t[1] = 0.0f - t[1];
    t[2] = 0.0f - t[2];
    t[3] = 0.0f - t[3];
    float w = (t[0] * cx + t[1] * cy) + t[2];
    float z = 1.0f / sqrtf((t[0] * t[0] + t[1] * t[1]) + 1.0f);
    float x = -t[0] * z;
    float y = -t[1] * z;
    if (!(isfinite(x) && isfinite(y) && isfinite(z))) {
        out[0] = 0.0f; out[1] = 0.0f; out[2] = 1.0f; out[3] = 0.0f;
        return;
    }
    out[0] = x; out[1] = y; out[2] = z; out[3] = w;
}

/* pmhstereo.cl:88-117, including the right-view quirk at :114-115 that
 * returns viewprop((1,0,0,0) - viewprop(s)) and drops the perturbation. */
void orc_state_refine(const float u[3], const float s[4], float cx, float cy,
                      float depthmax, int lr, float out[4])
{
    float st[4];
    if (!lr)
        orc_state_viewprop(s, st);
    else
        memcpy(st, s, sizeof(st));
    st[0] = 1.0f - st[0];
    st[1] = 0.0f - st[1];
    st[2] = 0.0f - st[2];
    st[3] = 0.0f - st[3];
    float z = (st[0] * cx + st[1] * cy) + st[2];
    float nz = 1.0f / sqrtf((st[0] * st[0] + st[1] * st[1]) + 1.0f);
    float nx = -st[0] * nz;
    float ny = -st[1] * nz;

    z += (u[0] - 0.5f) * 2.0f;
    z = clampf(z, 0.0f, depthmax);
    nx += (u[1] - 0.5f) * 0.1f;
    ny += (u[2] - 0.5f) * 0.1f;
    nx = clampf(nx, -0.95f, 0.95f);
    ny = clampf(ny, -0.95f, 0.95f);

    float nfactor = fmaxf(sqrtf(nx * nx + ny * ny) + 0.001f, 1.0f);
    nx = nx / nfactor;
    ny = ny / nfactor;
    nz = sqrtf((1.0f - nx * nx) - ny * ny);

    float r[4];
    r[0] = -nx / nz;
    r[1] = -ny / nz;
    r[2] = (nx * cx + ny * cy) / nz + z;
    r[3] = 0.0f;
    r[0] = 1.0f - r[0];
    r[1] = 0.0f - r[1];
    r[2] = 0.0f - r[2];
    r[3] = 0.0f - r[3];
    if (!lr) {
        orc_state_viewprop(st, out);
        return;
    }
    memcpy(out, r, sizeof(r));
}

/* pmhstereo.cl:144-153 */
void orc_state_to_normal(const float s[4], float out[3])
{
    float t0 = 1.0f - s[0], t1 = 0.0f - s[1];
    float z = 1.0f / sqrtf((t0 * t0 + t1 * t1) + 1.0f);
    out[0] = -t0 * z;
    out[1] = -t1 * z;
    out[2] = z;
}

/* pmhstereo.cl:266-283 */
static void smooth_distance(const float a[4], const float b[4], const float sm[4],
                            float cx, float cy, float depthmax, int lr, float out[2])
{
    float A[4], B[4];
    orc_state_to_ref_normal_depth(a, cx, cy, lr, A);
    orc_state_to_ref_normal_depth(b, cx, cy, lr, B);
    const float dw = 1.0f / (depthmax * depthmax);
    if (!finite3(a) || !finite3(b)) {
        out[0] = 0.0f;
        out[1] = 0.0f;
        return;
    }
    float d0 = A[0] - sm[0], d1 = A[1] - sm[1], d2 = A[2] - sm[2], d3 = A[3] - sm[3];
    out[0] = ((d0 * (1.0f * d0) + d1 * (1.0f * d1)) + d2 * (0.0f * d2)) + d3 * (dw * d3);
    d0 = B[0] - sm[0]; d1 = B[1] - sm[1]; d2 = B[2] - sm[2]; d3 = B[3] - sm[3];
    out[1] = ((d0 * (1.0f * d0) + d1 * (1.0f * d1)) + d2 * (0.0f * d2)) + d3 * (dw * d3);
}

/* ------------------------------------------------------------------------ */
/* Pre-processing: gradient and Huber weight images.                        */
/* ------------------------------------------------------------------------ */

/* dot(read_imagef(img), grayWeight) with the zero border of CLK_ADDRESS_CLAMP
 * (pmhstereo.cl:751,758,800,807).  nc = 1: CL_INTENSITY (I, I, I, I); nc = 4:
 * CL_RGBA (IFormat.h:302,312).  The dot product is summed left to right. */
static inline float gray_at(const float* img, int W, int H, int nc, int x, int y)
{
    if (x < 0 || y < 0 || x >= W || y >= H)
        return 0.0f;
    const float* v = img + ((size_t)y * W + x) * nc;
    if (nc == 1)
        return ((v[0] * 0.2126f + v[0] * 0.7152f) + v[0] * 0.0722f) + v[0] * 0.0f;
    return ((v[0] * 0.2126f + v[1] * 0.7152f) + v[2] * 0.0722f) + v[3] * 0.0f;
}

/* pmhstereo.cl:785-831 */
void orc_gradxy(const float* img, int W, int H, int nc, float* grad)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
#define G(dx, dy) gray_at(img, W, H, nc, x + (dx), y + (dy))
            float dx = ((G(1, 0) - G(-1, 0)) * 0.5f + (G(1, -1) - G(-1, -1)) * 0.25f) +
                       (G(1, 1) - G(-1, 1)) * 0.25f;
            float dy = ((G(0, 1) - G(0, -1)) * 0.5f + (G(-1, 1) - G(-1, -1)) * 0.25f) +
                       (G(1, 1) - G(1, -1)) * 0.25f;
            float dxy = G(1, 1) - G(-1, -1);
            float dyx = G(-1, 1) - G(1, -1);
#undef G
            float* o = grad + ((size_t)y * W + x) * 4;
            o[0] = dx; o[1] = dy; o[2] = dxy; o[3] = dyx;
        }
    }
}

/* pmhstereo.cl:737-783, including dy = g(x,y-1) - g(x,y-1) at :769 */
void orc_weight(const float* img, int W, int H, int nc, float* weight)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            float dx = gray_at(img, W, H, nc, x + 1, y) - gray_at(img, W, H, nc, x - 1, y);
            float dy = gray_at(img, W, H, nc, x, y - 1) - gray_at(img, W, H, nc, x, y - 1);
            float s = sqrtf(dx * dx + dy * dy);
            float t = (float)pow((double)s, (double)0.8f);
            float m = -3.0f * t;
            float e = (float)exp((double)m);
            weight[(size_t)y * W + x] = e + 0.0001f;
        }
    }
}

/* Per-offset support-weight factor of pmhstereo.cl:199:
 * smoothstep(0, 26, |(dx,dy)|) * 1.5 * COLORWEIGHT + 5 */
void orc_factor_table(int P, float* table)
{
    const int S = 2 * P + 1;
    for (int dy = -P; dy <= P; dy++)
        for (int dx = -P; dx <= P; dx++) {
            float fx = (float)dx * 1.0f, fy = (float)dy * 1.0f;
            float len = sqrtf(fx * fx + fy * fy);
            float t = (len - 0.0f) / (26.0f - 0.0f);
            t = clampf(t, 0.0f, 1.0f);
            float ss = t * t * (3.0f - 2.0f * t);
            table[(dy + P) * S + (dx + P)] = ss * 1.5f * 26.0f + 5.0f;
        }
}

/* ------------------------------------------------------------------------ */
/* Window cost: patch_eval_color_grad_weighted (pmhstereo.cl:170-228).       */
/* nc = 1: CL_INTENSITY inputs (the three colour channels are identical, so  */
/* the colour sums are 3 x one channel); nc = 4: RGBA, colour sums over RGB  */
/* left to right.  An RGBA image whose R = G = B equals the gray image gives */
/* the same bits: |d| + |d| + |d| and 3 |d| round the same exact value.      */
/* ------------------------------------------------------------------------ */
static float cost_core(const float* I1, const float* G1, const float* I2, const float* G2,
                       int W, int H, int nc, int x, int y, const float s[4], int P,
                       const float* F)
{
    if (!finite3(s))
        return 1e5f;
    const int S = 2 * P + 1;
    const float Wf = (float)W;
    const float* c0 = I1 + ((size_t)y * W + x) * nc;
    float wsum = 0.0f, ret = 0.0f;
    for (int dy = -P; dy <= P; dy++) {
        const int qy = y + dy;
        if (qy < 0 || qy >= H)
            continue;
        const float rowb = fmaf(s[1], (float)qy, s[2]);
        const float* i1r = I1 + (size_t)qy * W * nc;
        const float* g1r = G1 + (size_t)qy * W * 4;
        const float* i2r = I2 + (size_t)qy * W * nc;
        const float* g2r = G2 + (size_t)qy * W * 4;
        const float* Fr = F + (dy + P) * S + P;
        for (int dx = -P; dx <= P; dx++) {
            const int qx = x + dx;
            if (qx < 0 || qx >= W)
                continue;
            const float* v1 = i1r + (size_t)qx * nc;
            const float sad = nc == 1 ? 3.0f * fabsf(c0[0] - v1[0])
                                      : (fabsf(c0[0] - v1[0]) + fabsf(c0[1] - v1[1])) +
                                            fabsf(c0[2] - v1[2]);
            const float wt = expf(-sad * Fr[dx]);
            const float xp = fmaf(s[0], (float)qx, rowb);
            if (!(xp >= 0.0f && xp < Wf))
                continue;
            wsum += wt;
            const int i0 = (int)xp;
            const int i1 = (i0 + 1 < W) ? i0 + 1 : W - 1;
            const float fa = xp - (float)i0;
            float cm[3];
            for (int c = 0; c < (nc == 1 ? 1 : 3); c++) {
                const float t0 = i2r[(size_t)i0 * nc + c];
                const float v2 = fmaf(fa, i2r[(size_t)i1 * nc + c] - t0, t0);
                cm[c] = fminf(fabsf(v1[c] - v2), 0.04f);
            }
            const float cterm = nc == 1 ? 3.0f * cm[0] : (cm[0] + cm[1]) + cm[2];
            float m[4];
            for (int c = 0; c < 4; c++) {
                const float g0 = g2r[i0 * 4 + c];
                const float g2 = fmaf(fa, g2r[i1 * 4 + c] - g0, g0);
                m[c] = fminf(fabsf(g1r[qx * 4 + c] - g2), 0.01f);
            }
            const float gs = ((m[0] + m[1]) + m[2]) + m[3];
            const float C = fmaf(0.95f, gs, 0.05f * cterm);
            ret = fmaf(wt, C, ret);
        }
    }
    if (wsum <= 1.1f)
        return 1e5f;
    return ret / wsum;
}

float orc_cost(const float* I1, const float* G1, const float* I2, const float* G2,
               int W, int H, int nc, int x, int y, const float s[4], int P)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_factor_table(P, F);
    float r = cost_core(I1, G1, I2, G2, W, H, nc, x, y, s, P, F);
    free(F);
    return r;
}

void orc_cost_field(const float* I1, const float* G1, const float* I2, const float* G2,
                    int W, int H, int nc, const float* states, int P, float* cost)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_factor_table(P, F);
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            cost[(size_t)y * W + x] =
                cost_core(I1, G1, I2, G2, W, H, nc, x, y, states + ((size_t)y * W + x) * 4, P, F);
    free(F);
}

/* FP64 window cost and its exact (almost-everywhere) gradient.  Same taps,
 * truncations and bilinear cells as cost_core, evaluated in double. */
static double factor_f64(int dx, int dy)
{
    double len = sqrt((double)dx * dx + (double)dy * dy);
    double t = len / 26.0;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    return t * t * (3.0 - 2.0 * t) * 1.5 * 26.0 + 5.0;
}

static inline double sgn(double v) { return (v > 0) - (v < 0); }

/* Colour channels of the window cost: nc = 1 (CL_INTENSITY) counts its one
 * channel three times, nc = 4 (RGBA) sums R, G, B once each. */
static inline int colour_channels(int nc) { return nc == 1 ? 1 : 3; }
static inline double colour_mult(int nc) { return nc == 1 ? 3.0 : 1.0; }

static double cost_grad_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                            int W, int H, int nc, int x, int y, const double s[3], int P,
                            double gs, double dplane[3], double* dI1, double* dG1, double* dI2,
                            double* dG2);

double orc_cost_grad_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                         int W, int H, int nc, int x, int y, const double s[3], int P,
                         double dplane[3], double* dI1, double* dG1, double* dI2, double* dG2)
{
    return cost_grad_f64(I1, G1, I2, G2, W, H, nc, x, y, s, P, 1.0, dplane, dI1, dG1, dI2, dG2);
}

/* Gradient of L = sum_p gout[p] * cost_p over a plane field, in double; the
 * image gradients are accumulated (callers zero them).  dplane: 3 per pixel;
 * dI1, dI2: nc per pixel (the alpha channel of RGBA gets none). */
void orc_cost_grad_field_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                             int W, int H, int nc, const float* states, int P, const float* gout,
                             double* dplane, double* dI1, double* dG1, double* dI2, double* dG2)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* st = states + ((size_t)y * W + x) * 4;
            double s[3] = {st[0], st[1], st[2]};
            cost_grad_f64(I1, G1, I2, G2, W, H, nc, x, y, s, P, gout[(size_t)y * W + x],
                          dplane + ((size_t)y * W + x) * 3, dI1, dG1, dI2, dG2);
        }
}

static double cost_grad_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                            int W, int H, int nc, int x, int y, const double s[3], int P,
                            double gs, double dplane[3], double* dI1, double* dG1, double* dI2,
                            double* dG2)
{
    if (dplane) dplane[0] = dplane[1] = dplane[2] = 0.0;
    if (!(isfinite(s[0]) && isfinite(s[1]) && isfinite(s[2])))
        return 1e5;
    const int ncc = colour_channels(nc);
    const double mult = colour_mult(nc);
    const size_t pc = ((size_t)y * W + x) * nc;
    double wsum = 0.0, ret = 0.0;
    /* first pass: value */
    for (int dy = -P; dy <= P; dy++) {
        int qy = y + dy;
        if (qy < 0 || qy >= H) continue;
        for (int dx = -P; dx <= P; dx++) {
            int qx = x + dx;
            if (qx < 0 || qx >= W) continue;
            const size_t qc = ((size_t)qy * W + qx) * nc;
            double sad = 0.0;
            for (int c = 0; c < ncc; c++) sad += fabs((double)I1[pc + c] - (double)I1[qc + c]);
            double wt = exp(-mult * sad * factor_f64(dx, dy));
            double xp = s[0] * qx + s[1] * qy + s[2];
            if (!(xp >= 0.0 && xp < (double)W)) continue;
            int i0 = (int)xp, i1 = i0 + 1 < W ? i0 + 1 : W - 1;
            double fa = xp - i0;
            const float* i2r = I2 + (size_t)qy * W * nc;
            const float* g2r = G2 + (size_t)qy * W * 4;
            const float* g1p = G1 + ((size_t)qy * W + qx) * 4;
            double C = 0.0;
            for (int c = 0; c < ncc; c++) {
                double v2 = (1.0 - fa) * i2r[(size_t)i0 * nc + c] + fa * i2r[(size_t)i1 * nc + c];
                C += 0.05 * mult * fmin(fabs(I1[qc + c] - v2), 0.04);
            }
            for (int c = 0; c < 4; c++) {
                double g2 = (1.0 - fa) * g2r[i0 * 4 + c] + fa * g2r[i1 * 4 + c];
                C += 0.95 * fmin(fabs(g1p[c] - g2), 0.01);
            }
            wsum += wt;
            ret += wt * C;
        }
    }
    if (wsum <= 1.1)
        return 1e5;
    const double cost = ret / wsum;
    if (!dplane && !dI1 && !dG1 && !dI2 && !dG2)
        return cost;
    /* second pass: gradient */
    for (int dy = -P; dy <= P; dy++) {
        int qy = y + dy;
        if (qy < 0 || qy >= H) continue;
        for (int dx = -P; dx <= P; dx++) {
            int qx = x + dx;
            if (qx < 0 || qx >= W) continue;
            size_t qi = (size_t)qy * W + qx;
            const size_t qc = qi * nc;
            double Fq = factor_f64(dx, dy);
            double sad = 0.0;
            for (int c = 0; c < ncc; c++) sad += fabs((double)I1[pc + c] - (double)I1[qc + c]);
            double wt = exp(-mult * sad * Fq);
            double xp = s[0] * qx + s[1] * qy + s[2];
            if (!(xp >= 0.0 && xp < (double)W)) continue;
            int i0 = (int)xp, i1 = i0 + 1 < W ? i0 + 1 : W - 1;
            double fa = xp - i0;
            const float* i2r = I2 + (size_t)qy * W * nc;
            const float* g2r = G2 + (size_t)qy * W * 4;
            const float* g1p = G1 + qi * 4;
            double C = 0.0, dC_dxp = 0.0;
            double dC_dv1[3] = {0.0, 0.0, 0.0};
            /* dC/dv1 and dC/dx' */
            for (int c = 0; c < ncc; c++) {
                const double t0 = i2r[(size_t)i0 * nc + c], t1 = i2r[(size_t)i1 * nc + c];
                double dv = I1[qc + c] - ((1.0 - fa) * t0 + fa * t1);
                C += 0.05 * mult * fmin(fabs(dv), 0.04);
                dC_dv1[c] = (fabs(dv) < 0.04) ? 0.05 * mult * sgn(dv) : 0.0;
                dC_dxp += -dC_dv1[c] * (t1 - t0);
            }
            double dC_dg1[4], dC_dg2[4];
            for (int c = 0; c < 4; c++) {
                double g2 = (1.0 - fa) * g2r[i0 * 4 + c] + fa * g2r[i1 * 4 + c];
                double dg = g1p[c] - g2;
                C += 0.95 * fmin(fabs(dg), 0.01);
                dC_dg1[c] = (fabs(dg) < 0.01) ? 0.95 * sgn(dg) : 0.0;
                dC_dg2[c] = -dC_dg1[c];
                dC_dxp += dC_dg2[c] * ((double)g2r[i1 * 4 + c] - (double)g2r[i0 * 4 + c]);
            }
            /* d cost / d C_q = wt / wsum ; d cost / d wt_q = (C_q - cost) / wsum */
            double kC = gs * wt / wsum;
            double kw = gs * (C - cost) / wsum;
            if (dplane) {
                dplane[0] += kC * dC_dxp * qx;
                dplane[1] += kC * dC_dxp * qy;
                dplane[2] += kC * dC_dxp;
            }
            if (dI1)
                for (int c = 0; c < ncc; c++) {
                    /* wt = exp(-mult sum_c |c0_c - v1_c| F) */
                    double dwt_dv1 = wt * mult * Fq * sgn((double)I1[pc + c] - (double)I1[qc + c]);
                    dI1[qc + c] += kC * dC_dv1[c] + kw * dwt_dv1;
                    dI1[pc + c] += -kw * dwt_dv1;
                }
            if (dG1)
                for (int c = 0; c < 4; c++) dG1[qi * 4 + c] += kC * dC_dg1[c];
            if (dI2) {
                size_t r0 = (size_t)qy * W;
                for (int c = 0; c < ncc; c++) {
                    dI2[(r0 + i0) * nc + c] += kC * (-dC_dv1[c]) * (1.0 - fa);
                    dI2[(r0 + i1) * nc + c] += kC * (-dC_dv1[c]) * fa;
                }
            }
            if (dG2) {
                size_t r0 = (size_t)qy * W;
                for (int c = 0; c < 4; c++) {
                    dG2[(r0 + i0) * 4 + c] += kC * dC_dg2[c] * (1.0 - fa);
                    dG2[(r0 + i1) * 4 + c] += kC * dC_dg2[c] * fa;
                }
            }
        }
    }
    return cost;
}

double orc_cost_f64(const float* I1, const float* G1, const float* I2, const float* G2,
                    int W, int H, int nc, int x, int y, const double s[3], int P)
{
    return orc_cost_grad_f64(I1, G1, I2, G2, W, H, nc, x, y, s, P, NULL, NULL, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* pmhstereo_init (pmhstereo.cl:230-248)                                    */
/* ------------------------------------------------------------------------ */
void orc_init(const float* I1, const float* G1, const float* I2, const float* G2,
              int W, int H, int nc, int P, float depthmax, int lr, float* out)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_factor_table(P, F);
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            uint32_t st[2];
            float u[3], r[4];
            orc_rng_init((uint64_t)y * W + x, 3, st);
            for (int i = 0; i < 3; i++) u[i] = orc_rng_float(st);
            orc_state_init(u, (float)x, (float)y, lr, 1.0f, depthmax, r);
            r[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, r, P, F);
            memcpy(out + ((size_t)y * W + x) * 4, r, sizeof(r));
        }
    free(F);
}

/* ------------------------------------------------------------------------ */
/* pmhstereo_propagate_view (pmhstereo.cl:286-404)                          */
/* The scatter into the other view's buffer (:391-401) is done by           */
/* orc_viewbuf_build so that slot order is the sequential one (ascending x).*/
/* ------------------------------------------------------------------------ */
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static inline int accept(const float cand[4], const float self[4], const float sm[4],
                         float cx, float cy, float depthmax, int lr, float theta)
{
    float sd[2];
    smooth_distance(cand, self, sm, cx, cy, depthmax, lr, sd);
    return cand[3] + theta * sd[0] <= self[3] + theta * sd[1];
}

void orc_propagate(const float* old, float* out,
                   const float* I1, const float* G1, const float* I2, const float* G2,
                   const float* smooth, float theta, int W, int H, int nc, int P,
                   float depthmax, int lr, int iter, const orc_viewbuf* viewin,
                   const float* viewsrc, const float* draws, int sweep)
{
    orc_propagate_ex(old, out, I1, G1, I2, G2, smooth, theta, W, H, nc, P, depthmax, lr, iter,
                     viewin, viewsrc, draws, sweep, 6, 2);
}

void orc_propagate_ex(const float* old, float* out,
                      const float* I1, const float* G1, const float* I2, const float* G2,
                      const float* smooth, float theta, int W, int H, int nc, int P,
                      float depthmax, int lr, int iter, const orc_viewbuf* viewin,
                      const float* viewsrc, const float* draws, int sweep, int nbhd,
                      int nref)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_factor_table(P, F);

#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            const int parity = (x + y) & 1;
            if (sweep == ORC_SWEEP_RED && parity != 0) continue;
            if (sweep == ORC_SWEEP_BLACK && parity != 1) continue;
            const float cx = (float)x, cy = (float)y;
            /* neighbour fetch: in the black pass red pixels are already new */
#define CUR(X, Y) (((sweep == ORC_SWEEP_BLACK) && ((((X) + (Y)) & 1) == 0)) \
                       ? out + ((size_t)(Y) * W + (X)) * 4                        \
                       : old + ((size_t)(Y) * W + (X)) * 4)
            float self[4], cand[4], sm[4];
            memcpy(self, old + ((size_t)y * W + x) * 4, sizeof(self));

            /* :320-323 */
            const float* sp = smooth + ((size_t)y * W + x) * 4;
            float sx = sp[0], sy = sp[1];
            float nfactor = fmaxf(sqrtf(sx * sx + sy * sy) + 0.001f, 1.0f);
            sx = sx / nfactor;
            sy = sy / nfactor;
            sm[0] = sx;
            sm[1] = sy;
            sm[2] = sqrtf((1.0f - sx * sx) - sy * sy);
            sm[3] = sp[2] * depthmax;

            /* :327 */
            float u[ORC_PROP_DRAWS];
            if (draws) {
                memcpy(u, draws + ((size_t)y * W + x) * ORC_PROP_DRAWS, sizeof(u));
            } else {
                uint32_t st[2];
                orc_rng_init((uint64_t)y * W + x + (uint64_t)iter, 45, st);
                for (int i = 0; i < ORC_PROP_DRAWS; i++) u[i] = orc_rng_float(st);
            }

            /* :330-342 spatial neighbours */
            for (int py = -1; py <= 1; py++)
                for (int px = -1; px <= 1; px++) {
                    if (px == 0 && py == 0) continue;
                    int nxp = clampi(x + px, 0, W - 1), nyp = clampi(y + py, 0, H - 1);
                    memcpy(cand, CUR(nxp, nyp), sizeof(cand));
                    cand[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, cand, P, F);
                    if (accept(cand, self, sm, cx, cy, depthmax, lr, theta))
                        memcpy(self, cand, sizeof(self));
                }

            /* :345-352 smooth prior */
            cand[0] = 1.0f - (-sm[0] / sm[2]);
            cand[1] = 0.0f - (-sm[1] / sm[2]);
            cand[2] = 0.0f - ((sm[0] * cx + sm[1] * cy) / sm[2] + sm[3]);
            cand[3] = 0.0f;
            if (!lr) {
                float t[4];
                orc_state_viewprop(cand, t);
                memcpy(cand, t, sizeof(t));
            }
            cand[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, cand, P, F);
            if (accept(cand, self, sm, cx, cy, depthmax, lr, theta))
                memcpy(self, cand, sizeof(self));

            /* :355-361 random neighbourhood (x offset drawn first); nbhd = 0 skips it */
            for (int i = 0; i < nbhd; i++) {
                int ox = (int)(u[2 * i] * 7.0f + 0.5f);
                int oy = (int)(u[2 * i + 1] * 7.0f + 0.5f);
                int nxp = clampi(x + ox, 0, W - 1), nyp = clampi(y + oy, 0, H - 1);
                memcpy(cand, CUR(nxp, nyp), sizeof(cand));
                cand[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, cand, P, F);
                if (accept(cand, self, sm, cx, cy, depthmax, lr, theta))
                    memcpy(self, cand, sizeof(self));
            }

            /* :364-368 random plane */
            orc_state_init(u + 12, cx, cy, lr, 2.0f, depthmax, cand);
            cand[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, cand, P, F);
            if (accept(cand, self, sm, cx, cy, depthmax, lr, theta))
                memcpy(self, cand, sizeof(self));

            /* :372-379 view propagation */
            if (viewin) {
                const orc_viewbuf* vb = viewin + (size_t)y * W + x;
                int nview = vb->n < ORC_VIEWSAMPLES ? vb->n : ORC_VIEWSAMPLES;
                for (int i = 0; i < nview; i++) {
                    orc_state_viewprop(viewsrc + ((size_t)y * W + vb->src[i]) * 4, cand);
                    cand[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, cand, P, F);
                    if (accept(cand, self, sm, cx, cy, depthmax, lr, theta))
                        memcpy(self, cand, sizeof(self));
                }
            }

            /* :383-389 randomized refinement; nref < 2 stops early */
            for (int i = 0; i < nref; i++) {
                orc_state_refine(u + 15 + 3 * i, self, cx, cy, depthmax, lr, cand);
                cand[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, cand, P, F);
                if (accept(cand, self, sm, cx, cy, depthmax, lr, theta))
                    memcpy(self, cand, sizeof(self));
            }
#undef CUR
            memcpy(out + ((size_t)y * W + x) * 4, self, sizeof(self));
        }
    }
    free(F);
}

/* ------------------------------------------------------------------------ */
/* Gradient plane refinement (no counterpart in the reference: the product's  */
/* directed refinement, restated from its algorithm, not its arithmetic).    */
/* Per pixel, in the local parameters (e = x'(x, y), a, b) of the window     */
/* cost: g = d cost / d(e, a, b) from the FP64 gradient (cost_grad_f64,      */
/* pinned by central differences) with d/da = d/da' - x d/dc and d/db =      */
/* d/db' - y d/dc; a step moves x' by at most r anywhere in the window       */
/* (|de| + P (|da| + |db|) = r); the trial plane's float cost (cost_core) is  */
/* kept by the reference's acceptance rule (cost + theta x smoothness        */
/* distance, <=, pmhstereo.cl:266-283); r halves every step; an accepted     */
/* trial's gradient drives the next step.  A pixel whose plane has no        */
/* finite cost (or no gradient) keeps it.  In place: a pixel reads only its  */
/* own plane.                                                                 */
/* ------------------------------------------------------------------------ */
void orc_grad_refine(float* states, const float* I1, const float* G1, const float* I2,
                     const float* G2, const float* smooth, int W, int H, int nc, int P,
                     float depthmax, int lr, float theta, int steps, float radius)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_factor_table(P, F);
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float* st = states + ((size_t)y * W + x) * 4;
            const float cx = (float)x, cy = (float)y;
            float self[4], sm[4];
            memcpy(self, st, sizeof(self));
            /* the smooth prior of pmhstereo.cl:320-323, as in orc_propagate_ex */
            const float* sp = smooth + ((size_t)y * W + x) * 4;
            float sx = sp[0], sy = sp[1];
            const float nfactor = fmaxf(sqrtf(sx * sx + sy * sy) + 0.001f, 1.0f);
            sx = sx / nfactor;
            sy = sy / nfactor;
            sm[0] = sx;
            sm[1] = sy;
            sm[2] = sqrtf((1.0f - sx * sx) - sy * sy);
            sm[3] = sp[2] * depthmax;
            double g[3];
            double s64[3] = {self[0], self[1], self[2]};
            int ok = finite3(self) &&
                     cost_core(I1, G1, I2, G2, W, H, nc, x, y, self, P, F) < 1e5f;
            if (ok)
                cost_grad_f64(I1, G1, I2, G2, W, H, nc, x, y, s64, P, 1.0, g, NULL, NULL, NULL,
                              NULL);
            float r = radius;
            for (int k = 0; k < steps && ok; k++, r *= 0.5f) {
                const double ge = g[2], ga = g[0] - (double)x * g[2], gb = g[1] - (double)y * g[2];
                const double norm = fabs(ge) + (double)P * (fabs(ga) + fabs(gb));
                if (!(norm > 0.0))
                    break;
                const double m = (double)r / norm;
                const double da = -m * ga, db = -m * gb, de = -m * ge;
                float trial[4];
                trial[0] = (float)((double)self[0] + da);
                trial[1] = (float)((double)self[1] + db);
                trial[2] = (float)((double)self[2] + de - da * (double)x - db * (double)y);
                trial[3] = cost_core(I1, G1, I2, G2, W, H, nc, x, y, trial, P, F);
                if (!accept(trial, self, sm, cx, cy, depthmax, lr, theta))
                    continue;
                memcpy(self, trial, sizeof(self));
                if (k == steps - 1)
                    break;
                ok = trial[3] < 1e5f;
                s64[0] = trial[0];
                s64[1] = trial[1];
                s64[2] = trial[2];
                if (ok)
                    cost_grad_f64(I1, G1, I2, G2, W, H, nc, x, y, s64, P, 1.0, g, NULL, NULL,
                                  NULL, NULL);
            }
            memcpy(st, self, sizeof(self));
        }
    free(F);
}

/* Scatter of :391-401 in sequential work-item order: for each target pixel the
 * count of sources and the first four sources by ascending x. */
void orc_viewbuf_build(const float* states, int W, int H, orc_viewbuf* out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++) {
        orc_viewbuf* row = out + (size_t)y * W;
        for (int x = 0; x < W; x++) row[x].n = 0;
        for (int x = 0; x < W; x++) {
            float xp = orc_state_transform(states + ((size_t)y * W + x) * 4, (float)x, (float)y);
            float t = xp + 0.5f;
            if (!(t > -1.0f && t < (float)W)) continue; /* (int)t in [0, W) */
            int disp = (int)t;
            int nold = row[disp].n++;
            if (nold < ORC_VIEWSAMPLES) row[disp].src[nold] = x;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* pmhstereo_consistency (pmhstereo.cl:428-464)                             */
/* ------------------------------------------------------------------------ */
static inline float acospi_f(float v)
{
    return (float)(acos((double)v) / 3.14159265358979323846);
}

static inline float dot3(const float a[3], const float b[3])
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* The L/R test of pmhstereo_consistency (:428-464) and pmhstereo_occmap
 * (:682-712) at (x, y): 1 if the plane passes. */
static int lr_check(const float* left, const float* right, int W, int x, int y,
                    float maxdispdiff, float maxdeg, int lr)
{
    const float* sl = left + ((size_t)y * W + x) * 4;
    float xp = orc_state_transform(sl, (float)x, (float)y);
    float sr[4];
    if (!(xp >= 0.0f && xp < (float)W)) {
        sr[0] = sr[1] = sr[2] = sr[3] = 1e5f;
    } else {
        int xr = (int)roundf(xp);
        xr = clampi(xr, 0, W - 1);
        memcpy(sr, right + ((size_t)y * W + xr) * 4, sizeof(sr));
    }
    float n1[3], n2[3], vp[4];
    float ndiff;
    if (lr) {
        orc_state_to_normal(sl, n1);
        orc_state_viewprop(sr, vp);
        orc_state_to_normal(vp, n2);
        ndiff = dot3(n1, n2);
    } else {
        orc_state_to_normal(sr, n1);
        orc_state_viewprop(sl, vp);
        orc_state_to_normal(vp, n2);
        ndiff = dot3(n1, n2);
    }
    float dmax = fabsf((float)x - orc_state_transform(sr, xp, (float)y));
    return !(dmax >= maxdispdiff || acospi_f(ndiff) >= maxdeg / 180.0f);
}

void orc_consistency(const float* left, const float* right, int W, int H,
                     float maxdispdiff, float maxdeg, int lr, float* out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float* o = out + ((size_t)y * W + x) * 4;
            if (lr_check(left, right, W, x, y, maxdispdiff, maxdeg, lr))
                memcpy(o, left + ((size_t)y * W + x) * 4, sizeof(float) * 4);
            else
                o[0] = o[1] = o[2] = o[3] = 0.0f;
        }
}

/* ------------------------------------------------------------------------ */
/* pmhstereo_fill_state (pmhstereo.cl:466-510)                              */
/* ------------------------------------------------------------------------ */
static inline float len4(const float* v)
{
    return sqrtf(((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3]);
}

static inline float len3(const float* v)
{
    return sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}

void orc_fill_state(const float* in, int W, int H, float depthmax, int lr, float* out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* st = in + ((size_t)y * W + x) * 4;
            float val[4];
            if (len4(st) < 1e-1f) {
                float l[4] = {0, 0, 0, 0}, r[4] = {0, 0, 0, 0};
                int xx = x - 1;
                while (len4(l) < 1e-1f && xx >= 0) {
                    memcpy(l, in + ((size_t)y * W + xx) * 4, sizeof(l));
                    xx--;
                }
                xx = x + 1;
                while (len4(r) < 1e-1f && xx < W) {
                    memcpy(r, in + ((size_t)y * W + xx) * 4, sizeof(r));
                    xx++;
                }
                if (len4(l) < 1e-1f) l[0] = l[1] = l[2] = l[3] = 1e5f;
                if (len4(r) < 1e-1f) r[0] = r[1] = r[2] = r[3] = 1e5f;
                l[3] = fabsf(orc_state_transform(l, (float)x, (float)y) - (float)x);
                r[3] = fabsf(orc_state_transform(r, (float)x, (float)y) - (float)x);
                if (l[3] < r[3]) memcpy(val, l, sizeof(val));
                else memcpy(val, r, sizeof(val));
            } else {
                memcpy(val, st, sizeof(val));
            }
            float nd[4];
            orc_state_to_ref_normal_depth(val, (float)x, (float)y, lr, nd);
            float* o = out + ((size_t)y * W + x) * 4;
            o[0] = nd[0];
            o[1] = nd[1];
            o[2] = nd[3] / depthmax;
            o[3] = 1.0f;
        }
}

/* ------------------------------------------------------------------------ */
/* Weighted Huber-ROF primal-dual step (PDHuberWeighted.cl:34-104) and      */
/* PDROF::apply (PDROF.h:50-96).  Channel 3 of the primal is forced to 1 by */
/* the kernel (:99) and channel 3 of the dual feeds only that channel, so   */
/* the dual is kept for channels 0..2 (pout holds 8 floats per pixel:       */
/* lo.xyzw, hi.xyzw, with w left 0).                                        */
/* ------------------------------------------------------------------------ */
#define PD_SIGMA (1.0f / sqrtf(8.0f))
#define PD_TAU (1.0f / sqrtf(8.0f))
#define PD_THETA 0.5f
#define PD_ALPHA 0.001f

/* dual update of :59-81 at (gx, gy), gx in [-1, W-1], gy in [-1, H-1].
 * perchannel = 0: the projection of PDHuberWeighted.cl:75-76 (x and y
 * channels jointly, z alone); 1: that of PDHuberWeightedInpaint.cl:76 (every
 * channel alone). */
static void pd_dual_p(const float* last, const float* plast, const float* w, int W, int H,
                      int gx, int gy, int perchannel, float lo[3], float hi[3])
{
    const float sigma = PD_SIGMA;
    int inside = gx >= 0 && gy >= 0 && gx < W && gy < H;
    float weight = inside ? w[(size_t)gy * W + gx] : 0.0f;
    int cx0 = clampi(gx, 0, W - 1), cy0 = clampi(gy, 0, H - 1);
    int cx1 = clampi(gx + 1, 0, W - 1), cy1 = clampi(gy + 1, 0, H - 1);
    const float* uc = last + ((size_t)cy0 * W + cx0) * 4;
    const float* ux = last + ((size_t)cy0 * W + cx1) * 4;
    const float* uy = last + ((size_t)cy1 * W + cx0) * 4;
    float den = 1.0f + sigma * PD_ALPHA / weight;
    float ptmp[3];
    for (int c = 0; c < 3; c++) {
        float dx = weight * (uc[c] - ux[c]);
        float dy = weight * (uc[c] - uy[c]);
        float pl = inside ? plast[((size_t)gy * W + gx) * 8 + c] : 0.0f;
        float ph = inside ? plast[((size_t)gy * W + gx) * 8 + 4 + c] : 0.0f;
        lo[c] = (pl + sigma * dx) / den;
        hi[c] = (ph + sigma * dy) / den;
        ptmp[c] = lo[c] * lo[c] + hi[c] * hi[c];
    }
    if (perchannel) {
        for (int c = 0; c < 3; c++) {
            float pc = fmaxf(1.0f, sqrtf(ptmp[c]));
            lo[c] /= pc;
            hi[c] /= pc;
        }
        return;
    }
    float pxy = fmaxf(1.0f, sqrtf(ptmp[0] + ptmp[1]));
    float pz = fmaxf(1.0f, sqrtf(ptmp[2]));
    lo[0] /= pxy; hi[0] /= pxy;
    lo[1] /= pxy; hi[1] /= pxy;
    lo[2] /= pz;  hi[2] /= pz;
}

static void pd_dual(const float* last, const float* plast, const float* w, int W, int H,
                    int gx, int gy, float lo[3], float hi[3])
{
    pd_dual_p(last, plast, w, W, H, gx, gy, 0, lo, hi);
}

void orc_pd_step(const float* last, const float* plast, const float* f, const float* w,
                 int W, int H, float lambda, float* out, float* pout)
{
    const float tau = PD_TAU;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float lo[3], hi[3], lol[3], hil[3], lou[3], hiu[3];
            pd_dual(last, plast, w, W, H, x, y, lo, hi);
            pd_dual(last, plast, w, W, H, x - 1, y, lol, hil);
            pd_dual(last, plast, w, W, H, x, y - 1, lou, hiu);
            const size_t i = (size_t)y * W + x;
            const float wt = w[i];
            float* o = out + i * 4;
            for (int c = 0; c < 3; c++) {
                float div = ((lo[c] - lol[c]) + hi[c]) - hiu[c];
                float u = last[i * 4 + c];
                float un = (u + tau * (lambda * f[i * 4 + c] - wt * div)) / (1.0f + tau * lambda);
                un += PD_THETA * (un - u);
                o[c] = un;
                pout[i * 8 + c] = lo[c];
                pout[i * 8 + 4 + c] = hi[c];
            }
            o[3] = 1.0f;
            pout[i * 8 + 3] = 0.0f;
            pout[i * 8 + 7] = 0.0f;
        }
}

/* PDROF.h:50-96: u is the warm start on entry and the result on exit. */
void orc_pdrof(float* u, const float* f, const float* w, int W, int H, float lambda, int iters)
{
    size_t n = (size_t)W * H;
    float* tmp = (float*)malloc(sizeof(float) * 4 * n);
    float* p0 = (float*)calloc(8 * n, sizeof(float));
    float* p1 = (float*)calloc(8 * n, sizeof(float));
    float* img[2] = {tmp, u};
    float* p[2] = {p0, p1};
    int swap = 0;
    for (int i = 0; i < iters; i++) {
        swap = i & 1;
        orc_pd_step(img[1 - swap], p[1 - swap], f, w, W, H, lambda, img[swap], p[swap]);
    }
    if (!swap) /* PDROF.h:83-95 */
        orc_pd_step(img[0], p[0], f, w, W, H, lambda, img[1], p[1]);
    free(tmp);
    free(p0);
    free(p1);
}


/* One PDHuberWeightedInpaint step (PDHuberWeightedInpaint.cl:34-104): the
 * dual is projected channel by channel, and the data term's lambda is scaled
 * by the mask's first channel (0 where the planes failed the L/R check, so
 * those pixels are filled by the smoother alone). */
void orc_pd_step_inpaint(const float* last, const float* plast, const float* f, const float* w,
                         const float* mask, int W, int H, float lambda, float* out, float* pout)
{
    const float tau = PD_TAU;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float lo[3], hi[3], lol[3], hil[3], lou[3], hiu[3];
            pd_dual_p(last, plast, w, W, H, x, y, 1, lo, hi);
            pd_dual_p(last, plast, w, W, H, x - 1, y, 1, lol, hil);
            pd_dual_p(last, plast, w, W, H, x, y - 1, 1, lou, hiu);
            const size_t i = (size_t)y * W + x;
            const float wt = w[i];
            const float lm = lambda * mask[i * 4];
            float* o = out + i * 4;
            for (int c = 0; c < 3; c++) {
                float div = wt * (((lo[c] - lol[c]) + hi[c]) - hiu[c]);
                float u = last[i * 4 + c];
                float un = (u + tau * (lm * f[i * 4 + c] - div)) / (1.0f + tau * lm);
                un += PD_THETA * (un - u);
                o[c] = un;
                pout[i * 8 + c] = lo[c];
                pout[i * 8 + 4 + c] = hi[c];
            }
            o[3] = 1.0f;
            pout[i * 8 + 3] = 0.0f;
            pout[i * 8 + 7] = 0.0f;
        }
}

/* PDROFInpaint::apply (PDROFInpaint.h:50-98), the same ping-pong as PDROF. */
void orc_pdrof_inpaint(float* u, const float* f, const float* w, const float* mask, int W, int H,
                       float lambda, int iters)
{
    size_t n = (size_t)W * H;
    float* tmp = (float*)malloc(sizeof(float) * 4 * n);
    float* p0 = (float*)calloc(8 * n, sizeof(float));
    float* p1 = (float*)calloc(8 * n, sizeof(float));
    float* img[2] = {tmp, u};
    float* p[2] = {p0, p1};
    int swap = 0;
    for (int i = 0; i < iters; i++) {
        swap = i & 1;
        orc_pd_step_inpaint(img[1 - swap], p[1 - swap], f, w, mask, W, H, lambda, img[swap],
                            p[swap]);
    }
    if (!swap) /* PDROFInpaint.h:83-96 */
        orc_pd_step_inpaint(img[0], p[0], f, w, mask, W, H, lambda, img[1], p[1]);
    free(tmp);
    free(p0);
    free(p1);
}

/* pmhstereo_occmap (pmhstereo.cl:682-712): 1 where the L/R check of
 * pmhstereo_consistency passes (|x - x''| < maxdiff, normals within maxdeg;
 * the kernel hard-codes 5 degrees), else 0; written as (v, v, v, 1). */
void orc_occmap(const float* left, const float* right, int W, int H, float maxdiff,
                float maxdeg, int lr, float* out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float v = lr_check(left, right, W, x, y, maxdiff, maxdeg, lr) ? 1.0f : 0.0f;
            float* o = out + ((size_t)y * W + x) * 4;
            o[0] = v; o[1] = v; o[2] = v; o[3] = 1.0f;
        }
}

/* pmhstereo_normal_depth (pmhstereo.cl:663-680): (nx, ny, d / depthmax, 1) of
 * every pixel's plane, without filling. */
void orc_normal_depth(const float* in, int W, int H, float depthmax, int lr, float* out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            float nd[4];
            orc_state_to_ref_normal_depth(in + i * 4, (float)x, (float)y, lr, nd);
            out[i * 4 + 0] = nd[0];
            out[i * 4 + 1] = nd[1];
            out[i * 4 + 2] = nd[3] / depthmax;
            out[i * 4 + 3] = 1.0f;
        }
}

/* ------------------------------------------------------------------------ */
/* pmhstereo_fill_depthmap / fill_normalmap (pmhstereo.cl:512-603)          */
/* ------------------------------------------------------------------------ */
static void fill_pick(const float* in, int W, int y, int x, float l[4], float r[4])
{
    memset(l, 0, 4 * sizeof(float));
    memset(r, 0, 4 * sizeof(float));
    int xx = x - 1;
    while (len3(l) < 1e-1f && xx >= 0) {
        memcpy(l, in + ((size_t)y * W + xx) * 4, 4 * sizeof(float));
        xx--;
    }
    xx = x + 1;
    while (len3(r) < 1e-1f && xx < W) {
        memcpy(r, in + ((size_t)y * W + xx) * 4, 4 * sizeof(float));
        xx++;
    }
    if (len3(l) < 1e-1f) l[0] = l[1] = l[2] = -1e5f;
    if (len3(r) < 1e-1f) r[0] = r[1] = r[2] = -1e5f;
    l[3] = -(orc_state_transform(l, (float)x, (float)y) - (float)x);
    r[3] = -(orc_state_transform(r, (float)x, (float)y) - (float)x);
}

void orc_fill_depthmap(const float* in, int W, int H, float scale, float* dmap)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* st = in + ((size_t)y * W + x) * 4;
            float v;
            if (len3(st) < 1e-1f) {
                float l[4], r[4];
                fill_pick(in, W, y, x, l, r);
                v = fminf(l[3], r[3]);
            } else {
                v = fabsf(orc_state_transform(st, (float)x, (float)y) - (float)x);
            }
            dmap[(size_t)y * W + x] = v * scale;
        }
}

void orc_fill_normalmap(const float* in, int W, int H, float* nmap)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* st = in + ((size_t)y * W + x) * 4;
            float n[3];
            if (len3(st) < 1e-1f) {
                float l[4], r[4];
                fill_pick(in, W, y, x, l, r);
                if (l[3] < r[3]) orc_state_to_normal(l, n);
                else orc_state_to_normal(r, n);
            } else {
                orc_state_to_normal(st, n);
            }
            float* o = nmap + ((size_t)y * W + x) * 4;
            o[0] = n[0]; o[1] = n[1]; o[2] = n[2]; o[3] = 0.0f;
        }
}

/* ------------------------------------------------------------------------ */
/* PMHuberStereo::depthMap (PMHuberStereo.cpp:64-291)                       */
/* ------------------------------------------------------------------------ */
static float smoothstep01(float x) /* Math.h:88-92 */
{
    float t = (x < 0.0f) ? 0.0f : (x > 1.0f) ? 1.0f : x;
    return t * t * (3.0f - 2.0f * t);
}

void orc_depthmap(const float* left, const float* right, int W, int H, int nc, int P,
                  float depthmax,
                  int iterations, float dscale, const orc_opts* opts,
                  float* dmap, float* normalmap, float* lstate_out, float* rstate_out)
{
    const size_t n = (size_t)W * H;
    const float maxdispdiff = opts ? opts->maxdispdiff : 0.5f;
    const float maxangle = opts ? opts->maxanglediff : 10.0f;
    const int pd_iters = opts ? opts->pd_iters : 250;
    const int rb = opts ? opts->sweep : 0;
    const int nbhd = opts ? opts->random_neighbours : 6;
    const int nref = opts ? opts->random_refine : 2;
    const float thetascale = 50.0f;
    float theta = 0.0f;
    if (dscale <= 0.0f)
        dscale = 1.0f / depthmax;

    float* lg = (float*)malloc(sizeof(float) * 4 * n);
    float* rg = (float*)malloc(sizeof(float) * 4 * n);
    float* lw = (float*)malloc(sizeof(float) * n);
    float* rw = (float*)malloc(sizeof(float) * n);
    float* ls = (float*)calloc(4 * n, sizeof(float));
    float* rs = (float*)calloc(4 * n, sizeof(float));
    float* tmp = (float*)malloc(sizeof(float) * 4 * n);
    float* tmp2 = (float*)malloc(sizeof(float) * 4 * n);
    float* L[2] = {(float*)malloc(sizeof(float) * 4 * n), (float*)malloc(sizeof(float) * 4 * n)};
    float* R[2] = {(float*)malloc(sizeof(float) * 4 * n), (float*)malloc(sizeof(float) * 4 * n)};
    orc_viewbuf* vb1 = (orc_viewbuf*)calloc(n, sizeof(orc_viewbuf));
    orc_viewbuf* vb2 = (orc_viewbuf*)calloc(n, sizeof(orc_viewbuf));

    orc_gradxy(left, W, H, nc, lg);
    orc_gradxy(right, W, H, nc, rg);
    orc_weight(left, W, H, nc, lw);
    orc_weight(right, W, H, nc, rw);
    orc_init(left, lg, right, rg, W, H, nc, P, depthmax, 1, L[0]);
    orc_init(right, rg, left, lg, W, H, nc, P, depthmax, 0, R[0]);
    const float* vb2src = R[0]; /* vb2 is empty at iteration 0 */

    for (int iter = 0; iter < iterations; iter++) {
        int swap = iter & 1;
        if (rb) {
            orc_propagate_ex(L[swap], L[1 - swap], left, lg, right, rg, ls, theta, W, H, nc, P,
                          depthmax, 1, iter, vb2, vb2src, NULL, ORC_SWEEP_RED, nbhd, nref);
            orc_propagate_ex(L[swap], L[1 - swap], left, lg, right, rg, ls, theta, W, H, nc, P,
                          depthmax, 1, iter, vb2, vb2src, NULL, ORC_SWEEP_BLACK, nbhd, nref);
        } else {
            orc_propagate_ex(L[swap], L[1 - swap], left, lg, right, rg, ls, theta, W, H, nc, P,
                          depthmax, 1, iter, vb2, vb2src, NULL, ORC_SWEEP_ALL, nbhd, nref);
        }
        orc_viewbuf_build(L[1 - swap], W, H, vb1);
        if (rb) {
            orc_propagate_ex(R[swap], R[1 - swap], right, rg, left, lg, rs, theta, W, H, nc, P,
                          depthmax, 0, iter, vb1, L[1 - swap], NULL, ORC_SWEEP_RED, nbhd, nref);
            orc_propagate_ex(R[swap], R[1 - swap], right, rg, left, lg, rs, theta, W, H, nc, P,
                          depthmax, 0, iter, vb1, L[1 - swap], NULL, ORC_SWEEP_BLACK, nbhd, nref);
        } else {
            orc_propagate_ex(R[swap], R[1 - swap], right, rg, left, lg, rs, theta, W, H, nc, P,
                          depthmax, 0, iter, vb1, L[1 - swap], NULL, ORC_SWEEP_ALL, nbhd, nref);
        }
        orc_viewbuf_build(R[1 - swap], W, H, vb2);
        vb2src = R[1 - swap];

        orc_consistency(L[1 - swap], R[1 - swap], W, H, maxdispdiff, maxangle, 1, tmp);
        orc_fill_state(tmp, W, H, depthmax, 1, tmp2);
        orc_pdrof(ls, tmp2, lw, W, H, theta * thetascale + 5.0f, pd_iters);

        orc_consistency(R[1 - swap], L[1 - swap], W, H, maxdispdiff, maxangle, 0, tmp);
        orc_fill_state(tmp, W, H, depthmax, 0, tmp2);
        orc_pdrof(rs, tmp2, rw, W, H, theta * thetascale + 5.0f, pd_iters);

        if (iter >= 5)
            theta = smoothstep01(((float)iter - 5.0f) / ((float)iterations - 5.0f)) * 1.0f;
    }

    int fin = 1 - ((iterations - 1) & 1);
    if (iterations <= 0) fin = 0;
    orc_consistency(L[fin], R[fin], W, H, maxdispdiff, maxangle, 1, tmp);
    orc_fill_depthmap(tmp, W, H, dscale, dmap);
    if (normalmap)
        orc_fill_normalmap(tmp, W, H, normalmap);
    if (lstate_out) memcpy(lstate_out, L[fin], sizeof(float) * 4 * n);
    if (rstate_out) memcpy(rstate_out, R[fin], sizeof(float) * 4 * n);

    free(lg); free(rg); free(lw); free(rw); free(ls); free(rs); free(tmp); free(tmp2);
    free(L[0]); free(L[1]); free(R[0]); free(R[1]); free(vb1); free(vb2);
}

/* ---- PMHuberStereo::depthMapInpaint (PMHuberStereo.cpp:293-507, without the
 * debug block :376-409 that saves PNGs and waits on getchar) ----
 * opts: maxdispdiff / maxanglediff are the occlusion map's tolerances (the
 * reference: 0.5 px, 5 degrees), pd_iters the inpainting smoother's steps
 * (100), sweep the propagation schedule.  The final check uses the plane
 * fields of index 1 with 1 px / 5 degrees and writes disparities in px. */
void orc_depthmap_inpaint(const float* left, const float* right, int W, int H, int nc, int P,
                          float depthmax, int iterations, const orc_opts* opts, float* dmap,
                          float* lstate_out, float* rstate_out)
{
    const size_t n = (size_t)W * H;
    const float occdiff = opts ? opts->maxdispdiff : 0.5f;
    const float occdeg = opts ? opts->maxanglediff : 5.0f;
    const int pd_iters = opts ? opts->pd_iters : 100;
    const int rb = opts ? opts->sweep : 0;
    const int nbhd = opts ? opts->random_neighbours : 6;
    const int nref = opts ? opts->random_refine : 2;
    float theta = 0.0f;

    float* lg = (float*)malloc(sizeof(float) * 4 * n);
    float* rg = (float*)malloc(sizeof(float) * 4 * n);
    float* lw = (float*)malloc(sizeof(float) * n);
    float* rw = (float*)malloc(sizeof(float) * n);
    float* ls = (float*)calloc(4 * n, sizeof(float));
    float* rs = (float*)calloc(4 * n, sizeof(float));
    float* tmp = (float*)malloc(sizeof(float) * 4 * n);
    float* occ = (float*)malloc(sizeof(float) * 4 * n);
    float* L[2] = {(float*)malloc(sizeof(float) * 4 * n), (float*)calloc(4 * n, sizeof(float))};
    float* R[2] = {(float*)malloc(sizeof(float) * 4 * n), (float*)calloc(4 * n, sizeof(float))};
    orc_viewbuf* vb1 = (orc_viewbuf*)calloc(n, sizeof(orc_viewbuf));
    orc_viewbuf* vb2 = (orc_viewbuf*)calloc(n, sizeof(orc_viewbuf));

    orc_gradxy(left, W, H, nc, lg);        /* :313-320 */
    orc_gradxy(right, W, H, nc, rg);
    orc_weight(left, W, H, nc, lw);        /* :322-329 */
    orc_weight(right, W, H, nc, rw);
    orc_init(left, lg, right, rg, W, H, nc, P, depthmax, 1, L[0]);   /* :348-366 */
    orc_init(right, rg, left, lg, W, H, nc, P, depthmax, 0, R[0]);
    const float* vb2src = R[0];            /* vb2 is empty at iteration 0 (:368-372) */

    for (int iter = 0; iter < iterations; iter++) {   /* :374-489 */
        int swap = iter & 1;
        for (int pass = 0; pass < (rb ? 2 : 1); pass++)
            orc_propagate_ex(L[swap], L[1 - swap], left, lg, right, rg, ls, theta, W, H, nc, P,
                          depthmax, 1, iter, vb2, vb2src, NULL,
                          rb ? (pass ? ORC_SWEEP_BLACK : ORC_SWEEP_RED) : ORC_SWEEP_ALL, nbhd, nref);
        orc_viewbuf_build(L[1 - swap], W, H, vb1);
        for (int pass = 0; pass < (rb ? 2 : 1); pass++)
            orc_propagate_ex(R[swap], R[1 - swap], right, rg, left, lg, rs, theta, W, H, nc, P,
                          depthmax, 0, iter, vb1, L[1 - swap], NULL,
                          rb ? (pass ? ORC_SWEEP_BLACK : ORC_SWEEP_RED) : ORC_SWEEP_ALL, nbhd, nref);
        orc_viewbuf_build(R[1 - swap], W, H, vb2);
        vb2src = R[1 - swap];

        orc_occmap(L[1 - swap], R[1 - swap], W, H, occdiff, occdeg, 1, occ);    /* :451-456 */
        orc_normal_depth(L[1 - swap], W, H, depthmax, 1, tmp);                  /* :459-464 */
        orc_pdrof_inpaint(ls, tmp, lw, occ, W, H, theta * 20.0f + 5.0f, pd_iters); /* :466 */
        orc_occmap(R[1 - swap], L[1 - swap], W, H, occdiff, occdeg, 0, occ);    /* :469-474 */
        orc_normal_depth(R[1 - swap], W, H, depthmax, 0, tmp);                  /* :477-482 */
        orc_pdrof_inpaint(rs, tmp, rw, occ, W, H, theta * 20.0f + 5.0f, pd_iters); /* :484 */

        if (iter >= 4)                                                          /* :487-488 */
            theta = smoothstep01(((float)iter - 4.0f) / ((float)iterations - 4.0f)) * 1.0f;
    }

    orc_consistency(L[1], R[1], W, H, 1.0f, 5.0f, 1, tmp);   /* :491-497 */
    orc_fill_depthmap(tmp, W, H, 1.0f, dmap);                 /* :500-504 */
    if (lstate_out) memcpy(lstate_out, L[1], sizeof(float) * 4 * n);
    if (rstate_out) memcpy(rstate_out, R[1], sizeof(float) * 4 * n);

    free(lg); free(rg); free(lw); free(rw); free(ls); free(rs); free(tmp); free(occ);
    free(L[0]); free(L[1]); free(R[0]); free(R[1]); free(vb1); free(vb2);
}
