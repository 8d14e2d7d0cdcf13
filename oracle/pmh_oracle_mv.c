/*
 * pmh_oracle_mv.c -- CPU restatement of the multi-view slanted-plane
 * PatchMatch of this build (TEST INFRASTRUCTURE ONLY; see pmh_oracle.h).
 *
 * The reference holds a draft of the method that does not compile
 * (cvt/cl/kernel/pmmvstereo.cl:97-304 with cvt/cl/kernel/Plane.cl:60-86:
 * undefined normmul / self / lr / iter, a missing ';' in Plane.cl:28, an
 * inverted inside-the-image test at pmmvstereo.cl:127,142, and a FIXME on the
 * sign of the plane distance at Plane.cl:29).  The method restated here keeps
 * the draft's structure and constants and fixes its conventions:
 *
 *   - a particle is (nx, ny, d, cost): the plane n.X = d in the reference
 *     camera frame, n = (nx, ny, sqrt(1 - nx^2 - ny^2)) (pmmv_particle_to_plane,
 *     pmmvstereo.cl:79-87, with the distance stored where the draft stores
 *     its "depth");
 *   - cameras are K (3x3) and T = [R | t] (world -> camera, 3x4), row-major;
 *     for a source view, R = R_s R_r^T, t = t_s - R t_r, and the plane-induced
 *     homography (planef_homography, Plane.cl:60-86, with the sign of n.X = d)
 *     is H = K_s (R + t n^T / d) K_r^-1 = M + e (K_r^-T n)^T / d with the view
 *     constants M = K_s R K_r^-1 and e = K_s t (computed in double, rounded);
 *   - the window cost (pmmv3_patch_eval, pmmvstereo.cl:97-167) sums, for
 *     every tap inside the reference image whose ray meets the plane in front
 *     of the reference camera (q . (x, y, 1) > 0 with q = K_r^-T n / d), and
 *     every source view that sees that point in front of it (Z > 0) inside
 *     its image,
 *     the support weight and the weighted truncated colour+gradient cost; the
 *     support factor uses smoothstep(0, 28, |offset|) * 1.5 * 26 + 5
 *     (pmmvstereo.cl:119); the source views are sampled bilinearly in x and y
 *     (clamp to edge) from per-texel cells (orc_mv_cells) at x' = X / Z
 *     evaluated as X * (1 / Z); gray inputs (the colour sum is 3 x one
 *     channel); 1e5 if the weight sum is at most 1.1 or the particle is not
 *     finite;
 *   - a pass (pmmv3_iteration, pmmvstereo.cl:190-304) tries, in order, the 8
 *     neighbours, 3 random neighbours at offsets (int)(u 7 + 0.5) in [0, 7]^2
 *     (x drawn first), one random particle (normal spread 2, as the stereo
 *     path's random try) and `nref` randomized refinements of the current
 *     particle (the draft's NUMRNDTRIES = 1 means none; this build's default
 *     is 2), accepting a candidate whose cost is strictly lower (:245).
 *     Draws: RNG_init(y W + x + iter, 33) (:237; 15 used), supplied draws
 *     replace them.  Sweeps: all pixels from the old field (the draft) or red
 *     / black passes (a black pixel reads its red neighbours' new particles);
 *   - a random particle at a pixel (pmmv_particle_init, :52-63): depth
 *     z = dmin + u0 (dmax - dmin) at the pixel, nx, ny = (u - 0.5) normmul
 *     0.95 clamped to +-0.95 and scaled back into the unit disc as
 *     nd_state_init does (pmhstereo.cl:50-68), d = z (n . K_r^-1 (x, y, 1));
 *   - a refinement (pmmv_particle_refine, :65-77): the depth at the pixel
 *     moves by (u0 - 0.5) (dmax - dmin) / 32, clamped to the range, the normal
 *     by (u - 0.5) 0.2 per component (NORMALREFINEMUL), clamped and scaled
 *     back as above; d follows the new depth;
 *   - the depth of a particle at a pixel is d / (n . K_r^-1 (x, y, 1)) (0 if
 *     that is not a finite positive number: the ray meets the plane behind
 *     the camera or not at all).
 *
 * Float arithmetic follows the device code operation for operation
 * (-ffp-contract=off, explicit fmaf), so homographies, projections and
 * particles match the GPU's bits; weights use expf (the GPU's exp2 of a
 * folded exponent differs in the last bits, hence the 1e-5 cost tolerance).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pmh_oracle.h"

#define MV_DRAWS 33
#define MV_NBHD 3
#define MV_MAXVIEWS 4

static inline float clampf_mv(float v, float lo, float hi)
{
    return fminf(fmaxf(v, lo), hi);
}

/* 3x3 inverse in double (adjugate / determinant). */
static void inv3(const double* a, double* r)
{
    const double c00 = a[4] * a[8] - a[5] * a[7];
    const double c01 = a[5] * a[6] - a[3] * a[8];
    const double c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    r[0] = c00 / det;
    r[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    r[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    r[3] = c01 / det;
    r[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    r[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    r[6] = c02 / det;
    r[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    r[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

static void mul3(const double* a, const double* b, double* r)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

/* View constants: kinv[9] = K_r^-1, and per source view M[9], e[3] (float). */
void orc_mv_constants(const float* Kr, const float* Tr, int nsrc, const float* Ks,
                      const float* Ts, float* kinv, float* M, float* e)
{
    double K[9], Ki[9];
    for (int i = 0; i < 9; i++) K[i] = Kr[i];
    inv3(K, Ki);
    for (int i = 0; i < 9; i++) kinv[i] = (float)Ki[i];
    double Rr[9], tr[3];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) Rr[i * 3 + j] = Tr[i * 4 + j];
        tr[i] = Tr[i * 4 + 3];
    }
    for (int v = 0; v < nsrc; v++) {
        const float* Kv = Ks + 9 * v;
        const float* Tv = Ts + 12 * v;
        double Rs[9], ts[3], RrT[9], R[9], t[3], KvD[9], KR[9], Mv[9];
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                Rs[i * 3 + j] = Tv[i * 4 + j];
                RrT[i * 3 + j] = Rr[j * 3 + i];
            }
            ts[i] = Tv[i * 4 + 3];
        }
        for (int i = 0; i < 9; i++) KvD[i] = Kv[i];
        mul3(Rs, RrT, R);
        for (int i = 0; i < 3; i++)
            t[i] = ts[i] - (R[i * 3] * tr[0] + R[i * 3 + 1] * tr[1] + R[i * 3 + 2] * tr[2]);
        mul3(KvD, R, KR);
        mul3(KR, Ki, Mv);
        for (int i = 0; i < 9; i++) M[9 * v + i] = (float)Mv[i];
        for (int i = 0; i < 3; i++)
            e[3 * v + i] =
                (float)(KvD[i * 3] * t[0] + KvD[i * 3 + 1] * t[1] + KvD[i * 3 + 2] * t[2]);
    }
}

/* The normal of a particle (n_z from the unit length). */
static inline void mv_normal(const float p[4], float n[3])
{
    n[0] = p[0];
    n[1] = p[1];
    n[2] = sqrtf((1.0f - p[0] * p[0]) - p[1] * p[1]);
}

/* n . K_r^-1 (x, y, 1): the ray through (x, y) met by the plane normal. */
static inline float mv_ray_dot(const float n[3], const float* ki, float x, float y)
{
    const float rx = fmaf(ki[0], x, fmaf(ki[1], y, ki[2]));
    const float ry = fmaf(ki[3], x, fmaf(ki[4], y, ki[5]));
    const float rz = fmaf(ki[6], x, fmaf(ki[7], y, ki[8]));
    return fmaf(n[0], rx, fmaf(n[1], ry, n[2] * rz));
}

/* q = K_r^-T n / d: q . (x, y, 1) = 1 / lambda where the ray through (x, y)
 * meets the plane at lambda K_r^-1 (x, y, 1). */
static inline void mv_plane_q(const float p[4], const float* kinv, float q[3])
{
    float n[3];
    mv_normal(p, n);
    for (int j = 0; j < 3; j++) {
        const float m = fmaf(n[0], kinv[j], fmaf(n[1], kinv[3 + j], n[2] * kinv[6 + j]));
        q[j] = m / p[2];
    }
}

/* H of a particle for one source view: H_ij = M_ij + e_i q_j. */
static inline void mv_homography_q(const float q[3], const float* M, const float* e, float H[9])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H[i * 3 + j] = fmaf(e[i], q[j], M[i * 3 + j]);
}

void orc_mv_homography(const float p[4], const float* kinv, const float* M, const float* e,
                       float H[9])
{
    float q[3];
    mv_plane_q(p, kinv, q);
    mv_homography_q(q, M, e, H);
}

float orc_mv_depth(const float p[4], const float* kinv, float x, float y)
{
    float n[3];
    mv_normal(p, n);
    const float z = p[2] / mv_ray_dot(n, kinv, x, y);
    return (isfinite(z) && z > 0.0f) ? z : 0.0f;
}

/* Support factors of the multi-view cost (pmmvstereo.cl:119). */
void orc_mv_factor_table(int P, float* table)
{
    const int S = 2 * P + 1;
    for (int dy = -P; dy <= P; dy++)
        for (int dx = -P; dx <= P; dx++) {
            const float fx = (float)dx, fy = (float)dy;
            const float len = sqrtf(fx * fx + fy * fy);
            float t = len / 28.0f;
            t = fminf(fmaxf(t, 0.0f), 1.0f);
            const float ss = t * t * (3.0f - 2.0f * t);
            table[(dy + P) * S + (dx + P)] = ss * 1.5f * 26.0f + 5.0f;
        }
}

/* Bilinear sample at (x, y) inside the image of channel c of an nch-channel
 * image, from its cell (t, dx, dy, dxy) at i0 = (int)x, j0 = (int)y:
 * dx = T10 - T00, dy = T01 - T00, dxy = (T11 - T01) - dx (T with the last
 * column and row repeated), value fmaf(b, fmaf(a, dxy, dy), fmaf(a, dx, t)).
 * This is T00 + a (T10 - T00) + b (T01 - T00) + a b (T11 - T01 - T10 + T00),
 * the bilinear interpolant, evaluated from one stored cell per texel. */
static inline void cell_at(const float* img, int W, int H, int nch, int c, int i0, int j0,
                           float cell[4])
{
    const int i1 = i0 + 1 < W ? i0 + 1 : W - 1;
    const int j1 = j0 + 1 < H ? j0 + 1 : H - 1;
    const float t00 = img[((size_t)j0 * W + i0) * nch + c];
    const float t10 = img[((size_t)j0 * W + i1) * nch + c];
    const float t01 = img[((size_t)j1 * W + i0) * nch + c];
    const float t11 = img[((size_t)j1 * W + i1) * nch + c];
    const float dx = t10 - t00;
    cell[0] = t00;
    cell[1] = dx;
    cell[2] = t01 - t00;
    cell[3] = (t11 - t01) - dx;
}

static inline float bil(const float* img, int W, int H, int nch, int c, float x, float y)
{
    const int i0 = (int)x, j0 = (int)y;
    const float a = x - (float)i0, b = y - (float)j0;
    float cl[4];
    cell_at(img, W, H, nch, c, i0, j0, cl);
    return fmaf(b, fmaf(a, cl[3], cl[2]), fmaf(a, cl[1], cl[0]));
}

/* The cells of one image (W*H*4 floats, channel c of an nch-channel image). */
void orc_mv_cells(const float* img, int W, int H, int nch, int c, float* cells)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) cell_at(img, W, H, nch, c, x, y, cells + ((size_t)y * W + x) * 4);
}

float orc_mv_bilinear(const float* img, int W, int H, int nch, int c, float x, float y)
{
    return bil(img, W, H, nch, c, x, y);
}

static float mv_cost_core(const float* Ir, const float* Gr, int nsrc, const float* const* Is,
                          const float* const* Gs, int W, int H, int x, int y, const float p[4],
                          int P, const float* F, const float* kinv, const float* M,
                          const float* e)
{
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2])))
        return 1e5f;
    float q[3], Hm[MV_MAXVIEWS][9];
    mv_plane_q(p, kinv, q);
    for (int v = 0; v < nsrc; v++) mv_homography_q(q, M + 9 * v, e + 3 * v, Hm[v]);
    const int S = 2 * P + 1;
    const float Wf = (float)W, Hf = (float)H;
    const float c0 = Ir[(size_t)y * W + x];
    float wsum = 0.0f, ret = 0.0f;
    for (int dy = -P; dy <= P; dy++) {
        const int qy = y + dy;
        if (qy < 0 || qy >= H)
            continue;
        for (int dx = -P; dx <= P; dx++) {
            const int qx = x + dx;
            if (qx < 0 || qx >= W)
                continue;
            const float v1 = Ir[(size_t)qy * W + qx];
            const float* g1 = Gr + ((size_t)qy * W + qx) * 4;
            const float wt = expf(-(3.0f * fabsf(c0 - v1)) * F[(dy + P) * S + (dx + P)]);
            const float fqx = (float)qx, fqy = (float)qy;
            /* the plane point of this tap lies in front of the reference camera */
            if (!(fmaf(q[0], fqx, fmaf(q[1], fqy, q[2])) > 0.0f))
                continue;
            for (int v = 0; v < nsrc; v++) {
                const float* h = Hm[v];
                const float X = fmaf(h[0], fqx, fmaf(h[1], fqy, h[2]));
                const float Y = fmaf(h[3], fqx, fmaf(h[4], fqy, h[5]));
                const float Z = fmaf(h[6], fqx, fmaf(h[7], fqy, h[8]));
                if (!(Z > 0.0f))
                    continue;
                const float r = 1.0f / Z;
                const float xp = X * r, yp = Y * r;
                if (!(xp >= 0.0f && xp < Wf && yp >= 0.0f && yp < Hf))
                    continue;
                wsum += wt;
                const float v2 = bil(Is[v], W, H, 1, 0, xp, yp);
                const float cterm = 3.0f * fminf(fabsf(v1 - v2), 0.04f);
                float m[4];
                for (int c = 0; c < 4; c++)
                    m[c] = fminf(fabsf(g1[c] - bil(Gs[v], W, H, 4, c, xp, yp)), 0.01f);
                const float gs = ((m[0] + m[1]) + m[2]) + m[3];
                const float C = fmaf(0.95f, gs, 0.05f * cterm);
                ret = fmaf(wt, C, ret);
            }
        }
    }
    if (wsum <= 1.1f)
        return 1e5f;
    return ret / wsum;
}

float orc_mv_cost(const float* Ir, const float* Gr, int nsrc, const float* Is, const float* Gs,
                  int W, int H, int x, int y, const float p[4], int P, const float* kinv,
                  const float* M, const float* e)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_mv_factor_table(P, F);
    const float* is[MV_MAXVIEWS];
    const float* gs[MV_MAXVIEWS];
    for (int v = 0; v < nsrc; v++) {
        is[v] = Is + (size_t)v * W * H;
        gs[v] = Gs + (size_t)v * W * H * 4;
    }
    const float r = mv_cost_core(Ir, Gr, nsrc, is, gs, W, H, x, y, p, P, F, kinv, M, e);
    free(F);
    return r;
}

/* A random particle at pixel (x, y) from draws u[0..2] (normal spread nm). */
void orc_mv_particle_init(const float u[3], float x, float y, float nm, float dmin, float dmax,
                          const float* kinv, float out[4])
{
    const float z = fmaf(u[0], dmax - dmin, dmin);
    float nx = ((u[1] - 0.5f) * nm) * 0.95f;
    float ny = ((u[2] - 0.5f) * nm) * 0.95f;
    nx = clampf_mv(nx, -0.95f, 0.95f);
    ny = clampf_mv(ny, -0.95f, 0.95f);
    const float nf = fmaxf(sqrtf(nx * nx + ny * ny) + 0.001f, 1.0f);
    nx = nx / nf;
    ny = ny / nf;
    float n[3] = {nx, ny, 0.0f};
    n[2] = sqrtf((1.0f - nx * nx) - ny * ny);
    out[0] = nx;
    out[1] = ny;
    out[2] = z * mv_ray_dot(n, kinv, x, y);
    out[3] = 0.0f;
}

/* A randomized refinement of particle p at pixel (x, y) from draws u[0..2]. */
void orc_mv_particle_refine(const float u[3], const float p[4], float x, float y, float dmin,
                            float dmax, const float* kinv, float out[4])
{
    float n[3];
    mv_normal(p, n);
    const float z0 = p[2] / mv_ray_dot(n, kinv, x, y);
    const float z = clampf_mv(fmaf(u[0] - 0.5f, (dmax - dmin) / 32.0f, z0), dmin, dmax);
    float nx = fmaf(u[1] - 0.5f, 0.2f, p[0]);
    float ny = fmaf(u[2] - 0.5f, 0.2f, p[1]);
    nx = clampf_mv(nx, -0.95f, 0.95f);
    ny = clampf_mv(ny, -0.95f, 0.95f);
    const float nf = fmaxf(sqrtf(nx * nx + ny * ny) + 0.001f, 1.0f);
    nx = nx / nf;
    ny = ny / nf;
    float m[3] = {nx, ny, 0.0f};
    m[2] = sqrtf((1.0f - nx * nx) - ny * ny);
    out[0] = nx;
    out[1] = ny;
    out[2] = z * mv_ray_dot(m, kinv, x, y);
    out[3] = 0.0f;
}

/* pmmv3_init (pmmvstereo.cl:170-188): a random particle per pixel and its cost. */
void orc_mv_init(const float* Ir, const float* Gr, int nsrc, const float* Is, const float* Gs,
                 int W, int H, int P, float dmin, float dmax, const float* kinv, const float* M,
                 const float* e, float* out)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_mv_factor_table(P, F);
    const float* is[MV_MAXVIEWS];
    const float* gs[MV_MAXVIEWS];
    for (int v = 0; v < nsrc; v++) {
        is[v] = Is + (size_t)v * W * H;
        gs[v] = Gs + (size_t)v * W * H * 4;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            uint32_t st[2];
            float u[3], r[4];
            orc_rng_init((uint64_t)y * W + x, 3, st);
            for (int i = 0; i < 3; i++) u[i] = orc_rng_float(st);
            orc_mv_particle_init(u, (float)x, (float)y, 1.0f, dmin, dmax, kinv, r);
            r[3] = mv_cost_core(Ir, Gr, nsrc, is, gs, W, H, x, y, r, P, F, kinv, M, e);
            memcpy(out + ((size_t)y * W + x) * 4, r, sizeof(r));
        }
    free(F);
}

/* pmmv3_iteration (pmmvstereo.cl:190-304): one pass.  sweep 0: every pixel
 * from `old`; 1: red pixels ((x + y) even); 2: black pixels, reading the red
 * neighbours from `out`.  draws: NULL (RNG streams) or MV_DRAWS per pixel. */
void orc_mv_propagate(const float* old, float* out, const float* Ir, const float* Gr, int nsrc,
                      const float* Is, const float* Gs, int W, int H, int P, float dmin,
                      float dmax, int iter, const float* kinv, const float* M, const float* e,
                      const float* draws, int sweep, int nref)
{
    const int S = 2 * P + 1;
    float* F = (float*)malloc(sizeof(float) * S * S);
    orc_mv_factor_table(P, F);
    const float* is[MV_MAXVIEWS];
    const float* gs[MV_MAXVIEWS];
    for (int v = 0; v < nsrc; v++) {
        is[v] = Is + (size_t)v * W * H;
        gs[v] = Gs + (size_t)v * W * H * 4;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const int par = (x + y) & 1;
            if ((sweep == 1 && par != 0) || (sweep == 2 && par != 1))
                continue;
            float u[MV_DRAWS];
            if (draws) {
                memcpy(u, draws + ((size_t)y * W + x) * MV_DRAWS, sizeof(u));
            } else {
                uint32_t st[2];
                orc_rng_init((uint64_t)y * W + x + (uint64_t)iter, MV_DRAWS, st);
                for (int i = 0; i < MV_DRAWS; i++) u[i] = orc_rng_float(st);
            }
            float self[4], c[4];
            memcpy(self, old + ((size_t)y * W + x) * 4, sizeof(self));
            for (int k = 0; k < 8 + MV_NBHD + 1 + nref; k++) {
                if (k < 8 + MV_NBHD) {
                    int nx, ny;
                    if (k < 8) {
                        const int n = k < 4 ? k : k + 1;
                        nx = x + n % 3 - 1;
                        ny = y + n / 3 - 1;
                    } else {
                        const int i = k - 8;
                        nx = x + (int)(u[2 * i] * 7.0f + 0.5f);
                        ny = y + (int)(u[2 * i + 1] * 7.0f + 0.5f);
                    }
                    nx = nx < 0 ? 0 : (nx >= W ? W - 1 : nx);
                    ny = ny < 0 ? 0 : (ny >= H ? H - 1 : ny);
                    const float* src = (sweep == 2 && ((nx + ny) & 1) == 0) ? out : old;
                    memcpy(c, src + ((size_t)ny * W + nx) * 4, sizeof(c));
                } else if (k == 8 + MV_NBHD) {
                    orc_mv_particle_init(u + 2 * MV_NBHD, (float)x, (float)y, 2.0f, dmin, dmax,
                                         kinv, c);
                } else {
                    const int i = k - (9 + MV_NBHD);
                    orc_mv_particle_refine(u + 2 * MV_NBHD + 3 + 3 * i, self, (float)x, (float)y,
                                           dmin, dmax, kinv, c);
                }
                c[3] = mv_cost_core(Ir, Gr, nsrc, is, gs, W, H, x, y, c, P, F, kinv, M, e);
                if (c[3] < self[3])
                    memcpy(self, c, sizeof(self));
            }
            memcpy(out + ((size_t)y * W + x) * 4, self, sizeof(self));
        }
    free(F);
}

/* init + iterations passes (red-black if rb, else Jacobi), depth map out. */
void orc_mv_depthmap(const float* Ir, const float* Gr, int nsrc, const float* Is,
                     const float* Gs, int W, int H, int P, float dmin, float dmax,
                     int iterations, int rb, int nref, const float* kinv, const float* M,
                     const float* e, float* depth, float* particles)
{
    const size_t n = (size_t)W * H;
    float* a = (float*)malloc(sizeof(float) * 4 * n);
    float* b = (float*)malloc(sizeof(float) * 4 * n);
    orc_mv_init(Ir, Gr, nsrc, Is, Gs, W, H, P, dmin, dmax, kinv, M, e, a);
    for (int it = 0; it < iterations; it++) {
        if (rb) {
            memcpy(b, a, sizeof(float) * 4 * n);
            orc_mv_propagate(a, b, Ir, Gr, nsrc, Is, Gs, W, H, P, dmin, dmax, it, kinv, M, e,
                             NULL, 1, nref);
            orc_mv_propagate(a, b, Ir, Gr, nsrc, Is, Gs, W, H, P, dmin, dmax, it, kinv, M, e,
                             NULL, 2, nref);
        } else {
            orc_mv_propagate(a, b, Ir, Gr, nsrc, Is, Gs, W, H, P, dmin, dmax, it, kinv, M, e,
                             NULL, 0, nref);
        }
        float* t = a;
        a = b;
        b = t;
    }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            depth[(size_t)y * W + x] = orc_mv_depth(a + ((size_t)y * W + x) * 4, kinv, (float)x,
                                                    (float)y);
    if (particles) memcpy(particles, a, sizeof(float) * 4 * n);
    free(a);
    free(b);
}
