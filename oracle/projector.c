/*
 * projector.c -- oracle restatement of the rotation-based projectors.
 * TEST INFRASTRUCTURE ONLY (see hrrt_oracle.h).
 *
 * Forward:  forward_proj3d_view1_thread PJ:1728-1978 + scatter PJ:1397-1436.
 * Back:     back_proj3d2 PJ:1457-1620 with back_proj3d_view_n1 PJ:2927-3086.
 * Rotation: rotateimage2 PJ:1137-1186.
 *
 * The reference works on float4 "lanes"; here a lane is an explicit index
 * 0..3 and every float operation is issued in the reference's order.
 *
 * Values (image, sinogram, buffers, sums) are og_real: float in the oracle,
 * double in its FP64 build.  Positions, axial fractions and scale factors stay
 * float in both, so the FP64 build applies the same discretisation (the same
 * system matrix) and differs only by the rounding of the arithmetic.
 */
#include "hrrt_oracle.h"

#include <stdlib.h>
#include <string.h>

/* Debug counter: reads of the never-written plane z_pixels of the paired
 * sample columns (the reference reads uninitialised stack there; this
 * restatement reads zero). */
long og_fp_pad_reads = 0;

#define PB(g, buf, row, k) ((buf) + ((size_t)(row) * ((g)->th_pixels / 2 + 1) + (k)) * (size_t)(g)->P * 4 + 4)
/* PB(...)[4*yr + lane] addresses plane yr (yr = -1 .. z4) of a prjbuf row */

long og_prjbuf_floats(const og_geom *g)
{
    return (long)(g->N + 1) * (g->th_pixels / 2 + 1) * g->P * 4;
}

static inline const og_real *vox(const og_geom *g, const og_real *img, int x, int y)
{
    return img + ((size_t)x * g->N + y) * g->z4;
}

/* one bilinear sample of the full z column at the rotated position of (xx,yy)
 * (PJ:1821-1828, 1865-1892).  Returns 0 when the sample is rejected. */
static int sample_column(const og_geom *g, const og_real *image, float cosv, float sinv,
                         int xx, int yy, og_real *out)
{
    const int c = g->N / 2, N = g->N;
    float xx0 = (xx - c) * cosv - (yy - c) * sinv + c;
    float yy0 = (xx - c) * sinv + (yy - c) * cosv + c;
    int xi, yi, z;
    float x1, y1;
    og_real w1, w2, w3, w4;
    const og_real *r1, *r2, *r3, *r4;
    if (xx0 < 0 || xx0 > N - 2 || yy0 < 0 || xx0 > N - 2) return 0;
    xi = (int)xx0; yi = (int)yy0;
    x1 = xx0 - xi; y1 = yy0 - yi;
    r1 = vox(g, image, xi, yi);     r2 = vox(g, image, xi + 1, yi);
    r3 = vox(g, image, xi, yi + 1); r4 = vox(g, image, xi + 1, yi + 1);
    w1 = ((og_real)1 - x1) * ((og_real)1 - y1); w2 = (og_real)x1 * ((og_real)1 - y1);
    w3 = (og_real)y1 * ((og_real)1 - x1); w4 = (og_real)y1 * x1;
    for (z = 0; z < g->z4; z++)
        out[z] = (r1[z] * w1 + r2[z] * w2) + (r3[z] * w3 + r4[z] * w4);
    return 1;
}

static void fp_view(const og_geom *g, const og_real *image, og_real *prjbuf, int view, int xs, int xe)
{
    const int N = g->N, c = N / 2, Z = g->z_pixels, Z4 = g->z4, ZS = g->zshift;
    const int L = g->groupmax * 2 * ZS;
    const float cosv = g->cos_psi[view], sinv = g->sin_psi[view];
    og_real *rim = (og_real *)calloc((size_t)8 * Z4, sizeof(og_real));
    /* pam/pam2: float4 per plane, one zero pad plane beyond z_pixels */
    og_real *pam = (og_real *)calloc((size_t)(Z4 + 2) * 4, sizeof(og_real));
    og_real *pam2 = (og_real *)calloc((size_t)(Z4 + 2) * 4, sizeof(og_real));
    int x, y, i, z, z2, l, th, th2, yr, yri;

    for (x = xs; x < xe; x++) {
        og_real *pp1[64], *pp2[64];
        if (x == 0) continue;
        for (i = 0; i <= g->th_pixels / 2; i++) {
            pp1[i] = PB(g, prjbuf, x, i);
            pp2[i] = PB(g, prjbuf, N / 2 + x, i);
        }
        if (g->cylwiny[2 * x] == g->cylwiny[2 * x + 1]) continue;
        if (g->norm_zero_mask[view * (N / 2 + 1) + x] == 0) continue;

        /* centre row samples; a rejected sample keeps its previous column (PJ:1822) */
        {
            const int cx[4] = {x, N - x, c, c}, cy[4] = {c, c, x, N - x};
            for (i = 0; i < 4; i++) {
                float xx0 = (cx[i] - c) * cosv * 1.0f - (cy[i] - c) * sinv + c;
                float yy0 = (cx[i] - c) * sinv * 1.0f + (cy[i] - c) * cosv + c;
                if (xx0 < 0 || xx0 > N - 2 || yy0 < 0 || xx0 > N - 2) continue;
                sample_column(g, image, cosv, sinv, cx[i], cy[i], rim + i * Z4);
            }
        }
        for (z = 0; z < Z; z++)
            for (l = 0; l < 4; l++) pp1[0][4 * z + l] = rim[l * Z4 + z];
        for (th = 2; th < g->th_pixels; th += 2)
            for (yr = g->yr_bottom[th]; yr <= g->yr_top[th]; yr++)
                for (l = 0; l < 4; l++) {
                    pp1[th / 2][4 * yr + l] = rim[l * Z4 + yr];
                    pp2[th / 2][4 * yr + l] = rim[l * Z4 + yr];
                }

        for (y = g->cylwiny[2 * x]; y < c; y++) {
            /* eight symmetric samples PJ:1856-1864 */
            const int sx[8] = {x, N - x, N - y, N - y, x, N - x, y, y};
            const int sy[8] = {y, y, x, N - x, N - y, N - y, x, N - x};
            const og_flookup *lp1 = &g->flookup[(size_t)y * L];
            const og_flookup *lp2 = &g->flookup[(size_t)(N - y) * L];
            for (i = 0; i < 8; i++)
                if (!sample_column(g, image, cosv, sinv, sx[i], sy[i], rim + i * Z4))
                    memset(rim + i * Z4, 0, sizeof(og_real) * Z4);
            for (z = 0; z < Z; z++)
                for (l = 0; l < 4; l++) {
                    pam[4 * z + l] = rim[l * Z4 + z];
                    pam2[4 * z + l] = rim[(4 + l) * Z4 + z];
                }
            /* theta = 0 PJ:1907-1910 */
            for (z = 0; z < Z; z++)
                for (l = 0; l < 4; l++)
                    pp1[0][4 * z + l] = pp1[0][4 * z + l] + (pam[4 * z + l] + pam2[4 * z + l]);
            /* oblique segments PJ:1916-1951 */
            for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
                og_real *p1 = pp1[th / 2], *p2 = pp2[th / 2];
                float coef1, coef2, zz0;
                yri = th2 * ZS * 2;
                if (lp1[yri].zzi < 0) continue;
                if (lp2[yri + ZS * 2 - 1].zzi > Z - 1) continue;
#define FP_ACC(yr_, z_, z2_, f1_, f2_)                                                                  \
    do {                                                                                                \
        if ((z_) + 1 >= Z || (z2_) + 1 >= Z) og_fp_pad_reads++;                                         \
        for (l = 0; l < 4; l++) {                                                                       \
            p1[4 * (yr_) + l] = p1[4 * (yr_) + l] +                                                     \
                ((pam2[4 * ((z2_) + 1) + l] + pam[4 * (z_) + l]) * (f2_) +                              \
                 (pam[4 * ((z_) + 1) + l] + pam2[4 * (z2_) + l]) * (f1_));                              \
            p2[4 * (yr_) + l] = p2[4 * (yr_) + l] +                                                     \
                ((pam[4 * ((z2_) + 1) + l] + pam2[4 * (z_) + l]) * (f2_) +                              \
                 (pam2[4 * ((z_) + 1) + l] + pam[4 * (z2_) + l]) * (f1_));                              \
        }                                                                                               \
    } while (0)
                for (yr = g->yr_bottom[th]; yr < g->yr_bottom[th] + ZS; yr++, yri++)
                    FP_ACC(yr, lp1[yri].zzi, lp2[yri].zzi, lp1[yri].zzz0, lp2[yri].zzz0);
                zz0 = yr + (y - c) * g->tantheta[th2 * Z + yr];
                z = (int)zz0; coef1 = zz0 - z;
                zz0 = yr + (c - y) * g->tantheta[th2 * Z + yr];
                z2 = (int)zz0; coef2 = zz0 - z2;
                for (; yr <= g->yr_top[th] - ZS; yr++, z++, z2++)
                    FP_ACC(yr, z, z2, coef1, coef2);
                for (yr = g->yr_top[th] - ZS + 1; yr <= g->yr_top[th]; yr++, yri++)
                    FP_ACC(yr, lp1[yri].zzi, lp2[yri].zzi, lp1[yri].zzz0, lp2[yri].zzz0);
#undef FP_ACC
            }
        }
        /* axial/oblique scaling PJ:1953-1975 */
        for (yr = 0; yr < Z; yr++)
            for (l = 0; l < 4; l++) pp1[0][4 * yr + l] = pp1[0][4 * yr + l] * g->y_size;
        for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
            for (yr = g->yr_bottom[th]; yr <= g->yr_top[th]; yr++) {
                float s = g->rescalearray[th2 * Z + yr];
                for (l = 0; l < 4; l++) {
                    pp1[th / 2][4 * yr + l] = pp1[th / 2][4 * yr + l] * s;
                    pp2[th / 2][4 * yr + l] = pp2[th / 2][4 * yr + l] * s;
                }
            }
        }
    }
    free(rim); free(pam); free(pam2);
}

void og_forward_proj(const og_geom *g, const og_real *image, og_real *prj, int view, og_real *prjbuf)
{
    const int N = g->N, S = g->segmentsize, K = g->th_pixels / 2;
    int xr, xri, yr, k, y1, y2;
    fp_view(g, image, prjbuf, view, 0, N / 2 + 1);
    /* scatter into [2S][N] PJ:1397-1436 */
    for (xr = 1; xr <= N / 2; xr++) {
        const og_real *ptr = PB(g, prjbuf, xr, 0);
        if (g->cylwiny[2 * xr] == 0) continue;
        xri = N - xr;
        for (yr = 0; yr < g->z_pixels; yr++) {
            prj[(size_t)yr * N + xr] = ptr[4 * yr];
            prj[(size_t)yr * N + xri] = ptr[4 * yr + 1];
            prj[(size_t)(S + yr) * N + xr] = ptr[4 * yr + 2];
            prj[(size_t)(S + yr) * N + xri] = ptr[4 * yr + 3];
        }
    }
    for (k = 1; k <= K; k++) {
        for (xr = 0; xr <= N / 2; xr++) {
            const og_real *ptr;
            if (g->cylwiny[2 * xr] == 0) continue;
            xri = N - xr;
            ptr = PB(g, prjbuf, xr, k);
            y1 = g->segment_offset[2 * k]; y2 = y1 + S;
            for (yr = g->yr_bottom[2 * k]; yr <= g->yr_top[2 * k]; yr++, y1++, y2++) {
                prj[(size_t)y1 * N + xr] = ptr[4 * yr];
                prj[(size_t)y1 * N + xri] = ptr[4 * yr + 1];
                prj[(size_t)y2 * N + xr] = ptr[4 * yr + 2];
                prj[(size_t)y2 * N + xri] = ptr[4 * yr + 3];
            }
            ptr = PB(g, prjbuf, N / 2 + xr, k);
            y1 = g->segment_offset[2 * k - 1]; y2 = y1 + S;
            for (yr = g->yr_bottom[2 * k]; yr <= g->yr_top[2 * k]; yr++, y1++, y2++) {
                prj[(size_t)y1 * N + xr] = ptr[4 * yr];
                prj[(size_t)y1 * N + xri] = ptr[4 * yr + 1];
                prj[(size_t)y2 * N + xr] = ptr[4 * yr + 2];
                prj[(size_t)y2 * N + xri] = ptr[4 * yr + 3];
            }
        }
    }
}

/* in-place rescale of both views' rows by cos^3(theta)*span*ring_spacing/(2R) (PJ:1496-1507) */
static void bp_rescale(const og_geom *g, og_real *prj)
{
    const int N = g->N, S = g->segmentsize;
    int k, yr, yr2 = 0, xr;
    for (k = 0; k < g->th_pixels; k++) {
        float rescale = g->cos_theta[k] * g->cos_theta[k] * g->cos_theta[k] * g->p.span * g->ring_spacing /
                        (2 * g->p.crystal_radius);
        for (yr = g->yr_bottom[k]; yr <= g->yr_top[k]; yr++, yr2++) {
            for (xr = 0; xr < N; xr++) {
                prj[(size_t)yr2 * N + xr] = prj[(size_t)yr2 * N + xr] * rescale;
                prj[(size_t)(yr2 + S) * N + xr] = prj[(size_t)(yr2 + S) * N + xr] * rescale;
            }
        }
    }
}

/* transpose one view of the pair into prjbuf lanes for the direct BP (PJ:1511-1540, 1556-1583) */
static void bp_direct_transpose(const og_geom *g, const og_real *prj, og_real *prjbuf, int base)
{
    const int N = g->N, K = g->th_pixels / 2;
    int xr, yr, k, y1, y2;
    for (xr = 1; xr <= N / 2; xr++) {
        og_real *ptr = PB(g, prjbuf, xr, 0);
        for (yr = 0; yr < g->z_pixels; yr++) {
            ptr[4 * yr] = prj[(size_t)(base + yr) * N + xr];
            ptr[4 * yr + 2] = prj[(size_t)(base + yr) * N + N - xr];
            ptr[4 * yr + 1] = ptr[4 * yr + 3] = 0;
        }
    }
    for (k = 1; k <= K; k++) {
        for (xr = 1; xr <= N / 2; xr++) {
            og_real *ptr = PB(g, prjbuf, xr, k);
            int l;
            y1 = g->segment_offset[k * 2] + base;
            y2 = g->segment_offset[k * 2 - 1] + base;
            for (yr = g->yr_bottom[k * 2]; yr <= g->yr_top[k * 2]; yr++, y1++, y2++) {
                ptr[4 * yr] = prj[(size_t)y1 * N + xr];
                ptr[4 * yr + 1] = prj[(size_t)y2 * N + xr];
                ptr[4 * yr + 2] = prj[(size_t)y1 * N + N - xr];
                ptr[4 * yr + 3] = prj[(size_t)y2 * N + N - xr];
            }
            /* guard planes are zero in the reference's buffers (S3:2764-2769) */
            for (l = 0; l < 4; l++) {
                ptr[4 * (g->yr_bottom[k * 2] - 1) + l] = 0;
                ptr[4 * (g->yr_top[k * 2] + 1) + l] = 0;
            }
        }
    }
}

static inline og_real *voxw(const og_geom *g, og_real *img, int x, int y)
{
    return img + ((size_t)x * g->N + y) * g->z4;
}

/* back_proj3d_view_n1 PJ:2927-3086 for x in [xs, xe) */
static void bp_direct_pass(const og_geom *g, og_real *image, og_real *prjbuf, int f90, int view, int xs, int xe)
{
    const int N = g->N, c = N / 2, Z = g->z_pixels, Z4 = g->z4;
    og_real *i1 = (og_real *)malloc(sizeof(og_real) * 4 * Z4);
    og_real *i2 = (og_real *)malloc(sizeof(og_real) * 4 * Z4);
    int x, y, z, z2, yr, yr2, th, th2;
    for (x = xs; x < xe; x++) {
        const og_real *p0 = PB(g, prjbuf, x, 0);
        int ystart = g->cylwiny[2 * x], yend = g->cylwiny[2 * x + 1];
        int mask = g->norm_zero_mask[view * (c + 1) + x] == 0;
        if (ystart == 0 && yend == 0) continue;
        for (y = ystart; y < c; y++) {
            og_real *o1, *o2, *o3, *o4;
            if (!mask) {
                const og_blookup *l1 = &g->blookup[(size_t)y * g->groupmax];
                const og_blookup *l2 = &g->blookup[(size_t)(N - y) * g->groupmax];
                memcpy(i1, p0, sizeof(og_real) * 4 * Z);
                memcpy(i2, p0, sizeof(og_real) * 4 * Z);
                for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
                    const og_real *p = PB(g, prjbuf, x, th / 2);
                    float coef1 = l1[th2].yr2a, coef2 = l2[th2].yr2a;
                    int l;
                    z2 = l2[th2].zsa; yr2 = l2[th2].yra;
                    for (z = l1[th2].zsa, yr = l1[th2].yra; z <= l1[th2].zea; z++, yr++, z2++, yr2++)
                        for (l = 0; l < 4; l++) {
                            i1[4 * z + l] = i1[4 * z + l] + (p[4 * yr + l] + coef1 * (p[4 * (yr + 1) + l] - p[4 * yr + l]));
                            i2[4 * z2 + l] = i2[4 * z2 + l] + (p[4 * yr2 + l] + coef2 * (p[4 * (yr2 + 1) + l] - p[4 * yr2 + l]));
                        }
                }
            }
            if (f90 == 0) {
                o1 = voxw(g, image, x, y); o2 = voxw(g, image, x, N - y);
                o3 = voxw(g, image, N - x, y); o4 = voxw(g, image, N - x, N - y);
                if (mask) {
                    memset(o1, 0, sizeof(og_real) * Z4); memset(o2, 0, sizeof(og_real) * Z4);
                    if (x != c) { memset(o3, 0, sizeof(og_real) * Z4); memset(o4, 0, sizeof(og_real) * Z4); }
                } else {
                    for (z = 0; z < Z; z++) {
                        o1[z] = i1[4 * z] + i2[4 * z + 1];
                        o2[z] = i1[4 * z + 1] + i2[4 * z];
                        if (x != c) {
                            o3[z] = i1[4 * z + 2] + i2[4 * z + 3];
                            o4[z] = i1[4 * z + 3] + i2[4 * z + 2];
                        }
                    }
                }
            } else if (!mask) {
                o1 = voxw(g, image, N - y, x); o2 = voxw(g, image, y, x);
                o3 = voxw(g, image, N - y, N - x); o4 = voxw(g, image, y, N - x);
                for (z = 0; z < Z; z++) {
                    o1[z] = o1[z] + (i1[4 * z] + i2[4 * z + 1]);
                    o2[z] = o2[z] + (i1[4 * z + 1] + i2[4 * z]);
                    if (x != c) {
                        o3[z] = o3[z] + (i1[4 * z + 2] + i2[4 * z + 3]);
                        o4[z] = o4[z] + (i1[4 * z + 3] + i2[4 * z + 2]);
                    }
                }
            }
        }
        /* centre row PJ:3048-3083 (not masked in the reference) */
        memcpy(i1, p0, sizeof(og_real) * 4 * Z);
        for (th = 2; th < g->th_pixels; th += 2) {
            const og_real *p = PB(g, prjbuf, x, th / 2);
            int l;
            for (z = g->yr_bottom[th]; z <= g->yr_top[th]; z++)
                for (l = 0; l < 4; l++) i1[4 * z + l] = i1[4 * z + l] + p[4 * z + l];
        }
        {
            og_real *o1, *o4;
            if (f90 == 0) { o1 = voxw(g, image, x, c); o4 = voxw(g, image, N - x, c); }
            else { o1 = voxw(g, image, c, x); o4 = voxw(g, image, c, N - x); }
            for (z = 0; z < Z; z++) {
                og_real a = i1[4 * z] + i1[4 * z + 1];
                og_real b = i1[4 * z + 2] + i1[4 * z + 3];
                if (f90 == 0) {
                    o1[z] = a;
                    if (x != c) o4[z] = b;
                } else {
                    o1[z] = o1[z] + a;
                    if (x != c) o4[z] = o4[z] + b;
                }
            }
        }
    }
    free(i1); free(i2);
}

void og_back_proj_direct_rotframe(const og_geom *g, og_real *prj, og_real *imagebuf, int view, og_real *prjbuf)
{
    const int N = g->N;
    bp_rescale(g, prj);
    bp_direct_transpose(g, prj, prjbuf, 0);
    bp_direct_pass(g, imagebuf, prjbuf, 0, view, 0, N / 2 + 1);
    bp_direct_transpose(g, prj, prjbuf, g->segmentsize);
    bp_direct_pass(g, imagebuf, prjbuf, 1, view, 0, N / 2 + 1);
}

void og_back_proj_direct(const og_geom *g, og_real *prj, og_real *ima, int view, og_real *imagebuf, og_real *prjbuf)
{
    og_back_proj_direct_rotframe(g, prj, imagebuf, view, prjbuf);
    og_rotate_add(g, imagebuf, ima, view, 0, g->N);
}

void og_rotate_add(const og_geom *g, const og_real *imagebuf, og_real *out, int view, int xs, int xe)
{   /* rotateimage2 PJ:1137-1186 (zoom == 1) */
    const int N = g->N, Z4 = g->z4;
    const float cosv = g->cos_psi[view] / 1.0f, sinv = -g->sin_psi[view] / 1.0f;
    const float xxx0 = -N / 2 * cosv + N / 2 * sinv + N / 2;
    const float yyy0 = -N / 2 * sinv - N / 2 * cosv + N / 2;
    int x, y, z;
    for (x = xs; x < xe; x++) {
        int ystart = g->cylwiny[2 * x], yend = g->cylwiny[2 * x + 1];
        float xx0 = xxx0 - ystart * sinv + (float)(unsigned)x * cosv;
        float yy0 = yyy0 + ystart * cosv + (float)(unsigned)x * sinv;
        for (y = ystart; y < yend; y++) {
            int xx = (int)xx0, yy = (int)yy0;
            float x1, y1;
            og_real w1, w2, w3, w4;
            const og_real *im1, *im2, *im3, *im4;
            og_real *o;
            if (xx < 0 || xx > N - 2 || yy < 0 || yy > N - 2) {
                xx0 -= sinv; yy0 += cosv;
                continue;
            }
            x1 = xx0 - xx; y1 = yy0 - yy;
            im1 = vox(g, imagebuf, xx, yy); im2 = vox(g, imagebuf, xx + 1, yy);
            im3 = vox(g, imagebuf, xx, yy + 1); im4 = vox(g, imagebuf, xx + 1, yy + 1);
            o = out + ((size_t)x * N + y) * Z4;
            w1 = ((og_real)1 - x1) * ((og_real)1 - y1); w2 = (og_real)x1 * ((og_real)1 - y1);
            w3 = (og_real)y1 * ((og_real)1 - x1); w4 = (og_real)y1 * x1;
            for (z = 0; z < Z4; z++)
                o[z] = o[z] + ((im1[z] * w1 + im2[z] * w2) + (im3[z] * w3 + im4[z] * w4));
            xx0 -= sinv; yy0 += cosv;
        }
    }
}
