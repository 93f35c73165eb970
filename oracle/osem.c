/*
 * osem.c -- oracle restatement of the incremental back projector, the
 * per-bin/per-voxel OSEM steps and the subset drivers.
 * TEST INFRASTRUCTURE ONLY (see hrrt_oracle.h).
 */
#include "hrrt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define PB(g, buf, row, k) ((buf) + ((size_t)(row) * ((g)->th_pixels / 2 + 1) + (k)) * (size_t)(g)->P * 4 + 4)

static void bp_rescale(const og_geom *g, og_real *prj)
{   /* PJ:1638-1649 */
    const int N = g->N, S = g->segmentsize;
    int k, yr, yr2 = 0, xr;
    for (k = 0; k < g->th_pixels; k++) {
        float rescale = g->cos_theta[k] * g->cos_theta[k] * g->cos_theta[k] * g->p.span * g->ring_spacing /
                        (2 * g->p.crystal_radius);
        for (yr = g->yr_bottom[k]; yr <= g->yr_top[k]; yr++, yr2++)
            for (xr = 0; xr < N; xr++) {
                prj[(size_t)yr2 * N + xr] = prj[(size_t)yr2 * N + xr] * rescale;
                prj[(size_t)(yr2 + S) * N + xr] = prj[(size_t)(yr2 + S) * N + xr] * rescale;
            }
    }
}

/* ------------------------------------------------------------------ */
/* Incremental back projection: back_proj3d_thread1 PJ:1622-1699 and   */
/* back_proj3d_view_n1thread PJ:4162-4668.                             */
/* ------------------------------------------------------------------ */
static void bp_inc_transpose(const og_geom *g, const og_real *prj, og_real *prjbuf)
{   /* PJ:1654-1693 */
    const int N = g->N, S = g->segmentsize, K = g->th_pixels / 2;
    int xr, yr, k, l;
    for (xr = 1; xr <= N / 2; xr++) {
        og_real *ptr = PB(g, prjbuf, xr, 0);
        int xr2 = N - xr;
        for (yr = 0; yr < g->z_pixels; yr++) {
            ptr[4 * yr] = prj[(size_t)yr * N + xr];
            ptr[4 * yr + 1] = prj[(size_t)yr * N + xr2];
            ptr[4 * yr + 2] = prj[(size_t)(S + yr) * N + xr];
            ptr[4 * yr + 3] = prj[(size_t)(S + yr) * N + xr2];
        }
    }
    for (k = 1; k <= K; k++) {
        for (xr = 1; xr <= N / 2; xr++) {
            og_real *ptr = PB(g, prjbuf, xr, k), *ptr2 = PB(g, prjbuf, xr + N / 2, k);
            int y1 = g->segment_offset[k * 2], y2 = g->segment_offset[k * 2 - 1];
            int y3 = y1 + S, y4 = y2 + S, xr2 = N - xr;
            for (yr = g->yr_bottom[k * 2]; yr <= g->yr_top[k * 2]; yr++, y1++, y2++, y3++, y4++) {
                ptr[4 * yr] = prj[(size_t)y1 * N + xr];
                ptr[4 * yr + 1] = prj[(size_t)y1 * N + xr2];
                ptr[4 * yr + 2] = prj[(size_t)y3 * N + xr];
                ptr[4 * yr + 3] = prj[(size_t)y3 * N + xr2];
                ptr2[4 * yr] = prj[(size_t)y2 * N + xr];
                ptr2[4 * yr + 1] = prj[(size_t)y2 * N + xr2];
                ptr2[4 * yr + 2] = prj[(size_t)y4 * N + xr];
                ptr2[4 * yr + 3] = prj[(size_t)y4 * N + xr2];
            }
            for (l = 0; l < 4; l++) {
                ptr[4 * (g->yr_bottom[k * 2] - 1) + l] = ptr2[4 * (g->yr_bottom[k * 2] - 1) + l] = 0;
                ptr[4 * (g->yr_top[k * 2] + 1) + l] = ptr2[4 * (g->yr_top[k * 2] + 1) + l] = 0;
            }
        }
    }
}

static inline og_real *voxw(const og_geom *g, og_real *img, int x, int y)
{
    return img + ((size_t)x * g->N + y) * g->z4;
}

static void bp_inc_view(const og_geom *g, og_real *image, og_real *prjbuf, int view, int xs, int xe)
{
    const int N = g->N, c = N / 2, Z = g->z_pixels, Z4 = g->z4, GM = g->groupmax;
    og_real *s1 = (og_real *)calloc((size_t)4 * (Z4 + 2), sizeof(og_real));
    og_real *s2 = (og_real *)calloc((size_t)4 * (Z4 + 2), sizeof(og_real));
    og_real *d1 = (og_real *)calloc((size_t)4 * (Z4 + 2), sizeof(og_real));
    og_real *d2 = (og_real *)calloc((size_t)4 * (Z4 + 2), sizeof(og_real));
    int x, y, z, z2, z3, z4, yr, yr2, yr3, yr4, th, th2, l;
#define LOOP4(stmt) for (l = 0; l < 4; l++) { stmt; }
    for (x = xs; x < xe; x++) {
        og_real *pp1[64], *pp2[64];
        int ystart = g->cylwiny[2 * x], yend = g->cylwiny[2 * x + 1];
        int mask = 0;
        og_real *i1f, *i2f, *i3f, *i4f, *i5f, *i6f, *i7f, *i8f;
        for (th = 0; th <= g->th_pixels / 2; th++) {
            pp1[th] = PB(g, prjbuf, x, th);
            pp2[th] = PB(g, prjbuf, x + c, th);
        }
        if (ystart == 0 && yend == 0) continue;
        i1f = voxw(g, image, x, c); i2f = voxw(g, image, N - x, c);
        i3f = voxw(g, image, c, x); i4f = voxw(g, image, c, N - x);
        if (g->norm_zero_mask[view * (c + 1) + x] == 0) {
            mask = 1;
            memset(i1f, 0, sizeof(og_real) * Z4);
            if (x != c) {
                memset(i2f, 0, sizeof(og_real) * Z4); memset(i3f, 0, sizeof(og_real) * Z4); memset(i4f, 0, sizeof(og_real) * Z4);
            }
            y = c - 1;
            i1f = voxw(g, image, x, y); i2f = voxw(g, image, N - x, y);
            i3f = voxw(g, image, N - y, x); i4f = voxw(g, image, N - y, N - x);
            i5f = voxw(g, image, x, N - y); i6f = voxw(g, image, N - x, N - y);
            i7f = voxw(g, image, y, x); i8f = voxw(g, image, y, N - x);
            if (x != c) {
                if (x == y) {
                    memset(i1f, 0, sizeof(og_real) * Z4); memset(i2f, 0, sizeof(og_real) * Z4);
                    memset(i5f, 0, sizeof(og_real) * Z4); memset(i6f, 0, sizeof(og_real) * Z4);
                } else if (x < y) {
                    memset(i1f, 0, sizeof(og_real) * Z4); memset(i2f, 0, sizeof(og_real) * Z4);
                    memset(i3f, 0, sizeof(og_real) * Z4); memset(i4f, 0, sizeof(og_real) * Z4);
                    memset(i5f, 0, sizeof(og_real) * Z4); memset(i6f, 0, sizeof(og_real) * Z4);
                    memset(i7f, 0, sizeof(og_real) * Z4); memset(i8f, 0, sizeof(og_real) * Z4);
                }
            }
            y = c - 2;          /* the y loop below starts here and jumps to the write-out */
        } else {
            /* centre voxels PJ:4265-4285 */
            memcpy(s1, pp1[0], sizeof(og_real) * 4 * Z);
            for (th = 2; th < g->th_pixels; th += 2) {
                og_real *p1 = pp1[th / 2], *p2 = pp2[th / 2];
                for (yr = g->yr_bottom[th]; yr <= g->yr_top[th]; yr++)
                    LOOP4(s1[4 * yr + l] = s1[4 * yr + l] + (p1[4 * yr + l] + p2[4 * yr + l]));
            }
            for (z = 0; z < Z; z++) {
                if (x != c) {
                    i1f[z] = s1[4 * z]; i2f[z] = s1[4 * z + 1]; i3f[z] = s1[4 * z + 2]; i4f[z] = s1[4 * z + 3];
                } else {
                    i1f[z] = s1[4 * z] + s1[4 * z + 2];
                }
            }
            /* first row y = c-1: direct interpolation, then convert prjbuf to differences PJ:4287-4335 */
            y = c - 1;
            memset(d1, 0, sizeof(og_real) * 4 * (Z4 + 2));
            memset(d2, 0, sizeof(og_real) * 4 * (Z4 + 2));
            memcpy(s1, pp1[0], sizeof(og_real) * 4 * Z);
            memcpy(s2, pp1[0], sizeof(og_real) * 4 * Z);
            for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
                og_real *p1 = pp1[th / 2], *p2 = pp2[th / 2];
                const og_blookup *l1 = &g->blookup[(size_t)y * GM + th2];
                const og_blookup *l2 = &g->blookup[(size_t)(N - y) * GM + th2];
                const og_blutacc *bc = &g->blutacc[(size_t)(c + 1) * GM + th2];
                float coef1 = l1->yr2a, coef2 = l2->yr2a, tn;
                yr2 = l2->yra; z2 = l2->zsa;
                for (z = l1->zsa, yr = l1->yra; z <= l1->zea; z++, yr++, z2++, yr2++) {
                    LOOP4(s1[4 * z + l] = s1[4 * z + l] + (p1[4 * yr + l] + coef1 * (p1[4 * (yr + 1) + l] - p1[4 * yr + l])));
                    LOOP4(s2[4 * z + l] = s2[4 * z + l] + (p2[4 * yr + l] + coef1 * (p2[4 * (yr + 1) + l] - p2[4 * yr + l])));
                    LOOP4(s1[4 * z2 + l] = s1[4 * z2 + l] + (p2[4 * yr2 + l] + coef2 * (p2[4 * (yr2 + 1) + l] - p2[4 * yr2 + l])));
                    LOOP4(s2[4 * z2 + l] = s2[4 * z2 + l] + (p1[4 * yr2 + l] + coef2 * (p1[4 * (yr2 + 1) + l] - p1[4 * yr2 + l])));
                }
                tn = -g->sin_theta[th] / g->cos_theta[th];
                for (yr = g->yr_bottom[th] - 1; yr <= g->yr_top[th]; yr++) {
                    LOOP4(p1[4 * yr + l] = (p1[4 * (yr + 1) + l] - p1[4 * yr + l]) * tn);
                    LOOP4(p2[4 * yr + l] = (p2[4 * (yr + 1) + l] - p2[4 * yr + l]) * tn);
                    LOOP4(d1[4 * yr + l] = d1[4 * yr + l] + p1[4 * yr + l] * 1.0f);
                    LOOP4(d2[4 * yr + l] = d2[4 * yr + l] + p2[4 * yr + l] * 1.0f);
                }
                for (z = bc->z_start, yr = bc->yr_start; z <= bc->z_end; yr++, z++) {
                    LOOP4(d1[4 * z + l] = d1[4 * z + l] - p2[4 * yr + l] * 1.0f);
                    LOOP4(d2[4 * z + l] = d2[4 * z + l] - p1[4 * yr + l] * 1.0f);
                }
            }
            /* write-out of row c-1 happens through the common code below */
            goto write_out;
        }
        for (; y >= ystart; y--) {
            if (mask) goto write_out;
            for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
                og_real *p1 = pp1[th / 2], *p2 = pp2[th / 2];
                const og_blutacc *ba = &g->blutacc[(size_t)y * GM + th2];
                const og_blutacc *bb = &g->blutacc[(size_t)(N - y) * GM + th2];
                float tn, coef1;
                og_real m1;
                if (g->blutacc_flag[(size_t)y * GM + th2] == 0) continue;
                tn = -g->sin_theta[th] / g->cos_theta[th];
                tn = 1 - tn * (float)ba->num - ba->coef_pre;
                coef1 = -tn * g->cos_theta[th] / g->sin_theta[th];
                z = ba->z_start_pre; z2 = bb->z_start_pre; yr = ba->yr_start_pre; yr2 = bb->yr_start_pre;
                z3 = ba->z_start; z4 = bb->z_start; yr3 = ba->yr_start; yr4 = bb->yr_start;
                if ((z3 - z) != -1) {
                    for (; z <= ba->z_end_pre; z++, z2++, yr++, yr2++) {
                        LOOP4(s1[4 * z + l] = s1[4 * z + l] + p1[4 * yr + l] * coef1);
                        LOOP4(s2[4 * z2 + l] = s2[4 * z2 + l] - p1[4 * yr2 + l] * coef1);
                        LOOP4(s1[4 * z2 + l] = s1[4 * z2 + l] - p2[4 * yr2 + l] * coef1);
                        LOOP4(s2[4 * z + l] = s2[4 * z + l] + p2[4 * yr + l] * coef1);
                        LOOP4(d1[4 * z + l] = d1[4 * z + l] - p1[4 * yr + l] * 1.0f);
                        LOOP4(d1[4 * z2 + l] = d1[4 * z2 + l] + p2[4 * yr2 + l] * 1.0f);
                        LOOP4(d2[4 * z2 + l] = d2[4 * z2 + l] + p1[4 * yr2 + l] * 1.0f);
                        LOOP4(d2[4 * z + l] = d2[4 * z + l] - p2[4 * yr + l] * 1.0f);
                    }
                    for (; z3 <= ba->z_end; z3++, z4++, yr3++, yr4++) {
                        LOOP4(s1[4 * z3 + l] = s1[4 * z3 + l] - p1[4 * yr3 + l] * coef1);
                        LOOP4(s2[4 * z4 + l] = s2[4 * z4 + l] + p1[4 * yr4 + l] * coef1);
                        LOOP4(s1[4 * z4 + l] = s1[4 * z4 + l] + p2[4 * yr4 + l] * coef1);
                        LOOP4(s2[4 * z3 + l] = s2[4 * z3 + l] - p2[4 * yr3 + l] * coef1);
                        LOOP4(d1[4 * z3 + l] = d1[4 * z3 + l] + p1[4 * yr3 + l] * 1.0f);
                        LOOP4(d2[4 * z4 + l] = d2[4 * z4 + l] - p1[4 * yr4 + l] * 1.0f);
                        LOOP4(d2[4 * z3 + l] = d2[4 * z3 + l] + p2[4 * yr3 + l] * 1.0f);
                        LOOP4(d1[4 * z4 + l] = d1[4 * z4 + l] - p2[4 * yr4 + l] * 1.0f);
                    }
                } else {
                    LOOP4(s1[4 * z3 + l] = s1[4 * z3 + l] - p1[4 * yr + l] * coef1);
                    LOOP4(s2[4 * z3 + l] = s2[4 * z3 + l] - p2[4 * yr + l] * coef1);
                    LOOP4(d1[4 * z3 + l] = d1[4 * z3 + l] + p1[4 * yr + l] * 1.0f);
                    LOOP4(d2[4 * z3 + l] = d2[4 * z3 + l] + p2[4 * yr + l] * 1.0f);
                    LOOP4(s2[4 * z2 + l] = s2[4 * z2 + l] - p1[4 * yr + l] * coef1);
                    LOOP4(s1[4 * z2 + l] = s1[4 * z2 + l] - p2[4 * yr + l] * coef1);
                    LOOP4(d2[4 * z2 + l] = d2[4 * z2 + l] + p1[4 * yr + l] * 1.0f);
                    LOOP4(d1[4 * z2 + l] = d1[4 * z2 + l] + p2[4 * yr + l] * 1.0f);
                    z2++;
                    for (; z < ba->z_end_pre; yr++, z++, z2++) {
                        LOOP4(m1 = p1[4 * yr + l] - p1[4 * (yr + 1) + l];
                              s1[4 * z + l] = s1[4 * z + l] + m1 * coef1;
                              s2[4 * z2 + l] = s2[4 * z2 + l] + m1 * coef1;
                              d1[4 * z + l] = d1[4 * z + l] - m1;
                              d2[4 * z2 + l] = d2[4 * z2 + l] - m1;
                              m1 = p2[4 * yr + l] - p2[4 * (yr + 1) + l];
                              s1[4 * z2 + l] = s1[4 * z2 + l] + m1 * coef1;
                              s2[4 * z + l] = s2[4 * z + l] + m1 * coef1;
                              d2[4 * z + l] = d2[4 * z + l] - m1;
                              d1[4 * z2 + l] = d1[4 * z2 + l] - m1);
                    }
                    LOOP4(s1[4 * z + l] = s1[4 * z + l] + p1[4 * yr + l] * coef1);
                    LOOP4(s2[4 * z + l] = s2[4 * z + l] + p2[4 * yr + l] * coef1);
                    LOOP4(d1[4 * z + l] = d1[4 * z + l] - p1[4 * yr + l] * 1.0f);
                    LOOP4(d2[4 * z + l] = d2[4 * z + l] - p2[4 * yr + l] * 1.0f);
                    LOOP4(s2[4 * z2 + l] = s2[4 * z2 + l] + p1[4 * yr + l] * coef1);
                    LOOP4(s1[4 * z2 + l] = s1[4 * z2 + l] + p2[4 * yr + l] * coef1);
                    LOOP4(d2[4 * z2 + l] = d2[4 * z2 + l] - p1[4 * yr + l] * 1.0f);
                    LOOP4(d1[4 * z2 + l] = d1[4 * z2 + l] - p2[4 * yr + l] * 1.0f);
                }
            }
            for (z = 0; z < Z; z++)
                LOOP4(s1[4 * z + l] = s1[4 * z + l] + d1[4 * z + l]; s2[4 * z + l] = s2[4 * z + l] + d2[4 * z + l]);
        write_out:
            i1f = voxw(g, image, x, y); i2f = voxw(g, image, N - x, y);
            i3f = voxw(g, image, N - y, x); i4f = voxw(g, image, N - y, N - x);
            i5f = voxw(g, image, x, N - y); i6f = voxw(g, image, N - x, N - y);
            i7f = voxw(g, image, y, x); i8f = voxw(g, image, y, N - x);
            if (x != c) {
                if (x == y) {
                    if (mask) {
                        memset(i1f, 0, sizeof(og_real) * Z4); memset(i2f, 0, sizeof(og_real) * Z4);
                        memset(i5f, 0, sizeof(og_real) * Z4); memset(i6f, 0, sizeof(og_real) * Z4);
                    } else {
                        for (z = 0; z < Z; z++) {
                            i1f[z] = s1[4 * z] + s2[4 * z + 2];
                            i2f[z] = s1[4 * z + 1] + s1[4 * z + 2];
                            i5f[z] = s2[4 * z] + s2[4 * z + 3];
                            i6f[z] = s1[4 * z + 3] + s2[4 * z + 1];
                        }
                    }
                } else if (x < y) {
                    if (mask) {
                        memset(i1f, 0, sizeof(og_real) * Z4); memset(i2f, 0, sizeof(og_real) * Z4);
                        memset(i3f, 0, sizeof(og_real) * Z4); memset(i4f, 0, sizeof(og_real) * Z4);
                        memset(i5f, 0, sizeof(og_real) * Z4); memset(i6f, 0, sizeof(og_real) * Z4);
                        memset(i7f, 0, sizeof(og_real) * Z4); memset(i8f, 0, sizeof(og_real) * Z4);
                    } else {
                        for (z = 0; z < Z; z++) {
                            i1f[z] = s1[4 * z]; i2f[z] = s1[4 * z + 1]; i3f[z] = s1[4 * z + 2]; i4f[z] = s1[4 * z + 3];
                            i5f[z] = s2[4 * z]; i6f[z] = s2[4 * z + 1]; i7f[z] = s2[4 * z + 2]; i8f[z] = s2[4 * z + 3];
                        }
                    }
                } else if (!mask) {
                    for (z = 0; z < Z; z++) {
                        i1f[z] += s1[4 * z]; i2f[z] += s1[4 * z + 1]; i3f[z] += s1[4 * z + 2]; i4f[z] += s1[4 * z + 3];
                        i5f[z] += s2[4 * z]; i6f[z] += s2[4 * z + 1]; i7f[z] += s2[4 * z + 2]; i8f[z] += s2[4 * z + 3];
                    }
                }
            } else if (!mask) {
                for (z = 0; z < Z; z++) {
                    i1f[z] += s1[4 * z]; i3f[z] += s1[4 * z + 2]; i5f[z] += s2[4 * z]; i7f[z] += s2[4 * z + 2];
                }
            }
        }
        /* restore zero guards PJ:4657-4665 */
        for (th = 2; th < g->th_pixels; th += 2) {
            LOOP4(pp1[th / 2][4 * (g->yr_bottom[th] - 1) + l] = 0; pp1[th / 2][4 * (g->yr_top[th] + 1) + l] = 0;
                  pp2[th / 2][4 * (g->yr_bottom[th] - 1) + l] = 0; pp2[th / 2][4 * (g->yr_top[th] + 1) + l] = 0);
        }
    }
#undef LOOP4
    free(s1); free(s2); free(d1); free(d2);
}

void og_back_proj_incremental(const og_geom *g, og_real *prj, og_real *imagebuf, int view, og_real *prjbuf)
{
    bp_rescale(g, prj);
    bp_inc_transpose(g, prj, prjbuf);
    bp_inc_view(g, imagebuf, prjbuf, view, 0, g->N / 2 + 1);
}

/* ------------------------------------------------------------------ */
/* per-bin and per-voxel steps                                         */
/* ------------------------------------------------------------------ */
void og_ratio_w3(const og_geom *g, og_real *est, const og_real *add, const int16_t *prompt)
{   /* pt_update_est_w3 S3:2493-2505 */
    const float eps3 = g->p.epsilon3;
    size_t i, n = (size_t)2 * g->segmentsize * g->N;
    for (i = 0; i < n; i++) {
        og_real den = est[i] + add[i];
        if (prompt[i] == 0) est[i] = 0;
        else if (den > eps3) est[i] = prompt[i] / den;
        else est[i] = 1.0f;
    }
}

void og_ratio_w012(const og_geom *g, og_real *est, const og_real *data)
{   /* pt_update_est_w012 S3:2436-2447 (UW, AW and ANW weighting) */
    const float eps3 = g->p.epsilon3;
    size_t i, n = (size_t)2 * g->segmentsize * g->N;
    for (i = 0; i < n; i++) {
        if (est[i] > eps3) est[i] = data[i] / est[i];
        else est[i] = 0.0f;
    }
}

void og_update_image(const og_geom *g, og_real *image, og_real *correction, const og_real *normfac)
{   /* pt_cal_updateimage S3:2033-2056 */
    const int N = g->N, Z4 = g->z4;
    const float amax = 100.00f, amin = 0.01f, coef = g->p.image_max;
    int x, y, z;
    for (x = 0; x < N; x++)
        for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++) {
            size_t o = ((size_t)x * N + y) * Z4;
            for (z = 0; z < Z4; z++) {
                og_real cr = correction[o + z];
                og_real cl = amin > cr ? amin : cr;     /* _mm_max_ps(amin, corr) */
                og_real t;
                cl = amax < cl ? amax : cl;           /* _mm_min_ps(amax, .)    */
                t = image[o + z] * (cl * normfac[o + z]);
                image[o + z] = coef < t ? coef : t;   /* _mm_min_ps(image_max, .) */
                correction[o + z] = 0;
            }
        }
}

void og_normfac_invert(const og_geom *g, const og_real *sens, og_real *out)
{   /* pt_normfac_calculate S3:1870-1889 */
    const int N = g->N, Z4 = g->z4;
    int x, y, z;
    for (x = 0; x < N; x++)
        for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++) {
            size_t o = ((size_t)x * N + y) * Z4;
            for (z = 0; z < g->z_pixels; z++)
                out[o + z] = (sens[o + z] < 0 ? -sens[o + z] : sens[o + z]) > g->p.epsilon1 ? (1.0f / sens[o + z]) : 0.0f;
        }
}

void og_init_image(const og_geom *g, og_real *image)
{   /* Init_Image S3:3997-4005 */
    const int N = g->N, Z4 = g->z4;
    int x, y, z;
    memset(image, 0, sizeof(og_real) * (size_t)N * N * Z4);
    for (x = 0; x < N; x++) {
        if (g->cylwiny[2 * x] == g->cylwiny[2 * x + 1]) continue;
        for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++)
            for (z = g->z_min; z < g->z_max; z++) image[((size_t)x * N + y) * Z4 + z] = 1.0f;
    }
}

void og_end_iteration(const og_geom *g, og_real *image)
{   /* S3:4288-4297 */
    const int N = g->N, Z4 = g->z4;
    int x, y, z;
    for (x = 0; x < N; x++)
        for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++)
            for (z = g->z_min; z < g->z_max; z++) {
                og_real *v = &image[((size_t)x * N + y) * Z4 + z];
                if (*v < 1e-16) *v = 0;
            }
}

/* ------------------------------------------------------------------ */
/* drivers                                                              */
/* ------------------------------------------------------------------ */
void og_gather_pair_f(const og_geom *g, const og_real *sino, int view, og_real *out)
{   /* largeprj[v][j] = largeprj3[j][v], largeprj[v][j+S] = largeprj3[j][v+views/2] (S3:2149-2154) */
    const int N = g->N, S = g->segmentsize, V = g->views;
    int j;
    for (j = 0; j < S; j++) {
        memcpy(out + (size_t)j * N, sino + ((size_t)j * V + view) * N, sizeof(og_real) * N);
        memcpy(out + (size_t)(j + S) * N, sino + ((size_t)j * V + view + V / 2) * N, sizeof(og_real) * N);
    }
}

void og_gather_pair_s(const og_geom *g, const int16_t *sino, int view, int16_t *out)
{
    const int N = g->N, S = g->segmentsize, V = g->views;
    int j;
    for (j = 0; j < S; j++) {
        memcpy(out + (size_t)j * N, sino + ((size_t)j * V + view) * N, sizeof(int16_t) * N);
        memcpy(out + (size_t)(j + S) * N, sino + ((size_t)j * V + view + V / 2) * N, sizeof(int16_t) * N);
    }
}

static void scatter_pair_f(const og_geom *g, og_real *sino, int view, const og_real *in)
{
    const int N = g->N, S = g->segmentsize, V = g->views;
    int j;
    for (j = 0; j < S; j++) {
        memcpy(sino + ((size_t)j * V + view) * N, in + (size_t)j * N, sizeof(og_real) * N);
        memcpy(sino + ((size_t)j * V + view + V / 2) * N, in + (size_t)(j + S) * N, sizeof(og_real) * N);
    }
}

/* simple fork/join over tasks */
typedef struct {
    void (*fn)(void *ctx, int task);
    void *ctx;
    int ntask, nworker, id;
} pool_arg;

static void *pool_main(void *a)
{
    pool_arg *p = (pool_arg *)a;
    int t;
    for (t = p->id; t < p->ntask; t += p->nworker) p->fn(p->ctx, t);
    return NULL;
}

static void run_tasks(void (*fn)(void *, int), void *ctx, int ntask, int workers)
{
    pthread_t th[64];
    pool_arg args[64];
    int i, nw = workers < 1 ? 1 : (workers > 64 ? 64 : workers);
    if (nw > ntask) nw = ntask;
    if (nw <= 1) {
        for (i = 0; i < ntask; i++) fn(ctx, i);
        return;
    }
    for (i = 0; i < nw; i++) {
        args[i].fn = fn; args[i].ctx = ctx; args[i].ntask = ntask; args[i].nworker = nw; args[i].id = i;
        pthread_create(&th[i], NULL, pool_main, &args[i]);
    }
    for (i = 0; i < nw; i++) pthread_join(th[i], NULL);
}

typedef struct {
    const og_geom *g;
    int *views;          /* view of each task */
    int *incremental;    /* 1: *_thread1 variant, 0: *2 variant */
    og_real **est, **imagebuf, **prjbuf;
    const og_real *add;    /* OP: additive term; UW/AW/ANW: the corrected data (prompt == NULL) */
    const int16_t *prompt;
    og_real *invnorm;      /* sensitivity mode: rows rescaled in place */
    const og_real *image;
    int sens;
} pair_job;

static void pair_task(void *vctx, int t)
{
    pair_job *j = (pair_job *)vctx;
    const og_geom *g = j->g;
    const size_t n2 = (size_t)2 * g->segmentsize * g->N;
    int view = j->views[t];
    og_real *est = j->est[t];
    memset(j->imagebuf[t], 0, sizeof(og_real) * (size_t)g->N * g->N * g->z4);
    if (j->sens) {
        og_gather_pair_f(g, j->invnorm, view, est);
    } else {
        og_real *add = (og_real *)malloc(sizeof(og_real) * n2);
        og_gather_pair_f(g, j->add, view, add);
        og_forward_proj(g, j->image, est, view, j->prjbuf[t]);
        if (j->prompt) {   /* pt_osem3d_proj_thread1 S3:1915-1924 */
            int16_t *pr = (int16_t *)malloc(sizeof(int16_t) * n2);
            og_gather_pair_s(g, j->prompt, view, pr);
            og_ratio_w3(g, est, add, pr);
            free(pr);
        } else {
            og_ratio_w012(g, est, add);
        }
        free(add);
    }
    if (j->incremental[t]) og_back_proj_incremental(g, est, j->imagebuf[t], view, j->prjbuf[t]);
    else og_back_proj_direct_rotframe(g, est, j->imagebuf[t], view, j->prjbuf[t]);
}

typedef struct { const og_geom *g; og_real **bufs; int *views; int nbuf; og_real *out; int workers; } rot_job;

static void rot_task(void *vctx, int t)
{   /* each task owns a range of x and adds all buffers in order */
    rot_job *r = (rot_job *)vctx;
    int N = r->g->N, per = (N + r->workers - 1) / r->workers, xs = t * per, xe = xs + per, b;
    if (xe > N) xe = N;
    for (b = 0; b < r->nbuf; b++) og_rotate_add(r->g, r->bufs[b], r->out, r->views[b], xs, xe);
}

/* Run the view pairs of subset `isubset`, adding their rotated BPs into acc. */
static void run_subset_pairs(const og_geom *g, int isubset, int nthreads, int workers, int sens,
                             og_real *invnorm, const og_real *image, const og_real *add, const int16_t *prompt,
                             og_real *acc)
{
    const int npairs = g->sviews / 2;
    const size_t n2 = (size_t)2 * g->segmentsize * g->N;
    int views[256], inc[256], i;
    og_real *est[256], *ib[256], *pb[256];
    pair_job job;
    rot_job rj;
    int nfull = 0;
    if (npairs > 256) return;
    /* the reference runs *_thread1 in chunks of nthreads and *2 for the remainder (S3:3796-3825, 4174-4218) */
    if (nthreads > 0) nfull = npairs - npairs % nthreads;
    for (i = 0; i < npairs; i++) {
        views[i] = g->vieworder[isubset * g->sviews + i];
        inc[i] = i < nfull;
        est[i] = (og_real *)calloc(n2, sizeof(og_real));
        ib[i] = (og_real *)calloc((size_t)g->N * g->N * g->z4, sizeof(og_real));
        pb[i] = (og_real *)calloc((size_t)og_prjbuf_floats(g), sizeof(og_real));
    }
    job.g = g; job.views = views; job.incremental = inc; job.est = est; job.imagebuf = ib; job.prjbuf = pb;
    job.add = add; job.prompt = prompt; job.invnorm = invnorm; job.image = image; job.sens = sens;
    run_tasks(pair_task, &job, npairs, workers);
    rj.g = g; rj.bufs = ib; rj.views = views; rj.nbuf = npairs; rj.out = acc; rj.workers = workers < 1 ? 1 : workers;
    run_tasks(rot_task, &rj, rj.workers, workers);
    for (i = 0; i < npairs; i++) {
        if (sens) scatter_pair_f(g, invnorm, views[i], est[i]);
        free(est[i]); free(ib[i]); free(pb[i]);
    }
}

void og_calculate_normfac(const og_geom *g, og_real *invnorm, og_real *normfac_all, int nthreads)
{   /* CalculateNormfac S3:3787-3855 */
    const size_t V = (size_t)g->N * g->N * g->z4;
    og_real *sens = (og_real *)malloc(sizeof(og_real) * V);
    int s;
    for (s = 0; s < g->p.subsets; s++) {
        memset(sens, 0, sizeof(og_real) * V);
        run_subset_pairs(g, s, nthreads, nthreads > 0 ? nthreads : 1, 1, invnorm, NULL, NULL, NULL, sens);
        memset(normfac_all + (size_t)s * V, 0, sizeof(og_real) * V);
        og_normfac_invert(g, sens, normfac_all + (size_t)s * V);
    }
    free(sens);
}

void og_calculate_sensitivity(const og_geom *g, og_real *invnorm, og_real *sens_all, int nthreads)
{   /* CalculateNormfac S3:3787-3830 before the inversion: the back-projected weights per subset */
    const size_t V = (size_t)g->N * g->N * g->z4;
    int s;
    for (s = 0; s < g->p.subsets; s++) {
        memset(sens_all + (size_t)s * V, 0, sizeof(og_real) * V);
        run_subset_pairs(g, s, nthreads, nthreads > 0 ? nthreads : 1, 1, invnorm, NULL, NULL, NULL,
                         sens_all + (size_t)s * V);
    }
}

void og_osem_subset(const og_geom *g, og_real *image, og_real *correction, const og_real *normfac_subset,
                    const og_real *add, const int16_t *prompt, int isubset, int nthreads, int workers)
{   /* CalculateOsem3d S3:4174-4245 */
    run_subset_pairs(g, isubset, nthreads, workers, 0, NULL, image, add, prompt, correction);
    og_update_image(g, image, correction, normfac_subset);
}

void og_osem_subset_w012(const og_geom *g, og_real *image, og_real *correction, const og_real *normfac_subset,
                         const og_real *data, int isubset, int nthreads, int workers)
{   /* CalculateOsem3d S3:4174-4245 with update_estimate_w012 (-W 0, 1, 2) */
    run_subset_pairs(g, isubset, nthreads, workers, 0, NULL, image, data, NULL, correction);
    og_update_image(g, image, correction, normfac_subset);
}

void og_calculate_normfac_psf(const og_geom *g, const og_psf *p, og_real *invnorm, og_real *normfac_all, int nthreads)
{   /* CalculateNormfac with -B: convolve3d(normfac, normfac) before pt_normfac_calculate S3:3835-3840 */
    const size_t V = (size_t)g->N * g->N * g->z4;
    og_real *sens = (og_real *)malloc(sizeof(og_real) * V);
    og_real *buf = (og_real *)calloc(V, sizeof(og_real));
    int s;
    for (s = 0; s < g->p.subsets; s++) {
        memset(sens, 0, sizeof(og_real) * V);
        run_subset_pairs(g, s, nthreads, nthreads > 0 ? nthreads : 1, 1, invnorm, NULL, NULL, NULL, sens);
        og_convolve3d(g, p, sens, sens, buf);
        memset(normfac_all + (size_t)s * V, 0, sizeof(og_real) * V);
        og_normfac_invert(g, sens, normfac_all + (size_t)s * V);
    }
    free(buf);
    free(sens);
}

void og_osem_subset_psf(const og_geom *g, const og_psf *p, og_real *image, og_real *correction,
                        const og_real *normfac_subset, const og_real *add, const int16_t *prompt, int isubset,
                        int nthreads, int workers)
{   /* CalculateOsem3d with -B: image_psf for the forward projector S3:4172-4202, the
     * correction blurred before the update S3:4229 */
    const size_t V = (size_t)g->N * g->N * g->z4;
    og_real *image_psf = (og_real *)calloc(V, sizeof(og_real));
    og_real *buf = (og_real *)calloc(V, sizeof(og_real));
    og_convolve3d(g, p, image, image_psf, buf);
    run_subset_pairs(g, isubset, nthreads, workers, 0, NULL, image_psf, add, prompt, correction);
    og_convolve3d(g, p, correction, correction, buf);
    og_update_image(g, image, correction, normfac_subset);
    free(buf);
    free(image_psf);
}
