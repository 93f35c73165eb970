/*
 * geometry.c -- oracle restatement of the scanner/sinogram geometry tables.
 * TEST INFRASTRUCTURE ONLY (see hrrt_oracle.h).
 *
 * Follows GetScannerParameter (S3:3452-3551), get_info (PJ:159-191),
 * dependencies (PJ:199-800), calculatevieworder (PJ:144-152) and sequence
 * (S3:535-580).  Float expressions keep the reference's evaluation order;
 * libm calls use the float overloads that a C++ compile of the reference
 * resolves for float arguments (sinf/sqrtf), and double versions where the
 * reference passes a double.
 */
#include "hrrt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* PI is referenced by PJ:480 but defined nowhere in the reference tree;
 * the value of M_PI (double) is used. */
#define OG_PI M_PI

static void get_info(int nrings, int span, int group, int *zmin, int *zmax, int *imin, int *imin1)
{   /* PJ:159-191 (only the outputs used by dependencies) */
    int j;
    if (group == 0) {
        *zmin = 0; *zmax = 2 * nrings - 2; *imin = 0; *imin1 = 0;
        return;
    }
    *zmin = (span + 1) / 2 + (group - 1) * span;
    *zmax = 2 * nrings - 2 - *zmin;
    *imin = 2 * nrings - 1;
    for (j = 1; j < group; ++j)
        *imin += 2 * (2 * nrings - 1 - 2 * ((span + 1) / 2 + (j - 1) * span));
    *imin1 = *imin + (*zmax - *zmin) + 1;
}

int og_sequence(int sets, short *order)
{   /* S3:535-580: greedy maximum circular distance ordering */
    int *done = (int *)calloc((size_t)sets, sizeof(int));
    int i, j, k, d, ibest, dbest;
    if (!done) return -1;
    order[0] = 0;
    done[0] = 1;
    for (i = 1; i < sets; ++i) {
        ibest = 0; dbest = 0;
        for (j = 0; j < sets; ++j) {
            if (done[j]) continue;
            k = order[i - 1];
            d = abs(j - k);
            if (sets - abs(j - k) < d) d = sets - abs(j - k);
            if (d > dbest) { dbest = d; ibest = j; }
        }
        order[i] = (short)ibest;
        done[ibest] = 1;
    }
    free(done);
    return 1;
}

void og_destroy(og_geom *g)
{
    if (!g) return;
    free(g->yr_bottom); free(g->yr_top); free(g->seg_offset); free(g->segment_offset);
    free(g->sin_theta); free(g->cos_theta); free(g->sin_psi); free(g->cos_psi);
    free(g->cylwiny); free(g->vcos); free(g->vsin); free(g->tantheta); free(g->rescalearray);
    free(g->flookup); free(g->blookup); free(g->blutacc); free(g->blutacc_flag);
    free(g->vieworder); free(g->order); free(g->norm_zero_mask);
    free(g);
}

og_geom *og_create(const og_params *p)
{
    og_geom *g = (og_geom *)calloc(1, sizeof(og_geom));
    int N, x, y, v, th, th2, yr, yri, zz, group, zmin, zmax, iminp, iminm, zshift, i;
    float tmp_flt, zz0;
    if (!g) return NULL;
    g->p = *p;

    /* ---- GetScannerParameter S3:3488-3539 ---- */
    g->x_rebin = g->v_rebin = 1;
    if (p->x_pixels != 0 && p->x_pixels < p->def_elements)
        g->x_rebin = g->v_rebin = p->def_angles / p->x_pixels;
    g->sino_sampling = p->bin_size * g->x_rebin;
    g->z_pixels = p->rings * 2 - 1;
    g->z4 = ((g->z_pixels + 3) / 4) * 4;
    g->z_size = p->plane_sep;
    g->ring_spacing = 2.0f * g->z_size;
    g->groupmax = (p->maxdel - (p->span - 1) / 2) / p->span;
    N = p->def_elements / g->x_rebin;
    g->views = p->def_angles / g->v_rebin;
    g->z_min = p->z_min;
    g->z_max = p->z_max;
    if (g->z_max == 0) g->z_max = g->z_pixels - g->z_min;
    if (p->npskip > g->z_min) g->z_min = p->npskip;
    if (g->z_max > g->z_pixels - p->npskip) g->z_max = g->z_pixels - p->npskip;

    /* ---- dependencies PJ:221-300 (zoom == 1, nprojs == nviews == 0) ---- */
    if (p->x_pixels != 0 && p->x_pixels != N) { free(g); return NULL; }
    g->N = N;
    g->x_size = (g->sino_sampling * N) / (1.0f * N);
    g->y_size = g->x_size;
    /* dependencies overwrites z_size with ring_spacing/2 (PJ:262): same value */
    g->z_size = g->ring_spacing / 2.0f;
    if (g->views % p->subsets != 0 || N % 4 != 0) { free(g); return NULL; }
    g->sviews = g->views / p->subsets;
    g->th_pixels = 2 * g->groupmax + 1;
    g->rfov = p->rel_fov * g->x_size * N / 2.0f;
    g->zshift = p->span - 1;
    if (g->groupmax < 1 || g->zshift < 1) { free(g); return NULL; }

    g->yr_bottom = (int *)calloc((size_t)g->th_pixels, sizeof(int));
    g->yr_top = (int *)calloc((size_t)g->th_pixels, sizeof(int));
    g->seg_offset = (int *)calloc((size_t)g->th_pixels, sizeof(int));
    g->segment_offset = (int *)calloc((size_t)g->th_pixels, sizeof(int));
    g->sin_theta = (float *)calloc((size_t)g->th_pixels, sizeof(float));
    g->cos_theta = (float *)calloc((size_t)g->th_pixels, sizeof(float));
    g->sin_psi = (float *)calloc((size_t)g->views, sizeof(float));
    g->cos_psi = (float *)calloc((size_t)g->views, sizeof(float));
    g->cylwiny = (short *)calloc((size_t)N * 2, sizeof(short));
    g->vcos = (float *)calloc((size_t)g->groupmax * g->zshift, sizeof(float));
    g->vsin = (float *)calloc((size_t)g->groupmax * g->zshift, sizeof(float));
    g->norm_zero_mask = (short *)malloc((size_t)(g->views / 2 + 1) * (N / 2 + 1) * sizeof(short));
    for (i = 0; i < (g->views / 2 + 1) * (N / 2 + 1); i++) g->norm_zero_mask[i] = 1;

    /* FOV rows PJ:339-356 */
    {
        int x_off = N / 2, y_off = N / 2;
        for (y = 0; y < N; y++) {
            int flag = 0;
            g->cylwiny[2 * y] = g->cylwiny[2 * y + 1] = 0;
            for (x = 0; x < N; x++) {
                float dx = (x - x_off) * g->x_size * (x - x_off) * g->x_size;
                float dy = (y - y_off) * g->y_size * (y - y_off) * g->y_size;
                if (sqrtf(dx + dy) <= g->rfov) {
                    if (flag == 0) g->cylwiny[2 * y] = (short)x;
                    flag = 1;
                } else {
                    if (flag == 1) g->cylwiny[2 * y + 1] = (short)x;
                    flag = 0;
                }
            }
        }
    }

    /* segment angles PJ:424-444 */
    g->sin_theta[0] = 0.0f;
    g->cos_theta[0] = 1.0f;
    for (group = 1; group <= g->groupmax; group++) {
        tmp_flt = group * p->span * g->ring_spacing;
        tmp_flt = sqrtf(tmp_flt * tmp_flt + (2 * p->crystal_radius) * (2 * p->crystal_radius));
        g->sin_theta[2 * group - 1] = group * p->span * g->ring_spacing / tmp_flt;
        g->cos_theta[2 * group - 1] = (2.0f * p->crystal_radius) / tmp_flt;
        g->sin_theta[2 * group] = -g->sin_theta[2 * group - 1];
        g->cos_theta[2 * group] = g->cos_theta[2 * group - 1];
        for (zz = 0; zz < g->zshift; zz++) {
            int zm = p->span * group - (p->span - 1) / 2;
            int mean_rd = zm + (zz + 1) / 2;
            tmp_flt = mean_rd * g->ring_spacing;
            tmp_flt = sqrtf(tmp_flt * tmp_flt + (2 * p->crystal_radius) * (2 * p->crystal_radius));
            g->vcos[(group - 1) * g->zshift + zz] = (2.0f * p->crystal_radius) / tmp_flt;
            g->vsin[(group - 1) * g->zshift + zz] = mean_rd * g->ring_spacing / tmp_flt;
        }
    }

    /* plane ranges and offsets PJ:446-467 */
    g->yr_bottom[0] = 0;
    g->yr_top[0] = g->z_pixels - 1;
    g->seg_offset[0] = 0;
    for (group = 1; group <= g->groupmax; group++) {
        get_info(p->rings, p->span, group, &zmin, &zmax, &iminp, &iminm);
        g->yr_bottom[2 * group] = zmin;
        g->yr_top[2 * group] = zmax;
        g->seg_offset[2 * group] = iminp * N * g->views;
        g->yr_bottom[2 * group - 1] = zmin;
        g->yr_top[2 * group - 1] = zmax;
        g->seg_offset[2 * group - 1] = iminm * N * g->views;
    }
    g->segmentsize = 0;
    for (v = 0; v < g->th_pixels; v++) {
        g->segment_offset[v] = g->segmentsize;
        g->segmentsize += g->yr_top[v] - g->yr_bottom[v] + 1;
    }

    /* view angles PJ:476-486 */
    g->sin_psi[0] = 0.0f;
    g->cos_psi[0] = 1.0f;
    g->sin_psi[g->views / 2] = 1.0f;
    g->cos_psi[g->views / 2] = 0.0f;
    tmp_flt = (float)(OG_PI / g->views);
    for (v = 1; v < g->views / 2; v++) {
        g->sin_psi[v] = sinf(v * tmp_flt);
        g->cos_psi[v] = (float)sqrt(1. - g->sin_psi[v] * g->sin_psi[v]);
        g->sin_psi[g->views - v] = g->sin_psi[v];
        g->cos_psi[g->views - v] = -g->cos_psi[v];
    }

    /* subset view order PJ:144-152, subset order S3:535 */
    g->vieworder = (int *)calloc((size_t)p->subsets * g->sviews, sizeof(int));
    for (i = 0; i < p->subsets; i++)
        for (v = 0; v < g->sviews; v++) g->vieworder[i * g->sviews + v] = i + p->subsets * v;
    g->order = (short *)calloc((size_t)p->subsets, sizeof(short));
    og_sequence(p->subsets, g->order);

    /* axial tangents and rescale factors PJ:529-542 (zoom == 1) */
    g->tantheta = (float *)calloc((size_t)g->groupmax * g->z_pixels, sizeof(float));
    g->rescalearray = (float *)calloc((size_t)g->groupmax * g->z_pixels, sizeof(float));
    for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
        for (yr = g->yr_bottom[th]; yr <= g->yr_top[th]; yr++) {
            float *tt = &g->tantheta[th2 * g->z_pixels + yr];
            float *rs = &g->rescalearray[th2 * g->z_pixels + yr];
            *tt = g->sin_theta[th] / g->cos_theta[th];
            *rs = g->y_size / fabsf(g->cos_theta[th]);
            zshift = yr - g->yr_bottom[th];
            if (zshift >= g->zshift) zshift = g->yr_top[th] - yr;
            if (zshift < g->zshift) {
                *tt = g->vsin[th2 * g->zshift + zshift] / g->vcos[th2 * g->zshift + zshift];
                *tt = -*tt;
                *rs = g->y_size / fabsf(g->vcos[th2 * g->zshift + zshift]);
            }
            *tt = -*tt / 1.0f;
        }
    }

    /* end-plane forward lookup PJ:544-562 */
    {
        int L = g->groupmax * 2 * g->zshift, y_off = N / 2;
        g->flookup = (og_flookup *)calloc((size_t)N * L, sizeof(og_flookup));
        for (y = 0; y < N; y++) {
            for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
                og_flookup *f = &g->flookup[y * L + th2 * g->zshift * 2];
                yri = 0;
                for (yr = g->yr_bottom[th]; yr < g->yr_bottom[th] + g->zshift; yr++, yri++) {
                    zz0 = yr + ((y - y_off) * g->tantheta[th2 * g->z_pixels + yr]);
                    f[yri].zzi = (int)zz0;
                    f[yri].zzz0 = zz0 - f[yri].zzi;
                }
                for (yr = g->yr_top[th] - g->zshift + 1; yr <= g->yr_top[th]; yr++, yri++) {
                    zz0 = yr + ((y - y_off) * g->tantheta[th2 * g->z_pixels + yr]);
                    f[yri].zzi = (int)zz0;
                    f[yri].zzz0 = zz0 - f[yri].zzi;
                }
            }
        }
    }

    /* back-projection lookup PJ:670-700 */
    g->blookup = (og_blookup *)calloc((size_t)N * g->groupmax, sizeof(og_blookup));
    {
        int y_off = N / 2;
        for (y = 0; y < N; y++) {
            for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
                float tb = -g->sin_theta[th] / g->cos_theta[th];
                float yr0 = -(y - y_off) * tb;
                float yr1 = yr0;
                float yr2 = fabsf(yr0) - (int)fabsf(yr0);
                int yri2 = (int)yr1, zs, ze, yrs;
                if (yr1 < 0) { yr2 = 1 - yr2; yri2 -= 1; }
                zs = g->yr_bottom[th] - 1 - yri2;
                ze = g->yr_top[th] - yri2;
                yrs = g->yr_bottom[th] - 1;
                if (zs < 0) { yrs = yrs - zs; zs = 0; }
                if (ze > g->z_pixels - 1) ze = g->z_pixels - 1;
                g->blookup[y * g->groupmax + th2].zsa = zs;
                g->blookup[y * g->groupmax + th2].zea = ze;
                g->blookup[y * g->groupmax + th2].yra = yrs;
                g->blookup[y * g->groupmax + th2].yr2a = yr2;
            }
        }
    }

    /* incremental back-projection breakpoints PJ:718-784 */
    g->blutacc = (og_blutacc *)calloc((size_t)N * g->groupmax, sizeof(og_blutacc));
    g->blutacc_flag = (int *)calloc((size_t)N * g->groupmax, sizeof(int));
    for (y = 0; y < N * g->groupmax; y++) g->blutacc[y].num = -1;
    for (th = 2, th2 = 0; th < g->th_pixels; th += 2, th2++) {
        int cnt, prev_diff, dd, grp, c = N / 2, dir;
        og_blutacc *b;
        og_blookup *bl;
#define BL(yy) (&g->blookup[(yy) * g->groupmax + th2])
#define BA(yy) (&g->blutacc[(yy) * g->groupmax + th2])
        g->blutacc_flag[c * g->groupmax + th2] = 1;
        b = BA(c); bl = BL(c);
        b->z_start = bl->zsa; b->z_end = bl->zea; b->yr_start = bl->yra; b->coef = bl->yr2a; b->num = 0;
        for (dir = 0; dir < 2; dir++) {
            prev_diff = 0; cnt = 0; grp = c;
            if (dir == 1) g->blutacc_flag[c * g->groupmax + th2] = 0;
            for (y = (dir == 0 ? c + 1 : c - 1); dir == 0 ? y < N : y >= 0; y += (dir == 0 ? 1 : -1)) {
                g->blutacc_flag[y * g->groupmax + th2] = 0;
                dd = BL(y)->yra - BL(y)->zsa;
                if (prev_diff == dd) {
                    cnt++;
                } else {
                    b = BA(y); bl = BL(y);
                    g->blutacc_flag[y * g->groupmax + th2] = cnt;
                    b->z_start = bl->zsa; b->z_end = bl->zea; b->yr_start = bl->yra; b->coef = bl->yr2a;
                    b->num = cnt;
                    b->z_start_pre = BL(grp)->zsa; b->z_end_pre = BL(grp)->zea;
                    b->yr_start_pre = BL(grp)->yra; b->coef_pre = BL(grp)->yr2a;
                    grp = y; cnt = 0; prev_diff = dd;
                }
            }
        }
#undef BL
#undef BA
    }

    /* prjbuf rows carry one zero guard plane on each side of every segment
     * (S3:2767-2769 allocates yr_bottom-1 .. yr_top+11). */
    g->P = g->z4 + 2;
    (void)iminp;
    return g;
}
