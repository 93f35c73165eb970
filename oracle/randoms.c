/* Smoothed randoms from a coincidence histogram (gen_delays), restated for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never part of the product).
 *
 * Crystal geometry
 *   geometry_info.h:30-73            HRRT constants, block sizes, module pairs
 *   geometry_info.cpp:137-163        calc_det_to_phy: crystal centre, double math stored as float
 *   geometry_info.cpp:169-221        init_geometry_hrrt: pitch/diameter/thickness defaults, LR types
 * Segment tables
 *   segment_info_fromarchive.cpp:23-141  init_seginfo, init_seginfo2, init_segment_info
 * Module-pair LUT
 *   lor_sinogram_map.cpp:99-138      init_phy_to_pro
 *   lor_sinogram_map.cpp:141-221     init_sol: LUT, crystal z tables, segment-plane map
 *   lor_sinogram_map.cpp:305-340     save_lut_sol: LUT file layout
 * Singles from delayed coincidences
 *   gen_delays_lib.cpp:162-271       compute_drate, estimate_srate, compute_initial_srate,
 *                                    errtotal, compute_csings_from_drates
 * Smoothing
 *   gen_delays_fromarchive.cpp:147-233  compute_delays
 *   gen_delays_lib.cpp:360-541           module-pair rounds, plane-major output
 *
 * Where the current sources are undefined the restatement takes the evident intent, which
 * the earlier sources kept beside them spell out:
 *   - the A-crystal singles index (gen_delays_lib.cpp:117 adds ay*NUM_CRYSTALS_X_Y_HEADS and
 *     drops the head, leaving the array for ay > 1) is [alayer][ay][ahead][ax], as at
 *     gen_delays_fromarchive.cpp:191;
 *   - the segment tables (segment_info.cpp:40-93 writes through reserve()d vectors and sets
 *     m_segz0[0] for every segment) follow segment_info_fromarchive.cpp:35-141;
 *   - the crystal table index (geometry_info.cpp:194, never initialised) starts at 0, so
 *     det_to_phy returns calc_det_to_phy's position (the 2009 fix noted in
 *     geometry_info_fromarchive.cpp:6);
 *   - the error sum of the singles iteration (an int that overflows on large errors; it is
 *     only compared with 0) is accumulated in 64 bits;
 *   - the reference's four threads add into one array without synchronisation; the
 *     restatement adds sequentially, module pairs in the reference's round order;
 *   - a (segment, plane) outside the 63 x 207 map (reachable only for non-default geometry)
 *     is skipped instead of read out of bounds.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hrrt_oracle.h"

#define NH OD_NHEADS
#define NX OD_NX
#define NY OD_NY
#define XYH (NX * NY * NH)          /* crystals of one layer, [y][head][x] */
#define HX (NH * NX)                /* one crystal row over the heads */

/* HRRT_MPAIRS geometry_info.h:68-73 */
static const int MPAIRS[OD_NMP + 1][2] = {
    {-1, -1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {0, 6}, {1, 3}, {1, 4}, {1, 5}, {1, 6}, {1, 7},
    {2, 4}, {2, 5}, {2, 6}, {2, 7}, {3, 5}, {3, 6}, {3, 7}, {4, 6}, {4, 7}, {5, 7}};

/* conversiontable segment_info_fromarchive.cpp:23-26: span-3 segment -> span-9 segment */
static const int S3_TO_S9[45] = {0, 0, 0, 1, 2, 1, 2, 1, 2, 3, 4, 3, 4, 3, 4, 5, 6, 5, 6, 5, 6, 7, 8,
                                 7, 8, 7, 8, 9, 10, 9, 10, 9, 10, 11, 12, 11, 12, 11, 12, 13, 14, 13,
                                 14, 13, 14};

/* mporder gen_delays_lib.cpp:360-362, run as 5 rounds of 4 module pairs (:471-475) */
static const int MP_ROUNDS[OD_NMP] = {1, 6, 11, 15, 2, 7, 12, 16, 3, 8, 13, 17, 4, 9, 14, 19, 5, 10, 18, 20};

int og_delays_mpair(int mp, int *ahead, int *bhead)
{
    if (mp < 1 || mp > OD_NMP) return -1;
    *ahead = MPAIRS[mp][0];
    *bhead = MPAIRS[mp][1];
    return 0;
}

/* calc_det_to_phy geometry_info.cpp:137-163 (constants geometry_info.h:51-55 are floats) */
void og_delays_crystal(const og_delays *g, int head, int layer, int detx, int dety, float pos[3])
{
    const float csize = (float)0.22, cgap = (float)0.02, bgap = (float)0.05;
    const float bsize = (float)(8.0 * csize + (float)(7 * cgap));
    const float xhsize = (float)(9.0 * bsize + (float)(8 * bgap));
    const double sint = g->sin_head[head], cost = g->cos_head[head];
    double x, y, z, xpos, ypos;
    int bcrys = detx % 8, blk = detx / 8;
    float fx = (float)blk * (bsize + bgap);
    fx = fx + (float)bcrys * (csize + cgap);
    x = (double)fx - xhsize / 2.0 + csize / 2.0;
    y = g->crystal_radius + (double)(g->depth * (float)(1 - layer)) + 0.5 * g->depth;
    bcrys = dety % 8;
    blk = dety / 8;
    {
        float fz = (float)blk * (bsize + bgap);
        fz = fz + (float)bcrys * (csize + cgap);
        z = fz;
    }
    xpos = x * cost + y * sint;
    ypos = -x * sint + y * cost;
    pos[0] = (float)xpos;
    pos[1] = (float)ypos;
    pos[2] = (float)z;
}

/* init_phy_to_pro lor_sinogram_map.cpp:99-138 */
void og_delays_phy_to_pro(const og_delays *g, const float a[3], const float b[3], og_sol *sol)
{
    const float pi = (float)M_PI;
    const float fdy = a[1] - b[1], fdx = a[0] - b[0];
    const double dy = fdy, dx = fdx;
    double d = sqrt(dx * dx + dy * dy);
    const float cross = a[1] * b[0] - a[0] * b[1];
    float r = (float)(cross / d);
    float phi = (float)atan2(dx, dy);
    int view, bin;
    float z;
    if (phi < 0.0f) {
        phi = phi + pi;
        r = -r;
        d = -d;
    }
    if ((double)phi >= M_PI) {
        phi = 0.0f;
        r = -r;
        d = -d;
    }
    view = (int)((double)((float)g->nviews * phi) / M_PI);
    if (view >= 288) view = 0;
    if (view <= -1) view = 0;
    z = (float)((a[0] * dx + a[1] * dy) / (d * d));
    bin = (int)(g->nprojs / 2 + ((double)r / g->binsize + 0.5));
    sol->nsino = (bin < 0 || bin > g->nprojs - 1) ? -1 : bin + view * g->nprojs;
    sol->d = (float)(1.0 / d / g->d_tan_theta);
    sol->z = (float)(z / g->plane_sep);
}

/* init_seginfo / init_seginfo2 segment_info_fromarchive.cpp:35-116 for one span */
static int seginfo(int span, int maxrd, int *segz0, int *segzmax, int *segzoffset, int cap)
{
    const int nsegs = 2 * (maxrd / span) + 1, np = 2 * NY - 1, sp2 = (span + 1) / 2;
    int i, nplanes = 0, prev_off = 0, prev_nz = 0;
    if (nsegs > cap) return -1;
    for (i = 0; i < nsegs; i++) {
        const int z0 = i == 0 ? 0 : sp2 + span * ((i - 1) / 2);
        const int nz = np - 2 * z0;
        const int off = i == 0 ? 0 : prev_off + prev_nz;
        segz0[i] = z0;
        segzmax[i] = z0 + nz - 1;
        segzoffset[i] = -z0 + off;
        nplanes += nz;
        prev_off = off;
        prev_nz = nz;
    }
    return nplanes;
}

og_delays *og_delays_create(int span, int maxrd, int lr_type, float pitch, float diam, float thick)
{
    og_delays *g;
    int i, plane, mp;
    float pos[3];
    if ((span != 3 && span != 9) || maxrd < 0 || maxrd > 103 || lr_type < 0 || lr_type > 2) return NULL;
    if (2 * (maxrd / 3) + 1 > 45 && span == 9) return NULL;   /* conversion table has 45 entries */
    g = (og_delays *)calloc(1, sizeof(og_delays));
    g->span = span;
    g->maxrd = maxrd;
    g->lr_type = lr_type;
    /* init_geometry_hrrt geometry_info.cpp:169-221 */
    if (!(pitch > 0.0f)) pitch = 0.24375f;
    if (!(diam > 0.0f)) diam = 46.9f;
    if (!(thick > 0.0f)) thick = 1.0f;
    for (i = 0; i < NH; i++) {
        g->sin_head[i] = sin(i * 2.0 * M_PI / NH);
        g->cos_head[i] = cos(i * 2.0 * M_PI / NH);
    }
    g->crystal_radius = diam / 2.0;
    g->depth = thick;
    switch (lr_type) {
    case 0: g->nprojs = 256; g->nviews = 288; g->binsize = pitch / 2.0; g->plane_sep = pitch / 2.0; break;
    case 1: g->nprojs = 160; g->nviews = 144; g->binsize = 0.2f; g->plane_sep = pitch; break;
    default: g->nprojs = 128; g->nviews = 144; g->binsize = pitch; g->plane_sep = pitch; break;
    }
    g->nsino = g->nprojs * g->nviews;
    /* init_segment_info segment_info_fromarchive.cpp:118-141: the plane count is the span's,
       the plane map is span 3 (span-9 offsets through the conversion table) */
    {
        int z0[OD_NSEGMAP], zmax[OD_NSEGMAP], off[OD_NSEGMAP];
        g->nplanes = seginfo(span, maxrd, z0, zmax, off, OD_NSEGMAP);
        g->nsegs = 2 * (maxrd / 3) + 1;
        if (g->nplanes < 0 || seginfo(3, maxrd, g->segz0, g->segzmax, g->segzoffset, OD_NSEGMAP) < 0) {
            free(g);
            return NULL;
        }
        if (span == 9) {
            const int n9 = 2 * (maxrd / 9) + 1;
            for (i = 0; i < g->nsegs; i++) {
                if (S3_TO_S9[i] >= n9) {
                    free(g);
                    return NULL;
                }
                g->segzoffset[i] = off[S3_TO_S9[i]];
            }
        }
        g->d_tan_theta = 3 * g->plane_sep / g->crystal_radius;
    }
    /* init_sol lor_sinogram_map.cpp:141-221 */
    g->sol = (og_sol *)calloc((size_t)(OD_NMP + 1) * OD_NXD * OD_NXD, sizeof(og_sol));
    for (mp = 1; mp <= OD_NMP; mp++) {
        const int ahead = MPAIRS[mp][0], bhead = MPAIRS[mp][1];
        int al, ax, bl, bx;
        for (al = 0; al < OD_NDOIS; al++)
            for (ax = 0; ax < NX; ax++) {
                float pa[3];
                og_delays_crystal(g, ahead, al, ax, 0, pa);
                for (bl = 0; bl < OD_NDOIS; bl++)
                    for (bx = 0; bx < NX; bx++) {
                        float pb[3];
                        og_delays_crystal(g, bhead, bl, bx, 0, pb);
                        og_delays_phy_to_pro(g, pa, pb,
                                             &g->sol[((size_t)mp * OD_NXD + ax + NX * al) * OD_NXD + bx + NX * bl]);
                    }
            }
    }
    for (i = 0; i < NY; i++) {
        og_delays_crystal(g, 0, 0, 0, i, pos);
        g->czpos[i] = pos[2];
        g->czpos2[i] = (float)(pos[2] / g->plane_sep + 0.5);
    }
    for (i = 0; i < OD_NSEGMAP; i++)
        for (plane = 0; plane < OD_NPLANEMAP; plane++) {
            short v = -1;
            if (i < g->nsegs && plane >= g->segz0[i] && plane <= g->segzmax[i]) v = (short)(plane + g->segzoffset[i]);
            g->segplane[i][plane] = v;
        }
    return g;
}

void og_delays_destroy(og_delays *g)
{
    if (!g) return;
    free(g->sol);
    free(g);
}

void og_delays_dims(const og_delays *g, int *out)
{
    out[0] = g->nprojs;
    out[1] = g->nviews;
    out[2] = g->nplanes;
    out[3] = g->nsegs;
    out[4] = g->nsino;
    out[5] = g->span;
}

void og_delays_sol(const og_delays *g, int mp, int *nsino, float *d, float *z)
{
    const og_sol *s = g->sol + (size_t)mp * OD_NXD * OD_NXD;
    int i;
    for (i = 0; i < OD_NXD * OD_NXD; i++) {
        nsino[i] = s[i].nsino;
        d[i] = s[i].d;
        z[i] = s[i].z;
    }
}

void og_delays_tables(const og_delays *g, float *zpos, float *zpos2, short *segplane)
{
    memcpy(zpos, g->czpos, sizeof(g->czpos));
    memcpy(zpos2, g->czpos2, sizeof(g->czpos2));
    memcpy(segplane, g->segplane, sizeof(g->segplane));
}

/* Total singles rate of each head; crystals are [layer][y][head][x], summed layer, x, y
   (compute_drate gen_delays_lib.cpp:168-176) */
static void head_sums(const float *srate, float *hsum)
{
    int head, layer, cx, cy;
    for (head = 0; head < NH; head++) {
        hsum[head] = 0.0f;
        for (layer = 0; layer < OD_NDOIS; layer++)
            for (cx = 0; cx < NX; cx++)
                for (cy = 0; cy < NY; cy++) hsum[head] += srate[NX * head + cx + HX * cy + XYH * layer];
    }
}

/* the five heads in coincidence with `head` (gen_delays_lib.cpp:180-186) */
static float fan_sum(const float *hsum, int head)
{
    float s = 0.0f;
    int j;
    for (j = 0; j < 5; j++) s += hsum[(head + j + 2) % NH];
    return s;
}

/* compute_drate gen_delays_lib.cpp:162-196 */
void og_delays_drate(const float *srate, float tau, float *drate)
{
    float hsum[NH];
    int head, layer, cx, cy;
    head_sums(srate, hsum);
    for (head = 0; head < NH; head++) {
        const float fan = fan_sum(hsum, head);
        for (layer = 0; layer < OD_NDOIS; layer++)
            for (cx = 0; cx < NX; cx++)
                for (cy = 0; cy < NY; cy++) {
                    const int i = NX * head + cx + HX * cy + XYH * layer;
                    drate[i] = srate[i] * tau * fan;
                }
    }
}

/* estimate_srate gen_delays_lib.cpp:198-230 */
static void estimate_srate(const float *drate, float tau, float *srate)
{
    float hsum[NH];
    int head, layer, cx, cy;
    head_sums(srate, hsum);
    for (head = 0; head < NH; head++) {
        const float fan = fan_sum(hsum, head);
        for (layer = 0; layer < OD_NDOIS; layer++)
            for (cx = 0; cx < NX; cx++)
                for (cy = 0; cy < NY; cy++) {
                    const int i = NX * head + cx + HX * cy + XYH * layer;
                    srate[i] = (srate[i] + drate[i] / (tau * fan)) / 2.0f;
                }
    }
}

/* compute_csings_from_drates gen_delays_lib.cpp:232-271 (with compute_initial_srate, errtotal) */
int og_delays_singles(const int *dcoins, float tau, float dt, float *srates)
{
    float *drates = (float *)malloc(sizeof(float) * OD_NCRYS);
    float *xrates = (float *)malloc(sizeof(float) * OD_NCRYS);
    float dtotal = 0.0f, stotal;
    int i, iter;
    for (i = 0; i < OD_NCRYS; i++) drates[i] = (float)dcoins[i] / dt;
    for (i = 0; i < OD_NCRYS; i++) dtotal += drates[i];
    stotal = (float)(8.0 * sqrt(dtotal / (40. * tau)));
    for (i = 0; i < OD_NCRYS; i++) srates[i] = drates[i] * stotal / dtotal;
    for (iter = 0; iter < 100; iter++) {
        long long err = 0;
        og_delays_drate(srates, tau, xrates);
        for (i = 0; i < OD_NCRYS; i++) {
            const long long e = (int)(dcoins[i] - (0.5 + (double)(xrates[i] * dt)));
            err += e * e;
        }
        if (iter && err == 0) break;
        estimate_srate(drates, tau, srates);
    }
    free(drates);
    free(xrates);
    return iter;
}

/* compute_delays gen_delays_fromarchive.cpp:147-233 for one module pair; adds into
   out[plane][nsino]; returns the number of crystal pairs visited */
static long long smooth_mp(const og_delays *g, int mp, const float *csings, float tauftime, float *out)
{
    const int ahead = MPAIRS[mp][0], bhead = MPAIRS[mp][1];
    const og_sol *solmp = g->sol + (size_t)mp * OD_NXD * OD_NXD;
    const int window = g->maxrd + 6;
    long long pairs = 0;
    float dz2[NY];
    int alayer, ay, ax, blayer, bx, by;
    for (alayer = 0; alayer < OD_NDOIS; alayer++)
        for (ay = 0; ay < NY; ay++) {
            const float cay = g->czpos2[ay];
            int bs = 1000, be = -1000;
            for (by = 0; by < NY; by++) {
                int rd = ay - by;
                if (rd < 0) rd = -rd;
                dz2[by] = g->czpos[by] - g->czpos[ay];
                if (rd < window) {
                    if (bs > by) bs = by;
                    if (be < by) be = by;
                }
            }
            for (ax = 0; ax < NX; ax++) {
                const int axx = ax + NX * alayer;
                const float aw = csings[alayer * XYH + ay * HX + ahead * NX + ax];
                const float awt = aw * tauftime;
                if (awt == 0) continue;
                for (blayer = 0; blayer < OD_NDOIS; blayer++)
                    for (bx = 0; bx < NX; bx++) {
                        const og_sol *sol = &solmp[axx * OD_NXD + bx + NX * blayer];
                        float *col;
                        int nbd;
                        if (sol->nsino == -1) continue;
                        col = out + sol->nsino;
                        nbd = blayer * XYH + bhead * NX + bx + HX * bs;
                        for (by = bs; by <= be; by++, nbd += HX) {
                            const int plane = (int)(cay + sol->z * dz2[by]);
                            const float seg = (float)(0.5 + dz2[by] * sol->d);
                            int segnum = (int)seg;
                            segnum = seg < 0 ? 1 - 2 * segnum : 2 * segnum;
                            pairs++;
                            if (segnum < 0 || segnum >= OD_NSEGMAP || plane < 0 || plane >= OD_NPLANEMAP) continue;
                            if (g->segplane[segnum][plane] != -1)
                                col[(size_t)g->segplane[segnum][plane] * g->nsino] += csings[nbd] * awt;
                        }
                    }
            }
        }
    return pairs;
}

/* gen_delays gen_delays_lib.cpp:360-541: the 20 module pairs (or the nmp listed in mps, in that
   order) into out[nplanes][nviews][nprojs], zeroed first */
long long og_delays_smooth(const og_delays *g, const float *csings, float tau, float ftime, const int *mps, int nmp,
                           float *out)
{
    const float tauftime = tau * ftime;
    long long pairs = 0;
    int k;
    memset(out, 0, sizeof(float) * (size_t)g->nplanes * g->nsino);
    if (!mps) {
        mps = MP_ROUNDS;
        nmp = OD_NMP;
    }
    for (k = 0; k < nmp; k++) {
        if (mps[k] < 1 || mps[k] > OD_NMP) continue;
        pairs += smooth_mp(g, mps[k], csings, tauftime, out);
    }
    return pairs;
}
