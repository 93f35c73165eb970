/* List-mode histogramming (lmhistogram_mp), restated for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never part of the product).
 *
 *   LM_Reader_mp.hpp:39, LM_Reader_mp.cpp:171-198   packets of PACKET_SIZE 64-bit words
 *   histogram_mp.cpp:1026-1235   rebin_packet: word-pair walk (EWTYPES, geometry_info.h:75),
 *                                event decode, validity checks, emission branch, coincidence maps
 *   LM_Rebinner_mp.cpp:188-219   rebin_event: LUT, plane and segment -> sinogram address
 *   histogram_mp.cpp:1301-1398   histogram: TRU (+1 prompt, -1 delayed), PRO_RAN (prompts and
 *                                delayed apart; a delayed byte array saturates at 255), PRO
 *   histogram_mp.cpp:225-295     process_tagword: time tags, block singles
 *   histogram_mp.cpp:754-786, 988-1019   frame start and end at time tags
 *   histogram_mp.cpp:1240-1256   prompts / randoms counters (rebinned events)
 *
 * Restated with these choices:
 *   - the coincidence-map index is [doi][ring][head][x] (ring stride NHEADS*NXCRYS), as in
 *     histogram_mp_fromarchive.cpp:916, 1092; histogram_mp.cpp:1040 uses
 *     NUM_CRYSTALS_X_Y for the ring stride, which leaves the map;
 *   - transmission events (DOI 7) are counted and skipped (transmission mode is not
 *     restated);
 *   - the last, partial packet of the stream is processed (the reference's reader stops at
 *     end of file without it);
 *   - a (segment, plane) outside the segment-plane map is skipped, as in randoms.c.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hrrt_oracle.h"

#define LM_PACKET_WORDS (2 * 1024 * 1024)
#define NBLOCKS 936

/* EWTYPES geometry_info.h:75: 0 prompt, 1 delayed, 2 tag, 3 out of sync */
static const int EWT[16] = {3, 3, 1, 0, 3, 3, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3};

static int word_type(uint32_t ew1, uint32_t ew2)
{
    return EWT[((ew2 & 0xc0000000u) >> 30) | ((ew1 & 0xc0000000u) >> 28)];
}

/* rebin_event LM_Rebinner_mp.cpp:188-219: the sinogram address [plane][nsino] or -1 */
long long og_lm_rebin(const og_delays *g, int mp, int alayer, int ax, int ay, int blayer, int bx, int by)
{
    const og_sol *s = &g->sol[((size_t)mp * OD_NXD + ax + OD_NX * alayer) * OD_NXD + bx + OD_NX * blayer];
    float cay, dz2, seg;
    int plane, segnum;
    if (s->nsino == -1) return -1;
    cay = g->czpos2[ay];
    dz2 = g->czpos[by] - g->czpos[ay];
    plane = (int)(cay + s->z * dz2);
    seg = (float)(0.5 + dz2 * s->d);
    segnum = (int)seg;
    segnum = seg < 0 ? 1 - 2 * segnum : 2 * segnum;
    if (segnum >= g->nsegs) return -1;
    if (segnum < 0 || plane < 0 || plane >= OD_NPLANEMAP) return -1;
    if (g->segplane[segnum][plane] == -1) return -1;
    return (long long)(plane + g->segzoffset[segnum]) * g->nsino + s->nsino;
}

/* The walk over one stream: mode 0 TRU, 1 PRO_RAN, 2 PRO.  sino holds nplanes*nsino int16
   (2x that for PRO_RAN without a delayed array); counts are wrapped to int16 as the
   reference's short increments wrap.  stats: [0] prompts [1] randoms (rebinned events by
   type), [2] events decoded in the window, [3] tags in the window, [4] time tags in the
   window, [5] rejected events (bad module pair / crystal), [6] transmission events,
   [7] first word of the window (its start tag, or 0), [8] end of the window (the word after
   its end tag, or nwords), [9] last time tag (ms) in the window, -1 if none.
   block_total/nsamples/last: block singles tags in the window (NBLOCKS each). */
void og_lm_histogram(const og_delays *g, const uint32_t *w, long long nwords, int mode, int start_ms, int end_ms,
                     int ignore_border, int doi_processing, int16_t *sino, uint8_t *delayed, uint32_t *coin_p,
                     uint32_t *coin_d, long long *stats, long long *block_total, int *block_nsamples,
                     int *block_last)
{
    const long long S = (long long)g->nplanes * g->nsino;
    const int nvox = OD_NX * OD_NY * OD_NHEADS, rowstride = OD_NHEADS * OD_NX;
    const int xhigh = ignore_border ? OD_NX - 1 : OD_NX;
    int started = start_ms <= 0, done = 0;
    long long p0;
    memset(stats, 0, sizeof(long long) * 10);
    stats[7] = started ? 0 : -1;
    stats[8] = nwords;
    stats[9] = -1;
    for (p0 = 0; p0 < nwords && !done; p0 += LM_PACKET_WORDS) {
        const long long count = nwords - p0 < LM_PACKET_WORDS ? nwords - p0 : LM_PACKET_WORDS;
        const uint32_t *in = w + p0;
        long long pos = 0;
        while (pos + 1 < count) {
            const uint32_t ew1 = in[pos], ew2 = in[pos + 1];
            const int type = word_type(ew1, ew2);
            if (type == 3) {
                pos++;
                continue;
            }
            pos += 2;
            if (type == 2) {
                const uint32_t tag = (ew1 & 0xffff) | ((ew2 & 0xffff) << 16);
                if ((tag & 0xE0000000u) == 0x80000000u) {
                    const int t = (int)(tag & 0x3fffffff);
                    if (!started && t >= start_ms) {
                        started = 1;
                        stats[7] = p0 + pos - 2;
                    }
                    if (started) {
                        stats[4]++;
                        stats[9] = t;
                    }
                    if (started && end_ms > 0 && t >= end_ms) {
                        stats[3]++;
                        stats[8] = p0 + pos;        /* the end tag belongs to the frame */
                        done = 1;
                        break;
                    }
                } else if (started && (tag & 0xE0000000u) == 0xA0000000u) {
                    const int block = (int)((tag & 0x1ff80000u) >> 19), v = (int)(tag & 0x0007ffff);
                    if (block < NBLOCKS) {
                        block_total[block] += v;
                        block_nsamples[block]++;
                        block_last[block] = v;
                    }
                }
                if (started) stats[3]++;
                continue;
            }
            if (!started) continue;
            stats[2]++;
            {
                const int mpe = (int)(((ew1 & 0x00070000u) >> 16) | ((ew2 & 0x00070000u) >> 13));
                const int ax = (int)(ew1 & 0xff), ay = (int)((ew1 & 0xff00) >> 8);
                const int bx = (int)(ew2 & 0xff), by = (int)((ew2 & 0xff00) >> 8);
                int doia = 1, doib = 1;
                long long addr = -1;
                if (doi_processing) {
                    doia = (int)((ew1 & 0x01C00000u) >> 22);
                    doib = (int)((ew2 & 0x01C00000u) >> 22);
                }
                if (mpe < 1 || mpe > OD_NMP || ax >= OD_NX || ay >= OD_NY || bx >= OD_NX || by >= OD_NY) {
                    stats[5]++;
                    continue;
                }
                if (doia == 7 || doib == 7) {
                    stats[6]++;
                    continue;
                }
                if (doia > 1 || doib > 1) continue;
                if (ignore_border && (ax * bx == 0 || ax == xhigh || bx == xhigh)) addr = -1;
                else if (mode == 2 && type != 0) addr = -1;
                else addr = og_lm_rebin(g, mpe, doia, ax, ay, doib, bx, by);
                {
                    int ahead, bhead;
                    og_delays_mpair(mpe, &ahead, &bhead);
                    if (type == 0 && coin_p) {
                        coin_p[doia * nvox + ahead * OD_NX + ax + ay * rowstride]++;
                        coin_p[doib * nvox + bhead * OD_NX + bx + by * rowstride]++;
                    } else if (type == 1 && coin_d) {
                        coin_d[doia * nvox + ahead * OD_NX + ax + ay * rowstride]++;
                        coin_d[doib * nvox + bhead * OD_NX + bx + by * rowstride]++;
                    }
                }
                if (addr < 0) continue;
                if (type == 0) stats[0]++;
                else stats[1]++;
                if (mode == 0) {
                    sino[addr] = (int16_t)(uint16_t)((uint16_t)sino[addr] + (type == 0 ? 1u : 0xffffu));
                } else if (mode == 2 || type == 0) {
                    sino[addr] = (int16_t)(uint16_t)((uint16_t)sino[addr] + 1u);
                } else if (delayed) {
                    if (delayed[addr] < 255) delayed[addr]++;
                } else {
                    sino[addr + S] = (int16_t)(uint16_t)((uint16_t)sino[addr + S] + 1u);
                }
            }
        }
    }
    if (!started) stats[7] = stats[8] = -1;
}
