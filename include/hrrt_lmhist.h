/*
 * hrrt_lmhist.h -- C ABI of the MI355X (gfx950) list-mode histogrammer (lmhistogram_mp): a
 * stream of 64-bit HRRT list-mode words into sinograms and crystal coincidence histograms.
 *
 * The reference runs a reader thread (lm64_reader, LM_Reader_mp.cpp:124-198), rebinner threads
 * (rebin_packet histogram_mp.cpp:1026-1235, rebin_event LM_Rebinner_mp.cpp:188-219) and a
 * histogram loop (histogram histogram_mp.cpp:1301-1398, process_tagword :225-295) over
 * packets of PACKET_SIZE words (LM_Reader_mp.hpp:39), one frame at a time.  Here one call
 * histograms one frame of a stream already in device memory.  The module-pair LUT and
 * segment tables come from a hrrt_delays context (include/hrrt_randoms.h), as the
 * reference's rebinner and gen_delays share the lor_sinogram tables (LM_Rebinner_mp.cpp:138-142).
 *
 * Layouts:
 *   words      uint32 [nwords]: the .l64 file as 32-bit words (little endian)
 *   sinogram   int16 [nplanes][nviews][nprojs] (the prompts, or net trues for mode 0);
 *              mode 1 without a delayed array: [2][...] (prompts, then delayed)
 *   delayed    uint8 [nplanes][nviews][nprojs] (mode 1), saturating at 255
 *   coin maps  uint32 [2 layers][104 rings][8 heads][72] prompts / delayed: the .ch halves,
 *              the input of hrrt_delays_singles / hrrt_gen_delays
 * Counts wrap as the reference's short (int16) bins do.
 */
#ifndef HRRT_LMHIST_H
#define HRRT_LMHIST_H

#include <stdint.h>

#include "hrrt_randoms.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HRRT_LM_NBLOCKS 936         /* NBLOCKS, geometry_info.h:46 */
#define HRRT_LM_PACKET_WORDS 2097152 /* PACKET_SIZE 64-bit words, LM_Reader_mp.hpp:39 */

typedef struct {
    int mode;          /* HISTOGRAM_MODE (histogram_mp.hpp:41-46): 0 trues = prompts - delayed,
                          1 prompts and delayed apart, 2 prompts only                        */
    int start_ms;      /* frame start: the first time tag >= start_ms (<= 0: the first word)   */
    int end_ms;        /* frame end: the first later time tag >= end_ms, included (<= 0: none) */
    int ignore_border; /* IGNORE_BORDER_CRYSTAL (histogram_mp.hpp:130)                       */
    int no_doi;        /* NODOI_PROCESSING: every event on layer 1                           */
} hrrt_lm_options;

typedef struct {
    long long prompts, randoms;      /* rebinned prompt / delayed events (histogram_mp.cpp:1240-1256) */
    long long events;                /* event words decoded in the window                            */
    long long tags, time_tags;       /* tag words, time tags in the window                           */
    long long rejected;              /* events with an invalid module pair or crystal                */
    long long transmission;          /* events with a transmission (DOI 7) crystal, skipped          */
    long long first_word, end_word;  /* the window [first_word, end_word); -1, -1 if it never starts */
    int last_time_ms;                /* last time tag in the window, -1 if none                      */
    long long block_total[HRRT_LM_NBLOCKS];   /* block singles tags in the window (process_tagword) */
    int block_nsamples[HRRT_LM_NBLOCKS];
    int block_last[HRRT_LM_NBLOCKS];
} hrrt_lm_stats;

/* Histogram one frame of d_words into d_sino (and d_delayed in mode 1), with the prompt and
 * delayed coincidence maps (either may be NULL).  Outputs are overwritten.  stats (optional,
 * host) makes the call synchronous.  Mode 7 (transmission) returns HRRT_ERR_UNSUPPORTED.
 * The context keeps the call's scratch: two 4-byte key slots per two stream words (the binned
 * events, then sorted by sinogram tile) and the radix sort's storage; with coincidence maps also
 * 8 bytes per word of per-head crystal entries.  Streams up to 2^33 - 2^22 words, sinograms up
 * to 2^30 bins. */
int hrrt_lm_histogram(hrrt_delays *ctx, const uint32_t *d_words, long long nwords, const hrrt_lm_options *opt,
                      int16_t *d_sino, uint8_t *d_delayed, uint32_t *d_coin_p, uint32_t *d_coin_d,
                      hrrt_lm_stats *stats, void *stream);

#ifdef __cplusplus
}
#endif
#endif
