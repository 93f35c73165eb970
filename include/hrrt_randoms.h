/*
 * hrrt_randoms.h -- C ABI of the MI355X (gfx950) smoothed randoms ("fast random smoothing",
 * gen_delays_lib), the producer of the randoms input of hrrt_osem3d from a coincidence
 * histogram (.ch).
 *
 * The reference entry point is one function over process globals,
 *   int gen_delays(int is_inline, float scan_duration, float ***result, FILE *coins,
 *                  char *delays_file, int span, int maxrd)      (gen_delays_lib.hpp:15-23)
 * with geometry, tau and the LUT path in globals (gen_delays_lib.hpp:31-39, LUT_FILENAME in
 * lor_sinogram_map.h:25).  This ABI keeps its steps, argument meaning and layouts behind an
 * explicit context:
 *   hrrt_delays_create    init_geometry_hrrt + init_segment_info + init_lut_sol / init_sol
 *                         (gen_delays_lib.cpp:389-393; lor_sinogram_map.cpp:141-303)
 *   hrrt_delays_singles   compute_csings_from_drates (gen_delays_lib.cpp:254-271)
 *   hrrt_delays_smooth    the 20 compute_delays calls + the plane-major output
 *                         (gen_delays_lib.cpp:88-156, 466-541)
 *   hrrt_gen_delays       gen_delays(2, ...) with the histogram in host memory
 *   hrrt_delays_write     gen_delays(1, ...)'s output file (gen_delays_lib.cpp:519-527)
 *   hrrt_delays_save_lut  save_lut_sol (lor_sinogram_map.cpp:305-340)
 *
 * Layouts (contiguous, little endian):
 *   crystals   [2 layers][104 rings][8 heads][72 crystals]          (compute_drate, gen_delays_lib.cpp:173)
 *   .ch        int32 [2][crystals]: prompts, then delayed            (read_coincidence_sinogram_file :309-333)
 *   delays     float [nplanes][nviews][nprojs], segments in file order (gen_delays_lib.cpp:519-527)
 *   LUT file   20 module pairs x [144][144] {int nsino; float d; float z}, then float zpos[104],
 *              float zpos2[104]                                       (save_lut_sol)
 *
 * Errors are status codes of hrrt_osem.h (HRRT_OK, HRRT_ERR_*), with a message from
 * hrrt_last_error(); the reference exits instead.  Device pointers ("d_") are on the current
 * HIP device; `stream` is a hipStream_t (NULL = default stream); calls on one context must
 * not overlap.
 */
#ifndef HRRT_RANDOMS_H
#define HRRT_RANDOMS_H

#include <stdint.h>

#include "hrrt_osem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HRRT_DELAYS_CRYSTALS 119808   /* NUM_CRYSTALS_X_Y_HEADS_DOIS, geometry_info.h:39 */

typedef struct {
    int   span;          /* 3 or 9 (gen_delays -s span,maxrd; default 9)                  */
    int   maxrd;         /* maximum ring difference (default 67)                          */
    int   lr_type;       /* 0: 256 x 288 (default), 1: LR_20 160 x 144, 2: LR_24 128 x 144 */
    float pitch, diam, thick;  /* -g pitch,diam,thick in cm; 0 = 0.24375, 46.9, 1.0       */
    float tau;           /* -T coincidence time parameter in s; 0 = 6e-9                  */
    const char *lut_file;/* rebinner LUT (hrrt_rebinner.lut, LUT_FILENAME); NULL = computed
                            from the geometry as hrrt_rebinner_lut does (init_sol)          */
} hrrt_delays_params;

typedef struct hrrt_delays hrrt_delays;

/* Build the crystal geometry, segment tables and module-pair LUT on the host (no device
 * call; the first hrrt_delays_smooth uploads them to the then current device, which every
 * later call on the context must use). */
int  hrrt_delays_create(const hrrt_delays_params *params, hrrt_delays **out);
void hrrt_delays_destroy(hrrt_delays *ctx);

/* Sinogram sizes; pairs_full = crystal pairs a smoothing visits when no crystal is silent
 * (the unit of its throughput). */
int  hrrt_delays_get_dims(const hrrt_delays *ctx, int *nprojs, int *nviews, int *nplanes, int *nsegs,
                          long long *pairs_full);

/* One module pair's LUT (mp 1..20): nsino, d, z as [144][144] (A crystal x + 72*layer major). */
int  hrrt_delays_get_lut(const hrrt_delays *ctx, int mp, int32_t *nsino, float *d, float *z);
int  hrrt_delays_save_lut(const hrrt_delays *ctx, const char *path);

/* Crystal singles rates from the delayed half of a coincidence histogram (host arrays):
 * coins = the .ch contents (prompts then delayed), singles receives [crystals].  niter
 * (optional) receives the number of iterations.  Runs on the host. */
int  hrrt_delays_singles(const hrrt_delays *ctx, const int32_t *coins, float scan_duration, float *singles,
                         int *niter);

/* Smoothed randoms from crystal singles (d_singles, [crystals]) into d_delays
 * [nplanes][nviews][nprojs].  scan_duration <= 0 selects 1 s (the reference's -t default). */
int  hrrt_delays_smooth(hrrt_delays *ctx, const float *d_singles, float scan_duration, float *d_delays,
                        void *stream);

/* gen_delays(2, ...): singles from the histogram (host) then smoothing into d_delays.
 * singles_out (optional, host) receives the singles (-S). */
int  hrrt_gen_delays(hrrt_delays *ctx, const int32_t *coins, float scan_duration, float *d_delays,
                     float *singles_out, int *niter, void *stream);

/* gen_delays(1, ...)'s file: the delays as flat floats (synchronous). */
int  hrrt_delays_write(hrrt_delays *ctx, const float *d_delays, const char *path, void *stream);

#ifdef __cplusplus
}
#endif
#endif
