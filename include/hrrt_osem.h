/*
 * hrrt_osem.h -- C ABI of the MI355X (gfx950) OSEM hot path of hrrt_osem3d.
 *
 * The reference's projector interface is a set of C++ functions over
 * pointer-to-pointer arrays plus process globals (hrrt_osem3d/hrrt_osem3d.h:197-211).
 * This ABI keeps its operations, argument meaning and layouts, replaces the
 * globals by an explicit context, and reports errors through status codes
 * instead of exit(1) (the reference's LOG_EXIT, hrrt_common/my_spdlog.hpp:58-61).
 *
 * Citations: PJ = hrrt_osem3d/hrrt_osem3dv_sbrt_span3.cpp,
 *            S3 = hrrt_osem3d/hrrt_osem3d_span3.cpp,
 *            FIO = hrrt_osem3d/file_io_processor.cpp.
 *
 * Layouts (all contiguous, little endian):
 *   image        float [X][Y][Z4], Z4 = 4*ceil(z_pixels/4)           (matrix3dm128, nr_utils.cpp:248)
 *   file sino    [file plane][def_angles][def_elements]              (FIO:110-259; S3:3706-3719)
 *   estimate     float [2*S][N]: rows 0..S-1 view v, S..2S-1 view v+views/2 (S3:2142, 2145-2154)
 *   pair-column  device staging layout of a sinogram:
 *                [views/2 pairs][2 halves][N columns][spp padded planes]
 *                (half 0 = view v, half 1 = view v+views/2; segment s plane yr at
 *                 hrrt_dims.seg_data[s] + yr - yr_bottom[s]; zero guard plane on each side)
 *   normfac      float [subsets][X][Y][Z4]                           (normfac.i, FIO:360-427)
 *
 * All "d_" pointers are device pointers on the current HIP device; `stream`
 * is a hipStream_t (NULL = default stream).  Calls are asynchronous with
 * respect to the host unless stated otherwise; a context is used by one
 * stream at a time (calls on one context must not overlap).
 */
#ifndef HRRT_OSEM_H
#define HRRT_OSEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HRRT_OK = 0,
    HRRT_ERR_ARG = 1,        /* invalid argument / geometry                      */
    HRRT_ERR_HIP = 2,        /* HIP runtime failure                              */
    HRRT_ERR_NOMEM = 3,      /* device or host allocation failed                 */
    HRRT_ERR_UNSUPPORTED = 4 /* option the reference supports but this build does not */
};

/* Scanner model + reconstruction options (GLOBAL_OSEM3DPAR S3:387-445,
 * ScannerModel scanner_model.h:3-16). */
typedef struct {
    int   rings;           /* crystal rings (104)                        */
    float crystal_radius;  /* mm (234.5)                                 */
    float plane_sep;       /* mm (1.21875)                               */
    float bin_size;        /* mm (1.21875)                               */
    int   def_elements;    /* radial bins in the files (256)             */
    int   def_angles;      /* views in the files (288)                   */
    int   x_pixels;        /* image size -X (0 = def_elements)           */
    int   span;            /* -m span                                    */
    int   maxdel;          /* -m ,maxdel                                 */
    int   subsets;         /* -S                                         */
    float rel_fov;         /* -f rel_fov (0.95)                          */
    int   npskip;          /* planes killed at segment ends (3)          */
    int   z_min, z_max;    /* -f ,,z_min,z_max (6, 200)                  */
    float epsilon1, epsilon2, epsilon3, image_max; /* -e, -u              */
} hrrt_params;

typedef struct {
    int N;             /* radial bins = image X = Y                       */
    int views;         /* views (pairs = views/2)                         */
    int z_pixels, z4;  /* 207, 208                                        */
    int segmentsize;   /* S: planes of all segments (2209 span 9, 6367 span 3) */
    int th_pixels;     /* segments                                        */
    int sviews;        /* views per subset                                */
    int subsets;
    int spp;           /* padded planes per pair-column                   */
    int x_rebin, v_rebin;
    long long image_floats;       /* X*Y*Z4                               */
    long long sino_bins;          /* S*views*N (internal sinogram)        */
    long long file_bins;          /* S*def_angles*def_elements (file)     */
    long long paircol_elems;      /* (views/2)*2*N*spp                    */
    long long fov_voxels;         /* (x, y) image columns inside the FOV (cylwiny, PJ:339-356) */
    int fov_bins;                 /* radial bins 1..N-1 whose image row holds FOV pixels */
} hrrt_dims;

/* Create a context: GetScannerParameter (S3:3452) + dependencies (PJ:199).
 * Builds every geometry table on the host and uploads it to the current
 * device.  Returns HRRT_ERR_ARG for geometries the reference does not
 * support (views % subsets != 0, zoom != 1, groupmax < 1) or for which the
 * rotated samples could leave the image (FOV radius + 1 >= N/2 - 1). */
typedef struct hrrt_ctx hrrt_ctx;
int  hrrt_create(const hrrt_params *params, hrrt_ctx **out);
void hrrt_destroy(hrrt_ctx *ctx);
int  hrrt_get_dims(const hrrt_ctx *ctx, hrrt_dims *out);
/* Segment tables ([th_pixels] each): plane range (yr_bottom/yr_top, PJ:446-457),
 * plane offset in the [S] internal order (segment_offset) and start of the
 * segment in a pair-column (seg_data). */
int  hrrt_get_segments(const hrrt_ctx *ctx, int *yr_bottom, int *yr_top, int *segment_offset, int *seg_data);
/* Subset order (sequence S3:535) and view order (calculatevieworder PJ:144). */
int  hrrt_get_order(const hrrt_ctx *ctx, int *order /* [subsets] */, int *vieworder /* [subsets][sviews] */);
/* The same dimensions and tables as hrrt_get_dims/_order/_segments, computed
 * on the host without a context or a device, so that a caller can size its
 * buffers before it picks a GPU.  Any output pointer may be NULL; the table
 * outputs need th_pixels / subsets / subsets*sviews ints (read dims first). */
int  hrrt_query_geometry(const hrrt_params *params, hrrt_dims *dims, int *order, int *vieworder,
                         int *yr_bottom, int *yr_top, int *segment_offset, int *seg_data);
/* norm_zero_mask (S3:864-909): [views/2][N/2+1] int8, 0 = the (v, v+views/2, x, N-x)
 * columns of the norm are zero in every plane.  Staging computes it; a driver
 * that prepares the norm itself sets it here. */
int  hrrt_set_norm_zero_mask(hrrt_ctx *ctx, const int8_t *mask);
int  hrrt_get_norm_zero_mask(hrrt_ctx *ctx, int8_t *mask);
/* Thread-local message for the last non-zero status. */
const char *hrrt_last_error(void);

/* ---------------- staging (file_io_processor + prepare_*) ---------------- */
/* prepare_osem3dw3 (S3:1649-1818, OP weighting, short prompts, Atten3D) and
 * prepare_norm's 1/(N*A) (S3:827-1003, norm_ANW S3:796): reads whole flat
 * files that are already in device memory (file layout; NULL = file absent,
 * default values as FIO:119-123) and writes pair-column buffers:
 *   d_prompt  int16  prompts, zeroed where the norm is 0
 *   d_add     float  (d*N + s)*A
 *   d_invnorm float  N>0 ? (N*max(A,1) > eps2 ? 1/(N*max(A,1)) : 0) : N
 * Rebinning (-X 128: x_rebin = v_rebin = 2), norm harmonic rebin with gap
 * mask and npskip end-plane killing follow FIO:29-180.  Also computes the
 * norm zero-column mask (S3:864-909) kept in the context. */
int hrrt_stage_sinograms(hrrt_ctx *ctx, const int16_t *d_prompt_file, const float *d_delayed_file,
                         const float *d_norm_file, const float *d_scatter_file, const float *d_atten_file,
                         int16_t *d_prompt, float *d_add, float *d_invnorm, void *stream);

/* prepare_osem3dw0 / w1 / w2 (S3:1190-1639) and prepare_norm's weights for that weighting
 * (S3:827-1003: norm_UW :756, norm_AW :774, norm_ANW :796), for the reference's -W 0 (UW),
 * 1 (AW) and 2 (ANW, its default).  The data are d_data16_file (short prompts, -p, or trues,
 * -t) or d_dataf_file (a float sinogram, -Q; a fully precorrected one with precorrected = 1,
 * -P); exactly one is given.  d_delayed_file is subtracted when given (prompt - delayed,
 * readsinogram S3:1066-1134).  Outputs, pair-column layout:
 *   d_data    float  the numerator of update_estimate_w012, clamped at 0 (setnegative2zero):
 *                    UW  ((p-d)*N [- N with scatter, as the reference's S3:1233]) * A
 *                    AW  (p-d - s/N where N > 0, else 0) * N
 *                    ANW  p-d - s/N where N > 0, else 0
 *                    precorrected: UW t, AW t/A (A > 0), ANW t/N (0 where N = 0), then /A
 *   d_invnorm float  UW N > 0 ? 1 : N;  AW N > 0 ? 1/max(A,1) : 0;  ANW as hrrt_stage_sinograms
 * The norm is read with the end-plane killing and gap mask of hrrt_stage_sinograms, and the
 * norm zero-column mask is computed the same way.  HRRT_ERR_UNSUPPORTED for -W 0 with scatter
 * and no norm: the reference then subtracts a buffer it never filled. */
int hrrt_stage_sinograms_w012(hrrt_ctx *ctx, int weighting, int precorrected, const int16_t *d_data16_file,
                              const float *d_dataf_file, const float *d_delayed_file, const float *d_norm_file,
                              const float *d_scatter_file, const float *d_atten_file, float *d_data,
                              float *d_invnorm, void *stream);

/* The same two stagings from files in host memory, as the reference reads them
 * (read_flat_float / read_flat_short per segment, FIO:110-259): the file pointers are host
 * memory (pageable, or pinned -- hipHostMalloc, torch pin_memory -- which is copied
 * without a bounce), the outputs device memory.  Each segment's block of every file is
 * copied on the context's own copy stream into a device copy of that file kept in the
 * context (through a ring of pinned buffers for pageable memory), and that segment's
 * staging launch on `stream` waits for its blocks only, so later copies overlap earlier
 * staging.  Returns when every copy has completed: the host files may be reused then;
 * the staging itself may still be running on `stream`.  Results equal the device-memory
 * entry points bit for bit. */
int hrrt_stage_sinograms_host(hrrt_ctx *ctx, const int16_t *prompt_file, const float *delayed_file,
                              const float *norm_file, const float *scatter_file, const float *atten_file,
                              int16_t *d_prompt, float *d_add, float *d_invnorm, void *stream);
int hrrt_stage_sinograms_w012_host(hrrt_ctx *ctx, int weighting, int precorrected, const int16_t *data16_file,
                                   const float *dataf_file, const float *delayed_file, const float *norm_file,
                                   const float *scatter_file, const float *atten_file, float *d_data,
                                   float *d_invnorm, void *stream);

/* ---------------- sensitivity (CalculateNormfac S3:3760-3865) ---------------- */
/* For every subset: back-project the subset's view pairs of d_invnorm (BP
 * rescale applied on the fly, d_invnorm itself is not modified), rotate and
 * accumulate, then d_normfac[subset] = |sens| > eps1 ? 1/sens : 0
 * (pt_normfac_calculate S3:1852).  d_normfac: [subsets][X][Y][Z4]. */
int hrrt_calculate_normfac(hrrt_ctx *ctx, const float *d_invnorm, float *d_normfac, void *stream);

/* ---------------- OSEM (CalculateOsem3d S3:4098-4341) ---------------- */
int hrrt_init_image(hrrt_ctx *ctx, float *d_image, void *stream);              /* Init_Image S3:3997     */
/* One subset update: forward projection of every view pair of `isubset`
 * (forward_proj3d_thread1 PJ:1373), ratio (update_estimate_w3 S3:2474),
 * back projection (back_proj3d2 PJ:1457) and rotation (rotate_image_buf
 * PJ:1701) into the correction, then the image update (pt_cal_updateimage
 * S3:2018) with d_normfac_subset.  The correction is never materialised. */
int hrrt_osem_subset(hrrt_ctx *ctx, float *d_image, const float *d_normfac_subset,
                     const int16_t *d_prompt, const float *d_add, int isubset, void *stream);
int hrrt_end_iteration(hrrt_ctx *ctx, float *d_image, void *stream);           /* S3:4288-4297           */
/* The same subset update for the UW/AW/ANW weightings (-W 0, 1, 2): the forward projection
 * of each pair goes through update_estimate_w012 (S3:1918-1919, 2436-2447),
 * est = est > eps3 ? data / est : 0, with d_data from hrrt_stage_sinograms_w012, and
 * d_normfac_subset from hrrt_calculate_normfac of that weighting's d_invnorm. */
int hrrt_osem_subset_w012(hrrt_ctx *ctx, float *d_image, const float *d_normfac_subset, const float *d_data,
                          int isubset, void *stream);

/* ---------------- batched dynamic frames (no counterpart in the reference) ----------------
 * The reference reconstructs one frame per process run; a dynamic study runs the
 * same subset update (hrrt_osem_subset) for every frame.  These entry points do
 * it for nframes frames at once, one launch per kernel, frames at fixed strides:
 *   d_images  [nframes][X][Y][Z4]            (stride hrrt_dims.image_floats)
 *   d_prompts [nframes][pair-column] int16   (stride hrrt_dims.paircol_elems)
 *   d_adds    [nframes][pair-column] float   (stride hrrt_dims.paircol_elems)
 * The sensitivity images are shared (same geometry, norm and attenuation).  Each
 * frame's result is bit-identical to hrrt_osem_subset on that frame alone.
 * Per-frame scratch (~2 rotated-frame volumes per view pair per frame) is kept by
 * the context. */
int hrrt_osem_subset_frames(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac_subset,
                            const int16_t *d_prompts, const float *d_adds, int isubset, void *stream);
/* One full iteration (every subset in sequence order, S3:4174-4245, then the
 * end-of-iteration threshold S3:4288-4297) for all frames.  The subsets' launches
 * are captured once into one HIP graph per subset and replayed; the graphs are
 * rebuilt when nframes, a buffer pointer, the stream or a launch shape changes.
 * d_normfac: [subsets][X][Y][Z4].  `stream` must not be the default stream. */
int hrrt_osem_iteration_frames(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac,
                               const int16_t *d_prompts, const float *d_adds, void *stream);

/* ---------------- precision form (no counterpart in the reference) ----------------
 * The reference keeps every sinogram in float32 and sums in float32.  This form
 * stores the per-bin float inputs of the hot loop -- the additive term
 * (d*N + s)*A and the sensitivity weights 1/(N*A) -- as IEEE FP16 (bits in
 * uint16_t, round to nearest even), keeps prompts exact int16 and computes in FP32.
 * The sensitivity build accumulates its back projection over segment groups with
 * compensated (Kahan) summation; every form sums the view pairs of the rotation with
 * compensation; the subset updates back-project under the -T schedule
 * (hrrt_set_backproject_schedule, incremental by default) as the FP32 path does.  FP64
 * is used only for the Poisson log-likelihood of each subset and the global count
 * moments.  Results are within tolerance of an FP64 projector, not bit-identical to
 * the FP32 path. */
int hrrt_to_half(const float *d_in, uint16_t *d_out, long long n, void *stream);
/* hrrt_to_half that also reports the range of the input (no counterpart in the reference):
 * d_counts[0] = finite inputs that become +-inf (|x| >= 65520), d_counts[1] = non-zero
 * finite inputs that become FP16 subnormals or zero (|x| < 2^-14); d_counts is two
 * uint64 in device memory, cleared by the call. */
int hrrt_to_half_range(const float *d_in, uint16_t *d_out, long long n, unsigned long long *d_counts,
                       void *stream);
/* CalculateNormfac (S3:3760) from FP16 1/(N*A), Kahan accumulation. */
int hrrt_calculate_normfac_mixed(hrrt_ctx *ctx, const uint16_t *d_invnorm16, float *d_normfac, void *stream);
/* hrrt_osem_subset with FP16 additive term.  When d_loglik is not NULL, *d_loglik
 * (device double) receives sum over the subset's bins of y*log(mu) - mu with
 * mu = (est + add) * w, w = FP16 1/(N*A) (d_invnorm16, shared by all frames), at the
 * image before the update, summed in FP64 in a fixed order. */
int hrrt_osem_subset_mixed(hrrt_ctx *ctx, float *d_image, const float *d_normfac_subset, const int16_t *d_prompt,
                           const uint16_t *d_add16, const uint16_t *d_invnorm16, int isubset, double *d_loglik,
                           void *stream);
/* The same for nframes frames (d_loglik: [nframes] or NULL). */
int hrrt_osem_subset_frames_mixed(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac_subset,
                                  const int16_t *d_prompts, const uint16_t *d_adds16, const uint16_t *d_invnorm16,
                                  int isubset, double *d_loglik, void *stream);
/* Graph-replayed iteration of the precision form (d_loglik: [subsets][nframes] or NULL). */
int hrrt_osem_iteration_frames_mixed(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac,
                                     const int16_t *d_prompts, const uint16_t *d_adds16,
                                     const uint16_t *d_invnorm16, double *d_loglik, void *stream);
/* Global count moments in FP64 over the whole sinogram, forward projecting d_image
 * for every view pair: d_out[0] = sum prompts, d_out[1] = sum w*add,
 * d_out[2] = sum w*est, w = 1/(N*A).  The count normalisation of an image is
 * (d_out[0] - d_out[1]) / d_out[2]. */
int hrrt_count_moments(hrrt_ctx *ctx, const float *d_image, const int16_t *d_prompt, const float *d_add,
                       const float *d_invnorm, double *d_out, void *stream);

/* Kernel timing of hrrt_osem_subset (no counterpart in the reference, whose
 * only timing is clock() around loops, S3:4219-4285).  When enabled, HIP
 * events are recorded on the launch stream around the three kernels of every
 * subset update; hrrt_timing_read waits for them and returns the summed
 * milliseconds and launch counts per kernel: [0] forward projection + ratio,
 * [1] back projection, [2] rotation + image update.  Reading resets. */
#define HRRT_NKERNELS 3
/* Launch shape of the forward projector (no counterpart in the reference; set
 * by the autotuner): the planes of one ray column are split over
 * `fp_workgroups_per_ray` workgroups (0 = as few as fit, at most 1024 threads
 * each), with `fp_tasks_per_thread` (1..4) planes per thread.  Results do not
 * depend on the shape. */
int hrrt_set_launch_shape(hrrt_ctx *ctx, int fp_tasks_per_thread, int fp_workgroups_per_ray);
/* Workgroup size of the back projector in 64-lane waves (1..16; no counterpart in
 * the reference).  Each wave evaluates one voxel row of the column the
 * workgroup has staged in LDS; results do not depend on the shape. */
int hrrt_set_backproject_shape(hrrt_ctx *ctx, int waves_per_workgroup);
int hrrt_timing_enable(hrrt_ctx *ctx, int enable);
int hrrt_timing_read(hrrt_ctx *ctx, double *ms /* [HRRT_NKERNELS] */, int *count /* [HRRT_NKERNELS] */);

/* ---------------- per-view-pair projector calls ---------------- */
/* forward_proj3d_thread1 / forward_proj3d2 (PJ:1271, 1373): estimate of view
 * pair `view` in the [2S][N] layout.  Bins the reference does not write
 * (xr = 0, columns outside the FOV, masked columns) are left untouched. */
int hrrt_forward_proj3d(hrrt_ctx *ctx, const float *d_image, float *d_prj, int view, void *stream);
/* No counterpart in the reference.  hrrt_forward_proj3d for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same result, bit for bit, with no hipMalloc, hipFree or allocating
 * library call.  Its pair-column staging lives in a file-scope device array of
 * HRRT_NOALLOC_SCRATCH_FLOATS floats (13.6 MB, every GPU the library runs on holds one for
 * the life of the process); geometries whose pair-column (2*N*spp floats, 3.35 M for HRRT
 * span 3 at 256) exceeds it get HRRT_ERR_ARG.  All calls on one GPU share that array: calls
 * that overlap in time on one GPU (other streams, other contexts, other threads) race and
 * must be serialised by the caller. */
#define HRRT_NOALLOC_SCRATCH_FLOATS (2 * 256 * 6656)
int hrrt_forward_proj3d_noalloc(hrrt_ctx *ctx, const float *d_image, float *d_prj, int view, void *stream);
/* back_proj3d2 (PJ:1457): rescales d_prj in place (as the reference does),
 * back-projects both views of the pair and adds the rotated result into d_ima. */
int hrrt_back_proj3d(hrrt_ctx *ctx, float *d_prj, float *d_ima, int view, void *stream);
/* back_proj3d_thread1 (PJ:1622-1699, back_proj3d_view_n1thread PJ:4162-4668): the
 * incremental back projector, the reference's default for most view pairs (the
 * *2 variant serves the remainder, S3:4198-4218).  Same arguments and effect as
 * hrrt_back_proj3d; the two differ by float rounding (about 1e-6 relative).
 * HRRT_ERR_UNSUPPORTED when a breakpoint of the geometry moves the plane range
 * by other than one plane (none does for the HRRT models). */
int hrrt_back_proj3d_incremental(hrrt_ctx *ctx, float *d_prj, float *d_ima, int view, void *stream);
/* back_proj3d_thread1 (PJ:1622-1699) as the reference's driver uses it: rescales d_prj in
 * place and back-projects both views of the pair into d_imagebuf [X][Y][Z4], in the rotated
 * frame of `view`, without rotating.  The planes z < z_pixels of the FOV columns are
 * overwritten; the rest of d_imagebuf is left untouched, as the reference leaves it (it also
 * zeroes the pad planes of the columns it skips for a zero norm; its driver keeps every pad
 * plane zero, so the two agree there).  incremental = 1 takes the incremental projector (back_proj3d_thread1),
 * 0 the direct one (the rotated-frame half of back_proj3d2). */
int hrrt_back_proj3d_rotframe(hrrt_ctx *ctx, float *d_prj, float *d_imagebuf, int view, int incremental,
                              void *stream);
/* One buffer of rotate_image_buf (PJ:1701-1725, pt_rotate3d2): d_ima += rotate(d_imagebuf,
 * view), the bilinear rotation of the back projectors, over the FOV columns of d_ima. */
int hrrt_rotate_add(hrrt_ctx *ctx, const float *d_imagebuf, float *d_ima, int view, void *stream);
/* Back-projection schedule of the subset updates (hrrt_osem_subset, *_w012, *_frames, *_mixed) and of
 * hrrt_calculate_normfac, the reference's -T nthreads (S3:3796-3825, 4174-4218): the first
 * npairs - npairs % nthreads view pairs of each subset take the incremental projector
 * (back_proj3d_thread1), the rest the direct one (back_proj3d2); 0 = direct for every pair.
 * Default 1 (every pair incremental) when the geometry allows it, else 0.  The precision
 * form's subset updates follow the same schedule; its sensitivity build takes the direct
 * projector with compensated sums. */
int hrrt_set_backproject_schedule(hrrt_ctx *ctx, int nthreads);
int hrrt_get_backproject_schedule(const hrrt_ctx *ctx, int *nthreads);
/* PSF resolution model, the reference's -B fwhm1,fwhm2,ratio (S3:3327-3330, 4521-4540;
 * psf.cpp:20-70): a kernel per axis, the sum of two Gaussians of FWHM fwhm1 and fwhm2 (mm)
 * weighted 1 : ratio, sampled at the x, y and z pixel sizes; fwhm1 == 0 selects the HRRT
 * defaults.  When enabled, hrrt_osem_subset forward-projects the blurred image and blurs
 * the correction before the update (S3:4172-4229), and hrrt_calculate_normfac blurs each
 * sensitivity image before inverting it (S3:3835-3840).  enable = 0 turns it off.
 * The batched-frame and precision forms return HRRT_ERR_UNSUPPORTED while it is on. */
int hrrt_set_psf(hrrt_ctx *ctx, int enable, float fwhm1, float fwhm2, float ratio);
/* Kernel taps of axis 0/1/2 (x/y/z) into taps (host, may be NULL); *size receives the count. */
int hrrt_get_psf_kernel(const hrrt_ctx *ctx, int axis, float *taps, int *size);
/* convolve3d (psf.cpp:441-476) of an image [X][Y][Z4]; d_out may be d_in.  Only voxels
 * inside the FOV are written. */
int hrrt_convolve3d(hrrt_ctx *ctx, const float *d_in, float *d_out, void *stream);
/* ---------------- image and sensitivity files ----------------
 * Flat image files hold planes [z_pixels][N][N] float32, x fastest (SaveImage S3:4023-4035,
 * write_flat_image FIO:300-333).  flip = the reference's output/input flip (x <-> y).  Only
 * FOV voxels are copied; the image side outside the FOV and the pad plane become 0. */
int hrrt_image_to_flat(hrrt_ctx *ctx, const float *d_image, float *d_flat, int flip, void *stream);
/* Init_Image with -I (S3:3970-3993): the inverse of hrrt_image_to_flat. */
int hrrt_image_from_flat(hrrt_ctx *ctx, const float *d_flat, float *d_image, int flip, void *stream);
/* Header contents of write_image_header (write_image_header.cpp:67-319); NULL strings are
 * "not used".  recontype: 0 UW, 1 AW, 2 ANW, 3 OP (the -W weighting). */
typedef struct {
    int recontype;
    int iterations;
    const char *true_file, *prompt_file, *delayed_file, *norm_file, *atten_file, *scatter_file;
    const char *sw_version, *sw_build_id;
} hrrt_image_info;
/* SaveImage (S3:4016-4096): writes `path` (flat planes) and `path`.hdr (Interfile), carrying
 * forward the lines of the true or prompt file's .hdr; synchronises `stream`. */
int hrrt_write_image(hrrt_ctx *ctx, const float *d_image, const char *path, int flip, const hrrt_image_info *info,
                     void *stream);
/* Reads an initial image file (-I) into d_image; synchronises `stream`. */
int hrrt_read_image(hrrt_ctx *ctx, const char *path, float *d_image, int flip, void *stream);
/* The sensitivity cache normfac.i: subsets x [X][Y][Z4] float32 (write_norm / read_norm
 * FIO:360-427); synchronise `stream`. */
int hrrt_write_normfac(hrrt_ctx *ctx, const float *d_normfac, const char *path, void *stream);
int hrrt_read_normfac(hrrt_ctx *ctx, const char *path, float *d_normfac, void *stream);

/* update_estimate_w3 (S3:2474): est = prompt==0 ? 0 : (est+add > eps3 ? prompt/(est+add) : 1),
 * over the [2S][N] rows of one view pair. */
int hrrt_update_estimate_w3(hrrt_ctx *ctx, float *d_est, const float *d_add, const int16_t *d_prompt, void *stream);
/* update_estimate_w012 (S3:2452): est = est > eps3 ? data/est : 0 over the [2S][N] rows of one
 * view pair (UW, AW and ANW weighting). */
int hrrt_update_estimate_w012(hrrt_ctx *ctx, float *d_est, const float *d_data, void *stream);
/* pt_cal_updateimage (S3:2018): image = min(image_max, image*clamp(corr,0.01,100)*normfac); corr = 0. */
int hrrt_cal_updateimage(hrrt_ctx *ctx, float *d_image, float *d_correction, const float *d_normfac, void *stream);
/* pt_normfac_calculate (S3:1852): out = |sens|>eps1 ? 1/sens : 0 over the FOV, z < z_pixels. */
int hrrt_normfac_calculate(hrrt_ctx *ctx, const float *d_sens, float *d_out, void *stream);

/* ---------------- host-buffer shims (reference layouts, synchronous) ---------------- */
/* Same operations on host memory in the reference's own layouts; each call
 * copies in, runs on the device and copies out.  These are what a driver
 * written against the reference's functions binds to. */
int hrrt_forward_proj3d_host(hrrt_ctx *ctx, const float *image, float *prj, int view);
int hrrt_back_proj3d_host(hrrt_ctx *ctx, float *prj, float *ima, int view);
int hrrt_back_proj3d_rotframe_host(hrrt_ctx *ctx, float *prj, float *imagebuf, int view, int incremental);
int hrrt_rotate_add_host(hrrt_ctx *ctx, const float *imagebuf, float *ima, int view);

#ifdef __cplusplus
}
#endif
#endif /* HRRT_OSEM_H */
