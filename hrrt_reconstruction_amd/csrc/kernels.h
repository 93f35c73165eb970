// kernels.h -- device-side tables and kernel launch interface (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>

namespace hrrt {

// Geometry tables resident on the device (uploaded once per context).
struct DevGeom {
    int N, c, views, Z, Z4, NQ, S, th, K, GM, ZS, spp, sviews, subsets;
    int x_rebin, v_rebin, def_elements, def_angles, npskip;
    float y_size, eps1, eps2, eps3, image_max;
    const int *cyl_lo, *cyl_hi;          // [N]
    const float *sin_psi, *cos_psi;      // [views]
    const int *yb, *yt;                  // [th] plane ranges
    const int *seg_data;                 // [th] padded pair-column start
    const int *segment_offset;           // [th] plane offsets in [S]
    const int *file_seg;                 // [th] file plane offset (swapped segment, FIO:116) in file planes
    const float *bp_rescale;             // [th]
    const float *rescalearray;           // [GM][Z]
    const int4 *bp_lut;                  // [N][GM]: {zsa, zea, off = yra - zsa, bits(frac)}
    const int4 *bp_lut2;                 // [N][GM]: {zsa | zea << 16, LDS index of plane 0 of the shifted
                                         //  segment 2k, of segment 2k-1, bits(frac)}
    const float2 *rot;                   // [views/2][N][N] rotation sample positions (rotateimage2 order)
    const int4 *fp_planes;               // [fp_np] forward-projector planes, axial first, then by (k, yr):
                                         //  {k | yr << 8 | delta << 16, bits(base plane), bits(tan), ymin}
    int fp_np;
    int fp_fast;                         // 1: every forward position of a row the kernel projects is >= 0
    const int16_t *pseg;                 // [spp] padded plane -> segment (-1 on guards)
    const int16_t *plane_seg;            // [S] internal plane -> segment
    const int *bp_src;                   // [bp_lds + 1] LDS entry -> pair-column plane (-1 = zero)
    int bp_lds;                          // entries of the zero-padded back-projection column
    int8_t *nzmask;                      // [views/2][N/2+1], 0 = masked column
    // incremental back projector (back_proj3d_view_n1thread PJ:4162-4668)
    int inc_ok;                          // 1 when every breakpoint of the geometry is of the kind handled
    int inc_lds;                         // LDS floats of one lane's column (segment 0, then one window per group)
    int inc_shm;                         // LDS bytes of the incremental back projector's workgroup
    const int *inc_src;                  // [inc_lds] LDS float -> pair-column plane (-1 = zero)
    const int2 *inc_conv;                // [inc_lds] {bits(-tan) of the difference step (0 = none), second step}
    const int4 *inc_group;               // [K][4] per group: window, first-row lookups, row c+1 breakpoint
    const int4 *inc_sched;               // [n+1] breakpoints of rows c-2 down to 1, groups ascending
    const int *inc_nrow;                 // [c+1] breakpoints of row y
    int fp_maxt, fp_nsplit;              // forward-projector launch shape (tasks/thread, workgroups/(pair,x))
    int bp_waves;                        // back-projector waves per workgroup
};

// One batch of view pairs processed together (a subset, or a single pair), for one
// frame or for nframes frames of a dynamic study laid out at fixed strides.
struct PairBatch {
    int view0, view_stride, npairs;      // view of local pair lp = view0 + view_stride*lp
    int nframes = 1;
    long long img_stride = 0;            // floats between frames' images
    long long sino_stride = 0;           // elements between frames' pair-column prompts / additive terms
    long long ratio_stride = 0;          // floats between frames' ratio buffers
    long long buf_stride = 0;            // floats between frames' rotated-frame buffers
};

// Per-bin inputs of the forward projector's epilogue.
//   FP_PLAIN  : projection only (out = scaled estimate)
//   FP_RATIO  : OP ratio with float additive term (bit-exact path)
//   FP_RATIO16: OP ratio with FP16 additive term; with ll_part set, also the FP64
//               Poisson log-likelihood of the subset's bins (needs the FP16 1/(N*A))
//   FP_RATIO_W012: UW/AW/ANW ratio data / estimate (add holds the corrected data)
enum FpMode { FP_PLAIN = 0, FP_RATIO = 1, FP_RATIO16 = 2, FP_RATIO_W012 = 3 };
struct FpIO {
    const int16_t *prompt = nullptr;
    const void *add = nullptr;           // float (FP_RATIO, FP_RATIO_W012) or FP16 bits (FP_RATIO16)
    const uint16_t *invnorm = nullptr;   // FP16 1/(N*A) (FP_RATIO16 log-likelihood)
    double *ll_part = nullptr;           // one partial sum per workgroup (FP_RATIO16)
};

hipError_t launch_forward(const DevGeom &g, const PairBatch &b, const float *image, float *out_paircol,
                          const FpIO &io, int mode, hipStream_t s);
// launch shape the forward projector uses (workgroups per ray = nsplit, planes per thread = maxt)
void forward_shape(const DevGeom &g, int &maxt, int &nsplit, int &threads, int &items);
// src: float pair-column, or FP16 bits when src_half; kahan: compensated accumulation over groups
hipError_t launch_backproject(const DevGeom &g, const PairBatch &b, const void *src_paircol, int src_half,
                              int src_global, int apply_rescale, int kahan, float *bufA, float *bufB, hipStream_t s);
// incremental back projection of every view pair of the batch into the rotated-frame halves
hipError_t launch_backproject_incremental(const DevGeom &g, const PairBatch &b, const float *src, int src_global,
                                          int apply_rescale, float *bufA, float *bufB, hipStream_t s);
hipError_t launch_rotate_update(const DevGeom &g, const PairBatch &b, const float *bufA, const float *bufB,
                                float *image, const float *normfac, int kahan, hipStream_t s);
hipError_t launch_rotate_sens(const DevGeom &g, const PairBatch &b, const float *bufA, const float *bufB,
                              float *normfac_out, int kahan, hipStream_t s);
// per-frame sums of the forward projector's log-likelihood partials (fixed order, deterministic)
hipError_t launch_ll_reduce(const double *part, int per_frame, int nframes, double *out, hipStream_t s);
hipError_t launch_to_half(const float *in, uint16_t *out, size_t n, hipStream_t s);
// the same, counting inputs that overflow to +-inf ([0]) and non-zero inputs that become FP16
// subnormals or zero ([1]) into counts (cleared first by a kernel, so it may be captured)
hipError_t launch_to_half_range(const float *in, uint16_t *out, size_t n, unsigned long long *counts, hipStream_t s);
// FP64 sums over a pair-column sinogram: [0] sum prompt, [1] sum w*add, [2] sum w*est, w = 1/(N*A)
hipError_t launch_count_moments(const DevGeom &g, const int16_t *prompt, const float *add, const float *invnorm,
                                const float *est, int view0, int view_stride, int npairs, double *part,
                                double *out, hipStream_t s);
hipError_t launch_halves_sum(const DevGeom &g, const float *bufA, const float *bufB, float *out, hipStream_t s);
hipError_t launch_rotate_add(const DevGeom &g, int view, const float *bufA, const float *bufB, float *ima,
                             hipStream_t s);
hipError_t launch_paircol_to_estimate(const DevGeom &g, int view, const float *paircol, float *prj, hipStream_t s);
hipError_t launch_estimate_to_paircol(const DevGeom &g, float *prj, float *paircol, hipStream_t s);
hipError_t launch_ratio_w3(const DevGeom &g, float *est, const float *add, const int16_t *prompt, hipStream_t s);
hipError_t launch_ratio_w012(const DevGeom &g, float *est, const float *data, hipStream_t s);
hipError_t launch_update_image(const DevGeom &g, float *image, float *corr, const float *nf, hipStream_t s);
hipError_t launch_normfac_invert(const DevGeom &g, const float *sens, float *out, hipStream_t s);
// image [X][Y][Z4] <-> flat planes [Z][N][N] (to_flat = 1: image -> flat); flip swaps x and y
hipError_t launch_flat_image(const DevGeom &g, const float *src, float *dst, int to_flat, int flip, hipStream_t s);
// rotation of the batch's pairs summed into corr (FOV voxels), no update
hipError_t launch_rotate_sum(const DevGeom &g, const PairBatch &b, const float *bufA, const float *bufB, float *corr,
                             hipStream_t s);
// PSF blur convolve3d: x pass in -> buf, y/z pass buf -> out (out may be in); taps of the x and
// y kernels with half-widths hx, hy; normz [Z4]; the z pass keeps planes inside [z_min, z_max)
hipError_t launch_convolve3d(const DevGeom &g, const float *in, float *out, float *buf, const float *taps_x, int hx,
                             const float *taps_y, int hy, const float *normz, int z_min, int z_max, hipStream_t s);
hipError_t launch_init_image(const DevGeom &g, float *image, int z_min, int z_max, hipStream_t s);
hipError_t launch_end_iteration(const DevGeom &g, float *image, int z_min, int z_max, hipStream_t s);
// Segments [th0, th0 + nth), whose internal planes are [p_lo, p_hi) (their file blocks must be
// in place); finish: derive the zero-column mask (after the last segment).
hipError_t launch_stage(const DevGeom &g, const int16_t *prompt_f, const float *delayed_f, const float *norm_f,
                        const float *scatter_f, const float *atten_f, int16_t *prompt, float *add, float *invnorm,
                        uint8_t *gapmask, int *colmask, int th0, int nth, int p_lo, int p_hi, bool finish,
                        hipStream_t s);
// staging for the UW/AW/ANW weightings (weighting 0, 1, 2): data_f16 (short prompts or trues)
// or data_ff (float, or fully precorrected when precor) -> data (the w012 numerator) and the
// sensitivity weights of that weighting
hipError_t launch_stage_w012(const DevGeom &g, int weighting, int precor, const int16_t *data_f16,
                             const float *data_ff, const float *delayed_f, const float *norm_f, const float *scatter_f,
                             const float *atten_f, float *data, float *invnorm, uint8_t *gapmask, int *colmask,
                             int th0, int nth, int p_lo, int p_hi, bool finish, hipStream_t s);

// Records msg as hrrt_last_error()'s message of this thread and returns code.
int set_last_error(int code, const char *msg);

}  // namespace hrrt
