/*
 * hrrt_legacy.h -- the reference projector's own entry points, with its C++ signatures
 * (hrrt_osem3d.h:197-211), exported by libhrrt_osem3d_legacy.so as marshalling shims over
 * libhrrt_osem.so: images float***[X][Y] -> Z4 floats per column, estimates float**[2S] -> N
 * floats per row, all in host memory.  A driver written against hrrt_osem3dv_sbrt_span3.cpp
 * links this library in place of that file and adds one call, hrrt_legacy_set_params,
 * before dependencies().
 *
 * Every call is synchronous and calls are serialised.  Each returns 1 on success, as the
 * reference does, and 0 with a message on stderr on failure (the reference has no failure
 * return for these).  numthread, prjbuf, and imagebuf where the reference uses it as scratch
 * only, are accepted and ignored: the device holds its own scratch.
 */
#ifndef HRRT_LEGACY_H
#define HRRT_LEGACY_H

#include "hrrt_osem.h"

#ifdef __cplusplus
extern "C" {
#endif
/* No counterpart in the reference, whose dependencies() reads the globals that
 * GetScannerParameter (S3:3452-3551) and the command line set: the same values as a
 * hrrt_params.  Call before dependencies(). */
int hrrt_legacy_set_params(const hrrt_params *params);
/* No counterpart in the reference, whose projectors read the global norm_zero_mask
 * (short [views/2][N/2+1], H:263, set by prepare_norm S3:897-909; 0 = the column's
 * normalisation is zero in every plane, and the projectors skip it).  Call after
 * dependencies() and prepare_norm(); until then every column counts. */
int hrrt_legacy_set_norm_zero_mask(short **mask);
/* The context dependencies() created (NULL before), for callers that mix in the device API. */
hrrt_ctx *hrrt_legacy_context(void);
#ifdef __cplusplus
}

/* dependencies (PJ:199-601): builds the geometry from the parameters set above; nprojs and
 * nviews must match them (N = def_elements / x_rebin, views = def_angles / v_rebin). */
int dependencies(int nprojs, int nviews, int verbose);
void free_dependencies();
/* forward_proj3d2 (PJ:1271) and forward_proj3d_thread1 (PJ:1373): estimate rows of view pair
 * `view` from ima.  Bins the reference does not write keep their values. */
int forward_proj3d2(float ***ima, float **prj, int view, int numthread, float ***imagebuf, float ***prjbuf);
int forward_proj3d_thread1(float ***ima, float **prj, int view, int numthread, float ***imagebuf, float ***prjbuf);
/* back_proj3d2 (PJ:1457): rescales prj in place, back-projects (direct projector) and adds
 * the rotated result into ima. */
int back_proj3d2(float **prj, float ***ima, int view, int numthread, float ***imagebuf, float ***prjbuf);
/* back_proj3d_thread1 (PJ:1622): rescales prj in place and back-projects (incremental
 * projector) into the FOV columns of imagebuf (planes z < z_pixels), in the rotated frame of
 * `view`; ima is not used.  The reference's driver follows it with rotate_image_buf. */
int back_proj3d_thread1(float **prj, float ***ima, int view, int numthread, float ***imagebuf, float ***prjbuf);
/* rotate_image_buf (PJ:1701-1725): for i in 0..numthread-1, ima += rotate(imagebuf[i], view[i]). */
int rotate_image_buf(float ****imagebuf, float ***ima, int numthread, int *view);
#endif
#endif
