/*
 * staging.c -- oracle restatement of the flat-file readers, the sinogram
 * staging of every weighting (prepare_osem3dw0-3, prepare_norm) and the flat
 * image files.  These read and write the reference's float / short files, so
 * they are float in every build of the oracle.
 * TEST INFRASTRUCTURE ONLY (see hrrt_oracle.h).
 */
#include "hrrt_oracle.h"

#include <stdlib.h>
#include <string.h>
/* ------------------------------------------------------------------ */
/* staging                                                              */
/* ------------------------------------------------------------------ */
static int swap_theta(int theta)
{   /* FIO:116-118 */
    if (theta % 2 == 1) return theta + 1;
    if (theta == 0) return 0;
    return theta - 1;
}

/* sino_rebin FIO:29-74 for one plane, float */
static void rebin_plane_f(const og_geom *g, float *tmp, float *out, int norm_flag)
{
    const int N = g->N, V = g->views, xrb = g->x_rebin, vrb = g->v_rebin;
    const int r_size = N * xrb;
    int v, vr, i, xr;
    for (v = 0; v < V; v++) {
        float *dst = tmp + (size_t)v * vrb * r_size;
        for (vr = 0; vr < vrb; vr++) {
            float *src = dst + (size_t)vr * r_size;
            if (!norm_flag) {
                if (vr == 0) continue;
                for (i = 0; i < r_size; i++) dst[i] += src[i];
            } else if (vr == 0) {
                for (i = 0; i < r_size; i++) if (src[i] > 0.0f) dst[i] = 1.0f / src[i];
            } else {
                for (i = 0; i < r_size; i++) if (src[i] > 0.0f) dst[i] += 1.0f / src[i];
            }
        }
    }
    for (v = 0; v < V; v++) {
        float *dst = out + (size_t)v * N;
        const float *src = tmp + (size_t)v * vrb * r_size;
        for (i = 0; i < N; i++, dst++) {
            *dst = *src++;
            for (xr = 1; xr < xrb; xr++) *dst += *src++;
        }
    }
    if (norm_flag)
        for (i = 0; i < N * V; i++) if (out[i] > 0.0f) out[i] = 1.0f / out[i];
}

void og_read_flat_float(const og_geom *g, const float *file, float *out, int theta, float def_val, int norm_flag, int have_file)
{   /* read_flat_float_proj_ori_theta FIO:110-180 */
    const int N = g->N, V = g->views, rebin = g->x_rebin * g->v_rebin;
    const long plane_out = (long)N * V, file_plane = (long)N * V * rebin;
    int plane, nplanes, i;
    theta = swap_theta(theta);
    nplanes = g->yr_top[theta] - g->yr_bottom[theta] + 1;
    if (!have_file) {
        for (i = 0; i < nplanes * plane_out; i++) out[i] = def_val;
        return;
    }
    file += (size_t)g->seg_offset[theta] * rebin;
    if (rebin == 1) {
        for (plane = 0; plane < nplanes; plane++) {
            if (norm_flag && (plane < g->p.npskip || plane + g->p.npskip >= nplanes))
                memset(out + plane * file_plane, 0, sizeof(float) * file_plane);
            else
                memcpy(out + plane * file_plane, file + plane * file_plane, sizeof(float) * file_plane);
        }
    } else {
        unsigned char *mask = (unsigned char *)calloc((size_t)file_plane, 1);
        float *buf = (float *)malloc(sizeof(float) * file_plane);
        for (plane = 0; plane < nplanes; plane++) {
            memcpy(buf, file + plane * file_plane, sizeof(float) * file_plane);
            if (norm_flag && (plane < g->p.npskip || plane + g->p.npskip >= nplanes)) {
                memset(out + plane * plane_out, 0, sizeof(float) * plane_out);
                continue;
            }
            if (norm_flag)
                for (i = 0; i < file_plane; i++) if (buf[i] > 0) mask[i] = 1;
            rebin_plane_f(g, buf, out + plane * plane_out, norm_flag);
        }
        if (norm_flag) {
            /* sino_rebin on the mask (FIO:169): view sum in place, then radial sum */
            const int r_size = N * g->x_rebin;
            unsigned char *mr = (unsigned char *)calloc((size_t)plane_out, 1);
            int v, vr, xr;
            for (v = 0; v < V; v++) {
                unsigned char *dst = mask + (size_t)v * g->v_rebin * r_size;
                for (vr = 1; vr < g->v_rebin; vr++)
                    for (i = 0; i < r_size; i++) dst[i] = (unsigned char)(dst[i] + dst[vr * r_size + i]);
            }
            for (v = 0; v < V; v++) {
                const unsigned char *src = mask + (size_t)v * g->v_rebin * r_size;
                for (i = 0; i < N; i++) {
                    unsigned char s = *src++;
                    for (xr = 1; xr < g->x_rebin; xr++) s = (unsigned char)(s + *src++);
                    mr[v * N + i] = s;
                }
            }
            for (plane = 0; plane < nplanes; plane++)
                for (i = 0; i < plane_out; i++) if (mr[i] < rebin) out[plane * plane_out + i] = 0.0f;
            free(mr);
        }
        free(mask); free(buf);
    }
}

void og_read_flat_short(const og_geom *g, const int16_t *file, int16_t *out, int theta)
{   /* read_flat_short_proj_ori_to_short_array_theta FIO:226-259 */
    const int N = g->N, V = g->views, rebin = g->x_rebin * g->v_rebin;
    const long plane_out = (long)N * V, file_plane = (long)N * V * rebin;
    int plane, nplanes, i, v, vr, xr;
    theta = swap_theta(theta);
    nplanes = g->yr_top[theta] - g->yr_bottom[theta] + 1;
    file += (size_t)g->seg_offset[theta] * rebin;
    if (rebin == 1) {
        memcpy(out, file, sizeof(int16_t) * nplanes * file_plane);
        return;
    }
    {
        const int r_size = N * g->x_rebin;
        int16_t *buf = (int16_t *)malloc(sizeof(int16_t) * file_plane);
        for (plane = 0; plane < nplanes; plane++) {
            memcpy(buf, file + plane * file_plane, sizeof(int16_t) * file_plane);
            for (v = 0; v < V; v++) {
                int16_t *dst = buf + (size_t)v * g->v_rebin * r_size;
                for (vr = 1; vr < g->v_rebin; vr++)
                    for (i = 0; i < r_size; i++) dst[i] = (int16_t)(dst[i] + dst[vr * r_size + i]);
            }
            for (v = 0; v < V; v++) {
                const int16_t *src = buf + (size_t)v * g->v_rebin * r_size;
                int16_t *dst = out + plane * plane_out + (size_t)v * N;
                for (i = 0; i < N; i++) {
                    int16_t s = *src++;
                    for (xr = 1; xr < g->x_rebin; xr++) s = (int16_t)(s + *src++);
                    dst[i] = s;
                }
            }
        }
        free(buf);
    }
}

static float *theta_tmp(const og_geom *g)
{
    return (float *)malloc(sizeof(float) * (size_t)g->z_pixels * g->views * g->N);
}

void og_prepare_norm(og_geom *g, const float *norm_file, const float *atten_file, float *out)
{   /* prepare_norm S3:827-1003 (nFlag, Atten3D, ANW/OP weighting) */
    const int N = g->N, V = g->views;
    const long pl = (long)N * V;
    float *tmp = theta_tmp(g);
    unsigned char *mask = (unsigned char *)calloc((size_t)pl, 1);
    int theta, yr, v, xr;
    for (theta = 0; theta < g->th_pixels; theta++) {
        int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
        og_read_flat_float(g, norm_file, tmp, theta, 1.0f, 1, norm_file != NULL);
        memcpy(out + (size_t)g->segment_offset[theta] * pl, tmp, sizeof(float) * np * pl);
    }
    /* zero-column mask S3:864-909 */
    for (yr = 0; yr < g->segmentsize; yr++)
        for (v = 0; v < V; v++)
            for (xr = 0; xr < N; xr++)
                if (out[(size_t)yr * pl + (size_t)v * N + xr] != 0) mask[v * N + xr] = 1;
    for (v = 0; v < V / 2; v++)
        for (xr = 1; xr <= N / 2; xr++)
            g->norm_zero_mask[v * (N / 2 + 1) + xr] =
                (mask[v * N + xr] == 0 && mask[(v + V / 2) * N + xr] == 0 && mask[v * N + N - xr] == 0 &&
                 mask[(v + V / 2) * N + N - xr] == 0) ? 0 : 1;
    /* 1/(N*max(A,1)) norm_ANW S3:796-821 */
    for (theta = 0; theta < g->th_pixels; theta++) {
        int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
        long i;
        float *o = out + (size_t)g->segment_offset[theta] * pl;
        og_read_flat_float(g, atten_file, tmp, theta, 1.0f, 0, atten_file != NULL);
        for (i = 0; i < np * pl; i++) {
            float a = tmp[i];
            if (a < 1.0f) a = 1.0f;
            if (o[i] > 0.0f) {
                float den = o[i] * a;
                o[i] = den > g->p.epsilon2 ? 1.0f / den : 0.0f;
            }
        }
    }
    free(tmp); free(mask);
}

void og_prepare_norm_w(og_geom *g, int weighting, const float *norm_file, const float *atten_file, float *out)
{   /* prepare_norm S3:827-1003 for -W 0 (norm_UW S3:756-771), 1 (norm_AW S3:774-791) and
       2/3 (norm_ANW S3:796-821); Atten3D */
    const long pl = (long)g->N * g->views;
    float *tmp;
    int theta;
    long i;
    if (weighting >= 2) {
        og_prepare_norm(g, norm_file, atten_file, out);
        return;
    }
    /* the norm (or 1 everywhere) and the zero-column mask are the same for every weighting */
    og_prepare_norm(g, norm_file, NULL, out);
    if (weighting == 0 && !norm_file) {   /* S3:913-919: the inverse of 1 is 1 */
        for (i = 0; i < (long)g->segmentsize * pl; i++) out[i] = 1.0f;
        return;
    }
    /* og_prepare_norm(NULL atten) left 1/N where N > 0 and N elsewhere: read the norm again */
    tmp = theta_tmp(g);
    for (theta = 0; theta < g->th_pixels; theta++) {
        int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
        float *o = out + (size_t)g->segment_offset[theta] * pl;
        og_read_flat_float(g, norm_file, o, theta, 1.0f, 1, norm_file != NULL);
        if (weighting == 0) {
            for (i = 0; i < np * pl; i++) if (o[i] > 0.0f) o[i] = 1.0f;
        } else {
            og_read_flat_float(g, atten_file, tmp, theta, 1.0f, 0, atten_file != NULL);
            for (i = 0; i < np * pl; i++) {
                float a = tmp[i] < 1.0f ? 1.0f : tmp[i];     /* std::max(atten, 1.0f) */
                o[i] = o[i] > 0.0f ? 1.0f / a : 0.0f;
            }
        }
    }
    free(tmp);
}

/* read_flat_short_proj_ori_to_float_array_theta FIO:185-224: shorts to float, with the
   rebinning of sino_rebin_float FIO:77-104 (views summed as shorts, then bins in float) */
static void read_short_to_float(const og_geom *g, const int16_t *file, float *out, int theta)
{
    const int N = g->N, V = g->views, rebin = g->x_rebin * g->v_rebin;
    const long plane_out = (long)N * V, file_plane = (long)N * V * rebin;
    int plane, nplanes, i, v, vr, xr;
    theta = swap_theta(theta);
    nplanes = g->yr_top[theta] - g->yr_bottom[theta] + 1;
    file += (size_t)g->seg_offset[theta] * rebin;
    if (rebin == 1) {
        for (i = 0; i < nplanes * file_plane; i++) out[i] = file[i];
        return;
    }
    {
        const int r_size = N * g->x_rebin;
        int16_t *buf = (int16_t *)malloc(sizeof(int16_t) * file_plane);
        for (plane = 0; plane < nplanes; plane++) {
            memcpy(buf, file + plane * file_plane, sizeof(int16_t) * file_plane);
            for (v = 0; v < V; v++) {
                int16_t *dst = buf + (size_t)v * g->v_rebin * r_size;
                for (vr = 1; vr < g->v_rebin; vr++)
                    for (i = 0; i < r_size; i++) dst[i] = (int16_t)(dst[i] + dst[vr * r_size + i]);
            }
            for (v = 0; v < V; v++) {
                const int16_t *src = buf + (size_t)v * g->v_rebin * r_size;
                float *dst = out + plane * plane_out + (size_t)v * N;
                for (i = 0; i < N; i++) {
                    dst[i] = 0.0f;
                    for (xr = 0; xr < g->x_rebin; xr++) dst[i] += (float)(*src++);
                }
            }
        }
        free(buf);
    }
}

int og_prepare_w012(const og_geom *g, int weighting, int precorrected, const int16_t *data16_file,
                    const float *dataf_file, const float *delayed_file, const float *norm_file,
                    const float *scatter_file, const float *atten_file, float *data)
{   /* prepare_osem3dw0 S3:1190-1327, prepare_osem3dw1 S3:1339-1486, prepare_osem3dw2 S3:1497-1639
       (Atten3D, scatter as a full 3D sinogram) */
    const long pl = (long)g->N * g->views;
    float *tmp, *tmp2;
    int theta;
    long i;
    if (weighting < 0 || weighting > 2) return -1;
    if (!data16_file == !dataf_file) return -1;
    if (precorrected && !dataf_file) return -1;
    /* -W 0 with scatter and no norm subtracts a buffer the reference never filled (S3:1218-1235) */
    if (weighting == 0 && !precorrected && scatter_file && !norm_file) return -1;
    tmp = theta_tmp(g);
    tmp2 = theta_tmp(g);
    /* readsinogram S3:1066-1134 (prompt - delayed, or the true / precorrected sinogram) */
    for (theta = 0; theta < g->th_pixels; theta++) {
        float *o = data + (size_t)g->segment_offset[theta] * pl;
        if (data16_file) read_short_to_float(g, data16_file, o, theta);
        else og_read_flat_float(g, dataf_file, o, theta, 0.0f, 0, 1);
    }
    if (!precorrected && delayed_file)
        for (theta = 0; theta < g->th_pixels; theta++) {
            int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
            float *o = data + (size_t)g->segment_offset[theta] * pl;
            og_read_flat_float(g, delayed_file, tmp, theta, 0.0f, 0, 1);
            for (i = 0; i < np * pl; i++) o[i] = o[i] - tmp[i];          /* subprojection */
        }
    for (theta = 0; theta < g->th_pixels; theta++) {
        int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
        float *o = data + (size_t)g->segment_offset[theta] * pl;
        if (!precorrected) {
            if (weighting == 0) {
                if (norm_file) {                       /* S3:1218-1224, 1238-1245: t*N */
                    og_read_flat_float(g, norm_file, tmp, theta, 1.0f, 1, 1);
                    for (i = 0; i < np * pl; i++) o[i] = o[i] * tmp[i];
                    /* S3:1233: the reference subtracts tmp_prj1 -- the norm -- where the scatter
                       (read into tmp_prj2) was meant */
                    if (scatter_file)
                        for (i = 0; i < np * pl; i++) o[i] = o[i] - tmp[i];
                }
                if (atten_file) {                      /* S3:1247-1256 */
                    og_read_flat_float(g, atten_file, tmp, theta, 1.0f, 0, 1);
                    for (i = 0; i < np * pl; i++) o[i] = o[i] * tmp[i];
                }
            } else if (weighting == 1) {
                if (scatter_file) {                    /* S3:1370-1392: t - s/N, or t - s */
                    og_read_flat_float(g, scatter_file, tmp, theta, 0.0f, 0, 1);
                    if (norm_file) {
                        og_read_flat_float(g, norm_file, tmp2, theta, 1.0f, 1, 1);
                        for (i = 0; i < np * pl; i++) o[i] = tmp2[i] > 0.0f ? o[i] - tmp[i] / tmp2[i] : 0.0f;
                    } else {
                        for (i = 0; i < np * pl; i++) o[i] = o[i] - tmp[i];
                    }
                }
                if (norm_file) {                       /* S3:1398-1406: t*N */
                    og_read_flat_float(g, norm_file, tmp, theta, 1.0f, 1, 1);
                    for (i = 0; i < np * pl; i++) o[i] = o[i] * tmp[i];
                }
            } else if (scatter_file) {                 /* S3:1527-1540: t - s/N (N = 1 without a norm) */
                og_read_flat_float(g, scatter_file, tmp, theta, 0.0f, 0, 1);
                og_read_flat_float(g, norm_file, tmp2, theta, 1.0f, 1, norm_file != NULL);
                for (i = 0; i < np * pl; i++) o[i] = tmp2[i] > 0.0f ? o[i] - tmp[i] / tmp2[i] : 0.0f;
            }
        } else {
            if (weighting == 2) {                      /* S3:1557-1570: t/N */
                og_read_flat_float(g, norm_file, tmp, theta, 1.0f, 1, norm_file != NULL);
                for (i = 0; i < np * pl; i++) o[i] = tmp[i] > 0.0f ? o[i] / tmp[i] : 0.0f;
            }
            if (weighting >= 1 && atten_file) {        /* divideprojection simd_operation.cpp:24-38 */
                og_read_flat_float(g, atten_file, tmp, theta, 1.0f, 0, 1);
                for (i = 0; i < np * pl; i++) if (tmp[i] > 0.0f) o[i] = o[i] / tmp[i];
            }
        }
    }
    /* setnegative2zero simd_operation.cpp:78-96: _mm_max_ps(t, 0) keeps t only when t > 0 */
    for (i = 0; i < (long)g->segmentsize * pl; i++) data[i] = data[i] > 0.0f ? data[i] : 0.0f;
    free(tmp); free(tmp2);
    return 0;
}

void og_prepare_w3(const og_geom *g, const int16_t *prompt_file, const float *delayed_file,
                   const float *norm_file, const float *scatter_file, const float *atten_file,
                   int16_t *prompt, float *add)
{   /* prepare_osem3dw3 S3:1666-1770 (short prompts, float randoms, Atten3D) */
    const long pl = (long)g->N * g->views;
    float *tmp = theta_tmp(g);
    int theta;
    long i;
    for (theta = 0; theta < g->th_pixels; theta++)
        og_read_flat_short(g, prompt_file, prompt + (size_t)g->segment_offset[theta] * pl, theta);
    for (theta = 0; theta < g->th_pixels; theta++)
        og_read_flat_float(g, delayed_file, add + (size_t)g->segment_offset[theta] * pl, theta, 0.0f, 0,
                           delayed_file != NULL);
    for (theta = 0; theta < g->th_pixels; theta++) {
        int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
        float *a = add + (size_t)g->segment_offset[theta] * pl;
        int16_t *pr = prompt + (size_t)g->segment_offset[theta] * pl;
        og_read_flat_float(g, norm_file, tmp, theta, 1.0f, 1, norm_file != NULL);
        for (i = 0; i < np * pl; i++) {
            a[i] = a[i] * tmp[i];
            if (tmp[i] == 0.0f) pr[i] = 0;
        }
    }
    if (scatter_file)
        for (theta = 0; theta < g->th_pixels; theta++) {
            int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
            float *a = add + (size_t)g->segment_offset[theta] * pl;
            og_read_flat_float(g, scatter_file, tmp, theta, 0.0f, 0, 1);
            for (i = 0; i < np * pl; i++) a[i] = a[i] + tmp[i];
        }
    for (theta = 0; theta < g->th_pixels; theta++) {
        int np = g->yr_top[theta] - g->yr_bottom[theta] + 1;
        float *a = add + (size_t)g->segment_offset[theta] * pl;
        og_read_flat_float(g, atten_file, tmp, theta, 1.0f, 0, atten_file != NULL);
        for (i = 0; i < np * pl; i++) a[i] = a[i] * tmp[i];
    }
    free(tmp);
}

/* ------------------------------------------------------------------ */
/* Flat image files                                                    */
/* ------------------------------------------------------------------ */
void og_image_to_flat(const og_geom *g, const float *image, float *flat, int flip)
{   /* SaveImage S3:4023-4035: tmpimage[z][y][x] = image[x][y][z] (image[y][x][z] flipped), FOV only */
    const int N = g->N, Z4 = g->z4;
    int x, y, z;
    memset(flat, 0, sizeof(float) * (size_t)g->z_pixels * N * N);
    for (x = 0; x < N; x++)
        for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++)
            for (z = 0; z < g->z_pixels; z++)
                flat[((size_t)z * N + y) * N + x] = flip ? image[((size_t)y * N + x) * Z4 + z]
                                                         : image[((size_t)x * N + y) * Z4 + z];
}

void og_image_from_flat(const og_geom *g, const float *flat, float *image, int flip)
{   /* Init_Image -I S3:3970-3993: tmp[x][y] = plane z; image[y][x][z] = tmp[x][y] (image[x][y][z]
     * flipped), FOV only; the rest of the image is 0 here (the reference leaves it unset) */
    const int N = g->N, Z4 = g->z4;
    int x, y, z;
    memset(image, 0, sizeof(float) * (size_t)N * N * Z4);
    for (z = 0; z < g->z_pixels; z++)
        for (x = 0; x < N; x++)
            for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++) {
                const float v = flat[((size_t)z * N + x) * N + y];
                if (flip) image[((size_t)x * N + y) * Z4 + z] = v;
                else image[((size_t)y * N + x) * Z4 + z] = v;
            }
}
