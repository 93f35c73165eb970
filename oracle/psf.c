/* PSF resolution modelling (-B), restated from the reference for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never part of the product).
 *
 *   psf.cpp:20-70    make1d_psf_filter  kernel: double Gaussians, float storage
 *   psf.cpp:168-219  convolvex          x pass, renormalised by the taps inside the FOV row
 *   psf.cpp:324-439  convolveyz         y pass, then the z pass restricted to the support
 *                                       [z_min, z_max) and normalised by normz and 1/norm_y
 *   psf.cpp:441-476  convolve3d         x pass into a buffer, then the y/z pass into out
 *
 * Quirks kept (they change the result): the tap loops run mask = -h .. h-1, so the last
 * tap of every kernel is never used; the z pass uses the y kernel and the y kernel's
 * half-width, while normz sums the z kernel at the same indices; the x pass skips the
 * division when the used taps sum to exactly 1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hrrt_oracle.h"

/* make1d_psf_filter psf.cpp:20-70 (HRRT model "999" defaults when fwhm1 == 0) */
static int make_kernel(float fwhm1, float fwhm2, float ratio, float pixel_size, float **out)
{
    int bin, bin0, size;
    float norm_psf, sigma1, sigma2, *k;
    if (fwhm1 == 0.0f) {
        fwhm1 = (float)(0.9 * sqrt(8.0 * log(2.0)));
        fwhm2 = (float)(2.5 * sqrt(8.0 * log(2.0)));
        ratio = (float)0.05;
    }
    if (!(fwhm1 > 0.0f)) return -1;
    if (fwhm2 > 0.0f && ratio > 0.0f)
        size = 2 * ((int)(1.5 * fwhm2 * sqrt(log(10.0) / log(2.)) / pixel_size / 2.0)) + 1;
    else
        size = 2 * ((int)(1.5 * fwhm1 * sqrt(log(10.0) / log(2.)) / pixel_size / 2.0)) + 1;
    k = (float *)calloc((size_t)size, sizeof(float));
    sigma1 = (float)(fwhm1 / (sqrt(8.0 * log(2.0)) * pixel_size));
    sigma2 = (float)(fwhm2 / (sqrt(8.0 * log(2.0)) * pixel_size));
    bin0 = (size - 1) / 2;
    norm_psf = 0.0f;
    for (bin = 0; bin < size; bin++) {
        k[bin] = (float)exp(-0.5 * (pow((double)(bin - bin0) / sigma1, 2.0)));
        if (fwhm2 > 0.0f && ratio > 0.0f)
            k[bin] = (float)(k[bin] + ratio * exp(-0.5 * (pow((double)(bin - bin0) / sigma2, (double)2.0))));
        norm_psf += k[bin];
    }
    for (bin = 0; bin < size; bin++) k[bin] /= norm_psf;
    *out = k;
    return size;
}

og_psf *og_psf_create(const og_geom *g, float fwhm1, float fwhm2, float ratio)
{
    og_psf *p = (og_psf *)calloc(1, sizeof(og_psf));
    const float px[3] = {g->x_size, g->y_size, g->z_size};
    int a, z, mask;
    for (a = 0; a < 3; a++) {
        p->size[a] = make_kernel(fwhm1, fwhm2, ratio, px[a], &p->k[a]);
        if (p->size[a] < 1) {
            og_psf_destroy(p);
            return NULL;
        }
    }
    /* the z pass reads the z kernel at the y kernel's indices (psf.cpp:356-374) */
    if (p->size[2] < p->size[1]) {
        og_psf_destroy(p);
        return NULL;
    }
    p->normz = (float *)calloc((size_t)g->z4, sizeof(float));
    {
        const int h = (p->size[1] - 1) / 2;
        for (mask = -h; mask < h; mask++)
            for (z = 0; z < g->z4; z++) {
                int m2 = z + mask;
                if (m2 < g->z_min || m2 >= g->z_max) continue;
                p->normz[z] += p->k[2][mask + h];
            }
        for (z = 0; z < g->z4; z++)
            if (p->normz[z] > 0) p->normz[z] = 1 / p->normz[z];
    }
    return p;
}

void og_psf_destroy(og_psf *p)
{
    int a;
    if (!p) return;
    for (a = 0; a < 3; a++) free(p->k[a]);
    free(p->normz);
    free(p);
}

int og_psf_kernel(const og_psf *p, int axis, float *out)
{
    if (out) memcpy(out, p->k[axis], sizeof(float) * (size_t)p->size[axis]);
    return p->size[axis];
}

/* convolve3d psf.cpp:441-476: x pass image -> buf, then y/z pass buf -> out (out may be image) */
void og_convolve3d(const og_geom *g, const og_psf *p, const og_real *image, og_real *out, og_real *buf)
{
    const int N = g->N, Z4 = g->z4;
    og_real *acc = (og_real *)malloc(sizeof(og_real) * (size_t)Z4);
    int x, y, z, mask;
    {   /* convolvex psf.cpp:168-219 */
        const float *f = p->k[0];
        const int h = (p->size[0] - 1) / 2;
        for (x = 0; x < N; x++) {
            if (g->cylwiny[2 * x + 1] == 0) continue;
            for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++) {
                float norm = 0;
                og_real *o = buf + ((size_t)x * N + y) * Z4;
                memset(acc, 0, sizeof(og_real) * (size_t)Z4);
                for (mask = -h; mask < h; mask++) {
                    const int m2 = x + mask;
                    const og_real *src;
                    if (m2 < g->cylwiny[2 * y] || m2 >= g->cylwiny[2 * y + 1]) continue;
                    norm += f[mask + h];
                    src = image + ((size_t)m2 * N + y) * Z4;
                    for (z = 0; z < Z4; z++) acc[z] = acc[z] + src[z] * f[mask + h];
                }
                if (norm != 1 && norm != 0) {
                    const float coef = 1 / norm;
                    for (z = 0; z < Z4; z++) o[z] = acc[z] * coef;
                } else {
                    memcpy(o, acc, sizeof(og_real) * (size_t)Z4);
                }
            }
        }
    }
    {   /* convolveyz psf.cpp:324-439 */
        const float *f = p->k[1];
        const int h = (p->size[1] - 1) / 2;
        for (x = 0; x < N; x++) {
            if (g->cylwiny[2 * x + 1] == 0) continue;
            for (y = g->cylwiny[2 * x]; y < g->cylwiny[2 * x + 1]; y++) {
                float norm = 0;
                og_real *o = out + ((size_t)x * N + y) * Z4;
                memset(acc, 0, sizeof(og_real) * (size_t)Z4);
                for (mask = -h; mask < h; mask++) {
                    const int m2 = y + mask;
                    const og_real *src;
                    if (m2 < g->cylwiny[2 * x] || m2 >= g->cylwiny[2 * x + 1]) continue;
                    norm += f[mask + h];
                    src = buf + ((size_t)x * N + m2) * Z4;
                    for (z = 0; z < Z4; z++) acc[z] = acc[z] + src[z] * f[mask + h];
                }
                memset(o, 0, sizeof(og_real) * (size_t)Z4);
                if (norm != 0) {
                    float coef;
                    for (mask = -h; mask < h; mask++) {
                        int zstart = g->z_min - mask, zend = g->z_max - mask, z2;
                        if (zstart < 0) zstart = 0;
                        if (zend > g->z_pixels) zend = g->z_pixels;
                        z2 = mask + zstart;
                        for (z = zstart; z < zend; z++, z2++) o[z] = o[z] + acc[z2] * f[mask + h];
                    }
                    coef = 1 / norm;
                    for (z = 0; z < Z4; z++) o[z] = o[z] * (p->normz[z] * coef);
                }
            }
        }
    }
    free(acc);
}
