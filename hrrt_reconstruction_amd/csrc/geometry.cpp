// geometry.cpp -- host geometry tables (see geometry.h).
#include "geometry.h"

#include <cmath>

namespace hrrt {

namespace {

// PI is used by the reference (PJ:480) but defined nowhere in its tree; M_PI is used here.
constexpr double kPi = 3.14159265358979323846;

// get_info (PJ:159-191): plane range and file offsets of one group.
void group_info(int rings, int span, int group, int &zmin, int &zmax, int &imin, int &imin1) {
    if (group == 0) {
        zmin = 0; zmax = 2 * rings - 2; imin = imin1 = 0;
        return;
    }
    zmin = (span + 1) / 2 + (group - 1) * span;
    zmax = 2 * rings - 2 - zmin;
    imin = 2 * rings - 1;
    for (int j = 1; j < group; ++j) imin += 2 * (2 * rings - 1 - 2 * ((span + 1) / 2 + (j - 1) * span));
    imin1 = imin + (zmax - zmin) + 1;
}

// sequence (S3:535-580): each next subset maximises the circular distance to the previous one.
void subset_sequence(int sets, std::vector<int> &order) {
    std::vector<int> done(sets, 0);
    order.assign(sets, 0);
    done[0] = 1;
    for (int i = 1; i < sets; ++i) {
        int best = 0, dbest = 0;
        for (int j = 0; j < sets; ++j) {
            if (done[j]) continue;
            int d = std::abs(j - order[i - 1]);
            if (sets - d < d) d = sets - d;
            if (d > dbest) { dbest = d; best = j; }
        }
        order[i] = best;
        done[best] = 1;
    }
}

}  // namespace

bool Geometry::build(const hrrt_params &in, std::string &err) {
    p = in;
    if (p.rings < 2 || p.def_elements <= 0 || p.def_angles <= 0 || p.span < 1 || p.subsets < 1) {
        err = "invalid scanner parameters";
        return false;
    }
    // GetScannerParameter S3:3488-3539
    if (p.x_pixels != 0 && p.x_pixels < p.def_elements) x_rebin = v_rebin = p.def_angles / p.x_pixels;
    sino_sampling = p.bin_size * x_rebin;
    z_pixels = p.rings * 2 - 1;
    z4 = ((z_pixels + 3) / 4) * 4;
    ring_spacing = 2.0f * p.plane_sep;
    groupmax = (p.maxdel - (p.span - 1) / 2) / p.span;
    N = p.def_elements / x_rebin;
    views = p.def_angles / v_rebin;
    z_min = p.z_min;
    z_max = p.z_max;
    if (z_max == 0) z_max = z_pixels - z_min;
    if (p.npskip > z_min) z_min = p.npskip;
    if (z_max > z_pixels - p.npskip) z_max = z_pixels - p.npskip;
    if (p.x_pixels != 0 && p.x_pixels != N) {
        err = "x_pixels must equal def_elements or def_elements/rebin (zoom != 1 is not supported)";
        return false;
    }
    // dependencies PJ:247-300
    x_size = (sino_sampling * N) / (1.0f * N);
    y_size = x_size;
    z_size = ring_spacing / 2.0f;
    if (views % p.subsets != 0) { err = "views must be a multiple of subsets (PJ:270)"; return false; }
    if (N % 4 != 0 || N < 8) { err = "radial bins must be a multiple of 4"; return false; }
    if (z_pixels > 256) { err = "more than 128 rings is not supported (z_pixels > 256)"; return false; }
    if (groupmax < 1 || p.span < 2) { err = "need at least one oblique segment group and span >= 2"; return false; }
    sviews = views / p.subsets;
    th_pixels = 2 * groupmax + 1;
    K = th_pixels / 2;
    zshift = p.span - 1;
    rfov = p.rel_fov * x_size * N / 2.0f;
    {
        // every rotated sample stays inside [0, N-2] when the FOV radius (in pixels) <= N/2 - 2
        float rpix = rfov / x_size;
        if (!(rpix <= N / 2 - 2)) { err = "FOV radius too close to the image edge (rel_fov*N/2 > N/2-2)"; return false; }
    }

    // FOV rows PJ:339-356
    cyl_lo.assign(N, 0);
    cyl_hi.assign(N, 0);
    const int off = N / 2;
    for (int y = 0; y < N; y++) {
        bool in = false;
        for (int x = 0; x < N; x++) {
            float dx = (x - off) * x_size * (x - off) * x_size;
            float dy = (y - off) * y_size * (y - off) * y_size;
            if (std::sqrt(dx + dy) <= rfov) {
                if (!in) cyl_lo[y] = x;
                in = true;
            } else {
                if (in) cyl_hi[y] = x;
                in = false;
            }
        }
    }

    // segment angles PJ:424-444 (+ end-plane angles vcos/vsin)
    sin_theta.assign(th_pixels, 0.f);
    cos_theta.assign(th_pixels, 0.f);
    std::vector<float> vcos(groupmax * zshift), vsin(groupmax * zshift);
    sin_theta[0] = 0.0f;
    cos_theta[0] = 1.0f;
    const float two_r = 2 * p.crystal_radius;
    for (int g = 1; g <= groupmax; g++) {
        float t = g * p.span * ring_spacing;
        t = std::sqrt(t * t + two_r * two_r);
        sin_theta[2 * g - 1] = g * p.span * ring_spacing / t;
        cos_theta[2 * g - 1] = (2.0f * p.crystal_radius) / t;
        sin_theta[2 * g] = -sin_theta[2 * g - 1];
        cos_theta[2 * g] = cos_theta[2 * g - 1];
        for (int z = 0; z < zshift; z++) {
            int mean_rd = (p.span * g - (p.span - 1) / 2) + (z + 1) / 2;
            float u = mean_rd * ring_spacing;
            u = std::sqrt(u * u + two_r * two_r);
            vcos[(g - 1) * zshift + z] = (2.0f * p.crystal_radius) / u;
            vsin[(g - 1) * zshift + z] = mean_rd * ring_spacing / u;
        }
    }

    // plane ranges PJ:446-467
    yr_bottom.assign(th_pixels, 0);
    yr_top.assign(th_pixels, 0);
    seg_offset.assign(th_pixels, 0);
    segment_offset.assign(th_pixels, 0);
    yr_top[0] = z_pixels - 1;
    for (int g = 1; g <= groupmax; g++) {
        int zmin, zmax, ip, im;
        group_info(p.rings, p.span, g, zmin, zmax, ip, im);
        yr_bottom[2 * g] = yr_bottom[2 * g - 1] = zmin;
        yr_top[2 * g] = yr_top[2 * g - 1] = zmax;
        seg_offset[2 * g] = ip * N * views;
        seg_offset[2 * g - 1] = im * N * views;
    }
    segmentsize = 0;
    for (int s = 0; s < th_pixels; s++) {
        segment_offset[s] = segmentsize;
        segmentsize += nplanes(s);
    }
    if (yr_bottom[th_pixels - 1] < 1) { err = "invalid segment range"; return false; }

    // back-projection rescale per segment PJ:1638-1640
    bp_rescale.assign(th_pixels, 0.f);
    for (int s = 0; s < th_pixels; s++)
        bp_rescale[s] = cos_theta[s] * cos_theta[s] * cos_theta[s] * p.span * ring_spacing / (2 * p.crystal_radius);

    // view angles PJ:476-486
    sin_psi.assign(views, 0.f);
    cos_psi.assign(views, 0.f);
    cos_psi[0] = 1.0f;
    sin_psi[views / 2] = 1.0f;
    cos_psi[views / 2] = 0.0f;
    const float dpsi = static_cast<float>(kPi / views);
    for (int v = 1; v < views / 2; v++) {
        sin_psi[v] = std::sin(v * dpsi);
        cos_psi[v] = static_cast<float>(std::sqrt(1. - sin_psi[v] * sin_psi[v]));
        sin_psi[views - v] = sin_psi[v];
        cos_psi[views - v] = -cos_psi[v];
    }

    // subsets PJ:144-152, S3:535
    vieworder.assign(p.subsets * sviews, 0);
    for (int s = 0; s < p.subsets; s++)
        for (int k = 0; k < sviews; k++) vieworder[s * sviews + k] = s + p.subsets * k;
    subset_sequence(p.subsets, order);

    // tantheta / rescalearray PJ:529-542
    tantheta.assign(groupmax * z_pixels, 0.f);
    rescalearray.assign(groupmax * z_pixels, 0.f);
    for (int th = 2, k = 0; th < th_pixels; th += 2, k++) {
        for (int yr = yr_bottom[th]; yr <= yr_top[th]; yr++) {
            float t = sin_theta[th] / cos_theta[th];
            float r = y_size / std::fabs(cos_theta[th]);
            int zs = yr - yr_bottom[th];
            if (zs >= zshift) zs = yr_top[th] - yr;
            if (zs < zshift) {
                t = vsin[k * zshift + zs] / vcos[k * zshift + zs];
                t = -t;
                r = y_size / std::fabs(vcos[k * zshift + zs]);
            }
            tantheta[k * z_pixels + yr] = -t / 1.0f;
            rescalearray[k * z_pixels + yr] = r;
        }
    }

    // forward end-plane lookup PJ:544-562
    const int L = 2 * zshift;
    fl_frac.assign(static_cast<size_t>(N) * groupmax * L, 0.f);
    fl_z.assign(static_cast<size_t>(N) * groupmax * L, 0);
    for (int y = 0; y < N; y++) {
        for (int th = 2, k = 0; th < th_pixels; th += 2, k++) {
            size_t base = (static_cast<size_t>(y) * groupmax + k) * L;
            int i = 0;
            for (int yr = yr_bottom[th]; yr < yr_bottom[th] + zshift; yr++, i++) {
                float zz = yr + ((y - off) * tantheta[k * z_pixels + yr]);
                fl_z[base + i] = static_cast<int>(zz);
                fl_frac[base + i] = zz - fl_z[base + i];
            }
            for (int yr = yr_top[th] - zshift + 1; yr <= yr_top[th]; yr++, i++) {
                float zz = yr + ((y - off) * tantheta[k * z_pixels + yr]);
                fl_z[base + i] = static_cast<int>(zz);
                fl_frac[base + i] = zz - fl_z[base + i];
            }
        }
    }

    // backward lookup PJ:670-700
    bl_zsa.assign(static_cast<size_t>(N) * groupmax, 0);
    bl_zea.assign(bl_zsa.size(), 0);
    bl_yra.assign(bl_zsa.size(), 0);
    bl_frac.assign(bl_zsa.size(), 0.f);
    for (int y = 0; y < N; y++) {
        for (int th = 2, k = 0; th < th_pixels; th += 2, k++) {
            float tb = -sin_theta[th] / cos_theta[th];
            float y0 = -(y - off) * tb;
            float fr = std::fabs(y0) - static_cast<int>(std::fabs(y0));
            int yi = static_cast<int>(y0);
            if (y0 < 0) { fr = 1 - fr; yi -= 1; }
            int zs = yr_bottom[th] - 1 - yi, ze = yr_top[th] - yi, yr = yr_bottom[th] - 1;
            if (zs < 0) { yr = yr - zs; zs = 0; }
            if (ze > z_pixels - 1) ze = z_pixels - 1;
            size_t i = static_cast<size_t>(y) * groupmax + k;
            bl_zsa[i] = zs; bl_zea[i] = ze; bl_yra[i] = yr; bl_frac[i] = fr;
        }
    }

    // incremental back-projection breakpoints PJ:718-784: walking y away from the
    // centre row, a breakpoint is a row where floor((c-y)*tan) changes; it records
    // the lookup of that row and of the previous breakpoint row
    blut.assign(static_cast<size_t>(N) * groupmax, Blut{0, 0, 0, 0, 0, 0, 0.f, 0.f, -1});
    blut_flag.assign(static_cast<size_t>(N) * groupmax, 0);
    for (int k = 0; k < groupmax; k++) {
        const int c = N / 2;
        auto L = [&](int y) { return static_cast<size_t>(y) * groupmax + k; };
        blut_flag[L(c)] = 1;
        blut[L(c)] = Blut{bl_zsa[L(c)], bl_yra[L(c)], bl_zea[L(c)], 0, 0, 0, bl_frac[L(c)], 0.f, 0};
        for (int dir = 0; dir < 2; dir++) {
            int prev = 0, cnt = 0, grp = c;
            if (dir == 1) blut_flag[L(c)] = 0;
            for (int y = dir == 0 ? c + 1 : c - 1; dir == 0 ? y < N : y >= 0; y += dir == 0 ? 1 : -1) {
                blut_flag[L(y)] = 0;
                const int dd = bl_yra[L(y)] - bl_zsa[L(y)];
                if (prev == dd) {
                    cnt++;
                } else {
                    blut_flag[L(y)] = cnt;
                    blut[L(y)] = Blut{bl_zsa[L(y)], bl_yra[L(y)], bl_zea[L(y)], bl_zsa[L(grp)], bl_yra[L(grp)],
                                      bl_zea[L(grp)], bl_frac[L(y)], bl_frac[L(grp)], cnt};
                    grp = y;
                    cnt = 0;
                    prev = dd;
                }
            }
        }
    }

    // padded pair-column layout: each segment's planes start on a multiple of 4
    // with one zero guard plane on each side
    seg_data.assign(th_pixels, 0);
    int pos = 0;
    for (int s = 0; s < th_pixels; s++) {
        int d = ((pos + 1 + 3) / 4) * 4;
        seg_data[s] = d;
        pos = d + nplanes(s) + 1;
    }
    spp = ((pos + 3) / 4) * 4;
    return true;
}

}  // namespace hrrt
