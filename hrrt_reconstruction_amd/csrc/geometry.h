// geometry.h -- scanner/sinogram geometry and the lookup tables the HIP
// kernels consume.  Host code; restates GetScannerParameter
// (hrrt_osem3d/hrrt_osem3d_span3.cpp:3452-3551) and dependencies
// (hrrt_osem3d/hrrt_osem3dv_sbrt_span3.cpp:199-800) with the same float
// evaluation order, so every table value is identical to the reference's.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hrrt_osem.h"

namespace hrrt {

struct Geometry {
    hrrt_params p{};
    // derived scalars
    int x_rebin = 1, v_rebin = 1;
    float sino_sampling = 0, z_size = 0, ring_spacing = 0, x_size = 0, y_size = 0, rfov = 0;
    int z_pixels = 0, z4 = 0, groupmax = 0, th_pixels = 0, N = 0, views = 0, sviews = 0;
    int segmentsize = 0, zshift = 0, z_min = 0, z_max = 0, K = 0;
    // per segment
    std::vector<int> yr_bottom, yr_top, seg_offset, segment_offset;
    std::vector<float> sin_theta, cos_theta, bp_rescale;
    // per view
    std::vector<float> sin_psi, cos_psi;
    // FOV rows: cylwiny[y] = {first x, one past last x}
    std::vector<int> cyl_lo, cyl_hi;
    // [groupmax][z_pixels]
    std::vector<float> tantheta, rescalearray;
    // forward end-plane lookup [N][groupmax][2*zshift]
    std::vector<float> fl_frac;
    std::vector<int> fl_z;
    // backward lookup [N][groupmax]
    std::vector<int> bl_zsa, bl_zea, bl_yra;
    std::vector<float> bl_frac;
    // incremental back-projection breakpoints [N][groupmax] (PJ:718-784)
    struct Blut { int z_start, yr_start, z_end, z_start_pre, yr_start_pre, z_end_pre; float coef, coef_pre; int num; };
    std::vector<Blut> blut;
    std::vector<int> blut_flag;
    // subsets
    std::vector<int> vieworder;   // [subsets][sviews]
    std::vector<int> order;       // [subsets]
    // padded pair-column plane layout: segment s plane yr -> seg_data[s] + yr - yr_bottom[s];
    // zero guard planes at seg_data[s]-1 and seg_data[s]+nplanes(s)
    std::vector<int> seg_data;
    int spp = 0;                  // padded planes per column (multiple of 4)

    bool build(const hrrt_params &in, std::string &err);
    int nplanes(int s) const { return yr_top[s] - yr_bottom[s] + 1; }
};

}  // namespace hrrt
