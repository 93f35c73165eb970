// delays.h -- host context shared by the smoothed randoms (randoms.hip) and the list-mode
// histogrammer (lmhist.hip): the HRRT crystal geometry, the segment tables and the module-pair
// LUT that the reference's lor_sinogram tables hold for both (lor_sinogram_map.h:27-32).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/hrrt_randoms.h"

namespace hrrt_dl {

constexpr int NHEADS = 8, NDOIS = 2, NX = 72, NY = 104;
constexpr int NXD = NX * NDOIS;                   // 144 crystals across a head, both layers
constexpr int NCRYS = NHEADS * NDOIS * NX * NY;   // 119808
constexpr int XYH = NX * NY * NHEADS;             // one layer, [ring][head][x]
constexpr int HX = NHEADS * NX;                   // one ring over the heads
constexpr int NMP = 20;
constexpr int SEGMAP = 63;                        // segment rows of the plane map (lor_sinogram_map.cpp:154-158)

// HRRT_MPAIRS geometry_info.h:68-73
inline constexpr int kMpairs[NMP + 1][2] = {{-1, -1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {0, 6}, {1, 3}, {1, 4},
                                     {1, 5},   {1, 6}, {1, 7}, {2, 4}, {2, 5}, {2, 6}, {2, 7}, {3, 5},
                                     {3, 6},   {3, 7}, {4, 6}, {4, 7}, {5, 7}};
// span-3 segment -> span-9 segment (conversiontable, segment_info_fromarchive.cpp:23-26)
inline constexpr int kS3toS9[45] = {0, 0, 0, 1, 2, 1, 2, 1, 2, 3, 4, 3, 4, 3, 4, 5, 6, 5, 6, 5, 6, 7, 8,
                             7, 8, 7, 8, 9, 10, 9, 10, 9, 10, 11, 12, 11, 12, 11, 12, 13, 14, 13, 14, 13, 14};

struct Sol {       // SOL lor_sinogram_map.h:18-22 (also the LUT file record)
    int nsino;
    float d, z;
};
static_assert(sizeof(Sol) == 12, "LUT record is 12 bytes");


// Records msg as hrrt_last_error()'s message and returns code.
int fail(int code, const std::string &msg);

#define DTRY(expr)                                                                                              \
    do {                                                                                                        \
        hipError_t e_ = (expr);                                                                                 \
        if (e_ != hipSuccess) return hrrt_dl::fail(HRRT_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace hrrt_dl

struct hrrt_delays {
    using Sol = hrrt_dl::Sol;
    static constexpr int NHEADS = hrrt_dl::NHEADS, NY = hrrt_dl::NY, SEGMAP = hrrt_dl::SEGMAP;
    int span = 9, maxrd = 67, lr_type = 0;
    int nprojs = 256, nviews = 288, nsino = 0, nplanes = 0, nsegs = 0;
    float tau = 6e-9f, depth = 1.0f;
    double binsize = 0, plane_sep = 0, crystal_radius = 0, d_tan_theta = 0;
    double sin_head[NHEADS]{}, cos_head[NHEADS]{};
    std::vector<Sol> lut;                    // [21][144][144], module pair 0 unused
    float zpos[NY]{}, zpos2[NY]{};
    int segz0[SEGMAP]{}, segzmax[SEGMAP]{}, segoff[SEGMAP]{};
    long long pairs_full = 0;
    int nlor = 0, device = 0;
    std::vector<int4> h_lor;                 // LORs grouped by bin (device copy d_lor)
    std::vector<int> h_lor_off;
    int4 *d_lor = nullptr;
    int *d_lor_off = nullptr;
    int2 *d_seg = nullptr;
    float2 *d_z = nullptr;
    float *d_sings = nullptr;                // transposed singles [layer][head][x][ring], then column maxima
    float *d_cols = nullptr;                 // [nsino][nplanes rounded up to 64] columns
    // list-mode histogramming: the LUT by (module pair, A crystal, B crystal) on the device,
    // and its scratch (uploaded / allocated on first use)
    int4 *d_sol = nullptr;                   // [21][144][144] {nsino, bits(d), bits(z), 0}
    unsigned *d_lm_keys = nullptr;           // [2][lm_key_cap] key slots of the events (2*bin + delayed), then sorted
    long long lm_key_cap = 0;
    void *d_lm_sort_tmp = nullptr;           // radix sort scratch
    size_t lm_sort_bytes = 0;
    unsigned *d_lm_tiles = nullptr;          // [ntiles + 1] first sorted key of each sinogram tile
    unsigned short *d_lm_coin_seg = nullptr; // [lm_coin_cap][8 heads][1024] crystal entries of a chunk's events
    unsigned *d_lm_coin_cnt = nullptr;       // [lm_coin_cap][8 heads] entries per chunk and head
    long long lm_coin_cap = 0;
    unsigned *d_lm_coin_part = nullptr;      // per chunk range and head: [2 maps][2 layers][104][72] counts
    unsigned *d_lm_chunks = nullptr;         // [65][lm_chunk_cap]: per 2048-word chunk its walk summary and
                                             // entry parity, then its 64 out-of-sync mask words
    long long lm_chunk_cap = 0;
    unsigned long long *d_lm_stats = nullptr;

    // calc_det_to_phy geometry_info.cpp:137-163: block/crystal offsets in float, the head
    // rotation in double, stored as float (constants geometry_info.h:51-55 are floats)
    void crystal(int head, int layer, int detx, int dety, float pos[3]) const {
        const float csize = 0.22f, cgap = 0.02f, bgap = 0.05f;
        const float bsize = (float)(8.0 * (double)csize + (double)(7.0f * cgap));
        const float xhsize = (float)(9.0 * (double)bsize + (double)(8.0f * bgap));
        const float bx = (float)(detx / 8) * (bsize + bgap) + (float)(detx % 8) * (csize + cgap);
        const double x = (double)bx - (double)xhsize / 2.0 + (double)csize / 2.0;
        const double y = crystal_radius + (double)(depth * (float)(1 - layer)) + 0.5 * (double)depth;
        const float bz = (float)(dety / 8) * (bsize + bgap) + (float)(dety % 8) * (csize + cgap);
        const double s = sin_head[head], c = cos_head[head];
        pos[0] = (float)(x * c + y * s);
        pos[1] = (float)(-x * s + y * c);
        pos[2] = bz;
    }

    // init_phy_to_pro lor_sinogram_map.cpp:99-138
    Sol lor(const float a[3], const float b[3]) const {
        const double dy = (double)(a[1] - b[1]), dx = (double)(a[0] - b[0]);
        double d = std::sqrt(dx * dx + dy * dy);
        float r = (float)((double)(a[1] * b[0] - a[0] * b[1]) / d);
        float phi = (float)std::atan2(dx, dy);
        if (phi < 0.0f) {
            phi = phi + (float)M_PI;
            r = -r;
            d = -d;
        }
        if ((double)phi >= M_PI) {
            phi = 0.0f;
            r = -r;
            d = -d;
        }
        int view = (int)((double)((float)nviews * phi) / M_PI);
        if (view >= 288 || view < 0) view = 0;
        const float z = (float)(((double)a[0] * dx + (double)a[1] * dy) / (d * d));
        const int bin = (int)((double)(nprojs / 2) + ((double)r / binsize + 0.5));
        Sol s;
        s.nsino = (bin < 0 || bin > nprojs - 1) ? -1 : bin + view * nprojs;
        s.d = (float)(1.0 / d / d_tan_theta);
        s.z = (float)((double)z / plane_sep);
        return s;
    }

    // segment sizes of one span (init_seginfo segment_info_fromarchive.cpp:35-78): first
    // plane, last plane and file offset per segment; returns the plane total
    int segments(int sp, int *z0, int *zmax, int *off) const {
        const int n = 2 * (maxrd / sp) + 1, np = 2 * NY - 1, half = (sp + 1) / 2;
        int total = 0;
        for (int i = 0; i < n; i++) {
            z0[i] = i == 0 ? 0 : half + sp * ((i - 1) / 2);
            const int nz = np - 2 * z0[i];
            zmax[i] = z0[i] + nz - 1;
            off[i] = total - z0[i];
            total += nz;
        }
        return total;
    }
};


namespace hrrt_dl {
// Device copies of the tables (d_lor, d_lor_off, d_seg, d_z, d_sings) on the current device,
// made once; later calls must run with the same device current.
int upload(hrrt_delays &c);
}  // namespace hrrt_dl
