// randoms.hip -- smoothed randoms from crystal singles (gen_delays_lib) on gfx950, and its
// C ABI (include/hrrt_randoms.h).
//
// Host side (once per context): HRRT crystal geometry, segment tables, the module-pair LUT
// (each transaxial crystal pair of 20 head pairs -> sinogram bin, 1/(d tan), closest-approach
// z), the LUT grouped by sinogram bin, and the singles iteration.
// Device side (per smoothing):
//   sings_prepare_kernel    singles [layer][ring][head][x] -> [layer][head][x][ring], so that a
//                           crystal's 104 rings are contiguous, and each column's largest magnitude
//   delays_column_kernel    one workgroup per sinogram bin: every crystal pair of every LOR that
//                           lands in the bin, accumulated into the bin's column of planes in
//                           LDS (64-bit fixed point, ds_add_u64), then stored into the
//                           plane-major output (span 9) or as a contiguous column (span 3)
//   column_transpose_kernel span 3: [bin][plane] columns -> [plane][view][bin], 64 x 64 tiles
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hrrt_randoms.h"
#include "delays.h"
#include "kernels.h"

using namespace hrrt_dl;

namespace {

struct Tables {    // device tables of one context
    const int4 *lor;      // [nlor] {A ring-column offset, B ring-column offset, bits(d), bits(z)}
    const int *lor_off;   // [nsino + 1] LORs of bin b: lor[lor_off[b] .. lor_off[b+1])
    const int2 *seg;      // [128] {first | last << 16, file offset} by 2*(int)seg + (seg < 0) + 64
    const float2 *z;      // [104] {ring z (cm), ring z in planes + 0.5}
    int nsino, nplanes, window;
};

// ---------------------------------------------------------------- kernels

// One workgroup: singles [layer][ring][head][x] -> [layer][head][x][ring] (a crystal's 104 rings
// contiguous), then the largest |singles| of each such crystal column into out[NCRYS + column],
// which sets the fixed-point scale of every sinogram bin.
constexpr int PREP_THREADS = 1024;
constexpr int NCOLS = NCRYS / NY;   // 1152 crystal columns

__global__ __launch_bounds__(PREP_THREADS) void sings_prepare_kernel(const float *__restrict__ in,
                                                                     float *__restrict__ out) {
    for (int i = threadIdx.x; i < NCRYS; i += PREP_THREADS) {
        const int ring = i % NY, x = (i / NY) % NX, head = (i / (NY * NX)) % NHEADS, layer = i / (NY * NX * NHEADS);
        out[i] = in[layer * XYH + ring * HX + head * NX + x];
    }
    for (int j = threadIdx.x; j < NCOLS; j += PREP_THREADS) {
        const int x = j % NX, head = (j / NX) % NHEADS, layer = j / (NX * NHEADS);
        const float *src = in + layer * XYH + head * NX + x;
        float m = 0.0f;
        for (int ring = 0; ring < NY; ring++) m = fmaxf(m, fabsf(src[ring * HX]));
        out[NCRYS + j] = m;
    }
}

// Exponent k of a bin's fixed-point unit 2^-k.  A term s_A*tau*T*s_B of the bin is at most
// about m = max over its LORs of colmax(A)*colmax(B)*tau*T < 2^E, and an element of the bin
// sums at most nlor*1664 terms (104 A rings, each with at most 16 B rings on one segment and
// plane), so terms below 2^(62-bits) with 2^bits >= nlor*1664 keep every sum inside int64.
__device__ inline int fixed_point_exponent(float m, int nlor) {
    int e = 0;
    if (m > 0.0f && m <= 3.0e38f) frexpf(m, &e);
    const int bits = 32 - __clz(nlor * 1664);
    return max(min(62 - bits - e, 127), -126);
}

// Term -> int64 in units of 2^-k, rounded to nearest: x = v*2^k is exact (a power of two), and
// so are its split into x = hi*2^32 + lo (hi integral) and lo's rounding below 2^32.
// (x - hi*2^32 is exact: below 2^32 hi is 0, above it x is a multiple of 2^9 and the difference
// below 2^32 has at most 23 bits, so one fused multiply-add gives it.)
__device__ inline unsigned long long to_fixed(float x) {
    const float hi = floorf(x * 0x1p-32f);
    const float lo = __builtin_fmaf(hi, -0x1p32f, x);
    return ((unsigned long long)(long long)(int)hi << 32) + (unsigned long long)(unsigned)rintf(lo);
}

// One sinogram bin per workgroup of 512 threads (8 waves).  A wave covers 64 B rings (wave
// parity: rings 0..63 or 64..103) for every fourth A ring (row = wave / 2).  The A singles
// and ring z of the LOR sit one per lane in registers and are read with v_readlane for the
// wave-uniform A ring, so the ring loop issues no memory instruction but the segment-table
// read and the LDS add.  Each crystal pair adds s_A*tau*T*s_B to plane segplane(code, plane)
// of the bin's column, with the reference's float operations (compute_delays
// gen_delays_fromarchive.cpp:205-222):
//   plane = (int)(cay + z*dz),  seg = 0.5 + dz*d,  code = seg < 0 ? 1 - 2(int)seg : 2(int)seg
// The column accumulates in 64-bit fixed point (ds_add_u64; LDS float atomics run about
// four times slower on gfx950), in units of 2^-k with k from the largest possible term:
// the sum is order-independent, hence reproducible, and each element differs from the
// exact sum of its float terms by at most n*2^-(k+1) for n terms before the final rounding
// to float.
constexpr int COL_THREADS = 512;
// Columns up to this many planes (span 9: 2209) are stored straight into the plane-major
// output; longer ones (span 3: 6367) go through bin-major columns and a tiled transpose,
// since 96 workgroups' worth of them per XCD no longer gather in its L2.
constexpr int DIRECT_PLANES = 4096;
// The bin-major column path's LDS column (6400 planes: span 3 at maxrd up to 68 has 6367).
constexpr int COLUMN_PLANES = 6400;

__device__ inline float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <bool DIRECT>
__global__ __launch_bounds__(COL_THREADS) void delays_column_kernel(Tables t, const float *__restrict__ sings,
                                                                    float tauftime, float *__restrict__ cols,
                                                                    int cstride) {
    // segment entry by e = 2*(int)seg + (seg < 0) + 64: {first plane | last plane << 16, 8 * file offset}
    __shared__ int2 seg[128];
    // the bin's column of planes (a static array: its adds take their LDS base as an immediate)
    __shared__ unsigned long long col[DIRECT ? DIRECT_PLANES : COLUMN_PLANES];
    // XCD-aware: workgroups are dealt to the 8 XCDs in turn, so XCD x gets bins
    // [x*nsino/8, (x+1)*nsino/8) in order and its L2 gathers whole lines of the plane-major output
    const int per = gridDim.x >> 3;
    const int bin = (blockIdx.x & 7) * per + (blockIdx.x >> 3), tid = threadIdx.x;
    for (int i = tid; i < t.nplanes; i += COL_THREADS) col[i] = 0ull;
    if (tid < 128) {
        const int2 e = t.seg[tid];
        seg[tid] = make_int2(e.x, 8 * e.y);
    }
    const int k0 = t.lor_off[bin], k1 = t.lor_off[bin + 1];
    float mlor = 0.0f;
    for (int k = k0; k < k1; k++) {
        const int4 L = t.lor[k];
        mlor = fmaxf(mlor, sings[NCRYS + L.x / NY] * sings[NCRYS + L.y / NY]);
    }
    const int kexp = fixed_point_exponent(mlor * tauftime, max(k1 - k0, 1));
    const float up = ldexpf(1.0f, kexp);
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int row = wave >> 1;
    const int by = (wave & 1) * 64 + lane;
    const bool live = by < NY;
    const int w = t.window;
    const int ay_lo = by - w, ay_hi = by + w;   // A rings inside the ring window of B ring by
    const float2 z0 = t.z[lane];                                          // A rings 0..63
    const float2 z1 = lane < NY - 64 ? t.z[64 + lane] : make_float2(0.0f, 0.0f);   // A rings 64..103
    const float zb = live ? t.z[by].x : 0.0f;
    __syncthreads();
    if (k0 == k1) goto store;
    {
        int4 L = t.lor[k0];
        float sb = live ? sings[L.y + by] : 0.0f;
        float sa0 = sings[L.x + lane];
        float sa1 = lane < NY - 64 ? sings[L.x + 64 + lane] : 0.0f;
        for (int k = k0; k < k1; k++) {
            const float sd = __int_as_float(L.z), sz = __int_as_float(L.w);
            // the term s_B*(s_A*tau*T) scaled by 2^k: scaling s_B first is exact and rounds the
            // same product (a power of two commutes with rounding)
            const float bu = sb * up, a0 = sa0 * tauftime, a1 = sa1 * tauftime;
            if (k + 1 < k1) {     // next LOR's loads overlap this LOR's rings
                L = t.lor[k + 1];
                sb = live ? sings[L.y + by] : 0.0f;
                sa0 = sings[L.x + lane];
                sa1 = lane < NY - 64 ? sings[L.x + 64 + lane] : 0.0f;
            }
            // one crystal pair per lane for A ring ay (A singles*tau*T awt, ring z za, plane
            // cay): the predicate, column index and fixed-point value of its add.  The
            // reference skips silent A crystals; their terms are exact zeros here.
            auto prepare = [&](int ay, float awt, float za, float cay, bool &ok, int &idx8, unsigned long long &q) {
                const float dz = zb - za;
                const int plane = (int)(cay + sz * dz);
                const float s = 0.5f + dz * sd;
                const int si = (int)s;
                // e - 64 = 2*(int)seg + (seg < 0) names the code (2*si for seg >= 0, 1 - 2*si
                // for seg < 0); e outside [0, 128) and impossible pairs map to no plane.  The
                // byte offset 8 (e - 64), clamped to [-512, 504], reads seg[e].
                const int e8 = min(max((si * 16) | (s < 0.0f ? 8 : 0), -512), 504);
                const long long ent = *reinterpret_cast<const long long *>(reinterpret_cast<const char *>(seg) + 512 + e8);
                const int first = (int)(ent & 0xffff), last = (int)((ent >> 16) & 0xffff);
                ok = live & (ay >= ay_lo) & (ay <= ay_hi) & (plane >= first) & (plane <= last);
                idx8 = (plane << 3) + (int)(ent >> 32);
                q = to_fixed(bu * awt);
            };
            auto add = [&](int idx8, unsigned long long q) {
                atomicAdd(reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(col) + idx8), q);
            };
            // A rings row, row+4, ..., two per step: 16 below ring 64 and 10 from it (row < 4)
#pragma unroll 2
            for (int i = 0; i < 16; i += 2) {
                const int ay = row + 4 * i;
                bool okA, okB;
                int iA, iB;
                unsigned long long qA, qB;
                prepare(ay, lane_value(a0, ay), lane_value(z0.x, ay), lane_value(z0.y, ay), okA, iA, qA);
                prepare(ay + 4, lane_value(a0, ay + 4), lane_value(z0.x, ay + 4), lane_value(z0.y, ay + 4), okB, iB, qB);
                if (okA) add(iA, qA);
                if (okB) add(iB, qB);
            }
#pragma unroll 1
            for (int i = 0; i < 10; i += 2) {
                const int ay = 64 + row + 4 * i, r = row + 4 * i;
                bool okA, okB;
                int iA, iB;
                unsigned long long qA, qB;
                prepare(ay, lane_value(a1, r), lane_value(z1.x, r), lane_value(z1.y, r), okA, iA, qA);
                prepare(ay + 4, lane_value(a1, r + 4), lane_value(z1.x, r + 4), lane_value(z1.y, r + 4), okB, iB, qB);
                if (okA) add(iA, qA);
                if (okB) add(iB, qB);
            }
        }
    }
store:
    __syncthreads();
    const double down = ldexp(1.0, -kexp);
    if (DIRECT) {   // plane-major output [nplanes][nsino]
        for (int i = tid; i < t.nplanes; i += COL_THREADS) cols[(size_t)i * t.nsino + bin] = (float)((double)(long long)col[i] * down);
    } else {        // bin-major columns [nsino][cstride], cstride a multiple of 64 planes
        float *dst = cols + (size_t)bin * cstride;
        for (int i = tid; i < t.nplanes; i += COL_THREADS) dst[i] = (float)((double)(long long)col[i] * down);
    }
}

// [nsino][cstride] columns -> [nplanes][nsino]; 64 x 64 tiles through LDS, 256 threads (16
// elements each), TT tiles along the planes per workgroup with the next tile's loads in flight
// while the current one is stored.  A column stride of whole 256-byte lines keeps every tile
// row's read on line boundaries.
constexpr int TT = 4;

__global__ __launch_bounds__(256) void column_transpose_kernel(const float *__restrict__ cols, float *__restrict__ out,
                                                               int nsino, int nplanes, int cstride) {
    __shared__ float tile[2][64][65];
    const int b0 = blockIdx.y * 64, pt0 = blockIdx.x * TT * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    float v[16];
    auto load = [&](int p0) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int b = b0 + ty + 4 * k, p = p0 + tx;
            v[k] = (b < nsino && p < nplanes) ? cols[(size_t)b * cstride + p] : 0.0f;
        }
    };
    load(pt0);
    for (int i = 0; i < TT; i++) {
        const int p0 = pt0 + 64 * i;
        if (p0 >= nplanes) break;
        float(*T)[65] = tile[i & 1];
#pragma unroll
        for (int k = 0; k < 16; k++) T[ty + 4 * k][tx] = v[k];
        __syncthreads();
        if (i + 1 < TT && p0 + 64 < nplanes) load(p0 + 64);
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int p = p0 + ty + 4 * k, b = b0 + tx;
            if (p < nplanes && b < nsino) out[(size_t)p * nsino + b] = T[tx][ty + 4 * k];
        }
    }
}

}  // namespace

namespace {

int make_tables(hrrt_delays &c, const hrrt_delays_params &p, std::string &err) {
    if (p.span != 3 && p.span != 9) return err = "span must be 3 or 9", HRRT_ERR_ARG;
    if (p.maxrd < 0 || p.maxrd > NY - 1) return err = "maxrd out of range", HRRT_ERR_ARG;
    if (p.lr_type < 0 || p.lr_type > 2) return err = "lr_type must be 0, 1 or 2", HRRT_ERR_ARG;
    c.span = p.span;
    c.maxrd = p.maxrd;
    c.lr_type = p.lr_type;
    c.tau = p.tau > 0.0f ? p.tau : 6.0e-9f;
    // init_geometry_hrrt geometry_info.cpp:169-221
    const float pitch = p.pitch > 0.0f ? p.pitch : 0.24375f;
    const float diam = p.diam > 0.0f ? p.diam : 46.9f;
    c.depth = p.thick > 0.0f ? p.thick : 1.0f;
    for (int i = 0; i < NHEADS; i++) {
        c.sin_head[i] = std::sin(i * 2.0 * M_PI / NHEADS);
        c.cos_head[i] = std::cos(i * 2.0 * M_PI / NHEADS);
    }
    c.crystal_radius = diam / 2.0;
    if (c.lr_type == 0) {
        c.nprojs = 256, c.nviews = 288, c.binsize = pitch / 2.0, c.plane_sep = pitch / 2.0;
    } else if (c.lr_type == 1) {
        c.nprojs = 160, c.nviews = 144, c.binsize = 0.2f, c.plane_sep = pitch;
    } else {
        c.nprojs = 128, c.nviews = 144, c.binsize = pitch, c.plane_sep = pitch;
    }
    c.nsino = c.nprojs * c.nviews;
    // init_segment_info segment_info_fromarchive.cpp:118-141: planes of the span; the plane
    // map is span 3, with span-9 file offsets through the conversion table
    int z0[SEGMAP], zmax[SEGMAP], off[SEGMAP];
    c.nsegs = 2 * (c.maxrd / 3) + 1;
    if (c.nsegs > 45) return err = "maxrd above 68 leaves the segment tables", HRRT_ERR_ARG;
    c.nplanes = c.segments(c.span, z0, zmax, off);
    c.segments(3, c.segz0, c.segzmax, c.segoff);
    if (c.span == 9) {
        const int n9 = 2 * (c.maxrd / 9) + 1;
        for (int i = 0; i < c.nsegs; i++) {
            if (kS3toS9[i] >= n9) return err = "span-9 segment table too short for maxrd", HRRT_ERR_ARG;
            c.segoff[i] = off[kS3toS9[i]];
        }
    }
    c.d_tan_theta = 3 * c.plane_sep / c.crystal_radius;
    return HRRT_OK;
}

// init_sol lor_sinogram_map.cpp:141-209: the LUT and the ring z tables from the geometry
void compute_lut(hrrt_delays &c) {
    c.lut.assign((size_t)(NMP + 1) * NXD * NXD, Sol{-1, 0.0f, 0.0f});
    for (int mp = 1; mp <= NMP; mp++) {
        for (int al = 0; al < NDOIS; al++)
            for (int ax = 0; ax < NX; ax++) {
                float pa[3];
                c.crystal(kMpairs[mp][0], al, ax, 0, pa);
                Sol *row = &c.lut[((size_t)mp * NXD + ax + NX * al) * NXD];
                for (int bl = 0; bl < NDOIS; bl++)
                    for (int bx = 0; bx < NX; bx++) {
                        float pb[3];
                        c.crystal(kMpairs[mp][1], bl, bx, 0, pb);
                        row[bx + NX * bl] = c.lor(pa, pb);
                    }
            }
    }
    for (int i = 0; i < NY; i++) {
        float pos[3];
        c.crystal(0, 0, 0, i, pos);
        c.zpos[i] = pos[2];
        c.zpos2[i] = (float)((double)pos[2] / c.plane_sep + 0.5);
    }
}

// init_lut_sol lor_sinogram_map.cpp:224-303
int read_lut(hrrt_delays &c, const char *path, std::string &err) {
    FILE *f = fopen(path, "rb");
    if (!f) return err = std::string("cannot open LUT file ") + path, HRRT_ERR_ARG;
    c.lut.assign((size_t)(NMP + 1) * NXD * NXD, Sol{-1, 0.0f, 0.0f});
    const size_t n = (size_t)NMP * NXD * NXD;
    bool ok = fread(&c.lut[(size_t)NXD * NXD], sizeof(Sol), n, f) == n;
    ok = ok && fread(c.zpos, sizeof(float), NY, f) == (size_t)NY && fread(c.zpos2, sizeof(float), NY, f) == (size_t)NY;
    fclose(f);
    if (!ok) return err = std::string("short LUT file ") + path, HRRT_ERR_ARG;
    return HRRT_OK;
}



// Group the valid LORs by sinogram bin (host tables; uploaded by the first smoothing).
int group_lors(hrrt_delays &c) {
    std::vector<int> &count = c.h_lor_off;
    count.assign(c.nsino + 1, 0);
    for (int mp = 1; mp <= NMP; mp++)
        for (int i = 0; i < NXD * NXD; i++) {
            const Sol &s = c.lut[(size_t)mp * NXD * NXD + i];
            if (s.nsino < -1 || s.nsino >= c.nsino)
                return hrrt_dl::fail(HRRT_ERR_ARG, "LUT bin outside the sinogram (LUT built for another LR type?)");
            if (s.nsino >= 0) count[s.nsino + 1]++;
        }
    for (int b = 0; b < c.nsino; b++) count[b + 1] += count[b];
    c.nlor = count[c.nsino];
    std::vector<int4> &lor = c.h_lor;
    lor.assign(std::max(c.nlor, 1), make_int4(0, 0, 0, 0));
    std::vector<int> fill(count.begin(), count.end() - 1);
    // per A ring, the B rings inside |ay - by| < maxrd + 6 (gen_delays_fromarchive.cpp:175-184)
    const int window = c.maxrd + 5;
    long long ring_pairs = 0;
    for (int ay = 0; ay < NY; ay++) ring_pairs += std::min(NY - 1, ay + window) - std::max(0, ay - window) + 1;
    c.pairs_full = 0;
    for (int mp = 1; mp <= NMP; mp++) {
        const int ahead = kMpairs[mp][0], bhead = kMpairs[mp][1];
        for (int axx = 0; axx < NXD; axx++)
            for (int bxx = 0; bxx < NXD; bxx++) {
                const Sol &s = c.lut[((size_t)mp * NXD + axx) * NXD + bxx];
                if (s.nsino < 0) continue;
                const int al = axx / NX, ax = axx % NX, bl = bxx / NX, bx = bxx % NX;
                int4 e;
                e.x = ((al * NHEADS + ahead) * NX + ax) * NY;
                e.y = ((bl * NHEADS + bhead) * NX + bx) * NY;
                memcpy(&e.z, &s.d, 4);
                memcpy(&e.w, &s.z, 4);
                lor[fill[s.nsino]++] = e;
                c.pairs_full += ring_pairs;
            }
    }
    if (c.nplanes > COLUMN_PLANES) return hrrt_dl::fail(HRRT_ERR_UNSUPPORTED, "more sinogram planes than the LDS column holds");
    if (c.nsino % 8) return hrrt_dl::fail(HRRT_ERR_UNSUPPORTED, "sinogram bins not a multiple of 8");
    return HRRT_OK;
}

// Head totals of the singles (compute_drate gen_delays_lib.cpp:168-176): layer, x, ring order.
void head_sums(const float *srate, float *hsum) {
    for (int head = 0; head < NHEADS; head++) {
        float s = 0.0f;
        for (int layer = 0; layer < NDOIS; layer++)
            for (int cx = 0; cx < NX; cx++)
                for (int cy = 0; cy < NY; cy++) s += srate[NX * head + cx + HX * cy + XYH * layer];
        hsum[head] = s;
    }
}

// Sum over the five heads in coincidence with `head` (gen_delays_lib.cpp:180-186).
float fan_sum(const float *hsum, int head) {
    float s = 0.0f;
    for (int j = 0; j < 5; j++) s += hsum[(head + j + 2) % NHEADS];
    return s;
}

}  // namespace

int hrrt_dl::fail(int code, const std::string &msg) { return hrrt::set_last_error(code, msg.c_str()); }

// Device copies of the tables, on the current device, at the first smoothing.
int hrrt_dl::upload(hrrt_delays &c) {
    if (c.d_lor) return HRRT_OK;
    const std::vector<int4> &lor = c.h_lor;
    const std::vector<int> &count = c.h_lor_off;
    // segment entries by e = 2*(int)seg + (seg < 0) + 64 (the kernel's index): e - 64 even and
    // >= 0 is code e - 64, odd and <= 1 is code 2 - (e - 64); {first plane, last plane, offset}
    std::vector<int2> seg(128, make_int2(0x7fff, 0));            // first 32767 > last 0: no plane
    for (int e = 0; e < 128; e++) {
        const int ee = e - 64;
        const int code = (ee >= 0 && ee % 2 == 0) ? ee : ((ee <= 1 && (ee & 1)) ? 2 - ee : -1);
        if (code >= 0 && code < c.nsegs) seg[e] = make_int2(c.segz0[code] | (c.segzmax[code] << 16), c.segoff[code]);
    }
    std::vector<float2> z(NY);
    for (int i = 0; i < NY; i++) z[i] = make_float2(c.zpos[i], c.zpos2[i]);
    DTRY(hipGetDevice(&c.device));
    DTRY(hipMalloc(&c.d_lor, sizeof(int4) * lor.size()));
    DTRY(hipMalloc(&c.d_lor_off, sizeof(int) * count.size()));
    DTRY(hipMalloc(&c.d_seg, sizeof(int2) * 128));
    DTRY(hipMalloc(&c.d_z, sizeof(float2) * NY));
    DTRY(hipMalloc(&c.d_sings, sizeof(float) * (NCRYS + NCOLS)));
    DTRY(hipMemcpy(c.d_lor, lor.data(), sizeof(int4) * lor.size(), hipMemcpyHostToDevice));
    DTRY(hipMemcpy(c.d_lor_off, count.data(), sizeof(int) * count.size(), hipMemcpyHostToDevice));
    DTRY(hipMemcpy(c.d_seg, seg.data(), sizeof(int2) * 128, hipMemcpyHostToDevice));
    DTRY(hipMemcpy(c.d_z, z.data(), sizeof(float2) * NY, hipMemcpyHostToDevice));
    return HRRT_OK;
}

// ---------------------------------------------------------------- C ABI

extern "C" int hrrt_delays_create(const hrrt_delays_params *params, hrrt_delays **out) {
    if (!params || !out) return hrrt_dl::fail(HRRT_ERR_ARG, "null argument");
    *out = nullptr;
    hrrt_delays *c = new (std::nothrow) hrrt_delays();
    if (!c) return hrrt_dl::fail(HRRT_ERR_NOMEM, "host allocation failed");
    std::string err;
    int r = make_tables(*c, *params, err);
    if (r == HRRT_OK) {
        if (params->lut_file) r = read_lut(*c, params->lut_file, err);
        else compute_lut(*c);
    }
    if (r != HRRT_OK) {
        delete c;
        return hrrt_dl::fail(r, err);
    }
    if ((r = group_lors(*c)) != HRRT_OK) {
        delete c;
        return r;
    }
    *out = c;
    return HRRT_OK;
}

extern "C" void hrrt_delays_destroy(hrrt_delays *c) {
    if (!c) return;
    for (void *p : {(void *)c->d_lor, (void *)c->d_lor_off, (void *)c->d_seg, (void *)c->d_z, (void *)c->d_sings,
                    (void *)c->d_cols, (void *)c->d_sol, (void *)c->d_lm_keys, (void *)c->d_lm_sort_tmp, (void *)c->d_lm_tiles,
                    (void *)c->d_lm_coin_seg, (void *)c->d_lm_coin_cnt, (void *)c->d_lm_coin_part, (void *)c->d_lm_chunks,
                    (void *)c->d_lm_stats})
        if (p) (void)hipFree(p);
    delete c;
}

extern "C" int hrrt_delays_get_dims(const hrrt_delays *c, int *nprojs, int *nviews, int *nplanes, int *nsegs,
                                    long long *pairs_full) {
    if (!c) return hrrt_dl::fail(HRRT_ERR_ARG, "null context");
    if (nprojs) *nprojs = c->nprojs;
    if (nviews) *nviews = c->nviews;
    if (nplanes) *nplanes = c->nplanes;
    if (nsegs) *nsegs = c->nsegs;
    if (pairs_full) *pairs_full = c->pairs_full;
    return HRRT_OK;
}

extern "C" int hrrt_delays_get_lut(const hrrt_delays *c, int mp, int32_t *nsino, float *d, float *z) {
    if (!c) return hrrt_dl::fail(HRRT_ERR_ARG, "null context");
    if (mp < 1 || mp > NMP) return hrrt_dl::fail(HRRT_ERR_ARG, "module pair must be 1..20");
    const Sol *s = &c->lut[(size_t)mp * NXD * NXD];
    for (int i = 0; i < NXD * NXD; i++) {
        if (nsino) nsino[i] = s[i].nsino;
        if (d) d[i] = s[i].d;
        if (z) z[i] = s[i].z;
    }
    return HRRT_OK;
}

// save_lut_sol lor_sinogram_map.cpp:305-340
extern "C" int hrrt_delays_save_lut(const hrrt_delays *c, const char *path) {
    if (!c || !path) return hrrt_dl::fail(HRRT_ERR_ARG, "null argument");
    FILE *f = fopen(path, "wb");
    if (!f) return hrrt_dl::fail(HRRT_ERR_ARG, std::string("cannot create ") + path);
    const size_t n = (size_t)NMP * NXD * NXD;
    bool ok = fwrite(&c->lut[(size_t)NXD * NXD], sizeof(Sol), n, f) == n;
    ok = ok && fwrite(c->zpos, sizeof(float), NY, f) == (size_t)NY && fwrite(c->zpos2, sizeof(float), NY, f) == (size_t)NY;
    ok = (fclose(f) == 0) && ok;
    return ok ? HRRT_OK : hrrt_dl::fail(HRRT_ERR_ARG, std::string("write failed: ") + path);
}

// compute_csings_from_drates gen_delays_lib.cpp:254-271 with compute_initial_srate (:233-241),
// compute_drate (:162-196), errtotal (:244-252) and estimate_srate (:198-230).  The error
// sum is 64-bit (the reference's int overflows on large errors; only its zero test matters).
extern "C" int hrrt_delays_singles(const hrrt_delays *c, const int32_t *coins, float scan_duration, float *singles,
                                   int *niter) {
    if (!c || !coins || !singles) return hrrt_dl::fail(HRRT_ERR_ARG, "null argument");
    const float dt = scan_duration > 0.0f ? scan_duration : 1.0f, tau = c->tau;
    const int32_t *dcoins = coins + NCRYS;
    std::vector<float> drates(NCRYS), xrates(NCRYS);
    float dtotal = 0.0f;
    for (int i = 0; i < NCRYS; i++) drates[i] = (float)dcoins[i] / dt;
    for (int i = 0; i < NCRYS; i++) dtotal += drates[i];
    const float stotal = (float)(8.0 * std::sqrt((double)dtotal / (40. * (double)tau)));
    for (int i = 0; i < NCRYS; i++) singles[i] = drates[i] * stotal / dtotal;
    float hsum[NHEADS];
    int iter;
    for (iter = 0; iter < 100; iter++) {
        head_sums(singles, hsum);
        long long err = 0;
        for (int head = 0; head < NHEADS; head++) {
            const float fan = fan_sum(hsum, head);
            for (int layer = 0; layer < NDOIS; layer++)
                for (int cx = 0; cx < NX; cx++)
                    for (int cy = 0; cy < NY; cy++) {
                        const int i = NX * head + cx + HX * cy + XYH * layer;
                        xrates[i] = singles[i] * tau * fan;
                    }
        }
        for (int i = 0; i < NCRYS; i++) {
            const long long e = (int)((double)dcoins[i] - (0.5 + (double)(xrates[i] * dt)));
            err += e * e;
        }
        if (iter && err == 0) break;
        head_sums(singles, hsum);
        for (int head = 0; head < NHEADS; head++) {
            const float fan = fan_sum(hsum, head);
            for (int layer = 0; layer < NDOIS; layer++)
                for (int cx = 0; cx < NX; cx++)
                    for (int cy = 0; cy < NY; cy++) {
                        const int i = NX * head + cx + HX * cy + XYH * layer;
                        singles[i] = (singles[i] + drates[i] / (tau * fan)) / 2.0f;
                    }
        }
    }
    if (niter) *niter = iter;
    return HRRT_OK;
}

extern "C" int hrrt_delays_smooth(hrrt_delays *c, const float *d_singles, float scan_duration, float *d_delays,
                                  void *stream) {
    if (!c || !d_singles || !d_delays) return hrrt_dl::fail(HRRT_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    int r = upload(*c), dev = -1;
    if (r != HRRT_OK) return r;
    DTRY(hipGetDevice(&dev));
    if (dev != c->device) return hrrt_dl::fail(HRRT_ERR_ARG, "context tables live on another device");
    const float ftime = scan_duration > 0.0f ? scan_duration : 1.0f;
    const float tauftime = c->tau * ftime;
    Tables t{c->d_lor, c->d_lor_off, c->d_seg, c->d_z, c->nsino, c->nplanes, c->maxrd + 5};
    sings_prepare_kernel<<<1, PREP_THREADS, 0, s>>>(d_singles, c->d_sings);
    DTRY(hipGetLastError());
    if (c->nplanes <= DIRECT_PLANES) {
        delays_column_kernel<true><<<c->nsino, COL_THREADS, 0, s>>>(t, c->d_sings, tauftime, d_delays, 0);
        DTRY(hipGetLastError());
        return HRRT_OK;
    }
    const int cstride = (c->nplanes + 63) & ~63;
    if (!c->d_cols) DTRY(hipMalloc(&c->d_cols, sizeof(float) * (size_t)c->nsino * cstride));
    delays_column_kernel<false><<<c->nsino, COL_THREADS, 0, s>>>(t, c->d_sings, tauftime, c->d_cols, cstride);
    DTRY(hipGetLastError());
    column_transpose_kernel<<<dim3((c->nplanes + 64 * TT - 1) / (64 * TT), (c->nsino + 63) / 64), 256, 0, s>>>(c->d_cols, d_delays,
                                                                                               c->nsino, c->nplanes, cstride);
    DTRY(hipGetLastError());
    return HRRT_OK;
}

extern "C" int hrrt_gen_delays(hrrt_delays *c, const int32_t *coins, float scan_duration, float *d_delays,
                               float *singles_out, int *niter, void *stream) {
    if (!c || !coins || !d_delays) return hrrt_dl::fail(HRRT_ERR_ARG, "null argument");
    std::vector<float> sg(NCRYS);
    int r = hrrt_delays_singles(c, coins, scan_duration, sg.data(), niter);
    if (r != HRRT_OK) return r;
    if (singles_out) memcpy(singles_out, sg.data(), sizeof(float) * NCRYS);
    // the singles are staged in the head of the output (nplanes*nsino >= crystals), which the
    // smoothing reads into its transposed copy before its last kernel overwrites the output
    hipStream_t s = (hipStream_t)stream;
    DTRY(hipMemcpyAsync(d_delays, sg.data(), sizeof(float) * NCRYS, hipMemcpyHostToDevice, s));
    DTRY(hipStreamSynchronize(s));
    return hrrt_delays_smooth(c, d_delays, scan_duration, d_delays, stream);
}

extern "C" int hrrt_delays_write(hrrt_delays *c, const float *d_delays, const char *path, void *stream) {
    if (!c || !d_delays || !path) return hrrt_dl::fail(HRRT_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const size_t plane = (size_t)c->nsino;
    std::vector<float> h(plane * 64);
    FILE *f = fopen(path, "wb");
    if (!f) return hrrt_dl::fail(HRRT_ERR_ARG, std::string("cannot create ") + path);
    bool ok = true;
    for (int p0 = 0; ok && p0 < c->nplanes; p0 += 64) {
        const size_t n = plane * std::min(64, c->nplanes - p0);
        if (hipMemcpyAsync(h.data(), d_delays + (size_t)p0 * plane, sizeof(float) * n, hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            fclose(f);
            return hrrt_dl::fail(HRRT_ERR_HIP, "copy of the delays to the host failed");
        }
        ok = fwrite(h.data(), sizeof(float), n, f) == n;
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? HRRT_OK : hrrt_dl::fail(HRRT_ERR_ARG, std::string("write failed: ") + path);
}
