// kernels.hip -- CDNA4 (gfx950) kernels of the OSEM hot path.
//
// Every kernel issues its float operations in the order the reference does
// (no fused multiply-add: this file is compiled with -ffp-contract=off and
// the pragma below), so results are bit-identical to the reference
// arithmetic for the same inputs.  Citations: PJ = hrrt_osem3d/
// hrrt_osem3dv_sbrt_span3.cpp, S3 = hrrt_osem3d/hrrt_osem3d_span3.cpp,
// FIO = hrrt_osem3d/file_io_processor.cpp.
#include "kernels.h"

#include <type_traits>

#pragma clang fp contract(off)

namespace hrrt {

namespace {

__device__ __forceinline__ float4 f4(float a) { return make_float4(a, a, a, a); }
__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 mul4(float4 a, float s) {
    return make_float4(a.x * s, a.y * s, a.z * s, a.w * s);
}
__device__ __forceinline__ float half_at(const uint16_t *p, size_t i) { return __half2float(__ushort_as_half(p[i])); }

// Compensated (Kahan) summation: s accumulates, c carries the low-order bits lost so far.
__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ float lane(const float4 &v, int l) {
    return l == 0 ? v.x : (l == 1 ? v.y : (l == 2 ? v.z : v.w));
}

// Bilinear sample of the z-column chunk q (4 planes) at the rotated position of
// (xx, yy) (PJ:1865-1892).  Returns false when the reference rejects the sample.
__device__ __forceinline__ bool sample_chunk(const DevGeom &g, const float4 *__restrict__ image, float cosv,
                                             float sinv, int xx, int yy, int q, float4 &out) {
    const int c = g.c, N = g.N;
    float xx0 = (xx - c) * cosv - (yy - c) * sinv + c;
    float yy0 = (xx - c) * sinv + (yy - c) * cosv + c;
    if (xx0 < 0 || xx0 > N - 2 || yy0 < 0 || xx0 > N - 2) return false;
    int xi = (int)xx0, yi = (int)yy0;
    float x1 = xx0 - xi, y1 = yy0 - yi;
    float w1 = (1 - x1) * (1 - y1), w2 = x1 * (1 - y1), w3 = y1 * (1 - x1), w4 = y1 * x1;
    const size_t NQ = g.NQ;
    float4 r1 = image[((size_t)xi * N + yi) * NQ + q];
    float4 r2 = image[((size_t)(xi + 1) * N + yi) * NQ + q];
    float4 r3 = image[((size_t)xi * N + yi + 1) * NQ + q];
    float4 r4 = image[((size_t)(xi + 1) * N + yi + 1) * NQ + q];
    out = add4(add4(mul4(r1, w1), mul4(r2, w2)), add4(mul4(r3, w3), mul4(r4, w4)));
    return true;
}

__device__ __forceinline__ void put_lanes(float4 *buf, int q, int lanei, float4 v, int Z) {
    float *f = reinterpret_cast<float *>(buf);
    int z = 4 * q;
    if (z + 0 < Z) f[(z + 0) * 4 + lanei] = v.x;
    if (z + 1 < Z) f[(z + 1) * 4 + lanei] = v.y;
    if (z + 2 < Z) f[(z + 2) * 4 + lanei] = v.z;
    if (z + 3 < Z) f[(z + 3) * 4 + lanei] = v.w;
}

// The eight symmetric sample positions of (x, y) (PJ:1856-1864).  Samples 0..3
// feed "pam" (near half of the ray), 4..7 feed "pam2" (mirrored half).
// i: (x, y), (N-x, y), (N-y, x), (N-y, N-x), (x, N-y), (N-x, N-y), (y, x), (y, N-x).  Bit 0
// mirrors x, bit 1 swaps the two coordinates, bit 2 mirrors y; as selects rather than a branch
// tree, since the sample index differs between the 16-lane rows of a wave.
__device__ __forceinline__ void sym8(int i, int x, int y, int N, int &xx, int &yy) {
    const int X = (i & 1) ? N - x : x;
    const int Yd = (i & 4) ? N - y : y;   // y coordinate of the unswapped samples
    const int Ys = (i & 4) ? y : N - y;   // x coordinate of the swapped samples
    xx = (i & 2) ? Ys : X;
    yy = (i & 2) ? X : Yd;
}

__device__ __forceinline__ void sample8(const DevGeom &g, const float4 *__restrict__ image, float cosv, float sinv,
                                        int x, int y, float4 *pam, float4 *pam2) {
    const int NQ = g.NQ, N = g.N;
    for (int it = threadIdx.x; it < 8 * NQ; it += blockDim.x) {
        int i = it / NQ, q = it - i * NQ;
        int xx, yy;
        sym8(i, x, y, N, xx, yy);
        float4 v;
        if (!sample_chunk(g, image, cosv, sinv, xx, yy, q, v)) v = f4(0.f);
        put_lanes(i < 4 ? pam : pam2, q, i & 3, v, g.Z);
    }
}

// Two-phase version of sample8 for software pipelining: issue the global loads
// of the next y into registers, project the current y, then combine and store.
// Each thread owns ITEMS of the 8*NQ (sample, 4-plane chunk) items.
template <int ITEMS>
struct SampleRegs {
    float4 r1[ITEMS], r2[ITEMS], r3[ITEMS], r4[ITEMS];
    float w1[ITEMS], w2[ITEMS], w3[ITEMS], w4[ITEMS];
    bool ok[ITEMS];
};

template <int ITEMS>
__device__ __forceinline__ void sample8_load(const DevGeom &g, const float4 *__restrict__ image, float cosv,
                                             float sinv, int x, int y, SampleRegs<ITEMS> &s) {
    const int NQ = g.NQ, N = g.N, c = g.c;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int it = threadIdx.x + j * blockDim.x;
        s.ok[j] = false;
        if (it >= 8 * NQ) continue;
        const int i = it / NQ, q = it - i * NQ;
        int xx, yy;
        sym8(i, x, y, N, xx, yy);
        const float xx0 = (xx - c) * cosv - (yy - c) * sinv + c;
        const float yy0 = (xx - c) * sinv + (yy - c) * cosv + c;
        if (xx0 < 0 || xx0 > N - 2 || yy0 < 0 || xx0 > N - 2) continue;
        const int xi = (int)xx0, yi = (int)yy0;
        const float x1 = xx0 - xi, y1 = yy0 - yi;
        s.w1[j] = (1 - x1) * (1 - y1); s.w2[j] = x1 * (1 - y1); s.w3[j] = y1 * (1 - x1); s.w4[j] = y1 * x1;
        const float4 *p = image + ((size_t)xi * N + yi) * NQ + q;
        s.r1[j] = p[0];
        s.r2[j] = p[(size_t)N * NQ];
        s.r3[j] = p[NQ];
        s.r4[j] = p[(size_t)N * NQ + NQ];
        s.ok[j] = true;
    }
}

template <int ITEMS>
__device__ __forceinline__ void sample8_store(const DevGeom &g, const SampleRegs<ITEMS> &s, float4 *pam, float4 *pam2) {
    const int NQ = g.NQ;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int it = threadIdx.x + j * blockDim.x;
        if (it >= 8 * NQ) continue;
        const int i = it / NQ, q = it - i * NQ;
        const float4 v = s.ok[j] ? add4(add4(mul4(s.r1[j], s.w1[j]), mul4(s.r2[j], s.w2[j])),
                                        add4(mul4(s.r3[j], s.w3[j]), mul4(s.r4[j], s.w4[j])))
                                 : f4(0.f);
        put_lanes(i < 4 ? pam : pam2, q, i & 3, v, g.Z);
    }
}

// Transposed staging of a row (the FAST forward path).  A wave owns 16 chunks (4 planes
// each) of one half of the row's samples: lane l samples chunk 16 b + (l & 15) of sample
// 4 h + (l >> 4), so each row of 16 lanes loads 256 contiguous bytes of one image column.
// A 4x4 transpose across the four 16-lane rows (two permlane32 swaps, then two permlane16
// swaps) leaves lane l with plane 4 q + (l >> 4) of all four samples, stored with one
// 16-byte LDS write instead of four scattered 4-byte ones.
__device__ __forceinline__ int xpose_chunks(const DevGeom &g) { return 2 * ((g.NQ + 15) >> 4); }

__device__ __forceinline__ void xpose_item(const DevGeom &g, int it, int &h, int &q, int &i) {
    const int nq16 = (g.NQ + 15) >> 4, chunk = it >> 6, l = it & 63;
    h = chunk >= nq16;
    q = 16 * (chunk - h * nq16) + (l & 15);
    i = 4 * h + (l >> 4);
}

__device__ __forceinline__ void swap_pl32(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

__device__ __forceinline__ void swap_pl16(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

// v: planes 4q..4q+3 of sample (l >> 4) -> plane 4q + (l >> 4) of samples 0..3.
__device__ __forceinline__ float4 xpose4(float4 v) {
    swap_pl32(v.x, v.z);   // rows 0-1 receive plane 0 of samples 2-3, rows 2-3 plane 2 of samples 0-1
    swap_pl32(v.y, v.w);
    swap_pl16(v.x, v.y);   // odd rows' x <-> even rows' y
    swap_pl16(v.z, v.w);
    return v;
}

template <int ITEMS>
__device__ __forceinline__ void sample8x_load(const DevGeom &g, const float4 *__restrict__ image, float cosv,
                                              float sinv, int x, int y, SampleRegs<ITEMS> &s) {
    const int NQ = g.NQ, N = g.N, c = g.c;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int it = threadIdx.x + j * blockDim.x;
        s.ok[j] = false;
        if ((it >> 6) >= xpose_chunks(g)) continue;
        int h, q, i, xx, yy;
        xpose_item(g, it, h, q, i);
        if (q >= NQ) continue;
        sym8(i, x, y, N, xx, yy);
        const float xx0 = (xx - c) * cosv - (yy - c) * sinv + c;
        const float yy0 = (xx - c) * sinv + (yy - c) * cosv + c;
        if (xx0 < 0 || xx0 > N - 2 || yy0 < 0 || xx0 > N - 2) continue;
        const int xi = (int)xx0, yi = (int)yy0;
        const float x1 = xx0 - xi, y1 = yy0 - yi;
        s.w1[j] = (1 - x1) * (1 - y1); s.w2[j] = x1 * (1 - y1); s.w3[j] = y1 * (1 - x1); s.w4[j] = y1 * x1;
        // 32-bit chunk offsets from the uniform image base (an image holds < 2^32 chunks)
        const unsigned o = ((unsigned)xi * (unsigned)N + (unsigned)yi) * (unsigned)NQ + (unsigned)q;
        const unsigned dx = (unsigned)N * (unsigned)NQ;
        s.r1[j] = image[o];
        s.r2[j] = image[o + dx];
        s.r3[j] = image[o + (unsigned)NQ];
        s.r4[j] = image[o + dx + (unsigned)NQ];
        s.ok[j] = true;
    }
}

template <int ITEMS>
__device__ __forceinline__ void sample8x_store(const DevGeom &g, const SampleRegs<ITEMS> &s, float4 *pam,
                                               float4 *pam2) {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const int it = threadIdx.x + j * blockDim.x;
        if ((it >> 6) >= xpose_chunks(g)) continue;   // whole waves
        int h, q, i;
        xpose_item(g, it, h, q, i);
        float4 v = s.ok[j] ? add4(add4(mul4(s.r1[j], s.w1[j]), mul4(s.r2[j], s.w2[j])),
                                  add4(mul4(s.r3[j], s.w3[j]), mul4(s.r4[j], s.w4[j])))
                           : f4(0.f);
        v = xpose4(v);
        const int z = 4 * q + (i & 3);
        if (q < g.NQ && z < g.Z) (h ? pam2 : pam)[z] = v;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx k runs on XCD k mod 8),
// each with its own L2.  This maps a dispatch index to a logical index such that every XCD
// receives one contiguous range of logical indices, processed in order.
__device__ __forceinline__ int xcd_contiguous(int bid, int nb) {
    if (nb & 7) return bid;
    return (bid & 7) * (nb >> 3) + (bid >> 3);
}

// Rotation grid: voxel column x and group of 4 rows yg.  Each XCD owns a band of row
// groups and walks it in blocks of 16 columns, so the bilinear neighbourhoods that
// concurrent workgroups gather stay in that XCD's L2.
__device__ __forceinline__ void rot_block(int bid, int N, int NYG, int &x, int &yg) {
    if ((NYG & 7) == 0 && (N & 15) == 0) {
        const int xcd = bid & 7, r = bid >> 3;
        const int band = NYG >> 3, blk = 16 * band;
        x = (r / blk) * 16 + (r & 15);
        yg = xcd * band + (r % blk) / 16;
    } else {
        x = bid % N;
        yg = bid / N;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Forward projector (forward_proj3d_view1_thread PJ:1728-1978 + scatter
// PJ:1397-1436), optionally fused with the OP ratio (pt_update_est_w3
// S3:2493-2505) and the back-projection rescale (PJ:1638-1649).
//
// One workgroup per (view pair, x in [1, N/2], share of the planes).  The
// eight rotated z-columns of each row y are staged in LDS (double buffered,
// the next row's samples loaded into registers while the current row is
// projected); every thread owns up to MAXT planes (consecutive planes on
// consecutive lanes, so LDS reads are unit-stride) and accumulates each over
// y in the reference's order.  A plane's axial positions for row y are
// recomputed from per-plane constants with the float operations that build
// the reference's flookup table and middle-plane positions (PJ:544-562,
// 1929-1944), so every lane runs one code path.
// ---------------------------------------------------------------------------
// Write-out of one plane of one segment for the four lanes (v, x), (v, N-x),
// (v+90, x), (v+90, N-x) (PJ:1397-1436), optionally through the OP ratio
// (S3:2493-2505) and the back-projection rescale (PJ:1638-1649).
template <int MODE>
__device__ __forceinline__ void fp_write1(const DevGeom &g, float4 acc, int seg, int yr, int x, int view, int lp,
                                          float *__restrict__ out, const FpIO &io, double &ll) {
    const int N = g.N;
    const int pidx = g.seg_data[seg] + (yr - g.yb[seg]);
    const float resc = g.bp_rescale[seg];
#pragma unroll
    for (int l = 0; l < 4; l++) {
        if ((l & 1) && x == g.c) continue;   // at x == N/2 lanes 1/3 duplicate lanes 0/2
        const int col = (l & 1) ? N - x : x, h = l >> 1;
        float v = lane(acc, l);
        if (MODE == FP_RATIO_W012) {   // pt_update_est_w012 S3:2436-2447
            const size_t gi = ((size_t)(view * 2 + h) * N + col) * g.spp + pidx;
            const float dt = static_cast<const float *>(io.add)[gi];
            const float q = v > g.eps3 ? dt / v : 0.0f;
            v = q * resc;
        } else if (MODE != FP_PLAIN) {
            const size_t gi = ((size_t)(view * 2 + h) * N + col) * g.spp + pidx;
            const float a = MODE == FP_RATIO ? static_cast<const float *>(io.add)[gi]
                                             : half_at(static_cast<const uint16_t *>(io.add), gi);
            const float den = v + a;
            const int16_t pr = io.prompt[gi];
            float q;
            if (pr == 0) q = 0.0f;
            else if (den > g.eps3) q = (float)pr / den;
            else q = 1.0f;
            if (MODE == FP_RATIO16 && io.ll_part) {
                // Poisson log-likelihood of the bin (without log y!) at the mean (est + add) / (N*A)
                const float w = half_at(io.invnorm, gi);
                if (w > 0.0f && den > g.eps3) {
                    const double mu = (double)den * (double)w;
                    ll += (pr > 0 ? (double)pr * log(mu) : 0.0) - mu;
                }
            }
            v = q * resc;
        }
        out[((size_t)(lp * 2 + h) * N + col) * g.spp + pidx] = v;
    }
}

// FAST (HRRT's 207 planes, positions checked non-negative on the host): the staged
// column length is a compile-time constant, so every LDS read of a row takes its buffer
// and neighbour as immediate offsets, and the axial fraction is one v_fract.
constexpr int FP_PL_FAST = 216;

template <int MAXT, int ITEMS, int MODE, bool FAST>
__global__ void __launch_bounds__(1024) fp_kernel(const DevGeom g, const PairBatch b,
                                                   const float4 *__restrict__ image, float *__restrict__ out,
                                                   FpIO io, int nsplit) {
    extern __shared__ float4 lds[];
    const int PL = FAST ? FP_PL_FAST : g.Z4 + 8;   // staged planes per column (zero beyond z_pixels)
    float4 *rimc = lds;
    float4 *stage = lds + PL;  // [2][2][PL]
    // longest rays first (x = N/2 has the most y steps), then plane share, then view pair
    const int per_pair = g.c * nsplit;
    // each XCD: consecutive rays of one pair of one frame
    const int L0 = xcd_contiguous(blockIdx.x, per_pair * b.npairs * b.nframes);
    const int frame = L0 / (per_pair * b.npairs), L = L0 - frame * per_pair * b.npairs;
    image += frame * b.img_stride / 4;
    out += frame * b.ratio_stride;
    if (MODE == FP_RATIO || MODE == FP_RATIO16) io.prompt += frame * b.sino_stride;
    if (MODE == FP_RATIO || MODE == FP_RATIO_W012) io.add = static_cast<const float *>(io.add) + frame * b.sino_stride;
    else if (MODE == FP_RATIO16) io.add = static_cast<const uint16_t *>(io.add) + frame * b.sino_stride;
    double ll = 0.0;
    const int lp = L / per_pair, rem = L - lp * per_pair;
    const int x = g.c - rem / nsplit, split = rem % nsplit;
    const int p_lo = (int)(((long long)g.fp_np * split) / nsplit);
    const int p_hi = (int)(((long long)g.fp_np * (split + 1)) / nsplit);
    const int view = b.view0 + b.view_stride * lp;
    const int N = g.N, c = g.c;
    if (g.cyl_lo[x] == g.cyl_hi[x] || g.nzmask[view * (c + 1) + x] == 0) {
        if (MODE == FP_RATIO16 && io.ll_part && threadIdx.x == 0) io.ll_part[L0] = 0.0;
        return;
    }
    const float cosv = g.cos_psi[view], sinv = g.sin_psi[view];

    for (int i = threadIdx.x; i < 5 * PL; i += blockDim.x) lds[i] = f4(0.f);
    __syncthreads();
    // centre row samples (x,c),(N-x,c),(c,x),(c,N-x) PJ:1815-1829
    for (int it = threadIdx.x; it < 4 * g.NQ; it += blockDim.x) {
        int i = it / g.NQ, q = it - i * g.NQ;
        int xx = i == 0 ? x : (i == 1 ? N - x : c);
        int yy = i < 2 ? c : (i == 2 ? x : N - x);
        float4 v;
        if (!sample_chunk(g, image, cosv, sinv, xx, yy, q, v)) v = f4(0.f);
        put_lanes(rimc, q, i, v, g.Z);
    }
    const int ystart = g.cyl_lo[x];
    if (ystart < c) sample8(g, image, cosv, sinv, x, ystart, stage, stage + PL);
    __syncthreads();

    float4 acc1[MAXT], acc2[MAXT];
    int code[MAXT], ymin[MAXT], dbyte[MAXT];
    float pbase[MAXT], ptan[MAXT];
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        const int p = p_lo + threadIdx.x + t * blockDim.x;
        code[t] = -1; ymin[t] = c; pbase[t] = 0.f; ptan[t] = 0.f; dbyte[t] = 0;
        int yr = 0;
        if (p < p_hi) {
            const int4 e = g.fp_planes[p];
            code[t] = e.x; pbase[t] = __int_as_float(e.y); ptan[t] = __int_as_float(e.z); ymin[t] = e.w;
            yr = (e.x >> 8) & 0xff;
            dbyte[t] = (e.x >> 16) * (int)sizeof(float4);   // the plane shift delta, in LDS bytes
        }
        const float4 v = rimc[yr];   // centre row sample at plane yr (PJ:1830-1838)
        acc1[t] = v;
        acc2[t] = v;
    }

    // project the staged row y (halves pa, pb) into every plane this thread owns
    auto project = [&](const float4 *pa, const float4 *pb, int y) {
        const float fy2 = (float)(c - y);   // (y - c) * tan = -((c - y) * tan) exactly
#pragma unroll
        for (int t = 0; t < MAXT; t++) {
            if (code[t] < 0) continue;
            const int k = code[t] & 0xff;
            if (k == 0) {   // theta = 0 PJ:1907-1910
                const int yr = (code[t] >> 8) & 0xff;
                acc1[t] = add4(acc1[t], add4(pa[yr], pb[yr]));
                continue;
            }
            if (y < ymin[t]) continue;   // PJ:1920-1921
            const float pt = fy2 * ptan[t];
            const float zz1 = pbase[t] - pt, zz2 = pbase[t] + pt;
            const int i1 = (int)zz1, i2 = (int)zz2;
            float f1, f2;
            if (FAST) {        // zz >= 0: zz - trunc(zz) = zz - floor(zz)
                f1 = __builtin_amdgcn_fractf(zz1);
                f2 = __builtin_amdgcn_fractf(zz2);
            } else {
                f1 = zz1 - i1;
                f2 = zz2 - i2;
            }
            // planes z = i1 + delta and z2 = i2 + delta: the shift folded into the base address
            const float4 *qa = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(pa) + dbyte[t]);
            const float4 *qb = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(pb) + dbyte[t]);
            const float4 a0 = qa[i1], a1 = qa[i1 + 1], b0 = qb[i1], b1 = qb[i1 + 1];
            const float4 d0 = qb[i2], d1 = qb[i2 + 1], e0 = qa[i2], e1 = qa[i2 + 1];
            // PJ:1934-1944: (pam2[z2+1] + pam[z])*f2 + (pam[z+1] + pam2[z2])*f1 and its mirror
            acc1[t] = add4(acc1[t], add4(mul4(add4(d1, a0), f2), mul4(add4(a1, d0), f1)));
            acc2[t] = add4(acc2[t], add4(mul4(add4(e1, b0), f2), mul4(add4(b1, e0), f1)));
        }
    };
    // one row per barrier: project row y (buffer ST) while the next row's samples load
    auto row = [&](auto stc, int y) {
        constexpr int st = decltype(stc)::value;
        SampleRegs<ITEMS> nxt;
        if (y + 1 < c) {
            if (FAST) sample8x_load<ITEMS>(g, image, cosv, sinv, x, y + 1, nxt);
            else sample8_load<ITEMS>(g, image, cosv, sinv, x, y + 1, nxt);
        }
        const float4 *pa = stage + st * 2 * PL;
        project(pa, pa + PL, y);
        if (y + 1 < c) {
            float4 *na = stage + (st ^ 1) * 2 * PL;
            if (FAST) sample8x_store<ITEMS>(g, nxt, na, na + PL);
            else sample8_store<ITEMS>(g, nxt, na, na + PL);
        }
        __syncthreads();
    };
    for (int y = ystart; y < c; y += 2) {   // row ystart + 2j reads buffer 0, the next one buffer 1
        row(std::integral_constant<int, 0>{}, y);
        if (y + 1 < c) row(std::integral_constant<int, 1>{}, y + 1);
    }

    // scaling PJ:1953-1975 and write-out
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        if (code[t] < 0) continue;
        const int k = code[t] & 0xff, yr = (code[t] >> 8) & 0xff;
        const float sc = k == 0 ? g.y_size : g.rescalearray[(k - 1) * g.Z + yr];
        fp_write1<MODE>(g, mul4(acc1[t], sc), k == 0 ? 0 : 2 * k, yr, x, view, lp, out, io, ll);
        if (k != 0) fp_write1<MODE>(g, mul4(acc2[t], sc), 2 * k - 1, yr, x, view, lp, out, io, ll);
    }
    if (MODE == FP_RATIO16 && io.ll_part) {
        // workgroup sum of the log-likelihood in a fixed order (waves, then lanes)
        for (int o = 32; o > 0; o >>= 1) ll += __shfl_xor(ll, o);
        __syncthreads();
        double *red = reinterpret_cast<double *>(lds);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ll;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
            io.ll_part[L0] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// Back projector, direct variant (back_proj3d2 PJ:1457-1620 with
// back_proj3d_view_n1 PJ:2927-3086), as a gather.  For a rotated-frame voxel
// (X, Y) the reference's pass over view v gives A(X,Y) from sinogram column X
// and its pass over view v+90 gives B(X,Y) from column Y; R = A + B.  One
// workgroup per (column, role, view pair) stages the column in LDS and
// evaluates every voxel of the disk row/column that reads it; a wave owns one
// voxel and its lanes the planes z = lane + 64 j.
//
// LDS holds (p[i], p[i+1] - p[i]) per plane, the difference rounded as the
// reference rounds it, and every oblique segment sits inside a window of zero
// planes sized on the host from its rows' shifts.  Slots 0-2 (z < 192) are
// therefore evaluated without any range test: lanes outside the reference's
// range [zsa, zea] (PJ:2986-2995) interpolate between zeros and add exactly +0;
// slot 3 runs only when the range reaches it (a wave-uniform test).  The wave
// keeps its voxel row's lookup entries in VGPRs (lane k-1: row sl, lane 32+k-1:
// row N-sl) and takes entry k with readlane, so the loop over segment groups
// waits on nothing but its own LDS reads.
// ---------------------------------------------------------------------------
template <typename SRC_T, bool KAHAN>
__global__ void __launch_bounds__(1024) bp_kernel(const DevGeom g, const PairBatch b, const SRC_T *__restrict__ src,
                                                  int src_global, int apply_rescale, const int16_t *__restrict__ pseg,
                                                  float *__restrict__ bufA, float *__restrict__ bufB) {
    extern __shared__ float2 colpq[];
    const int col = blockIdx.x + 1, role = blockIdx.y, lp = blockIdx.z % b.npairs, frame = blockIdx.z / b.npairs;
    const int view = b.view0 + b.view_stride * lp;
    const int N = g.N, c = g.c, Z = g.Z, GM = g.GM, K = g.K;
    src += frame * b.ratio_stride;
    bufA += frame * b.buf_stride;
    bufB += frame * b.buf_stride;
    const int lo = g.cyl_lo[col], hi = g.cyl_hi[col];
    if (lo == hi) return;
    const int gp = src_global ? view : lp;
    const SRC_T *cp = src + ((size_t)(gp * 2 + role) * N + col) * g.spp;
    auto value = [&](int i) -> float {
        float v;
        if constexpr (sizeof(SRC_T) == 2) v = half_at(reinterpret_cast<const uint16_t *>(cp), i);
        else v = cp[i];
        return apply_rescale ? v * g.bp_rescale[pseg[i]] : v;
    };
    for (int i = threadIdx.x; i < g.bp_lds; i += blockDim.x) {
        const int s0 = g.bp_src[i], s1 = g.bp_src[i + 1];
        const float v0 = s0 >= 0 ? value(s0) : 0.0f, v1 = s1 >= 0 ? value(s1) : 0.0f;
        colpq[i] = make_float2(v0, v1 - v0);
    }
    __syncthreads();
    const int mcol = col <= c ? col : N - col;
    const bool masked = g.nzmask[view * (c + 1) + mcol] == 0;
    float *dst = role == 0 ? bufA : bufB;
    dst += (size_t)lp * N * N * g.Z4;
    const int lanei = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (lo == c && hi == c + 1) {
        // a one-pixel column: the incremental variant also writes rows c-1 and c+1, zero here
        for (int t = threadIdx.x; t < 2 * 64; t += blockDim.x) {
            const int V = (t >> 6) ? c + 1 : c - 1;
            const int X = role == 0 ? col : V, Y = role == 0 ? V : col;
            for (int z = t & 63; z < Z; z += 64) dst[((size_t)X * N + Y) * g.Z4 + z] = 0.0f;
        }
    }
    for (int Vv = lo + wave; Vv < hi; Vv += nw) {
        const int V = __builtin_amdgcn_readfirstlane(Vv);
        const int X = role == 0 ? col : V, Y = role == 0 ? V : col;
        float *o = dst + ((size_t)X * N + Y) * g.Z4;
        if (masked && V != c) {
            for (int z = lanei; z < Z; z += 64) o[z] = 0.0f;
            continue;
        }
        const int sl = V <= c ? V : N - V, ll = N - sl;
        const bool first_even = role == 0 ? (V <= c) : (V >= c);
        // lookup entries of rows sl (lanes 0..GM-1) and ll (lanes 32..32+GM-1), with the LDS base
        // of the segment each row reads already chosen
        int range = 0, base = 0, frac = 0;   // frac: float bits (readlane moves 32-bit integers)
        {
            const bool own = lanei < GM, other = lanei >= 32 && lanei - 32 < GM;
            if (own || other) {
                const int4 e = g.bp_lut2[(own ? sl : ll) * GM + (lanei & 31)];
                range = e.x;
                base = (own == first_even) ? e.y : e.z;
                frac = e.w;
            }
        }
        float i1[4], i2[4], c1[4], c2[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int z = lanei + 64 * j;
            const float t0 = z < Z ? colpq[z].x : 0.0f;
            i1[j] = first_even ? t0 : 0.0f;
            i2[j] = first_even ? 0.0f : t0;
            c1[j] = c2[j] = 0.0f;
        }
        for (int k = 1; k <= K; k++) {
            const int b1 = __builtin_amdgcn_readlane(base, k - 1), b2 = __builtin_amdgcn_readlane(base, 31 + k);
            const float f1 = __int_as_float(__builtin_amdgcn_readlane(frac, k - 1));
            const float f2 = __int_as_float(__builtin_amdgcn_readlane(frac, 31 + k));
            const float2 *q1 = colpq + b1 + lanei, *q2 = colpq + b2 + lanei;
            float2 a[3], bq[3];
#pragma unroll
            for (int j = 0; j < 3; j++) {
                a[j] = q1[64 * j];
                bq[j] = q2[64 * j];
            }
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if (KAHAN) {
                    kahan_add(i1[j], c1[j], a[j].x + f1 * a[j].y);
                    kahan_add(i2[j], c2[j], bq[j].x + f2 * bq[j].y);
                } else {
                    i1[j] = i1[j] + (a[j].x + f1 * a[j].y);
                    i2[j] = i2[j] + (bq[j].x + f2 * bq[j].y);
                }
            }
            const int r1 = __builtin_amdgcn_readlane(range, k - 1), r2 = __builtin_amdgcn_readlane(range, 31 + k);
            const int zs1 = r1 & 0xffff, ze1 = r1 >> 16, zs2 = r2 & 0xffff;
            if (ze1 >= 192) {
                const float2 t = q1[192];
                if (KAHAN) kahan_add(i1[3], c1[3], t.x + f1 * t.y);
                else i1[3] = i1[3] + (t.x + f1 * t.y);
            }
            if (zs2 + (ze1 - zs1) >= 192) {
                const float2 t = q2[192];
                if (KAHAN) kahan_add(i2[3], c2[3], t.x + f2 * t.y);
                else i2[3] = i2[3] + (t.x + f2 * t.y);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int z = lanei + 64 * j;
            if (z < Z) o[z] = i1[j] + i2[j];
        }
    }
}

// ---------------------------------------------------------------------------
// Back projector, incremental variant (back_proj3d_thread1 PJ:1622-1699 with
// back_proj3d_view_n1thread PJ:4162-4668), the reference's default.  Walking a
// column's rows away from the centre, each plane keeps running sums s1 (row y)
// and s2 (row N-y) and their per-row increments d1, d2; at the rows where a
// group's integer shift changes (the breakpoints of blutacc, PJ:718-784) the
// increments are corrected.  One workgroup per (column x <= N/2, lane, view
// pair); lane l is bin x or N-x (l & 1) of view v or v+90 (l >> 1); a thread
// owns one plane and applies, to its own accumulators, exactly the terms the
// reference's loops apply to that plane, in the reference's order.  Lanes 0/1
// write the rotated-frame half A, lanes 2/3 the half B, so each voxel of A and
// of B has one writer and R = A + B as for the direct variant.
//
// LDS: segment 0, then per group a window of float2 {segment 2k, 2k-1} by row
// (layout and the two in-place conversions are described with the host
// tables).  A breakpoint is then, for every plane, two clamped float2 reads
// and eight fixed operations (b-term first, then a-term, each one s += m*coef,
// d -= m for both sums); the reads that fall outside the breakpoint's rows hit
// the window's zero cells and add +0.  Breakpoint records come through the
// scalar cache one ahead of use; per-group constants sit in VGPRs (lane k-1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float rlf(int v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(v, lane)); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ void __launch_bounds__(256) bp_inc_kernel(const DevGeom g, const PairBatch b, const float *__restrict__ src,
                                                     int src_global, int apply_rescale,
                                                     const int16_t *__restrict__ pseg, float *__restrict__ bufA,
                                                     float *__restrict__ bufB) {
    extern __shared__ float P[];
    const float2 *W = reinterpret_cast<const float2 *>(P);
    const int x = blockIdx.x + 1, l = blockIdx.y, lp = blockIdx.z % b.npairs, frame = blockIdx.z / b.npairs;
    const int N = g.N, c = g.c, Z = g.Z, K = g.K;
    if (x == c && (l & 1)) return;          // at x == N/2 lanes 1/3 duplicate lanes 0/2
    const int lo = g.cyl_lo[x], hi = g.cyl_hi[x];
    if (lo == hi) return;
    const int view = b.view0 + b.view_stride * lp;
    const int h = l >> 1, col = (l & 1) ? N - x : x;
    const int z = threadIdx.x, lane = threadIdx.x & 63;
    const bool act = z < Z;
    float *dst = (l < 2 ? bufA : bufB) + frame * b.buf_stride + (size_t)lp * N * N * g.Z4;
    // s1 of row y goes to A(col, y) / B(N-y, col), s2 to A(col, N-y) / B(y, col)
    auto vox = [&](int y, bool second) -> float * {
        const int X = l < 2 ? col : (second ? y : N - y);
        const int Y = l < 2 ? (second ? N - y : y) : col;
        return dst + ((size_t)X * N + Y) * g.Z4;
    };
    const int ylast = lo < c - 1 ? lo : c - 1;                       // rows c-1 .. ylast are written
    if (g.nzmask[view * (c + 1) + x] == 0) {                          // masked column PJ:4236-4262
        if (act) {
            vox(c, false)[z] = 0.0f;
            for (int y = c - 1; y >= ylast; y--) {
                vox(y, false)[z] = 0.0f;
                vox(y, true)[z] = 0.0f;
            }
        }
        return;
    }
    src += frame * b.ratio_stride;
    const int gp = src_global ? view : lp;
    const float *cp = src + ((size_t)(gp * 2 + h) * N + col) * g.spp;
    for (int base = 0; base < g.inc_lds; base += 4 * (int)blockDim.x) {   // transpose PJ:1654-1693
        int si[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int i = base + threadIdx.x + j * blockDim.x;
            si[j] = i < g.inc_lds ? g.inc_src[i] : -1;
        }
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            v[j] = 0.0f;
            if (si[j] >= 0) v[j] = apply_rescale ? cp[si[j]] * g.bp_rescale[pseg[si[j]]] : cp[si[j]];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int i = base + threadIdx.x + j * blockDim.x;
            if (i < g.inc_lds) P[i] = v[j];
        }
    }
    // per-group constants of group lane+1
    int4 e0 = make_int4(0, 0, 0, 0), e1 = e0, e2 = e0, e3 = e0;
    if (lane < K) {
        const int4 *q = g.inc_group + 4 * lane;
        e0 = q[0]; e1 = q[1]; e2 = q[2]; e3 = q[3];
    }
    __syncthreads();
    const int zc = act ? z : 0;
    // centre voxels PJ:4265-4285
    {
        float sv = P[zc];
        for (int k = 0; k < K; k++) {
            const int gb = rl(e0.x, k), yb = rl(e0.y, k), yt = rl(e0.z, k);
            const float2 w = W[gb + clampi(z, yb, yt)];
            const float t = sv + (w.x + w.y);
            sv = (z >= yb && z <= yt) ? t : sv;
        }
        if (act) vox(c, false)[z] = sv;
    }
    // first row y = c-1 by direct interpolation PJ:4287-4312
    float s1 = P[zc], s2 = s1, d1 = 0.0f, d2 = 0.0f;
    for (int k = 0; k < K; k++) {
        const int gb = rl(e0.x, k);
        const int za = rl(e1.x, k), n = rl(e1.y, k), ya = rl(e1.z, k), zb = rl(e2.x, k), yb2 = rl(e2.y, k);
        const float c1 = rlf(e1.w, k), c2 = rlf(e2.z, k);
        const int ia = z - za, ib = z - zb, nm = n > 1 ? n - 1 : 0;
        const bool va = ia >= 0 && ia < n, vb = ib >= 0 && ib < n;
        const int ra = gb + ya + clampi(ia, 0, nm), rb = gb + yb2 + clampi(ib, 0, nm);
        const float2 pa = W[ra], qa = W[ra + 1], pb = W[rb], qb = W[rb + 1];
        const float a1 = pa.x + c1 * (qa.x - pa.x), a2 = pa.y + c1 * (qa.y - pa.y);
        const float b1 = pb.y + c2 * (qb.y - pb.y), b2 = pb.x + c2 * (qb.x - pb.x);
        if (zb <= za) {              // iteration ia <= ib: the row-y term first
            s1 = va ? s1 + a1 : s1; s2 = va ? s2 + a2 : s2;
            s1 = vb ? s1 + b1 : s1; s2 = vb ? s2 + b2 : s2;
        } else {
            s1 = vb ? s1 + b1 : s1; s2 = vb ? s2 + b2 : s2;
            s1 = va ? s1 + a1 : s1; s2 = va ? s2 + a2 : s2;
        }
    }
    __syncthreads();
    // in place, twice: the reference's scaled differences P' (PJ:4313-4324), then the values the
    // breakpoint terms read
    for (int pass = 0; pass < 2; pass++) {
        for (int base = 0; base < g.inc_lds; base += 4 * (int)blockDim.x) {
            float nv[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int i = base + threadIdx.x + j * blockDim.x;
                nv[j] = 0.0f;
                if (i < g.inc_lds) {
                    const int2 cv = g.inc_conv[i];
                    const float p = P[i];
                    if (pass == 0) nv[j] = cv.x ? (P[i + 2] - p) * __int_as_float(cv.x) : p;
                    else nv[j] = cv.y == 1 ? p - P[i + 2] : (cv.y == 2 ? -P[i + 2] : p);
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int i = base + threadIdx.x + j * blockDim.x;
                if (i < g.inc_lds) P[i] = nv[j];
            }
            __syncthreads();
        }
        if (pass == 1) break;
        for (int k = 0; k < K; k++) {                                // PJ:4325-4335
            const int gb = rl(e0.x, k), yb = rl(e0.y, k), yt = rl(e0.z, k);
            const int zs = rl(e3.x, k), ze = rl(e3.y, k), yr0 = rl(e3.z, k);
            const bool vd = z >= yb - 1 && z <= yt;
            const float2 w = W[gb + clampi(z, yb - 1, yt)];
            const float t1 = d1 + w.x, t2 = d2 + w.y;
            d1 = vd ? t1 : d1; d2 = vd ? t2 : d2;
            const int j = z - zs;
            const bool vj = j >= 0 && j <= ze - zs;
            const float2 u = W[gb + yr0 + clampi(j, 0, ze > zs ? ze - zs : 0)];
            const float u1 = d1 - u.y, u2 = d2 - u.x;
            d1 = vj ? u1 : d1; d2 = vj ? u2 : d2;
        }
        __syncthreads();
    }
    if (act) {
        vox(c - 1, false)[z] = s1;
        vox(c - 1, true)[z] = s2;
    }
    // rows c-2 .. ystart PJ:4337-4562: breakpoints, then the running sums advance by their
    // increments
    const int z8 = 8 * z;
    const char *Wb = reinterpret_cast<const char *>(P);
    const int4 *__restrict__ T = g.inc_sched;
    int e = 0;
    int4 cur = T[0];
    for (int y = c - 2; y >= lo; y--) {
        const int n = g.inc_nrow[y];
        for (int j = 0; j < n; j++) {
            const int4 nxt = T[e + 1];
            const float cf = __int_as_float(cur.x);
            const int lo8 = cur.w & 0xffff, hi8 = cur.w >> 16;
            // lo8 < hi8 always, so the clamp is max(lo, min(v, hi)): two VALU instead of a compare,
            // a select and a min
            const int ra = max(lo8, min(z8 + cur.y, hi8)), rb = max(lo8, min(z8 + cur.z, hi8));
            const float2 ma = *reinterpret_cast<const float2 *>(Wb + ra);
            const float2 mb = *reinterpret_cast<const float2 *>(Wb + rb);
            // b-term (PJ:4386-4391 for rows N-y): s1 += m2*coef, s2 += m1*coef, d1 -= m2, d2 -= m1
            s1 = s1 + mb.y * cf; d1 = d1 - mb.y;
            s2 = s2 + mb.x * cf; d2 = d2 - mb.x;
            // a-term (PJ:4382-4385): s1 += m1*coef, s2 += m2*coef, d1 -= m1, d2 -= m2
            s1 = s1 + ma.x * cf; d1 = d1 - ma.x;
            s2 = s2 + ma.y * cf; d2 = d2 - ma.y;
            cur = nxt;
            e++;
        }
        s1 = s1 + d1;
        s2 = s2 + d2;
        if (act) {
            vox(y, false)[z] = s1;
            vox(y, true)[z] = s2;
        }
    }
}

// ---------------------------------------------------------------------------
// Rotation (rotateimage2 PJ:1137-1186) of R = A + B of every pair of the
// batch, accumulated in pair order, fused with the image update
// (pt_cal_updateimage S3:2033-2056).  Block (64, 4): x lanes over z chunks of 4.
// ---------------------------------------------------------------------------
// One pair's bilinear neighbourhood of R = A + B at the rotated position of (x, y):
// the eight loads are issued here and consumed one pair later.
struct RotTap {
    float4 a1, a2, a3, a4, b1, b2, b3, b4;
    float x1, y1;
    bool ok;
};

__device__ __forceinline__ void rot_issue(const DevGeom &g, const float4 *__restrict__ A,
                                          const float4 *__restrict__ B, int lp, float2 pnt, int q, RotTap &t) {
    const int N = g.N, NQ = g.NQ;
    const int xi = (int)pnt.x, yi = (int)pnt.y;
    t.ok = !(xi < 0 || xi > N - 2 || yi < 0 || yi > N - 2);
    t.x1 = pnt.x - xi;
    t.y1 = pnt.y - yi;
    const int xc = t.ok ? xi : 0, yc = t.ok ? yi : 0;      // in-bounds reads for a skipped pair
    const size_t base = (size_t)lp * N * N * NQ + q;
    const size_t i1 = base + ((size_t)xc * N + yc) * NQ, i2 = base + ((size_t)(xc + 1) * N + yc) * NQ;
    t.a1 = A[i1]; t.b1 = B[i1];
    t.a2 = A[i2]; t.b2 = B[i2];
    t.a3 = A[i1 + NQ]; t.b3 = B[i1 + NQ];
    t.a4 = A[i2 + NQ]; t.b4 = B[i2 + NQ];
}

// Sum over the batch's pairs, in pair order, of the bilinear interpolation of R = A + B
// at the position the reference's rotation table gives (x, y) for the pair's view.
// Software-pipelined: the table entry two pairs ahead and the neighbourhood one pair
// ahead are in flight while a pair is accumulated.
template <bool KAHAN>
__device__ __forceinline__ float4 rotate_accumulate(const DevGeom &g, const PairBatch &b,
                                                    const float4 *__restrict__ A, const float4 *__restrict__ B,
                                                    int x, int y, int q) {
    const int N = g.N;
    const size_t vstride = (size_t)N * N * b.view_stride;
    const float2 *rt = g.rot + ((size_t)b.view0 * N + x) * N + y;
    float4 corr = f4(0.f), comp = f4(0.f);
    RotTap cur, nxt;
    float2 pn = b.npairs > 1 ? rt[vstride] : make_float2(0.f, 0.f);
    rot_issue(g, A, B, 0, rt[0], q, cur);
    for (int lp = 0; lp < b.npairs; lp++) {
        if (lp + 1 < b.npairs) {
            const float2 p = pn;
            if (lp + 2 < b.npairs) pn = rt[vstride * (lp + 2)];
            rot_issue(g, A, B, lp + 1, p, q, nxt);
        }
        if (cur.ok) {
            const float x1 = cur.x1, y1 = cur.y1;
            const float w1 = (1 - x1) * (1 - y1), w2 = x1 * (1 - y1), w3 = y1 * (1 - x1), w4 = y1 * x1;
            const float4 m1 = add4(cur.a1, cur.b1), m2 = add4(cur.a2, cur.b2);
            const float4 m3 = add4(cur.a3, cur.b3), m4 = add4(cur.a4, cur.b4);
            const float4 v = add4(add4(mul4(m1, w1), mul4(m2, w2)), add4(mul4(m3, w3), mul4(m4, w4)));
            if (KAHAN) {
                kahan_add(corr.x, comp.x, v.x);
                kahan_add(corr.y, comp.y, v.y);
                kahan_add(corr.z, comp.z, v.z);
                kahan_add(corr.w, comp.w, v.w);
            } else {
                corr = add4(corr, v);
            }
        }
        cur = nxt;
    }
    return corr;
}

__device__ __forceinline__ float update1(float img, float corr, float nf, float image_max) {
    const float amin = 0.01f, amax = 100.00f;
    float cl = amin > corr ? amin : corr;   // _mm_max_ps(amin, corr)
    cl = amax < cl ? amax : cl;             // _mm_min_ps(amax, .)
    const float t = img * (cl * nf);
    return image_max < t ? image_max : t;   // _mm_min_ps(image_max, .)
}

template <bool KAHAN>
__global__ void __launch_bounds__(256) rot_update_kernel(const DevGeom g, const PairBatch b,
                                                         const float4 *__restrict__ A, const float4 *__restrict__ B,
                                                         float4 *__restrict__ image, const float4 *__restrict__ nf) {
    int x, yg;
    const int frame = blockIdx.y;
    A += frame * b.buf_stride / 4;
    B += frame * b.buf_stride / 4;
    image += frame * b.img_stride / 4;
    rot_block(blockIdx.x, g.N, (g.N + 3) / 4, x, yg);
    const int q = threadIdx.x, y = yg * blockDim.y + threadIdx.y;
    if (q >= g.NQ || y < g.cyl_lo[x] || y >= g.cyl_hi[x]) return;
    const float4 corr = rotate_accumulate<KAHAN>(g, b, A, B, x, y, q);
    const size_t i = ((size_t)x * g.N + y) * g.NQ + q;
    const float4 im = image[i], n = nf[i];
    image[i] = make_float4(update1(im.x, corr.x, n.x, g.image_max), update1(im.y, corr.y, n.y, g.image_max),
                           update1(im.z, corr.z, n.z, g.image_max), update1(im.w, corr.w, n.w, g.image_max));
}

// Rotation sum only (rotate_image_buf into a zeroed correction, PJ:1701-1725), for the
// paths that filter the correction or the sensitivity before using it.
__global__ void __launch_bounds__(256) rot_sum_kernel(const DevGeom g, const PairBatch b,
                                                      const float4 *__restrict__ A, const float4 *__restrict__ B,
                                                      float4 *__restrict__ corr) {
    int x, yg;
    rot_block(blockIdx.x, g.N, (g.N + 3) / 4, x, yg);
    const int q = threadIdx.x, y = yg * blockDim.y + threadIdx.y;
    if (q >= g.NQ || y < g.cyl_lo[x] || y >= g.cyl_hi[x]) return;
    corr[((size_t)x * g.N + y) * g.NQ + q] = rotate_accumulate<false>(g, b, A, B, x, y, q);
}

// ---------------------------------------------------------------------------
// PSF resolution model (convolve3d psf.cpp:441-476).  The x pass (convolvex
// psf.cpp:168-219) sums, for a FOV voxel, the taps mask = -h .. h-1 whose column
// lies inside the row's FOV and divides by their sum unless it is 0 or 1.  The
// y/z pass (convolveyz psf.cpp:324-439) sums the y taps the same way into a
// z column kept in LDS, then every plane z < z_pixels takes, with the y kernel,
// the planes z + mask inside [z_min, z_max), and is scaled by normz(z) * (1/sum
// of the y taps).  Only FOV voxels are written, as in the reference.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) psf_x_kernel(const DevGeom g, const float4 *__restrict__ in,
                                                    float4 *__restrict__ out, const float *__restrict__ taps, int h) {
    int x, yg;
    rot_block(blockIdx.x, g.N, (g.N + 3) / 4, x, yg);
    const int q = threadIdx.x, y = yg * blockDim.y + threadIdx.y;
    if (q >= g.NQ || y < g.cyl_lo[x] || y >= g.cyl_hi[x]) return;
    const int lo = g.cyl_lo[y], hi = g.cyl_hi[y];
    float4 acc = f4(0.f);
    float norm = 0.0f;
    for (int mask = -h; mask < h; mask++) {
        const int m2 = x + mask;
        if (m2 < lo || m2 >= hi) continue;
        const float f = taps[mask + h];
        norm = norm + f;
        acc = add4(acc, mul4(in[((size_t)m2 * g.N + y) * g.NQ + q], f));
    }
    if (norm != 1.0f && norm != 0.0f) acc = mul4(acc, 1.0f / norm);
    out[((size_t)x * g.N + y) * g.NQ + q] = acc;
}

__global__ void __launch_bounds__(256) psf_yz_kernel(const DevGeom g, const float4 *__restrict__ in,
                                                     float *__restrict__ out, const float *__restrict__ taps, int h,
                                                     const float *__restrict__ normz, int z_min, int z_max) {
    __shared__ float col[4][256];
    int x, yg;
    rot_block(blockIdx.x, g.N, (g.N + 3) / 4, x, yg);
    const int q = threadIdx.x, ty = threadIdx.y, y = yg * blockDim.y + ty;
    const int lo = g.cyl_lo[x], hi = g.cyl_hi[x];
    const bool fov = y >= lo && y < hi;
    float norm = 0.0f;
    if (fov) {
        float4 acc = f4(0.f);
        for (int mask = -h; mask < h; mask++) {
            const int m2 = y + mask;
            if (m2 < lo || m2 >= hi) continue;
            const float f = taps[mask + h];
            norm = norm + f;
            if (q < g.NQ) acc = add4(acc, mul4(in[((size_t)x * g.N + m2) * g.NQ + q], f));
        }
        if (q < g.NQ) {
            col[ty][4 * q + 0] = acc.x;
            col[ty][4 * q + 1] = acc.y;
            col[ty][4 * q + 2] = acc.z;
            col[ty][4 * q + 3] = acc.w;
        }
    }
    __syncthreads();
    if (!fov) return;
    float *o = out + ((size_t)x * g.N + y) * g.Z4;
    for (int z = q; z < g.Z4; z += 64) {
        float v = 0.0f;
        if (norm != 0.0f) {
            for (int mask = -h; mask < h; mask++) {
                const int zz = z + mask;
                if (z < g.Z && zz >= z_min && zz < z_max) v = v + col[ty][zz] * taps[mask + h];
            }
            v = v * (normz[z] * (1.0f / norm));
        }
        o[z] = v;
    }
}

// ---------------------------------------------------------------------------
// Flat image files, planes [z][v][u] (SaveImage S3:4023-4035 with write_flat_image
// FIO:300-333; the -I read of Init_Image S3:3970-3993).  Image voxel (a, b) holds flat
// (v, u) = (b, a), or (a, b) when flipped; only FOV voxels are copied, the rest is 0.
// A workgroup transposes a 64 (u) x 64 (z) tile of one v through LDS, so both sides
// move whole 256-byte rows.
// ---------------------------------------------------------------------------
template <bool TO_FLAT>
__global__ void __launch_bounds__(256) flat_image_kernel(const DevGeom g, const float *__restrict__ src,
                                                         float *__restrict__ dst, int flip) {
    __shared__ float tile[64][65];
    const int v = blockIdx.y, u0 = blockIdx.x * 64, z0 = blockIdx.z * 64, N = g.N;
    const int t = threadIdx.x & 63, w = threadIdx.x >> 6;
    auto voxel = [&](int u, int &a, int &b) {
        a = flip ? v : u;
        b = flip ? u : v;
        return u < N && b >= g.cyl_lo[a] && b < g.cyl_hi[a];
    };
    if (TO_FLAT) {
        for (int i = w; i < 64; i += 4) {            // image rows: u0 + i, planes z0 + t
            int a, b;
            const int z = z0 + t;
            float val = 0.0f;
            if (voxel(u0 + i, a, b) && z < g.Z) val = src[((size_t)a * N + b) * g.Z4 + z];
            tile[i][t] = val;
        }
        __syncthreads();
        for (int i = w; i < 64; i += 4) {            // flat rows: plane z0 + i, columns u0 + t
            const int z = z0 + i, u = u0 + t;
            if (z < g.Z && u < N) dst[((size_t)z * N + v) * N + u] = tile[t][i];
        }
    } else {
        for (int i = w; i < 64; i += 4) {            // flat rows: plane z0 + i, columns u0 + t
            const int z = z0 + i, u = u0 + t;
            tile[t][i] = (z < g.Z && u < N) ? src[((size_t)z * N + v) * N + u] : 0.0f;
        }
        __syncthreads();
        for (int i = w; i < 64; i += 4) {            // image rows: u0 + i, planes z0 + t
            int a, b;
            const int z = z0 + t;
            if (u0 + i >= N || z >= g.Z4) continue;
            const bool in = voxel(u0 + i, a, b);       // (a, b) set either way; every voxel is written
            dst[((size_t)a * N + b) * g.Z4 + z] = (in && z < g.Z) ? tile[i][t] : 0.0f;
        }
    }
}

__device__ __forceinline__ float invert1(float s, float eps1) { return fabsf(s) > eps1 ? (1.0f / s) : 0.0f; }

template <bool KAHAN>
__global__ void __launch_bounds__(256) rot_sens_kernel(const DevGeom g, const PairBatch b,
                                                       const float4 *__restrict__ A, const float4 *__restrict__ B,
                                                       float4 *__restrict__ nf) {
    int x, yg;
    rot_block(blockIdx.x, g.N, (g.N + 3) / 4, x, yg);
    const int q = threadIdx.x, y = yg * blockDim.y + threadIdx.y;
    if (q >= g.NQ || y < g.cyl_lo[x] || y >= g.cyl_hi[x]) return;
    const float4 s = rotate_accumulate<KAHAN>(g, b, A, B, x, y, q);
    const int z = 4 * q;
    float4 o;
    o.x = z + 0 < g.Z ? invert1(s.x, g.eps1) : 0.0f;
    o.y = z + 1 < g.Z ? invert1(s.y, g.eps1) : 0.0f;
    o.z = z + 2 < g.Z ? invert1(s.z, g.eps1) : 0.0f;
    o.w = z + 3 < g.Z ? invert1(s.w, g.eps1) : 0.0f;
    nf[((size_t)x * g.N + y) * g.NQ + q] = o;
}

__global__ void __launch_bounds__(256) rot_add_kernel(const DevGeom g, int view, const float4 *__restrict__ A,
                                                      const float4 *__restrict__ B, float4 *__restrict__ ima) {
    int x, yg;
    rot_block(blockIdx.x, g.N, (g.N + 3) / 4, x, yg);
    const int q = threadIdx.x, y = yg * blockDim.y + threadIdx.y;
    if (q >= g.NQ || y < g.cyl_lo[x] || y >= g.cyl_hi[x]) return;
    PairBatch b{view, 0, 1};
    const float4 r = rotate_accumulate<false>(g, b, A, B, x, y, q);
    const size_t i = ((size_t)x * g.N + y) * g.NQ + q;
    ima[i] = add4(ima[i], r);
}

// The rotated-frame buffer of one view pair, R = A + B on the FOV columns: the reference's
// imagebuf after back_proj3d_thread1 (back_proj3d_view_n1thread PJ:4162-4668), whose two
// halves it writes with = and += (the sum is the same single addition).  Like the reference,
// it leaves the pad planes z >= z_pixels as they were.
__global__ void __launch_bounds__(256) halves_sum_kernel(const DevGeom g, const float4 *__restrict__ A,
                                                         const float4 *__restrict__ B, float4 *__restrict__ out) {
    int x, yg;
    rot_block(blockIdx.x, g.N, (g.N + 3) / 4, x, yg);
    const int q = threadIdx.x, y = yg * blockDim.y + threadIdx.y;
    if (q >= g.NQ || y < g.cyl_lo[x] || y >= g.cyl_hi[x]) return;
    const size_t i = ((size_t)x * g.N + y) * g.NQ + q;
    float4 r = add4(A[i], B[i]);
    if (4 * q + 3 >= g.Z) {
        const float4 o = out[i];
        const int z = 4 * q;
        r = make_float4(z + 0 < g.Z ? r.x : o.x, z + 1 < g.Z ? r.y : o.y, z + 2 < g.Z ? r.z : o.z,
                        z + 3 < g.Z ? r.w : o.w);
    }
    out[i] = r;
}

// ---------------------------------------------------------------------------
// Layout conversions for the per-view-pair API.
// ---------------------------------------------------------------------------
// pair-column -> estimate [2S][N], covered bins only (PJ:1397-1436).
__global__ void paircol_to_estimate_kernel(const DevGeom g, int view, const float *__restrict__ pc,
                                           float *__restrict__ prj) {
    const int col = blockIdx.x, h = blockIdx.y;
    const int N = g.N, c = g.c;
    if (col == 0) return;
    const int m = col <= c ? col : N - col;
    if (g.cyl_lo[m] == 0 || g.nzmask[view * (c + 1) + m] == 0) return;
    const float *src = pc + ((size_t)h * N + col) * g.spp;
    for (int p = threadIdx.x; p < g.S; p += blockDim.x) {
        int s = 0;
        while (s + 1 < g.th && g.segment_offset[s + 1] <= p) s++;
        const int i = p - g.segment_offset[s];
        prj[((size_t)h * g.S + p) * N + col] = src[g.seg_data[s] + i];
    }
}

// estimate [2S][N] -> pair-column with the back-projection rescale; the
// estimate itself is rescaled in place as back_proj3d2 does (PJ:1496-1507).
__global__ void estimate_to_paircol_kernel(const DevGeom g, float *__restrict__ prj, float *__restrict__ pc) {
    const int p = blockIdx.x, h = blockIdx.y;
    int s = 0;
    while (s + 1 < g.th && g.segment_offset[s + 1] <= p) s++;
    const int i = p - g.segment_offset[s];
    const float r = g.bp_rescale[s];
    for (int col = threadIdx.x; col < g.N; col += blockDim.x) {
        float *e = prj + ((size_t)h * g.S + p) * g.N + col;
        const float v = *e * r;
        *e = v;
        pc[((size_t)h * g.N + col) * g.spp + g.seg_data[s] + i] = v;
    }
}

// ---------------------------------------------------------------------------
// Elementwise kernels.
// ---------------------------------------------------------------------------
__global__ void ratio_w3_kernel(const DevGeom g, float *__restrict__ est, const float *__restrict__ add,
                                const int16_t *__restrict__ prompt, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float den = est[i] + add[i];
        const int16_t pr = prompt[i];
        if (pr == 0) est[i] = 0.0f;
        else if (den > g.eps3) est[i] = (float)pr / den;
        else est[i] = 1.0f;
    }
}

__global__ void ratio_w012_kernel(const DevGeom g, float *__restrict__ est, const float *__restrict__ data, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float e = est[i];
        est[i] = e > g.eps3 ? data[i] / e : 0.0f;
    }
}

__global__ void update_image_kernel(const DevGeom g, float *__restrict__ image, float *__restrict__ corr,
                                    const float *__restrict__ nf) {
    const int x = blockIdx.x;
    for (int y = g.cyl_lo[x] + threadIdx.y; y < g.cyl_hi[x]; y += blockDim.y)
        for (int z = threadIdx.x; z < g.Z4; z += blockDim.x) {
            const size_t i = ((size_t)x * g.N + y) * g.Z4 + z;
            image[i] = update1(image[i], corr[i], nf[i], g.image_max);
            corr[i] = 0.0f;
        }
}

__global__ void normfac_invert_kernel(const DevGeom g, const float *__restrict__ sens, float *__restrict__ out) {
    const int x = blockIdx.x;
    for (int y = g.cyl_lo[x] + threadIdx.y; y < g.cyl_hi[x]; y += blockDim.y)
        for (int z = threadIdx.x; z < g.Z; z += blockDim.x) {
            const size_t i = ((size_t)x * g.N + y) * g.Z4 + z;
            out[i] = invert1(sens[i], g.eps1);
        }
}

__global__ void init_image_kernel(const DevGeom g, float *__restrict__ image, int z_min, int z_max) {
    const int x = blockIdx.x;
    for (int y = threadIdx.y; y < g.N; y += blockDim.y)
        for (int z = threadIdx.x; z < g.Z4; z += blockDim.x) {
            const bool in = y >= g.cyl_lo[x] && y < g.cyl_hi[x] && z >= z_min && z < z_max;
            image[((size_t)x * g.N + y) * g.Z4 + z] = in ? 1.0f : 0.0f;
        }
}

__global__ void end_iteration_kernel(const DevGeom g, float *__restrict__ image, int z_min, int z_max) {
    const int x = blockIdx.x;
    for (int y = g.cyl_lo[x] + threadIdx.y; y < g.cyl_hi[x]; y += blockDim.y)
        for (int z = z_min + (int)threadIdx.x; z < z_max; z += blockDim.x) {
            float *v = image + ((size_t)x * g.N + y) * g.Z4 + z;
            if ((double)*v < 1e-16) *v = 0.0f;
        }
}

__global__ void fill_kernel(float4 *__restrict__ p, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void fill_u64_kernel(unsigned long long *__restrict__ p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Staging (prepare_osem3dw3 S3:1666-1770, prepare_norm S3:841-1003 with
// norm_ANW S3:796-821, read_flat_* FIO:110-259, sino_rebin FIO:29-74).
// ---------------------------------------------------------------------------
struct StageArgs {
    const int16_t *prompt_f;
    const float *delayed_f, *norm_f, *scatter_f, *atten_f;
    int16_t *prompt;
    float *add, *invnorm;
    const uint8_t *gapmask;   // [th][def_angles][def_elements] when rebinning the norm
    int *colmask;             // [views][N]
    const int16_t *plane_seg; // [S]
    // UW/AW/ANW staging (weighting 0, 1, 2; 3 = OP above): prompt_f holds short prompts or
    // trues, dataf_f a float sinogram (fully precorrected when precor); data is the output
    int weighting = 3, precor = 0;
    const float *dataf_f = nullptr;
    float *data = nullptr;
    int p_lo = 0, p_hi = 0;   // internal planes [p_lo, p_hi) staged by this launch
};

__device__ __forceinline__ float rebin_sum(const DevGeom &g, const float *f, size_t fplane, int v, int col) {
    const int DE = g.def_elements;
    if (g.x_rebin * g.v_rebin == 1) return f[fplane * (size_t)g.def_angles * DE + (size_t)v * DE + col];
    // sino_rebin FIO:38-66: views summed first (in place), then radial bins
    const float *p = f + fplane * (size_t)g.def_angles * DE;
    float out = 0.0f;
    for (int xr = 0; xr < g.x_rebin; xr++) {
        float s = p[(size_t)(v * g.v_rebin) * DE + col * g.x_rebin + xr];
        for (int vr = 1; vr < g.v_rebin; vr++) s += p[(size_t)(v * g.v_rebin + vr) * DE + col * g.x_rebin + xr];
        out = xr == 0 ? s : out + s;
    }
    return out;
}

__device__ __forceinline__ float rebin_norm(const DevGeom &g, const float *f, size_t fplane, int v, int col) {
    const int DE = g.def_elements;
    const float *p = f + fplane * (size_t)g.def_angles * DE;
    float out = 0.0f;
    for (int xr = 0; xr < g.x_rebin; xr++) {
        float s = p[(size_t)(v * g.v_rebin) * DE + col * g.x_rebin + xr];
        if (s > 0.0f) s = 1.0f / s;
        for (int vr = 1; vr < g.v_rebin; vr++) {
            const float a = p[(size_t)(v * g.v_rebin + vr) * DE + col * g.x_rebin + xr];
            if (a > 0.0f) s += 1.0f / a;
        }
        out = xr == 0 ? s : out + s;
    }
    if (out > 0.0f) out = 1.0f / out;
    return out;
}

// read_flat_short_proj_ori_to_float_array_theta FIO:185-224 with sino_rebin_float FIO:77-104:
// views summed as shorts (wrapping), then the bins added in float from 0
__device__ __forceinline__ float rebin_short_f(const DevGeom &g, const int16_t *f, size_t fplane, int v, int col) {
    const int DE = g.def_elements;
    const int16_t *p = f + fplane * (size_t)g.def_angles * DE;
    if (g.x_rebin * g.v_rebin == 1) return (float)p[(size_t)v * DE + col];
    float out = 0.0f;
    for (int xr = 0; xr < g.x_rebin; xr++) {
        int16_t s = p[(size_t)(v * g.v_rebin) * DE + col * g.x_rebin + xr];
        for (int vr = 1; vr < g.v_rebin; vr++)
            s = (int16_t)(s + p[(size_t)(v * g.v_rebin + vr) * DE + col * g.x_rebin + xr]);
        out = out + (float)s;
    }
    return out;
}

__device__ __forceinline__ int16_t rebin_short(const DevGeom &g, const int16_t *f, size_t fplane, int v, int col) {
    const int DE = g.def_elements;
    const int16_t *p = f + fplane * (size_t)g.def_angles * DE;
    int16_t out = 0;
    for (int xr = 0; xr < g.x_rebin; xr++) {
        int16_t s = p[(size_t)(v * g.v_rebin) * DE + col * g.x_rebin + xr];
        for (int vr = 1; vr < g.v_rebin; vr++)
            s = (int16_t)(s + p[(size_t)(v * g.v_rebin + vr) * DE + col * g.x_rebin + xr]);
        out = xr == 0 ? s : (int16_t)(out + s);
    }
    return out;
}

// mask of file (view, bin) with a positive norm in some non-killed plane of segment th (FIO:158-160)
__global__ void gapmask_kernel(const DevGeom g, const float *__restrict__ norm_f, uint8_t *__restrict__ gm, int th0) {
    const int th = th0 + blockIdx.y;
    const size_t fp = (size_t)g.def_angles * g.def_elements;
    const int np = g.yt[th] - g.yb[th] + 1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < fp; i += (size_t)gridDim.x * blockDim.x) {
        uint8_t m = 0;
        for (int pl = 0; pl < np; pl++) {
            if (pl < g.npskip || pl + g.npskip >= np) continue;
            if (norm_f[((size_t)g.file_seg[th] + pl) * fp + i] > 0.0f) m = 1;
        }
        gm[th * fp + i] = m;
    }
}

// One element (internal plane p, view v, bin col) of the staging: the corrected data or the
// additive term (o1), the sensitivity weight (o_inv) and, for OP, the prompt (o_pr); flags the
// column when its norm is non-zero.
__device__ __forceinline__ void stage_one(const DevGeom &g, const StageArgs &a, int p, int v, int col, float &o1,
                                          float &o_inv, int16_t &o_pr) {
    const int th = a.plane_seg[p];
    const int i = p - g.segment_offset[th];
    const int np = g.yt[th] - g.yb[th] + 1;
    const size_t fplane = (size_t)g.file_seg[th] + i;
    const int rebin = g.x_rebin * g.v_rebin;
    // norm (read with norm_flag = 1, FIO:110-180)
    float nrm = 1.0f;
    if (a.norm_f) {
        if (i < g.npskip || i + g.npskip >= np) {
            nrm = 0.0f;
        } else if (rebin == 1) {
            nrm = rebin_sum(g, a.norm_f, fplane, v, col);
        } else {
            nrm = rebin_norm(g, a.norm_f, fplane, v, col);
            const size_t fp = (size_t)g.def_angles * g.def_elements;
            int cnt = 0;
            for (int vr = 0; vr < g.v_rebin; vr++)
                for (int xr = 0; xr < g.x_rebin; xr++)
                    cnt += a.gapmask[th * fp + (size_t)(v * g.v_rebin + vr) * g.def_elements + col * g.x_rebin + xr];
            if (cnt < rebin) nrm = 0.0f;
        }
    }
    const float d = a.delayed_f ? rebin_sum(g, a.delayed_f, fplane, v, col) : 0.0f;
    const float at = a.atten_f ? rebin_sum(g, a.atten_f, fplane, v, col) : 1.0f;
    if (nrm != 0.0f) atomicOr(&a.colmask[v * g.N + col], 1);
    if (a.weighting < 3) {
        // prepare_osem3dw0 S3:1198-1256, w1 S3:1346-1407 and w2 S3:1505-1546; precorrected data
        // S3:1309-1323, 1410-1482, 1548-1636; then setnegative2zero simd_operation.cpp:78-96
        float t;
        if (!a.precor) {
            t = a.prompt_f ? rebin_short_f(g, a.prompt_f, fplane, v, col) : rebin_sum(g, a.dataf_f, fplane, v, col);
            if (a.delayed_f) t = t - d;                               // readsinogram S3:1124-1131
            const float sc = a.scatter_f ? rebin_sum(g, a.scatter_f, fplane, v, col) : 0.0f;
            if (a.weighting == 0) {
                if (a.norm_f) {
                    t = t * nrm;                                      // S3:1224, 1244
                    if (a.scatter_f) t = t - nrm;                     // S3:1233: subtracts the norm buffer
                }
                if (a.atten_f) t = t * at;                            // S3:1253
            } else if (a.weighting == 1) {
                if (a.scatter_f) {
                    if (a.norm_f) t = nrm > 0.0f ? t - sc / nrm : 0.0f;   // S3:1380-1382
                    else t = t - sc;                                      // S3:1389
                }
                if (a.norm_f) t = t * nrm;                                // S3:1405
            } else if (a.scatter_f) {
                t = nrm > 0.0f ? t - sc / nrm : 0.0f;                     // S3:1535-1537
            }
        } else {
            t = rebin_sum(g, a.dataf_f, fplane, v, col);
            if (a.weighting == 2) t = nrm > 0.0f ? t / nrm : 0.0f;        // S3:1564-1566
            if (a.weighting >= 1 && a.atten_f && at > 0.0f) t = t / at;  // divideprojection simd_operation.cpp:34
        }
        o1 = t > 0.0f ? t : 0.0f;
        float inv;
        if (a.weighting == 0) {                                           // norm_UW S3:756-771
            inv = nrm > 0.0f ? 1.0f : nrm;
        } else if (a.weighting == 1) {                                    // norm_AW S3:774-791
            inv = nrm > 0.0f ? 1.0f / (at < 1.0f ? 1.0f : at) : 0.0f;
        } else {                                                          // norm_ANW S3:796-821
            const float am = at < 1.0f ? 1.0f : at;
            if (nrm > 0.0f) {
                const float den = nrm * am;
                inv = den > g.eps2 ? 1.0f / den : 0.0f;
            } else {
                inv = nrm;
            }
        }
        o_inv = inv;
        return;
    }
    int16_t pr = 0;
    if (a.prompt_f) pr = rebin_short(g, a.prompt_f, fplane, v, col);
    float ad = d * nrm;                                     // mulprojection S3:1715
    if (nrm == 0.0f) pr = 0;                                // S3:1720-1724
    if (a.scatter_f) ad = ad + rebin_sum(g, a.scatter_f, fplane, v, col);  // addprojection S3:1753
    ad = ad * at;                                           // mulprojection S3:1769
    float inv;                                              // norm_ANW S3:809-817
    const float am = at < 1.0f ? 1.0f : at;
    if (nrm > 0.0f) {
        const float den = nrm * am;
        inv = den > g.eps2 ? 1.0f / den : 0.0f;
    } else {
        inv = nrm;
    }
    o_pr = pr;
    o1 = ad;
    o_inv = inv;
}

// Staging of internal planes [p_lo, p_hi): a workgroup computes a tile of 32 planes x 64 bins of
// one view in the files' layout (bins contiguous, coalesced reads), turns it in LDS, and writes
// it in the pair-column layout (planes contiguous, coalesced writes).
__global__ void __launch_bounds__(256) stage_kernel(const DevGeom g, StageArgs a) {
    __shared__ float t_o1[64][33], t_inv[64][33];
    __shared__ int16_t t_pr[64][34];
    const int N = g.N, V = g.views;
    const int p0 = a.p_lo + 32 * (int)blockIdx.x, v = blockIdx.y, c0 = 64 * (int)blockIdx.z;
    for (int e = threadIdx.x; e < 32 * 64; e += blockDim.x) {
        const int pl = e >> 6, cl = e & 63, p = p0 + pl, col = c0 + cl;
        if (p >= a.p_hi || col >= N) continue;
        float o1 = 0.0f, inv = 0.0f;
        int16_t pr = 0;
        stage_one(g, a, p, v, col, o1, inv, pr);
        t_o1[cl][pl] = o1;
        t_inv[cl][pl] = inv;
        t_pr[cl][pl] = pr;
    }
    __syncthreads();
    const int pair = v % (V / 2), h = v / (V / 2);
    for (int e = threadIdx.x; e < 32 * 64; e += blockDim.x) {
        const int cl = e >> 5, pl = e & 31, p = p0 + pl, col = c0 + cl;
        if (p >= a.p_hi || col >= N) continue;
        const int th = a.plane_seg[p];
        const size_t o = ((size_t)(pair * 2 + h) * N + col) * g.spp + g.seg_data[th] + (p - g.segment_offset[th]);
        if (a.weighting < 3) {
            a.data[o] = t_o1[cl][pl];
        } else {
            a.prompt[o] = t_pr[cl][pl];
            a.add[o] = t_o1[cl][pl];
        }
        a.invnorm[o] = t_inv[cl][pl];
    }
}

// norm_zero_mask S3:899-909
__global__ void nzmask_kernel(const DevGeom g, const int *__restrict__ colmask, int8_t *__restrict__ nz) {
    const int v = blockIdx.x, N = g.N, V = g.views;
    for (int xr = threadIdx.x; xr <= N / 2; xr += blockDim.x) {
        int8_t m = 1;
        if (xr >= 1) {
            const bool z = colmask[v * N + xr] == 0 && colmask[(v + V / 2) * N + xr] == 0 &&
                           colmask[v * N + N - xr] == 0 && colmask[(v + V / 2) * N + N - xr] == 0;
            m = z ? 0 : 1;
        }
        nz[v * (N / 2 + 1) + xr] = m;
    }
}

// ---------------------------------------------------------------------------
// FP64 reductions (log-likelihood, count normalisation) and FP16 conversion.
// Every sum runs in a fixed order, so results are reproducible run to run.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum_fixed(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    return t;   // valid in thread 0
}

__global__ void __launch_bounds__(256) ll_reduce_kernel(const double *__restrict__ part, int per, double *__restrict__ out) {
    __shared__ double red[4];
    const double *p = part + (size_t)blockIdx.x * per;
    double v = 0.0;
    for (int i = threadIdx.x; i < per; i += blockDim.x) v += p[i];
    const double t = block_sum_fixed(v, red);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

__global__ void to_half_kernel(const float *__restrict__ in, uint16_t *__restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = __half_as_ushort(__float2half_rn(in[i]));
}

// The same conversion, counting the finite inputs that become +-inf ([0]) and the non-zero
// inputs that become FP16 subnormals or zero ([1]); one atomic per wave and counter.
__global__ void to_half_range_kernel(const float *__restrict__ in, uint16_t *__restrict__ out, size_t n,
                                     unsigned long long *__restrict__ counts) {
    unsigned long long over = 0, small = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = in[i];
        const __half h = __float2half_rn(x);
        const uint16_t b = __half_as_ushort(h);
        out[i] = b;
        const bool finite_in = isfinite(x);
        over += (finite_in && (b & 0x7fffu) == 0x7c00u) ? 1 : 0;   // exponent all ones, mantissa 0
        small += (x != 0.0f && finite_in && (b & 0x7c00u) == 0) ? 1 : 0;   // exponent 0: subnormal or zero
    }
    for (int o = 32; o > 0; o >>= 1) {
        over += __shfl_xor(over, o);
        small += __shfl_xor(small, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (over) atomicAdd(&counts[0], over);
        if (small) atomicAdd(&counts[1], small);
    }
}

// moments of one pair-column block (pair p, half h, column col): sum prompt, sum w*add, sum w*est
__global__ void __launch_bounds__(256) count_moments_kernel(const DevGeom g, const int16_t *__restrict__ prompt,
                                                            const float *__restrict__ add,
                                                            const float *__restrict__ invnorm,
                                                            const float *__restrict__ est, int view0,
                                                            int view_stride, double *__restrict__ part) {
    __shared__ double red[4];
    const int col = blockIdx.x, h = blockIdx.y, lp = blockIdx.z;
    const int view = view0 + view_stride * lp;
    const size_t gbase = ((size_t)(view * 2 + h) * g.N + col) * g.spp;
    const size_t ebase = ((size_t)(lp * 2 + h) * g.N + col) * g.spp;
    double sy = 0.0, sa = 0.0, se = 0.0;
    for (int i = threadIdx.x; i < g.spp; i += blockDim.x) {
        const double w = invnorm[gbase + i];
        sy += prompt[gbase + i];
        sa += w * add[gbase + i];
        se += w * est[ebase + i];
    }
    const size_t bi = ((size_t)lp * 2 + h) * g.N + col;
    const size_t nb = (size_t)gridDim.x * gridDim.y * gridDim.z;
    double t = block_sum_fixed(sy, red);
    if (threadIdx.x == 0) part[bi] = t;
    t = block_sum_fixed(sa, red);
    if (threadIdx.x == 0) part[nb + bi] = t;
    t = block_sum_fixed(se, red);
    if (threadIdx.x == 0) part[2 * nb + bi] = t;
}

__global__ void __launch_bounds__(256) moments_reduce_kernel(const double *__restrict__ part, int nb,
                                                             double *__restrict__ out) {
    __shared__ double red[4];
    const double *p = part + (size_t)blockIdx.x * nb;
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v += p[i];
    const double t = block_sum_fixed(v, red);
    if (threadIdx.x == 0) out[blockIdx.x] = out[blockIdx.x] + t;   // accumulates over calls
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int MAXT, int ITEMS, bool FAST>
static hipError_t fp_launch_t(const DevGeom &g, const PairBatch &b, const float *image, float *out, const FpIO &io,
                              int mode, int nsplit, int threads, hipStream_t s) {
    dim3 grid(g.c * nsplit * b.npairs * b.nframes), block(threads);
    size_t shm = sizeof(float4) * 5 * (FAST ? FP_PL_FAST : g.Z4 + 8);
    const float4 *im = reinterpret_cast<const float4 *>(image);
    if (mode == FP_RATIO)
        hipLaunchKernelGGL((fp_kernel<MAXT, ITEMS, FP_RATIO, FAST>), grid, block, shm, s, g, b, im, out, io, nsplit);
    else if (mode == FP_RATIO_W012)
        hipLaunchKernelGGL((fp_kernel<MAXT, ITEMS, FP_RATIO_W012, FAST>), grid, block, shm, s, g, b, im, out, io, nsplit);
    else if (mode == FP_RATIO16)
        hipLaunchKernelGGL((fp_kernel<MAXT, ITEMS, FP_RATIO16, FAST>), grid, block, shm, s, g, b, im, out, io, nsplit);
    else
        hipLaunchKernelGGL((fp_kernel<MAXT, ITEMS, FP_PLAIN, FAST>), grid, block, shm, s, g, b, im, out, io, nsplit);
    return hipGetLastError();
}

// Launch shape: MAXT planes per thread; the planes of one ray column are split over
// nsplit workgroups; each workgroup has enough 64-lane waves (a multiple of 4, at most
// 16) for its share of the planes and for the sampling of a row.
void forward_shape(const DevGeom &g, int &maxt, int &nsplit, int &threads, int &items) {
    maxt = g.fp_maxt;
    nsplit = g.fp_nsplit;
    if (maxt < 1 || maxt > 4) maxt = 1;
    const int need = (g.fp_np + 1024 * maxt - 1) / (1024 * maxt);
    if (nsplit < need) nsplit = need;
    const int per = (g.fp_np + nsplit - 1) / nsplit;
    // waves in multiples of 4, one per SIMD each: with one workgroup per CU and a barrier per
    // row, a SIMD holding one wave more than the others would set the pace of every row
    int waves = ((per + maxt - 1) / maxt + 63) / 64;
    waves = (waves + 3) / 4 * 4;
    if (waves > 16) waves = 16;
    threads = 64 * waves;
    if ((8 * g.NQ + threads - 1) / threads > 2) threads = ((8 * g.NQ + 1) / 2 + 63) / 64 * 64;
    items = (8 * g.NQ + threads - 1) / threads;
}

hipError_t launch_forward(const DevGeom &g, const PairBatch &b, const float *image, float *out, const FpIO &io,
                          int mode, hipStream_t s) {
    int maxt, nsplit, threads, it2;
    forward_shape(g, maxt, nsplit, threads, it2);
#define FP_CASE(M)                                                                                    \
    case M:                                                                                           \
        return it2 <= 1 ? fp_launch_t<M, 1, false>(g, b, image, out, io, mode, nsplit, threads, s)    \
                        : fp_launch_t<M, 2, false>(g, b, image, out, io, mode, nsplit, threads, s);
    // the FAST path stages a row as 2 x ceil(NQ/16) whole waves of samples, one item each
    if (maxt == 4 && g.fp_fast && g.Z4 + 8 == FP_PL_FAST && 128 * ((g.NQ + 15) / 16) <= threads)
        return fp_launch_t<4, 1, true>(g, b, image, out, io, mode, nsplit, threads, s);
    switch (maxt) {
        FP_CASE(1)
        FP_CASE(2)
        FP_CASE(3)
        default:
            FP_CASE(4)
    }
#undef FP_CASE
}

hipError_t launch_backproject(const DevGeom &g, const PairBatch &b, const void *src, int src_half, int src_global,
                              int apply_rescale, int kahan, float *bufA, float *bufB, hipStream_t s) {
    dim3 grid(g.N - 1, 2, b.npairs * b.nframes), block(64 * g.bp_waves);
    size_t shm = sizeof(float2) * g.bp_lds;
    const float *sf = static_cast<const float *>(src);
    const uint16_t *sh = static_cast<const uint16_t *>(src);
    if (src_half) {
        if (kahan) hipLaunchKernelGGL((bp_kernel<uint16_t, true>), grid, block, shm, s, g, b, sh, src_global, apply_rescale, g.pseg, bufA, bufB);
        else hipLaunchKernelGGL((bp_kernel<uint16_t, false>), grid, block, shm, s, g, b, sh, src_global, apply_rescale, g.pseg, bufA, bufB);
    } else {
        if (kahan) hipLaunchKernelGGL((bp_kernel<float, true>), grid, block, shm, s, g, b, sf, src_global, apply_rescale, g.pseg, bufA, bufB);
        else hipLaunchKernelGGL((bp_kernel<float, false>), grid, block, shm, s, g, b, sf, src_global, apply_rescale, g.pseg, bufA, bufB);
    }
    return hipGetLastError();
}

static dim3 rot_grid(const DevGeom &g, int nframes = 1) { return dim3(g.N * ((g.N + 3) / 4), nframes); }

hipError_t launch_backproject_incremental(const DevGeom &g, const PairBatch &b, const float *src, int src_global,
                                          int apply_rescale, float *bufA, float *bufB, hipStream_t s) {
    if (!g.inc_ok) return hipErrorInvalidValue;   // tables failed the host's window checks
    dim3 grid(g.c, 4, b.npairs * b.nframes), block(256);
    hipLaunchKernelGGL(bp_inc_kernel, grid, block, g.inc_shm, s, g, b, src, src_global, apply_rescale,
                       g.pseg, bufA, bufB);
    return hipGetLastError();
}

hipError_t launch_rotate_update(const DevGeom &g, const PairBatch &b, const float *bufA, const float *bufB,
                                float *image, const float *normfac, int kahan, hipStream_t s) {
    const float4 *A = reinterpret_cast<const float4 *>(bufA), *B = reinterpret_cast<const float4 *>(bufB);
    float4 *im = reinterpret_cast<float4 *>(image);
    const float4 *nf = reinterpret_cast<const float4 *>(normfac);
    if (kahan) hipLaunchKernelGGL(rot_update_kernel<true>, rot_grid(g, b.nframes), dim3(64, 4), 0, s, g, b, A, B, im, nf);
    else hipLaunchKernelGGL(rot_update_kernel<false>, rot_grid(g, b.nframes), dim3(64, 4), 0, s, g, b, A, B, im, nf);
    return hipGetLastError();
}

hipError_t launch_rotate_sens(const DevGeom &g, const PairBatch &b, const float *bufA, const float *bufB,
                              float *normfac_out, int kahan, hipStream_t s) {
    size_t n4 = (size_t)g.N * g.N * g.NQ;
    hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, s, reinterpret_cast<float4 *>(normfac_out), n4);
    const float4 *A = reinterpret_cast<const float4 *>(bufA), *B = reinterpret_cast<const float4 *>(bufB);
    float4 *o = reinterpret_cast<float4 *>(normfac_out);
    if (kahan) hipLaunchKernelGGL(rot_sens_kernel<true>, rot_grid(g), dim3(64, 4), 0, s, g, b, A, B, o);
    else hipLaunchKernelGGL(rot_sens_kernel<false>, rot_grid(g), dim3(64, 4), 0, s, g, b, A, B, o);
    return hipGetLastError();
}

hipError_t launch_ll_reduce(const double *part, int per_frame, int nframes, double *out, hipStream_t s) {
    hipLaunchKernelGGL(ll_reduce_kernel, dim3(nframes), dim3(256), 0, s, part, per_frame, out);
    return hipGetLastError();
}

hipError_t launch_to_half(const float *in, uint16_t *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(to_half_kernel, dim3(2048), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_to_half_range(const float *in, uint16_t *out, size_t n, unsigned long long *counts, hipStream_t s) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, counts, 2);
    hipLaunchKernelGGL(to_half_range_kernel, dim3(2048), dim3(256), 0, s, in, out, n, counts);
    return hipGetLastError();
}

hipError_t launch_count_moments(const DevGeom &g, const int16_t *prompt, const float *add, const float *invnorm,
                                const float *est, int view0, int view_stride, int npairs, double *part,
                                double *out, hipStream_t s) {
    dim3 grid(g.N, 2, npairs);
    hipLaunchKernelGGL(count_moments_kernel, grid, dim3(256), 0, s, g, prompt, add, invnorm, est, view0, view_stride,
                       part);
    hipLaunchKernelGGL(moments_reduce_kernel, dim3(3), dim3(256), 0, s, part, (int)(grid.x * grid.y * grid.z), out);
    return hipGetLastError();
}

hipError_t launch_halves_sum(const DevGeom &g, const float *bufA, const float *bufB, float *out, hipStream_t s) {
    hipLaunchKernelGGL(halves_sum_kernel, rot_grid(g), dim3(64, 4), 0, s, g, reinterpret_cast<const float4 *>(bufA),
                       reinterpret_cast<const float4 *>(bufB), reinterpret_cast<float4 *>(out));
    return hipGetLastError();
}

hipError_t launch_rotate_add(const DevGeom &g, int view, const float *bufA, const float *bufB, float *ima,
                             hipStream_t s) {
    hipLaunchKernelGGL(rot_add_kernel, rot_grid(g), dim3(64, 4), 0, s, g, view,
                       reinterpret_cast<const float4 *>(bufA), reinterpret_cast<const float4 *>(bufB),
                       reinterpret_cast<float4 *>(ima));
    return hipGetLastError();
}

hipError_t launch_paircol_to_estimate(const DevGeom &g, int view, const float *paircol, float *prj, hipStream_t s) {
    hipLaunchKernelGGL(paircol_to_estimate_kernel, dim3(g.N, 2), dim3(256), 0, s, g, view, paircol, prj);
    return hipGetLastError();
}

hipError_t launch_estimate_to_paircol(const DevGeom &g, float *prj, float *paircol, hipStream_t s) {
    hipLaunchKernelGGL(estimate_to_paircol_kernel, dim3(g.S, 2), dim3(256), 0, s, g, prj, paircol);
    return hipGetLastError();
}

hipError_t launch_ratio_w3(const DevGeom &g, float *est, const float *add, const int16_t *prompt, hipStream_t s) {
    size_t n = (size_t)2 * g.S * g.N;
    hipLaunchKernelGGL(ratio_w3_kernel, dim3(1024), dim3(256), 0, s, g, est, add, prompt, n);
    return hipGetLastError();
}

hipError_t launch_ratio_w012(const DevGeom &g, float *est, const float *data, hipStream_t s) {
    const size_t n = (size_t)2 * g.S * g.N;
    hipLaunchKernelGGL(ratio_w012_kernel, dim3(1024), dim3(256), 0, s, g, est, data, n);
    return hipGetLastError();
}

hipError_t launch_update_image(const DevGeom &g, float *image, float *corr, const float *nf, hipStream_t s) {
    hipLaunchKernelGGL(update_image_kernel, dim3(g.N), dim3(64, 4), 0, s, g, image, corr, nf);
    return hipGetLastError();
}

hipError_t launch_rotate_sum(const DevGeom &g, const PairBatch &b, const float *bufA, const float *bufB, float *corr,
                             hipStream_t s) {
    hipLaunchKernelGGL(rot_sum_kernel, rot_grid(g), dim3(64, 4), 0, s, g, b, reinterpret_cast<const float4 *>(bufA),
                       reinterpret_cast<const float4 *>(bufB), reinterpret_cast<float4 *>(corr));
    return hipGetLastError();
}

hipError_t launch_convolve3d(const DevGeom &g, const float *in, float *out, float *buf, const float *taps_x, int hx,
                             const float *taps_y, int hy, const float *normz, int z_min, int z_max, hipStream_t s) {
    if (g.Z4 > 256 || g.NQ > 64 || z_min < 0 || z_max > g.Z) return hipErrorInvalidValue;   // one z column per wave
    hipLaunchKernelGGL(psf_x_kernel, rot_grid(g), dim3(64, 4), 0, s, g, reinterpret_cast<const float4 *>(in),
                       reinterpret_cast<float4 *>(buf), taps_x, hx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(psf_yz_kernel, rot_grid(g), dim3(64, 4), 0, s, g, reinterpret_cast<const float4 *>(buf), out,
                       taps_y, hy, normz, z_min, z_max);
    return hipGetLastError();
}

hipError_t launch_flat_image(const DevGeom &g, const float *src, float *dst, int to_flat, int flip, hipStream_t s) {
    dim3 grid((g.N + 63) / 64, g.N, (g.Z4 + 63) / 64), block(256);
    if (to_flat) hipLaunchKernelGGL(flat_image_kernel<true>, grid, block, 0, s, g, src, dst, flip);
    else hipLaunchKernelGGL(flat_image_kernel<false>, grid, block, 0, s, g, src, dst, flip);
    return hipGetLastError();
}

hipError_t launch_normfac_invert(const DevGeom &g, const float *sens, float *out, hipStream_t s) {
    hipLaunchKernelGGL(normfac_invert_kernel, dim3(g.N), dim3(64, 4), 0, s, g, sens, out);
    return hipGetLastError();
}

hipError_t launch_init_image(const DevGeom &g, float *image, int z_min, int z_max, hipStream_t s) {
    hipLaunchKernelGGL(init_image_kernel, dim3(g.N), dim3(64, 4), 0, s, g, image, z_min, z_max);
    return hipGetLastError();
}

hipError_t launch_end_iteration(const DevGeom &g, float *image, int z_min, int z_max, hipStream_t s) {
    hipLaunchKernelGGL(end_iteration_kernel, dim3(g.N), dim3(64, 4), 0, s, g, image, z_min, z_max);
    return hipGetLastError();
}

// Segments [th0, th0 + nth) with internal planes [p_lo, p_hi): their gap masks first when the
// norm is rebinned, then their planes; finish derives the zero-column mask from the column
// flags of every segment staged so far.
static void stage_launch(const DevGeom &g, StageArgs a, int th0, int nth, int p_lo, int p_hi, bool finish,
                         hipStream_t s) {
    a.p_lo = p_lo;
    a.p_hi = p_hi;
    if (a.norm_f && g.x_rebin * g.v_rebin > 1 && nth > 0)
        hipLaunchKernelGGL(gapmask_kernel, dim3(256, nth), dim3(256), 0, s, g, a.norm_f, const_cast<uint8_t *>(a.gapmask), th0);
    if (p_hi > p_lo)
        hipLaunchKernelGGL(stage_kernel, dim3((p_hi - p_lo + 31) / 32, g.views, (g.N + 63) / 64), dim3(256), 0, s, g, a);
    if (finish) hipLaunchKernelGGL(nzmask_kernel, dim3(g.views / 2), dim3(256), 0, s, g, a.colmask, g.nzmask);
}

hipError_t launch_stage(const DevGeom &g, const int16_t *prompt_f, const float *delayed_f, const float *norm_f,
                        const float *scatter_f, const float *atten_f, int16_t *prompt, float *add, float *invnorm,
                        uint8_t *gapmask, int *colmask, int th0, int nth, int p_lo, int p_hi, bool finish,
                        hipStream_t s) {
    StageArgs a{prompt_f, delayed_f, norm_f, scatter_f, atten_f, prompt, add, invnorm, gapmask, colmask, g.plane_seg};
    stage_launch(g, a, th0, nth, p_lo, p_hi, finish, s);
    return hipGetLastError();
}

hipError_t launch_stage_w012(const DevGeom &g, int weighting, int precor, const int16_t *data_f16,
                             const float *data_ff, const float *delayed_f, const float *norm_f, const float *scatter_f,
                             const float *atten_f, float *data, float *invnorm, uint8_t *gapmask, int *colmask,
                             int th0, int nth, int p_lo, int p_hi, bool finish, hipStream_t s) {
    StageArgs a{data_f16, delayed_f, norm_f, scatter_f, atten_f, nullptr, nullptr, invnorm, gapmask, colmask,
                g.plane_seg};
    a.weighting = weighting;
    a.precor = precor;
    a.dataf_f = data_ff;
    a.data = data;
    stage_launch(g, a, th0, nth, p_lo, p_hi, finish, s);
    return hipGetLastError();
}

}  // namespace hrrt
