// capi.hip -- C ABI of the OSEM hot path (include/hrrt_osem.h): context
// creation (geometry tables -> device), scratch management and the drivers
// that chain the kernels of kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hrrt_osem.h"
#include "geometry.h"
#include "kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(HRRT_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// Scratch of hrrt_forward_proj3d_noalloc: one pair-column ([2][N][spp] floats) in a
// file-scope device array, so the entry point allocates nothing.  Every GPU holds its own
// copy for the life of the process.  In host code `scratch` only names the symbol: its
// device address comes from hipGetSymbolAddress, taken once per device when a context is
// created on it (kept in g_scratch_ptr), and only that address is passed to kernels.
static __device__ __align__(256) float scratch[HRRT_NOALLOC_SCRATCH_FLOATS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// Device address of `scratch` on the current device (looked up once per device).
hipError_t scratch_address(int dev, float **out) {
    *out = nullptr;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev] && g_scratch_err[dev] == hipSuccess) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) g_scratch_err[dev] = e;
        else g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return g_scratch_ptr[dev] ? hipSuccess : g_scratch_err[dev];
}
}  // namespace

struct hrrt_ctx {
    hrrt::Geometry geo;
    hrrt::DevGeom dg{};
    std::vector<void *> allocs;      // geometry tables
    // scratch (allocated on first use, zero-initialised once)
    float *sub_ratio = nullptr;      // [sviews/2][2][N][spp]
    float *bufA = nullptr, *bufB = nullptr;   // [sviews/2][N][N][Z4]
    float *pc1 = nullptr;            // [2][N][spp]
    int *colmask = nullptr;          // [views][N]
    uint8_t *gapmask = nullptr;      // [th][def_angles][def_elements]
    float *h_img = nullptr, *h_prj = nullptr, *h_buf = nullptr;  // device staging for the host shims
    float *rot_zero = nullptr;       // a zero rotated-frame half, for rotating a caller's buffer
    int device = 0;
    float *noalloc_scratch = nullptr;   // this device's address of `scratch` (NULL: lookup failed)
    hipError_t noalloc_err = hipSuccess;
    int bp_threads = 0;              // back-projection schedule (the reference's -T): 0 = direct for every pair
    // PSF resolution model (-B): host kernels (x, y, z), normz, their device copies, and three
    // scratch volumes (the blurred image, the correction, the x-pass buffer), zeroed once so
    // that voxels outside the FOV, which the blur never writes, read as zero
    bool psf_on = false;
    std::vector<float> psf_k[3], psf_normz;
    float *d_psf = nullptr;          // [taps x][taps y][normz]
    float *psf_img = nullptr, *psf_corr = nullptr, *psf_buf = nullptr;
    float *flat = nullptr;           // [z_pixels][N][N] staging of the flat image files
    int ensure_psf_scratch() {
        int r;
        if ((r = zalloc((void **)&psf_img, sizeof(float) * vol_floats()))) return r;
        if ((r = zalloc((void **)&psf_corr, sizeof(float) * vol_floats()))) return r;
        if ((r = zalloc((void **)&psf_buf, sizeof(float) * vol_floats()))) return r;
        return HRRT_OK;
    }
    hipError_t convolve3d(const float *in, float *out, hipStream_t s) const {
        const int hx = ((int)psf_k[0].size() - 1) / 2, hy = ((int)psf_k[1].size() - 1) / 2;
        return hrrt::launch_convolve3d(dg, in, out, psf_buf, d_psf, hx, d_psf + psf_k[0].size(), hy,
                                       d_psf + psf_k[0].size() + psf_k[1].size(), geo.z_min, geo.z_max, s);
    }
    // optional per-kernel timing of hrrt_osem_subset (HIP events on the launch stream)
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_marks;   // (kernel id, index of start event)

    hipEvent_t next_event() {
        if (ev_used == ev_pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ev_pool.push_back(e);
        }
        return ev_pool[ev_used++];
    }
    void mark(int id, hipStream_t s) {
        if (!timing) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;   // eager launches only
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
            (void)hipGetLastError();
            return;
        }
        size_t i = ev_used;
        hipEvent_t e = next_event();
        if (e) {
            (void)hipEventRecord(e, s);
            if (id >= 0) ev_marks.push_back({id, i});
        }
    }

    template <typename T>
    int upload(const std::vector<T> &v, const T *&dst) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(T) * (v.empty() ? 1 : v.size())) != hipSuccess)
            return fail(HRRT_ERR_NOMEM, "hipMalloc of a geometry table failed");
        allocs.push_back(p);
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        dst = static_cast<const T *>(p);
        return HRRT_OK;
    }
    int zalloc(void **p, size_t bytes) {
        if (*p) return HRRT_OK;
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return fail(HRRT_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
        }
        HIP_TRY(hipMemset(*p, 0, bytes));
        return HRRT_OK;
    }
    size_t vol_floats() const { return (size_t)geo.N * geo.N * geo.z4; }
    size_t col_floats() const { return (size_t)2 * geo.N * geo.spp; }
    int npairs() const { return geo.sviews / 2; }
    int ensure_subset_scratch() {
        int r;
        if ((r = zalloc((void **)&sub_ratio, sizeof(float) * col_floats() * npairs()))) return r;
        if ((r = zalloc((void **)&bufA, sizeof(float) * vol_floats() * npairs()))) return r;
        if ((r = zalloc((void **)&bufB, sizeof(float) * vol_floats() * npairs()))) return r;
        return HRRT_OK;
    }
    int ensure_pair_scratch() {
        int r;
        if ((r = zalloc((void **)&pc1, sizeof(float) * col_floats()))) return r;
        if ((r = zalloc((void **)&bufA, sizeof(float) * vol_floats() * (npairs() > 0 ? npairs() : 1)))) return r;
        if ((r = zalloc((void **)&bufB, sizeof(float) * vol_floats() * (npairs() > 0 ? npairs() : 1)))) return r;
        return HRRT_OK;
    }
    // batched dynamic frames: per-frame scratch (zeroed once: voxels outside the FOV are
    // never written and must read as zero) and one instantiated graph per subset
    float *f_ratio = nullptr, *fA = nullptr, *fB = nullptr;
    int f_cap = 0;
    struct GraphKey {
        int nframes = 0;
        const void *img = nullptr, *nf = nullptr, *prompt = nullptr, *add = nullptr;
        void *stream = nullptr;
        bool mixed = false;
        const void *inv = nullptr, *ll = nullptr;
        bool operator==(const GraphKey &o) const {
            return nframes == o.nframes && img == o.img && nf == o.nf && prompt == o.prompt && add == o.add &&
                   stream == o.stream && mixed == o.mixed && inv == o.inv && ll == o.ll;
        }
    } gkey;
    double *ll_part = nullptr;           // FP64 partial sums (log-likelihood, count moments)
    size_t ll_cap = 0;
    int ensure_ll_part(size_t n) {
        if (n <= ll_cap) return HRRT_OK;
        if (ll_part) (void)hipFree(ll_part);
        ll_part = nullptr;
        ll_cap = 0;
        drop_graphs();
        int r;
        if ((r = zalloc((void **)&ll_part, sizeof(double) * n))) return r;
        ll_cap = n;
        return HRRT_OK;
    }
    std::vector<hipGraphExec_t> gexec;   // [subsets]
    int ensure_frame_scratch(int nframes) {
        if (nframes <= f_cap) return HRRT_OK;
        // instantiated graphs hold the old buffers in their kernel arguments
        drop_graphs();
        void *s[] = {f_ratio, fA, fB};
        for (void *p : s)
            if (p) (void)hipFree(p);
        f_ratio = fA = fB = nullptr;
        f_cap = 0;
        int r;
        const size_t nr = col_floats() * npairs() * nframes, nb = vol_floats() * npairs() * nframes;
        if ((r = zalloc((void **)&f_ratio, sizeof(float) * nr))) return r;
        if ((r = zalloc((void **)&fA, sizeof(float) * nb))) return r;
        if ((r = zalloc((void **)&fB, sizeof(float) * nb))) return r;
        f_cap = nframes;
        return HRRT_OK;
    }
    void drop_graphs() {
        for (hipGraphExec_t e : gexec)
            if (e) (void)hipGraphExecDestroy(e);
        gexec.clear();
        gkey = GraphKey{};
    }
    // staging from host memory: device copies of the flat files, a pinned bounce ring, a copy
    // stream, and the events that order them against the staging launches
    static constexpr int NRING = 4;
    static constexpr size_t RING_BYTES = 16u << 20;
    void *hf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // per file kind, 4 bytes per file bin
    std::vector<int> file_seg;                                     // [th] first file plane of segment th
    void *ring[NRING] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ring_ev[NRING] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_start = nullptr;
    std::vector<hipEvent_t> seg_ev;
    int ensure_host_staging() {
        if (!copy_stream) HIP_TRY(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
        if (!copy_start) HIP_TRY(hipEventCreateWithFlags(&copy_start, hipEventDisableTiming));
        for (int i = 0; i < NRING; i++) {
            if (!ring[i] && hipHostMalloc(&ring[i], RING_BYTES, hipHostMallocDefault) != hipSuccess) {
                ring[i] = nullptr;
                return fail(HRRT_ERR_NOMEM, "hipHostMalloc(%zu) of a staging buffer failed", RING_BYTES);
            }
            if (!ring_ev[i]) HIP_TRY(hipEventCreateWithFlags(&ring_ev[i], hipEventDisableTiming));
        }
        while ((int)seg_ev.size() < geo.th_pixels) {
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            seg_ev.push_back(e);
        }
        return HRRT_OK;
    }
    ~hrrt_ctx() {
        drop_graphs();
        if (copy_stream) (void)hipStreamSynchronize(copy_stream);
        for (void *p : hf)
            if (p) (void)hipFree(p);
        for (int i = 0; i < NRING; i++) {
            if (ring[i]) (void)hipHostFree(ring[i]);
            if (ring_ev[i]) (void)hipEventDestroy(ring_ev[i]);
        }
        for (hipEvent_t e : seg_ev) (void)hipEventDestroy(e);
        if (copy_start) (void)hipEventDestroy(copy_start);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        for (void *p : allocs) (void)hipFree(p);
        void *s[] = {sub_ratio, bufA, bufB, pc1, colmask, gapmask, h_img, h_prj, h_buf, rot_zero, (void *)dg.nzmask, f_ratio, fA, fB,
                     ll_part, d_psf, psf_img, psf_corr, psf_buf, flat};
        for (void *p : s)
            if (p) (void)hipFree(p);
    }
};

static inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
static inline int bits(float f) {
    int i;
    std::memcpy(&i, &f, 4);
    return i;
}

extern "C" const char *hrrt_last_error(void) { return g_err.c_str(); }

int hrrt::set_last_error(int code, const char *msg) {
    g_err = msg;
    return code;
}

extern "C" int hrrt_create(const hrrt_params *params, hrrt_ctx **out) {
    if (!params || !out) return fail(HRRT_ERR_ARG, "null argument");
    *out = nullptr;
    auto *ctx = new (std::nothrow) hrrt_ctx();
    if (!ctx) return fail(HRRT_ERR_NOMEM, "host allocation failed");
    std::string err;
    hrrt::Geometry &G = ctx->geo;
    if (!G.build(*params, err)) {
        delete ctx;
        return fail(HRRT_ERR_ARG, "%s", err.c_str());
    }
    (void)hipGetDevice(&ctx->device);
    // the no-alloc entry point's scratch on this device (its failure is reported by that call)
    ctx->noalloc_err = scratch_address(ctx->device, &ctx->noalloc_scratch);
    const int N = G.N, c = N / 2, GM = G.groupmax, ZS = G.zshift, Z = G.z_pixels;
    hrrt::DevGeom &d = ctx->dg;
    d.N = N; d.c = c; d.views = G.views; d.Z = Z; d.Z4 = G.z4; d.NQ = G.z4 / 4; d.S = G.segmentsize;
    d.th = G.th_pixels; d.K = G.K; d.GM = GM; d.ZS = ZS; d.spp = G.spp; d.sviews = G.sviews;
    d.subsets = G.p.subsets; d.x_rebin = G.x_rebin; d.v_rebin = G.v_rebin; d.def_elements = G.p.def_elements;
    d.def_angles = G.p.def_angles; d.npskip = G.p.npskip; d.y_size = G.y_size; d.eps1 = G.p.epsilon1;
    d.eps2 = G.p.epsilon2; d.eps3 = G.p.epsilon3; d.image_max = G.p.image_max;

    // forward-projector interpolation tables (flookup PJ:544-562 and the middle-plane
    // positions computed inline at PJ:1929-1940)
    std::vector<int4> mid((size_t)N * GM), endp((size_t)N * GM * 2 * ZS);
    std::vector<int> skip((size_t)N * GM, 1);
    for (int y = 1; y < N; y++) {
        for (int k = 0; k < GM; k++) {
            const int th = 2 * k + 2, ym = G.yr_bottom[th] + ZS;
            const float t = G.tantheta[k * Z + ym];
            float zz = ym + (y - c) * t;
            int z = (int)zz;
            float c1 = zz - z;
            zz = ym + (c - y) * t;
            int z2 = (int)zz;
            float c2 = zz - z2;
            mid[(size_t)y * GM + k] = make_int4(z, z2, bits(c1), bits(c2));
            const size_t a = ((size_t)y * GM + k) * 2 * ZS, bb = ((size_t)(N - y) * GM + k) * 2 * ZS;
            skip[(size_t)y * GM + k] = (G.fl_z[a] < 0 || G.fl_z[bb + 2 * ZS - 1] > Z - 1) ? 1 : 0;
            for (int i = 0; i < 2 * ZS; i++)
                endp[a + i] = make_int4(G.fl_z[a + i], bits(G.fl_frac[a + i]), G.fl_z[bb + i], bits(G.fl_frac[bb + i]));
        }
    }
    // every sample index the projectors will touch must lie inside the staged
    // columns: forward z, z+1 in [0, Z] (plane Z is the zero pad), and equal
    // lengths for the paired back-projection lookups (PJ:2986-2989)
    for (int y = 1; y < N; y++) {
        for (int k = 0; k < GM; k++) {
            const size_t li = (size_t)y * GM + k;
            const int th = 2 * k + 2;
            if (!skip[li] && y < c) {
                const int4 m = mid[li];
                const int nmid = G.yr_top[th] - ZS - (G.yr_bottom[th] + ZS);
                int lo = std::min(m.x, m.y), hi = std::max(m.x, m.y) + nmid + 1;
                for (int i = 0; i < 2 * ZS; i++) {
                    const int4 e = endp[li * 2 * ZS + i];
                    lo = std::min(lo, std::min(e.x, e.z));
                    hi = std::max(hi, std::max(e.x, e.z) + 1);
                }
                if (lo < 0 || hi > Z) {
                    delete ctx;
                    return fail(HRRT_ERR_ARG, "forward interpolation leaves the axial range (y=%d, group=%d)", y, k + 1);
                }
            }
            const size_t a = li, b2 = (size_t)(N - y) * GM + k;
            if (G.bl_zea[a] - G.bl_zsa[a] != G.bl_zea[b2] - G.bl_zsa[b2]) {
                delete ctx;
                return fail(HRRT_ERR_ARG, "unpaired back-projection lookup lengths (y=%d, group=%d)", y, k + 1);
            }
        }
    }
    std::vector<int4> blut((size_t)N * GM);
    for (size_t i = 0; i < blut.size(); i++)
        blut[i] = make_int4(G.bl_zsa[i], G.bl_zea[i], G.bl_yra[i] - G.bl_zsa[i], bits(G.bl_frac[i]));
    // rotation sample positions, accumulated along y exactly as rotateimage2 does (PJ:1148-1184)
    std::vector<float2> rot((size_t)(G.views / 2) * N * N, make_float2(-1.f, -1.f));
    for (int v = 0; v < G.views / 2; v++) {
        const float cosv = G.cos_psi[v] / 1.0f, sinv = -G.sin_psi[v] / 1.0f;
        const float xxx0 = -N / 2 * cosv + N / 2 * sinv + N / 2;
        const float yyy0 = -N / 2 * sinv - N / 2 * cosv + N / 2;
        for (int x = 0; x < N; x++) {
            const int ys = G.cyl_lo[x], ye = G.cyl_hi[x];
            float xx0 = xxx0 - ys * sinv + (float)(unsigned)x * cosv;
            float yy0 = yyy0 + ys * cosv + (float)(unsigned)x * sinv;
            for (int y = ys; y < ye; y++) {
                rot[((size_t)v * N + x) * N + y] = make_float2(xx0, yy0);
                xx0 -= sinv;
                yy0 += cosv;
            }
        }
    }
    // forward-projector planes.  A plane's axial sample positions for row y are
    // base + (y - c) * tan and base + (c - y) * tan, truncated, plus delta: end planes use
    // their own plane and angle (flookup PJ:544-562), middle planes the first middle plane
    // ym of the segment and its position advanced by yr - ym (PJ:1929-1944), so the device
    // repeats the tables' float operations exactly.  Rows below ymin are skipped
    // (PJ:1920-1921); the skip rule is checked to hold for exactly the rows below a threshold.
    std::vector<int4> fplanes;
    int fp_fast = 1;   // positions never negative: the forward projector may take fract for zz - (int)zz
    for (int yr = 0; yr < Z; yr++) fplanes.push_back(make_int4(yr << 8, 0, 0, 0));
    for (int k = 1; k <= G.K; k++) {
        const int th = 2 * k, yb = G.yr_bottom[th], yt = G.yr_top[th], ym = yb + ZS;
        int ymin = 1;
        for (int y = 1; y < c; y++)
            if (skip[(size_t)y * GM + k - 1]) ymin = y + 1;
        for (int y = ymin; y < c; y++)
            if (skip[(size_t)y * GM + k - 1]) {
                delete ctx;
                return fail(HRRT_ERR_UNSUPPORTED, "forward-projector row skip is not a threshold (group %d)", k);
            }
        for (int yr = yb; yr <= yt; yr++) {
            const bool end = yr < yb + ZS || yr > yt - ZS;
            const int base = end ? yr : ym, delta = end ? 0 : yr - ym;
            const float bf = (float)base;
            const float t = G.tantheta[(size_t)(k - 1) * Z + base];
            fplanes.push_back(make_int4(k | (yr << 8) | (delta << 16), bits(bf), bits(t), ymin));
            // the device recomputes what the tables hold: check it on the host for every row
            for (int y = ymin; y < c; y++) {
                const float zz1 = bf + (y - c) * t, zz2 = bf + (c - y) * t;
                const int z1 = (int)zz1, z2 = (int)zz2;
                if (zz1 < 0.0f || zz2 < 0.0f) fp_fast = 0;
                int rz, rz2;
                float rf1, rf2;
                const size_t li = (size_t)y * GM + k - 1;
                if (end) {
                    const int e = yr < yb + ZS ? yr - yb : ZS + (yr - (yt - ZS + 1));
                    const int4 q = endp[li * 2 * ZS + e];
                    rz = q.x; rf1 = __builtin_bit_cast(float, q.y); rz2 = q.z; rf2 = __builtin_bit_cast(float, q.w);
                } else {
                    const int4 m = mid[li];
                    rz = m.x + delta; rz2 = m.y + delta;
                    rf1 = __builtin_bit_cast(float, m.z); rf2 = __builtin_bit_cast(float, m.w);
                }
                if (z1 + delta != rz || z2 + delta != rz2 || bits(zz1 - z1) != bits(rf1) || bits(zz2 - z2) != bits(rf2)) {
                    delete ctx;
                    return fail(HRRT_ERR_UNSUPPORTED, "forward-projector positions disagree (y=%d, group=%d)", y, k);
                }
            }
        }
    }
    std::vector<int16_t> pseg(G.spp, -1), plane_seg(G.segmentsize, 0);
    for (int s = 0; s < G.th_pixels; s++)
        for (int i = 0; i < G.nplanes(s); i++) {
            pseg[G.seg_data[s] + i] = (int16_t)s;
            plane_seg[G.segment_offset[s] + i] = (int16_t)s;
        }
    std::vector<int> file_seg(G.th_pixels);
    for (int s = 0; s < G.th_pixels; s++) {
        const int sw = s % 2 == 1 ? s + 1 : (s == 0 ? 0 : s - 1);   // FIO:116-118
        file_seg[s] = G.seg_offset[sw] / (N * G.views);
    }
    ctx->file_seg = file_seg;
    std::vector<float> resc = G.rescalearray;
    // back-projection column in LDS: segment 0, then every oblique segment inside a window of
    // zero planes wide enough that the back projector's 64-plane slots 0-2 (z < 192) always,
    // and slot 3 when the range reaches z >= 192, read only that segment or zeros (the
    // reference's zero guard planes PJ:1689-1692, extended); neighbouring windows share zeros
    if (GM > 32) {
        delete ctx;
        return fail(HRRT_ERR_UNSUPPORTED, "more than 32 oblique segment groups");
    }
    std::vector<int> need_lo(G.th_pixels, 0), need_hi(G.th_pixels, 0);
    for (int k = 1; k <= G.K; k++)
        for (int y = 1; y < N; y++) {
            const int4 L = blut[(size_t)y * GM + k - 1];
            const int yb = G.yr_bottom[2 * k], yt = G.yr_top[2 * k];
            int top = 192 + L.z;                          // slots 0-2 read yr = z + off, z+1 <= 192
            if (L.y >= 192) top = std::max(top, 256 + L.z);
            for (int s = 2 * k - 1; s <= 2 * k; s++) {
                need_lo[s] = std::max(need_lo[s], yb - L.z);
                need_hi[s] = std::max(need_hi[s], top - yt);
            }
        }
    std::vector<int> padd(G.th_pixels, 0);
    int lds = G.nplanes(0);
    for (int s = 1; s < G.th_pixels; s++) {
        padd[s] = lds + std::max(need_lo[s], s > 1 ? need_hi[s - 1] : 0);
        lds = padd[s] + G.nplanes(s);
    }
    lds += std::max(need_hi[G.th_pixels - 1], 1);
    if ((size_t)lds * sizeof(float2) > 160 * 1024) {
        delete ctx;
        return fail(HRRT_ERR_UNSUPPORTED, "back-projection column (%d planes) exceeds LDS", lds);
    }
    d.bp_lds = lds;
    std::vector<int> bp_src(lds + 1, -1);
    for (int s = 0; s < G.th_pixels; s++)
        for (int i = 0; i < G.nplanes(s); i++) bp_src[padd[s] + i] = G.seg_data[s] + i;
    // per (row, group): {zsa | zea << 16, LDS index of plane 0 of segment 2k shifted by the
    // row's integer offset, same for 2k-1, bits(frac)}; every read must stay in its window
    std::vector<int4> blut2((size_t)N * GM);
    for (int y = 0; y < N; y++)
        for (int k = 1; k <= G.K; k++) {
            const int4 L = blut[(size_t)y * GM + k - 1];
            const int yb = G.yr_bottom[2 * k];
            blut2[(size_t)y * GM + k - 1] =
                make_int4(L.x | (L.y << 16), padd[2 * k] - yb + L.z, padd[2 * k - 1] - yb + L.z, L.w);
            if (y == 0) continue;
            for (int s = 2 * k - 1; s <= 2 * k; s++) {
                const int b = padd[s] - yb + L.z;
                const int last = (L.y >= 192 ? 256 : 192);          // last plane touched: z = last-1, p[z+1]
                const int wlo = padd[s] - need_lo[s], whi = padd[s] + G.nplanes(s) + need_hi[s];  // [wlo, whi)
                if (b < wlo || b + last >= whi || b + last >= lds) {
                    delete ctx;
                    return fail(HRRT_ERR_UNSUPPORTED, "back-projection read outside its LDS window (y=%d, group=%d)",
                                y, k);
                }
            }
        }

    // incremental back projector (back_proj3d_view_n1thread PJ:4162-4668).  LDS holds one lane's
    // column: segment 0 (Z floats, padded to an even count), then per group k a window of
    // float2 {segment 2k, segment 2k-1} for rows yr = yb-3 .. yt+1, where yb-1 and yt+1 are the
    // reference's zero guard planes (PJ:1689-1692) and yb-3, yb-2 two more zero cells.  The
    // window is turned in place into the reference's differences P'(yr) = (p(yr+1)-p(yr))*tan
    // (PJ:4313-4324), then into the values the breakpoint terms read: M(yr) = P'(yr)-P'(yr+1)
    // for yr in [yb-1, yt-1] (the middle terms PJ:4381-4394), M(yb-2) = -P'(yb-1) (the two
    // terms before the loop, PJ:4369-4380) and M(yt) = P'(yt) (the two after it, PJ:4395-4406),
    // with zeros at yb-3 and yt+1: a breakpoint then adds, to plane z, the terms read at rows
    // z + const clamped into the window, a clamped read giving exactly +0.
    int inc_ok = Z <= 256 && N <= 256 && G.K <= 32;
    std::vector<int> isrc;
    std::vector<int2> iconv;
    std::vector<int> gbase(G.K + 1, 0);
    for (int z = 0; z < Z; z++) {
        isrc.push_back(G.seg_data[0] + z);
        iconv.push_back(make_int2(0, 0));
    }
    if (Z & 1) {
        isrc.push_back(-1);
        iconv.push_back(make_int2(0, 0));
    }
    for (int k = 1; k <= G.K; k++) {
        const int th = 2 * k, yb = G.yr_bottom[th], yt = G.yr_top[th];
        if (yb != G.yr_bottom[th - 1] || yt != G.yr_top[th - 1] || yt < yb) inc_ok = 0;
        const float tn = -G.sin_theta[th] / G.cos_theta[th];
        gbase[k] = (int)isrc.size() / 2 - (yb - 3);                  // float2 index of row 0
        for (int yr = yb - 3; yr <= yt + 1; yr++)
            for (int e = 0; e < 2; e++) {
                const int sg = e == 0 ? th : th - 1;
                isrc.push_back(yr >= yb && yr <= yt ? G.seg_data[sg] + (yr - yb) : -1);
                const int second = yr == yb - 2 ? 2 : (yr >= yb - 1 && yr <= yt - 1 ? 1 : 0);
                iconv.push_back(make_int2(yr >= yb - 1 && yr <= yt ? bits(tn) : 0, second));
            }
    }
    const int inc_lds = (int)isrc.size();
    auto in_window = [&](int k, int first, int count) {      // rows first .. first+count of group k
        if (count < 0) return true;
        return first >= G.yr_bottom[2 * k] - 1 && first + count <= G.yr_top[2 * k] + 1;
    };
    // per group k (entry k-1): {float2 index of row 0 of the window, yb, yt, 0}, row c-1 lookup
    // {zsa, count, yra, bits(frac)}, row c+1 lookup {zsa, yra, bits(frac), 0}, row c+1 breakpoint
    // {z_start, z_end, yr_start, 0}; a range with no planes points at yb so that its (unused)
    // reads stay in the window
    std::vector<int4> igroup((size_t)std::max(G.K, 1) * 4, make_int4(0, 0, 0, 0));
    for (int k = 1; k <= G.K; k++) {
        const int th = 2 * k, yb = G.yr_bottom[th], yt = G.yr_top[th];
        const size_t l1 = (size_t)(c - 1) * GM + k - 1, l2 = (size_t)(c + 1) * GM + k - 1;
        const hrrt::Geometry::Blut &bc = G.blut[l2];
        const int n = G.bl_zea[l1] - G.bl_zsa[l1] + 1;
        int yra1 = G.bl_yra[l1], yra2 = G.bl_yra[l2], bcyr = bc.yr_start;
        if (n <= 0) yra1 = yra2 = yb;
        if (bc.z_end < bc.z_start) bcyr = yb;
        if (n > 0 && (!in_window(k, yra1, n) || !in_window(k, yra2, n))) inc_ok = 0;
        if (!in_window(k, bcyr, std::max(bc.z_end - bc.z_start, 0))) inc_ok = 0;
        int4 *e = &igroup[(size_t)(k - 1) * 4];
        e[0] = make_int4(gbase[k], yb, yt, 0);
        e[1] = make_int4(G.bl_zsa[l1], n, yra1, bits(G.bl_frac[l1]));
        e[2] = make_int4(G.bl_zsa[l2], yra2, bits(G.bl_frac[l2]), 0);
        e[3] = make_int4(bc.z_start, bc.z_end, bcyr, 0);
    }
    // breakpoints of rows c-2 down to 1, groups ascending: {bits(coef1), byte offsets (from 8z) of
    // the a-term row yr_start_pre + (z - z_start_pre) and of the b-term row yr_start_pre +
    // (z - z_start_pre(N-y) - 1), window bounds lo | hi << 16 in bytes}.  Handled: the range moved
    // down by one plane, it spans the segment's rows [yb-1, yt], and the b-term's iteration of a
    // plane is never after its a-term's (z_start_pre(N-y) + 1 > z_start_pre), so every plane takes
    // its b-term first (PJ:4381-4394)
    std::vector<int4> isched;
    std::vector<int> inrow(c + 1, 0);
    for (int y = c - 2; y >= 1; y--) {
        for (int k = 1; k <= G.K; k++) {
            const size_t li = (size_t)y * GM + k - 1;
            if (G.blut_flag[li] == 0) continue;
            const hrrt::Geometry::Blut &ba = G.blut[li], &bb = G.blut[(size_t)(N - y) * GM + k - 1];
            const int th = 2 * k, yb = G.yr_bottom[th], yt = G.yr_top[th];
            const int zp = ba.z_start_pre, z2p = bb.z_start_pre, nl = ba.z_end_pre - ba.z_start_pre;
            if (ba.z_start - zp != -1) inc_ok = 0;                     // the range moved by other than one plane
            if (ba.yr_start_pre != yb - 1 || ba.yr_start_pre + nl != yt) inc_ok = 0;
            if (z2p + 1 - zp <= 0) inc_ok = 0;
            float tn = -G.sin_theta[th] / G.cos_theta[th];
            tn = 1 - tn * (float)ba.num - ba.coef_pre;                    // PJ:4371-4373
            const float coef1 = -tn * G.cos_theta[th] / G.sin_theta[th];
            const int offa = 8 * (gbase[k] + ba.yr_start_pre - zp), offb = 8 * (gbase[k] + ba.yr_start_pre - z2p - 1);
            const int lo8 = 8 * (gbase[k] + yb - 3), hi8 = 8 * (gbase[k] + yt + 1);
            isched.push_back(make_int4(bits(coef1), offa, offb, lo8 | (hi8 << 16)));
            inrow[y]++;
        }
    }
    isched.push_back(make_int4(0, 0, 0, 0));                           // read ahead of the last one
    d.inc_ok = inc_ok;
    d.inc_lds = inc_lds;
    d.inc_shm = (int)(sizeof(float) * (inc_lds + 2));                // + the two floats the last conversion reads
    if (d.inc_shm > 64 * 1024) d.inc_ok = 0;

    int r = 0;
    if ((r = ctx->upload(G.cyl_lo, d.cyl_lo)) || (r = ctx->upload(G.cyl_hi, d.cyl_hi)) ||
        (r = ctx->upload(G.sin_psi, d.sin_psi)) || (r = ctx->upload(G.cos_psi, d.cos_psi)) ||
        (r = ctx->upload(G.yr_bottom, d.yb)) || (r = ctx->upload(G.yr_top, d.yt)) ||
        (r = ctx->upload(G.seg_data, d.seg_data)) || (r = ctx->upload(G.segment_offset, d.segment_offset)) ||
        (r = ctx->upload(file_seg, d.file_seg)) || (r = ctx->upload(G.bp_rescale, d.bp_rescale)) ||
        (r = ctx->upload(resc, d.rescalearray)) || (r = ctx->upload(blut, d.bp_lut)) ||
        (r = ctx->upload(rot, d.rot)) || (r = ctx->upload(fplanes, d.fp_planes)) || (r = ctx->upload(pseg, d.pseg)) ||
        (r = ctx->upload(plane_seg, d.plane_seg)) || (r = ctx->upload(bp_src, d.bp_src)) ||
        (r = ctx->upload(blut2, d.bp_lut2)) || (r = ctx->upload(isrc, d.inc_src)) ||
        (r = ctx->upload(iconv, d.inc_conv)) || (r = ctx->upload(igroup, d.inc_group)) ||
        (r = ctx->upload(isched, d.inc_sched)) || (r = ctx->upload(inrow, d.inc_nrow))) {
        delete ctx;
        return r;
    }
    d.fp_np = (int)fplanes.size();
    d.fp_fast = fp_fast;
    d.fp_maxt = 4;
    d.fp_nsplit = 0;
    d.bp_waves = 16;
    ctx->bp_threads = d.inc_ok ? 1 : 0;                                 // every pair incremental when possible
    {
        std::vector<int8_t> nz((size_t)(G.views / 2) * (c + 1), 1);
        void *p = nullptr;
        if (hipMalloc(&p, nz.size()) != hipSuccess) {
            delete ctx;
            return fail(HRRT_ERR_NOMEM, "hipMalloc failed");
        }
        d.nzmask = static_cast<int8_t *>(p);
        if (hipMemcpy(p, nz.data(), nz.size(), hipMemcpyHostToDevice) != hipSuccess) {
            delete ctx;
            return fail(HRRT_ERR_HIP, "hipMemcpy failed");
        }
    }
    *out = ctx;
    return HRRT_OK;
}

extern "C" void hrrt_destroy(hrrt_ctx *ctx) { delete ctx; }

static void fill_dims(const hrrt::Geometry &G, hrrt_dims *o) {
    o->N = G.N; o->views = G.views; o->z_pixels = G.z_pixels; o->z4 = G.z4; o->segmentsize = G.segmentsize;
    o->th_pixels = G.th_pixels; o->sviews = G.sviews; o->subsets = G.p.subsets; o->spp = G.spp;
    o->x_rebin = G.x_rebin; o->v_rebin = G.v_rebin;
    o->image_floats = (long long)G.N * G.N * G.z4;
    o->sino_bins = (long long)G.segmentsize * G.views * G.N;
    o->file_bins = (long long)G.segmentsize * G.p.def_angles * G.p.def_elements;
    o->paircol_elems = (long long)(G.views / 2) * 2 * G.N * G.spp;
    o->fov_voxels = 0;
    o->fov_bins = 0;
    for (int x = 0; x < G.N; x++) {
        o->fov_voxels += G.cyl_hi[x] - G.cyl_lo[x];
        if (x >= 1 && G.cyl_hi[x] > G.cyl_lo[x]) o->fov_bins++;
    }
}

static void copy_segments(const hrrt::Geometry &G, int *yb, int *yt, int *so, int *sd) {
    const size_t n = sizeof(int) * G.th_pixels;
    if (yb) std::memcpy(yb, G.yr_bottom.data(), n);
    if (yt) std::memcpy(yt, G.yr_top.data(), n);
    if (so) std::memcpy(so, G.segment_offset.data(), n);
    if (sd) std::memcpy(sd, G.seg_data.data(), n);
}

static void copy_order(const hrrt::Geometry &G, int *order, int *vieworder) {
    if (order) std::memcpy(order, G.order.data(), sizeof(int) * G.order.size());
    if (vieworder) std::memcpy(vieworder, G.vieworder.data(), sizeof(int) * G.vieworder.size());
}

extern "C" int hrrt_query_geometry(const hrrt_params *params, hrrt_dims *dims, int *order, int *vieworder,
                                   int *yb, int *yt, int *so, int *sd) {
    if (!params) return fail(HRRT_ERR_ARG, "null argument");
    hrrt::Geometry G;
    std::string err;
    if (!G.build(*params, err)) return fail(HRRT_ERR_ARG, "%s", err.c_str());
    if (dims) fill_dims(G, dims);
    copy_order(G, order, vieworder);
    copy_segments(G, yb, yt, so, sd);
    return HRRT_OK;
}

extern "C" int hrrt_get_dims(const hrrt_ctx *ctx, hrrt_dims *o) {
    if (!ctx || !o) return fail(HRRT_ERR_ARG, "null argument");
    fill_dims(ctx->geo, o);
    return HRRT_OK;
}

extern "C" int hrrt_get_segments(const hrrt_ctx *ctx, int *yb, int *yt, int *so, int *sd) {
    if (!ctx) return fail(HRRT_ERR_ARG, "null context");
    copy_segments(ctx->geo, yb, yt, so, sd);
    return HRRT_OK;
}

extern "C" int hrrt_get_order(const hrrt_ctx *ctx, int *order, int *vieworder) {
    if (!ctx) return fail(HRRT_ERR_ARG, "null context");
    copy_order(ctx->geo, order, vieworder);
    return HRRT_OK;
}

extern "C" int hrrt_set_norm_zero_mask(hrrt_ctx *ctx, const int8_t *mask) {
    if (!ctx || !mask) return fail(HRRT_ERR_ARG, "null argument");
    size_t n = (size_t)(ctx->geo.views / 2) * (ctx->geo.N / 2 + 1);
    HIP_TRY(hipMemcpy(ctx->dg.nzmask, mask, n, hipMemcpyHostToDevice));
    return HRRT_OK;
}

extern "C" int hrrt_get_norm_zero_mask(hrrt_ctx *ctx, int8_t *mask) {
    if (!ctx || !mask) return fail(HRRT_ERR_ARG, "null argument");
    size_t n = (size_t)(ctx->geo.views / 2) * (ctx->geo.N / 2 + 1);
    HIP_TRY(hipMemcpy(mask, ctx->dg.nzmask, n, hipMemcpyDeviceToHost));
    return HRRT_OK;
}

extern "C" int hrrt_stage_sinograms(hrrt_ctx *ctx, const int16_t *d_prompt_file, const float *d_delayed_file,
                                    const float *d_norm_file, const float *d_scatter_file, const float *d_atten_file,
                                    int16_t *d_prompt, float *d_add, float *d_invnorm, void *stream) {
    if (!ctx || !d_prompt || !d_add || !d_invnorm) return fail(HRRT_ERR_ARG, "null argument");
    const hrrt::Geometry &G = ctx->geo;
    int r;
    if ((r = ctx->zalloc((void **)&ctx->colmask, sizeof(int) * G.views * G.N))) return r;
    HIP_TRY(hipMemsetAsync(ctx->colmask, 0, sizeof(int) * G.views * G.N, S(stream)));
    if (d_norm_file && G.x_rebin * G.v_rebin > 1)
        if ((r = ctx->zalloc((void **)&ctx->gapmask, (size_t)G.th_pixels * G.p.def_angles * G.p.def_elements))) return r;
    HIP_TRY(hrrt::launch_stage(ctx->dg, d_prompt_file, d_delayed_file, d_norm_file, d_scatter_file, d_atten_file,
                               d_prompt, d_add, d_invnorm, ctx->gapmask, ctx->colmask, 0, G.th_pixels, 0,
                               G.segmentsize, true, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_stage_sinograms_w012(hrrt_ctx *ctx, int weighting, int precorrected, const int16_t *d_data16_file,
                                         const float *d_dataf_file, const float *d_delayed_file,
                                         const float *d_norm_file, const float *d_scatter_file,
                                         const float *d_atten_file, float *d_data, float *d_invnorm, void *stream) {
    if (!ctx || !d_data || !d_invnorm) return fail(HRRT_ERR_ARG, "null argument");
    if (weighting < 0 || weighting > 2) return fail(HRRT_ERR_ARG, "weighting %d is not 0 (UW), 1 (AW) or 2 (ANW)", weighting);
    if (!d_data16_file == !d_dataf_file) return fail(HRRT_ERR_ARG, "give exactly one of the short and the float data file");
    if (precorrected && !d_dataf_file) return fail(HRRT_ERR_ARG, "a precorrected sinogram is a float file");
    if (weighting == 0 && !precorrected && d_scatter_file && !d_norm_file)
        return fail(HRRT_ERR_UNSUPPORTED, "-W 0 with scatter and no norm: the reference subtracts a buffer it never "
                                          "filled (hrrt_osem3d_span3.cpp:1233)");
    const hrrt::Geometry &G = ctx->geo;
    int r;
    if ((r = ctx->zalloc((void **)&ctx->colmask, sizeof(int) * G.views * G.N))) return r;
    HIP_TRY(hipMemsetAsync(ctx->colmask, 0, sizeof(int) * G.views * G.N, S(stream)));
    if (d_norm_file && G.x_rebin * G.v_rebin > 1)
        if ((r = ctx->zalloc((void **)&ctx->gapmask, (size_t)G.th_pixels * G.p.def_angles * G.p.def_elements))) return r;
    HIP_TRY(hrrt::launch_stage_w012(ctx->dg, weighting, precorrected != 0, d_data16_file, d_dataf_file, d_delayed_file,
                                    d_norm_file, d_scatter_file, d_atten_file, d_data, d_invnorm, ctx->gapmask,
                                    ctx->colmask, 0, G.th_pixels, 0, G.segmentsize, true, S(stream)));
    return HRRT_OK;
}

// Staging from host memory (the reference reads the flat files segment by segment,
// FIO:110-259).  Every segment's block of each given file (file planes file_seg[th] ..
// + planes of th, contiguous) goes to the context's device copy of that file on a copy stream
// -- straight from the caller's memory when it is pinned, else through a ring of pinned
// buffers filled on this thread while earlier blocks are in flight -- and the staging launch
// of that segment waits for its blocks only, so copies of later segments overlap the staging
// of earlier ones.  Returns when every copy has completed (the host files may then be
// reused); the staging launches may still run on the caller's stream.
struct HostFile {
    const void *src;   // host memory of the whole flat file, or null (file absent)
    size_t elem;       // bytes per element
};

static int stage_host(hrrt_ctx *ctx, HostFile files[5], void *dev[5], hipStream_t s,
                      const std::function<hipError_t(int th, bool finish)> &launch) {
    const hrrt::Geometry &G = ctx->geo;
    int r;
    if ((r = ctx->ensure_host_staging())) return r;
    const size_t plane = (size_t)G.p.def_angles * G.p.def_elements;
    bool pinned[5];
    for (int f = 0; f < 5; f++) {
        dev[f] = nullptr;
        pinned[f] = false;
        if (!files[f].src) continue;
        const size_t bytes = (size_t)G.segmentsize * plane * sizeof(float);   // short or float files
        if (!ctx->hf[f] && hipMalloc(&ctx->hf[f], bytes) != hipSuccess) {
            ctx->hf[f] = nullptr;
            return fail(HRRT_ERR_NOMEM, "hipMalloc(%zu) of a device file copy failed", bytes);
        }
        dev[f] = ctx->hf[f];
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, files[f].src) == hipSuccess && at.type == hipMemoryTypeHost) pinned[f] = true;
        (void)hipGetLastError();   // pageable memory reports an error here: not one of ours
    }
    // the device copies may still be read by an earlier staging on the caller's stream
    HIP_TRY(hipEventRecord(ctx->copy_start, s));
    HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->copy_start, 0));
    int slot = 0;
    for (int th = 0; th < G.th_pixels; th++) {
        const size_t np = (size_t)(G.yr_top[th] - G.yr_bottom[th] + 1);
        for (int f = 0; f < 5; f++) {
            if (!files[f].src) continue;
            const size_t off = (size_t)ctx->file_seg[th] * plane * files[f].elem, n = np * plane * files[f].elem;
            const char *src = static_cast<const char *>(files[f].src) + off;
            char *dst = static_cast<char *>(dev[f]) + off;
            if (pinned[f]) {
                HIP_TRY(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, ctx->copy_stream));
                continue;
            }
            for (size_t done = 0; done < n; done += hrrt_ctx::RING_BYTES) {
                const size_t c = std::min(hrrt_ctx::RING_BYTES, n - done);
                HIP_TRY(hipEventSynchronize(ctx->ring_ev[slot]));   // the slot's previous copy is done
                std::memcpy(ctx->ring[slot], src + done, c);
                HIP_TRY(hipMemcpyAsync(dst + done, ctx->ring[slot], c, hipMemcpyHostToDevice, ctx->copy_stream));
                HIP_TRY(hipEventRecord(ctx->ring_ev[slot], ctx->copy_stream));
                slot = (slot + 1) % hrrt_ctx::NRING;
            }
        }
        HIP_TRY(hipEventRecord(ctx->seg_ev[th], ctx->copy_stream));
        HIP_TRY(hipStreamWaitEvent(s, ctx->seg_ev[th], 0));
        HIP_TRY(launch(th, th == G.th_pixels - 1));
    }
    HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    return HRRT_OK;
}

extern "C" int hrrt_stage_sinograms_host(hrrt_ctx *ctx, const int16_t *prompt_file, const float *delayed_file,
                                         const float *norm_file, const float *scatter_file, const float *atten_file,
                                         int16_t *d_prompt, float *d_add, float *d_invnorm, void *stream) {
    if (!ctx || !d_prompt || !d_add || !d_invnorm) return fail(HRRT_ERR_ARG, "null argument");
    const hrrt::Geometry &G = ctx->geo;
    int r;
    if ((r = ctx->zalloc((void **)&ctx->colmask, sizeof(int) * G.views * G.N))) return r;
    HIP_TRY(hipMemsetAsync(ctx->colmask, 0, sizeof(int) * G.views * G.N, S(stream)));
    if (norm_file && G.x_rebin * G.v_rebin > 1)
        if ((r = ctx->zalloc((void **)&ctx->gapmask, (size_t)G.th_pixels * G.p.def_angles * G.p.def_elements))) return r;
    HostFile files[5] = {{prompt_file, 2}, {delayed_file, 4}, {norm_file, 4}, {scatter_file, 4}, {atten_file, 4}};
    void *dev[5];
    return stage_host(ctx, files, dev, S(stream), [&](int th, bool finish) {
        const int p0 = G.segment_offset[th], p1 = p0 + G.yr_top[th] - G.yr_bottom[th] + 1;
        return hrrt::launch_stage(ctx->dg, (const int16_t *)dev[0], (const float *)dev[1], (const float *)dev[2],
                                  (const float *)dev[3], (const float *)dev[4], d_prompt, d_add, d_invnorm,
                                  ctx->gapmask, ctx->colmask, th, 1, p0, p1, finish, S(stream));
    });
}

extern "C" int hrrt_stage_sinograms_w012_host(hrrt_ctx *ctx, int weighting, int precorrected,
                                              const int16_t *data16_file, const float *dataf_file,
                                              const float *delayed_file, const float *norm_file,
                                              const float *scatter_file, const float *atten_file, float *d_data,
                                              float *d_invnorm, void *stream) {
    if (!ctx || !d_data || !d_invnorm) return fail(HRRT_ERR_ARG, "null argument");
    if (weighting < 0 || weighting > 2) return fail(HRRT_ERR_ARG, "weighting %d is not 0 (UW), 1 (AW) or 2 (ANW)", weighting);
    if (!data16_file == !dataf_file) return fail(HRRT_ERR_ARG, "give exactly one of the short and the float data file");
    if (precorrected && !dataf_file) return fail(HRRT_ERR_ARG, "a precorrected sinogram is a float file");
    if (weighting == 0 && !precorrected && scatter_file && !norm_file)
        return fail(HRRT_ERR_UNSUPPORTED, "-W 0 with scatter and no norm: the reference subtracts a buffer it never "
                                          "filled (hrrt_osem3d_span3.cpp:1233)");
    const hrrt::Geometry &G = ctx->geo;
    int r;
    if ((r = ctx->zalloc((void **)&ctx->colmask, sizeof(int) * G.views * G.N))) return r;
    HIP_TRY(hipMemsetAsync(ctx->colmask, 0, sizeof(int) * G.views * G.N, S(stream)));
    if (norm_file && G.x_rebin * G.v_rebin > 1)
        if ((r = ctx->zalloc((void **)&ctx->gapmask, (size_t)G.th_pixels * G.p.def_angles * G.p.def_elements))) return r;
    HostFile files[5] = {{data16_file ? (const void *)data16_file : (const void *)dataf_file, data16_file ? 2u : 4u},
                         {delayed_file, 4}, {norm_file, 4}, {scatter_file, 4}, {atten_file, 4}};
    void *dev[5];
    return stage_host(ctx, files, dev, S(stream), [&](int th, bool finish) {
        const int p0 = G.segment_offset[th], p1 = p0 + G.yr_top[th] - G.yr_bottom[th] + 1;
        return hrrt::launch_stage_w012(ctx->dg, weighting, precorrected != 0,
                                       data16_file ? (const int16_t *)dev[0] : nullptr,
                                       data16_file ? nullptr : (const float *)dev[0], (const float *)dev[1],
                                       (const float *)dev[2], (const float *)dev[3], (const float *)dev[4], d_data,
                                       d_invnorm, ctx->gapmask, ctx->colmask, th, 1, p0, p1, finish, S(stream));
    });
}

// Back projection of the pairs of a batch under the reference's thread schedule
// (CalculateNormfac S3:3796-3825, CalculateOsem3d S3:4174-4218): with T threads the first
// npairs - npairs % T pairs of a subset take the incremental projector (back_proj3d_thread1),
// the rest the direct one (back_proj3d2); T = 0 takes the direct one for every pair.
static hipError_t backproject_scheduled(const hrrt_ctx *ctx, const hrrt::PairBatch &b, const float *src,
                                        int src_global, int apply_rescale, float *bufA, float *bufB, hipStream_t s) {
    const hrrt::DevGeom &g = ctx->dg;
    const int T = ctx->bp_threads;
    const int nfull = T > 0 ? b.npairs - b.npairs % T : 0;
    if (nfull > 0) {
        hrrt::PairBatch bi = b;
        bi.npairs = nfull;
        hipError_t e = hrrt::launch_backproject_incremental(g, bi, src, src_global, apply_rescale, bufA, bufB, s);
        if (e != hipSuccess) return e;
    }
    if (nfull < b.npairs) {
        hrrt::PairBatch bd = b;
        bd.view0 = b.view0 + b.view_stride * nfull;
        bd.npairs = b.npairs - nfull;
        const size_t vol = (size_t)g.N * g.N * g.Z4;
        const float *sp = src_global ? src : src + (size_t)nfull * 2 * g.N * g.spp;
        return hrrt::launch_backproject(g, bd, sp, 0, src_global, apply_rescale, 0, bufA + nfull * vol,
                                        bufB + nfull * vol, s);
    }
    return hipSuccess;
}

extern "C" int hrrt_calculate_normfac(hrrt_ctx *ctx, const float *d_invnorm, float *d_normfac, void *stream) {
    if (!ctx || !d_invnorm || !d_normfac) return fail(HRRT_ERR_ARG, "null argument");
    int r;
    if ((r = ctx->ensure_subset_scratch())) return r;
    const hrrt::Geometry &G = ctx->geo;
    if (ctx->psf_on && (r = ctx->ensure_psf_scratch())) return r;
    for (int s = 0; s < G.p.subsets; s++) {
        hrrt::PairBatch b{s, G.p.subsets, ctx->npairs()};
        float *nf = d_normfac + (size_t)s * ctx->vol_floats();
        HIP_TRY(backproject_scheduled(ctx, b, d_invnorm, 1, 1, ctx->bufA, ctx->bufB, S(stream)));
        if (ctx->psf_on) {   // the sensitivity is blurred before inversion (S3:3835-3840)
            HIP_TRY(hrrt::launch_rotate_sum(ctx->dg, b, ctx->bufA, ctx->bufB, nf, S(stream)));
            HIP_TRY(ctx->convolve3d(nf, nf, S(stream)));
            HIP_TRY(hrrt::launch_normfac_invert(ctx->dg, nf, nf, S(stream)));
        } else {
            HIP_TRY(hrrt::launch_rotate_sens(ctx->dg, b, ctx->bufA, ctx->bufB, nf, 0, S(stream)));
        }
    }
    return HRRT_OK;
}

extern "C" int hrrt_init_image(hrrt_ctx *ctx, float *d_image, void *stream) {
    if (!ctx || !d_image) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_init_image(ctx->dg, d_image, ctx->geo.z_min, ctx->geo.z_max, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_end_iteration(hrrt_ctx *ctx, float *d_image, void *stream) {
    if (!ctx || !d_image) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_end_iteration(ctx->dg, d_image, ctx->geo.z_min, ctx->geo.z_max, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_osem_subset(hrrt_ctx *ctx, float *d_image, const float *d_normfac_subset,
                                const int16_t *d_prompt, const float *d_add, int isubset, void *stream) {
    if (!ctx || !d_image || !d_normfac_subset || !d_prompt || !d_add) return fail(HRRT_ERR_ARG, "null argument");
    const hrrt::Geometry &G = ctx->geo;
    if (isubset < 0 || isubset >= G.p.subsets) return fail(HRRT_ERR_ARG, "subset %d out of range", isubset);
    int r;
    if ((r = ctx->ensure_subset_scratch())) return r;
    if (ctx->psf_on && (r = ctx->ensure_psf_scratch())) return r;
    hrrt::PairBatch b{isubset, G.p.subsets, ctx->npairs()};
    ctx->mark(0, S(stream));
    hrrt::FpIO io;
    io.prompt = d_prompt;
    io.add = d_add;
    const float *fp_image = d_image;
    if (ctx->psf_on) {   // the forward projector reads the blurred image (S3:4172, 4202)
        HIP_TRY(ctx->convolve3d(d_image, ctx->psf_img, S(stream)));
        fp_image = ctx->psf_img;
    }
    HIP_TRY(hrrt::launch_forward(ctx->dg, b, fp_image, ctx->sub_ratio, io, hrrt::FP_RATIO, S(stream)));
    ctx->mark(1, S(stream));
    HIP_TRY(backproject_scheduled(ctx, b, ctx->sub_ratio, 0, 0, ctx->bufA, ctx->bufB, S(stream)));
    ctx->mark(2, S(stream));
    if (ctx->psf_on) {   // the correction is blurred before the update (S3:4229)
        HIP_TRY(hrrt::launch_rotate_sum(ctx->dg, b, ctx->bufA, ctx->bufB, ctx->psf_corr, S(stream)));
        HIP_TRY(ctx->convolve3d(ctx->psf_corr, ctx->psf_corr, S(stream)));
        HIP_TRY(hrrt::launch_update_image(ctx->dg, d_image, ctx->psf_corr, d_normfac_subset, S(stream)));
    } else {
        HIP_TRY(hrrt::launch_rotate_update(ctx->dg, b, ctx->bufA, ctx->bufB, d_image, d_normfac_subset, 0,
                                           S(stream)));
    }
    ctx->mark(-1, S(stream));
    return HRRT_OK;
}

extern "C" int hrrt_osem_subset_w012(hrrt_ctx *ctx, float *d_image, const float *d_normfac_subset, const float *d_data,
                                     int isubset, void *stream) {
    if (!ctx || !d_image || !d_normfac_subset || !d_data) return fail(HRRT_ERR_ARG, "null argument");
    const hrrt::Geometry &G = ctx->geo;
    if (isubset < 0 || isubset >= G.p.subsets) return fail(HRRT_ERR_ARG, "subset %d out of range", isubset);
    int r;
    if ((r = ctx->ensure_subset_scratch())) return r;
    if (ctx->psf_on && (r = ctx->ensure_psf_scratch())) return r;
    hrrt::PairBatch b{isubset, G.p.subsets, ctx->npairs()};
    ctx->mark(0, S(stream));
    hrrt::FpIO io;
    io.add = d_data;
    const float *fp_image = d_image;
    if (ctx->psf_on) {
        HIP_TRY(ctx->convolve3d(d_image, ctx->psf_img, S(stream)));
        fp_image = ctx->psf_img;
    }
    HIP_TRY(hrrt::launch_forward(ctx->dg, b, fp_image, ctx->sub_ratio, io, hrrt::FP_RATIO_W012, S(stream)));
    ctx->mark(1, S(stream));
    HIP_TRY(backproject_scheduled(ctx, b, ctx->sub_ratio, 0, 0, ctx->bufA, ctx->bufB, S(stream)));
    ctx->mark(2, S(stream));
    if (ctx->psf_on) {
        HIP_TRY(hrrt::launch_rotate_sum(ctx->dg, b, ctx->bufA, ctx->bufB, ctx->psf_corr, S(stream)));
        HIP_TRY(ctx->convolve3d(ctx->psf_corr, ctx->psf_corr, S(stream)));
        HIP_TRY(hrrt::launch_update_image(ctx->dg, d_image, ctx->psf_corr, d_normfac_subset, S(stream)));
    } else {
        HIP_TRY(hrrt::launch_rotate_update(ctx->dg, b, ctx->bufA, ctx->bufB, d_image, d_normfac_subset, 0,
                                           S(stream)));
    }
    ctx->mark(-1, S(stream));
    return HRRT_OK;
}

// ---------------------------------------------------------------------------
// Batched dynamic frames: every frame of a study runs the same subset in one
// launch per kernel (the frame is a grid dimension), frames at fixed strides.
// ---------------------------------------------------------------------------
static hrrt::PairBatch frame_batch(const hrrt_ctx *ctx, int isubset, int nframes) {
    const hrrt::Geometry &G = ctx->geo;
    hrrt::PairBatch b{isubset, G.p.subsets, ctx->npairs()};
    b.nframes = nframes;
    b.img_stride = (long long)ctx->vol_floats();
    b.sino_stride = (long long)(G.views / 2) * ctx->col_floats();
    b.ratio_stride = (long long)ctx->col_floats() * ctx->npairs();
    b.buf_stride = (long long)ctx->vol_floats() * ctx->npairs();
    return b;
}

// precision form of a launch: plain FP32 (bit-exact) or FP16 per-bin inputs with Kahan accumulation
struct Precision {
    bool mixed = false;
    const uint16_t *invnorm16 = nullptr;   // for the log-likelihood
    double *loglik = nullptr;              // [nframes] (NULL: not computed)
};

static int frames_subset_launch(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac_subset,
                                const int16_t *d_prompts, const void *d_adds, int isubset, hipStream_t s,
                                const Precision &pr = Precision()) {
    const hrrt::PairBatch b = frame_batch(ctx, isubset, nframes);
    hrrt::FpIO io;
    io.prompt = d_prompts;
    io.add = d_adds;
    if (pr.mixed && pr.loglik) {
        io.invnorm = pr.invnorm16;
        io.ll_part = ctx->ll_part;
    }
    ctx->mark(0, s);
    HIP_TRY(hrrt::launch_forward(ctx->dg, b, d_images, ctx->f_ratio, io, pr.mixed ? hrrt::FP_RATIO16 : hrrt::FP_RATIO, s));
    if (pr.mixed && pr.loglik) {
        int maxt, nsplit, threads, items;
        hrrt::forward_shape(ctx->dg, maxt, nsplit, threads, items);
        HIP_TRY(hrrt::launch_ll_reduce(ctx->ll_part, ctx->dg.c * nsplit * b.npairs, nframes, pr.loglik, s));
    }
    // the subset update back-projects under the schedule in both forms (incremental by default);
    // the precision form sums the view pairs of the rotation with compensation
    ctx->mark(1, s);
    HIP_TRY(backproject_scheduled(ctx, b, ctx->f_ratio, 0, 0, ctx->fA, ctx->fB, s));
    ctx->mark(2, s);
    HIP_TRY(hrrt::launch_rotate_update(ctx->dg, b, ctx->fA, ctx->fB, d_images, d_normfac_subset, pr.mixed, s));
    ctx->mark(-1, s);
    return HRRT_OK;
}

static int frames_check(hrrt_ctx *ctx, int nframes, const void *a, const void *b, const void *c, const void *d) {
    if (!ctx || !a || !b || !c || !d) return fail(HRRT_ERR_ARG, "null argument");
    if (nframes < 1) return fail(HRRT_ERR_ARG, "nframes must be >= 1");
    if (ctx->psf_on)
        return fail(HRRT_ERR_UNSUPPORTED, "the PSF model is available for the single-frame FP32 path only");
    const long long blocks = (long long)ctx->geo.N / 2 * 64 * ctx->npairs() * nframes;   // forward grid bound
    if (blocks > 0x7fffffffLL || (long long)ctx->npairs() * nframes > 65535)
        return fail(HRRT_ERR_ARG, "too many frames for one launch (%d)", nframes);
    int r;
    if ((r = ctx->ensure_frame_scratch(nframes))) return r;
    int maxt, nsplit, threads, items;
    hrrt::forward_shape(ctx->dg, maxt, nsplit, threads, items);
    return ctx->ensure_ll_part((size_t)ctx->dg.c * nsplit * ctx->npairs() * nframes);
}

extern "C" int hrrt_osem_subset_frames(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac_subset,
                                       const int16_t *d_prompts, const float *d_adds, int isubset, void *stream) {
    int r;
    if ((r = frames_check(ctx, nframes, d_images, d_normfac_subset, d_prompts, d_adds))) return r;
    if (isubset < 0 || isubset >= ctx->geo.p.subsets) return fail(HRRT_ERR_ARG, "subset %d out of range", isubset);
    return frames_subset_launch(ctx, nframes, d_images, d_normfac_subset, d_prompts, d_adds, isubset, S(stream));
}

static int iteration_frames(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac,
                            const int16_t *d_prompts, const void *d_adds, void *stream, bool mixed,
                            const uint16_t *d_invnorm16, double *d_loglik) {
    int r;
    if ((r = frames_check(ctx, nframes, d_images, d_normfac, d_prompts, d_adds))) return r;
    if (!stream) return fail(HRRT_ERR_ARG, "graph replay needs a non-default stream");
    if (d_loglik && !d_invnorm16) return fail(HRRT_ERR_ARG, "the log-likelihood needs the 1/(N*A) sinogram");
    const hrrt::Geometry &G = ctx->geo;
    hipStream_t s = S(stream);
    hrrt_ctx::GraphKey key;
    key.nframes = nframes; key.img = d_images; key.nf = d_normfac; key.prompt = d_prompts; key.add = d_adds;
    key.stream = stream; key.mixed = mixed; key.inv = d_invnorm16; key.ll = d_loglik;
    if (!(key == ctx->gkey) || (int)ctx->gexec.size() != G.p.subsets) {
        ctx->drop_graphs();
        for (int i = 0; i < G.p.subsets; i++) {
            const int isub = G.order[i];
            Precision pr;
            pr.mixed = mixed;
            pr.invnorm16 = d_invnorm16;
            pr.loglik = d_loglik ? d_loglik + (size_t)isub * nframes : nullptr;
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = frames_subset_launch(ctx, nframes, d_images, d_normfac + (size_t)isub * ctx->vol_floats(),
                                          d_prompts, d_adds, isub, s, pr);
            hipError_t e = hipStreamEndCapture(s, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                ctx->drop_graphs();
                return rc;
            }
            if (e != hipSuccess) {
                ctx->drop_graphs();
                return fail(HRRT_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            }
            hipGraphExec_t ex = nullptr;
            e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                ctx->drop_graphs();
                return fail(HRRT_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
            }
            ctx->gexec.push_back(ex);
        }
        ctx->gkey = key;
    }
    for (int i = 0; i < G.p.subsets; i++) HIP_TRY(hipGraphLaunch(ctx->gexec[i], s));
    for (int f = 0; f < nframes; f++)
        HIP_TRY(hrrt::launch_end_iteration(ctx->dg, d_images + (size_t)f * ctx->vol_floats(), G.z_min, G.z_max, s));
    return HRRT_OK;
}

extern "C" int hrrt_osem_iteration_frames(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac,
                                          const int16_t *d_prompts, const float *d_adds, void *stream) {
    return iteration_frames(ctx, nframes, d_images, d_normfac, d_prompts, d_adds, stream, false, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// Precision form: FP16 per-bin inputs, Kahan accumulation, FP64 log-likelihood
// and count normalisation.
// ---------------------------------------------------------------------------
extern "C" int hrrt_to_half(const float *d_in, uint16_t *d_out, long long n, void *stream) {
    if (!d_in || !d_out || n < 0) return fail(HRRT_ERR_ARG, "invalid argument");
    HIP_TRY(hrrt::launch_to_half(d_in, d_out, (size_t)n, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_to_half_range(const float *d_in, uint16_t *d_out, long long n, unsigned long long *d_counts,
                                  void *stream) {
    if (!d_in || !d_out || !d_counts || n < 0) return fail(HRRT_ERR_ARG, "invalid argument");
    HIP_TRY(hrrt::launch_to_half_range(d_in, d_out, (size_t)n, d_counts, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_calculate_normfac_mixed(hrrt_ctx *ctx, const uint16_t *d_invnorm16, float *d_normfac,
                                            void *stream) {
    if (!ctx || !d_invnorm16 || !d_normfac) return fail(HRRT_ERR_ARG, "null argument");
    if (ctx->psf_on)
        return fail(HRRT_ERR_UNSUPPORTED, "the PSF model is available for the single-frame FP32 path only");
    int r;
    if ((r = ctx->ensure_subset_scratch())) return r;
    const hrrt::Geometry &G = ctx->geo;
    for (int s = 0; s < G.p.subsets; s++) {
        hrrt::PairBatch b{s, G.p.subsets, ctx->npairs()};
        HIP_TRY(hrrt::launch_backproject(ctx->dg, b, d_invnorm16, 1, 1, 1, 1, ctx->bufA, ctx->bufB, S(stream)));
        HIP_TRY(hrrt::launch_rotate_sens(ctx->dg, b, ctx->bufA, ctx->bufB, d_normfac + (size_t)s * ctx->vol_floats(),
                                         1, S(stream)));
    }
    return HRRT_OK;
}

extern "C" int hrrt_osem_subset_frames_mixed(hrrt_ctx *ctx, int nframes, float *d_images,
                                             const float *d_normfac_subset, const int16_t *d_prompts,
                                             const uint16_t *d_adds16, const uint16_t *d_invnorm16, int isubset,
                                             double *d_loglik, void *stream) {
    int r;
    if ((r = frames_check(ctx, nframes, d_images, d_normfac_subset, d_prompts, d_adds16))) return r;
    if (isubset < 0 || isubset >= ctx->geo.p.subsets) return fail(HRRT_ERR_ARG, "subset %d out of range", isubset);
    if (d_loglik && !d_invnorm16) return fail(HRRT_ERR_ARG, "the log-likelihood needs the 1/(N*A) sinogram");
    Precision pr;
    pr.mixed = true;
    pr.invnorm16 = d_invnorm16;
    pr.loglik = d_loglik;
    return frames_subset_launch(ctx, nframes, d_images, d_normfac_subset, d_prompts, d_adds16, isubset, S(stream), pr);
}

extern "C" int hrrt_osem_subset_mixed(hrrt_ctx *ctx, float *d_image, const float *d_normfac_subset,
                                      const int16_t *d_prompt, const uint16_t *d_add16, const uint16_t *d_invnorm16,
                                      int isubset, double *d_loglik, void *stream) {
    return hrrt_osem_subset_frames_mixed(ctx, 1, d_image, d_normfac_subset, d_prompt, d_add16, d_invnorm16, isubset,
                                         d_loglik, stream);
}

extern "C" int hrrt_osem_iteration_frames_mixed(hrrt_ctx *ctx, int nframes, float *d_images, const float *d_normfac,
                                                const int16_t *d_prompts, const uint16_t *d_adds16,
                                                const uint16_t *d_invnorm16, double *d_loglik, void *stream) {
    return iteration_frames(ctx, nframes, d_images, d_normfac, d_prompts, d_adds16, stream, true, d_invnorm16,
                            d_loglik);
}

extern "C" int hrrt_count_moments(hrrt_ctx *ctx, const float *d_image, const int16_t *d_prompt, const float *d_add,
                                  const float *d_invnorm, double *d_out, void *stream) {
    if (!ctx || !d_image || !d_prompt || !d_add || !d_invnorm || !d_out) return fail(HRRT_ERR_ARG, "null argument");
    int r;
    if ((r = ctx->ensure_subset_scratch())) return r;
    const hrrt::Geometry &G = ctx->geo;
    if ((r = ctx->ensure_ll_part((size_t)3 * G.N * 2 * ctx->npairs()))) return r;
    HIP_TRY(hipMemsetAsync(d_out, 0, 3 * sizeof(double), S(stream)));
    hrrt::FpIO io;
    for (int s = 0; s < G.p.subsets; s++) {
        hrrt::PairBatch b{s, G.p.subsets, ctx->npairs()};
        HIP_TRY(hrrt::launch_forward(ctx->dg, b, d_image, ctx->sub_ratio, io, hrrt::FP_PLAIN, S(stream)));
        HIP_TRY(hrrt::launch_count_moments(ctx->dg, d_prompt, d_add, d_invnorm, ctx->sub_ratio, s, G.p.subsets,
                                           ctx->npairs(), ctx->ll_part, d_out, S(stream)));
    }
    return HRRT_OK;
}

extern "C" int hrrt_set_launch_shape(hrrt_ctx *ctx, int fp_tasks_per_thread, int fp_workgroups_per_ray) {
    if (!ctx) return fail(HRRT_ERR_ARG, "null context");
    if (fp_tasks_per_thread < 1 || fp_tasks_per_thread > 4 || fp_workgroups_per_ray < 0 || fp_workgroups_per_ray > 64)
        return fail(HRRT_ERR_ARG, "launch shape out of range");
    ctx->dg.fp_maxt = fp_tasks_per_thread;
    ctx->dg.fp_nsplit = fp_workgroups_per_ray;
    ctx->drop_graphs();
    return HRRT_OK;
}

extern "C" int hrrt_set_backproject_shape(hrrt_ctx *ctx, int waves_per_workgroup) {
    if (!ctx) return fail(HRRT_ERR_ARG, "null context");
    if (waves_per_workgroup < 1 || waves_per_workgroup > 16) return fail(HRRT_ERR_ARG, "launch shape out of range");
    ctx->dg.bp_waves = waves_per_workgroup;
    ctx->drop_graphs();
    return HRRT_OK;
}

extern "C" int hrrt_timing_enable(hrrt_ctx *ctx, int enable) {
    if (!ctx) return fail(HRRT_ERR_ARG, "null context");
    ctx->timing = enable != 0;
    ctx->ev_used = 0;
    ctx->ev_marks.clear();
    return HRRT_OK;
}

extern "C" int hrrt_timing_read(hrrt_ctx *ctx, double *ms, int *count) {
    if (!ctx || !ms || !count) return fail(HRRT_ERR_ARG, "null argument");
    for (int k = 0; k < HRRT_NKERNELS; k++) { ms[k] = 0.0; count[k] = 0; }
    for (auto &m : ctx->ev_marks) {
        HIP_TRY(hipEventSynchronize(ctx->ev_pool[m.second + 1]));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, ctx->ev_pool[m.second], ctx->ev_pool[m.second + 1]));
        ms[m.first] += t;
        count[m.first] += 1;
    }
    ctx->ev_used = 0;
    ctx->ev_marks.clear();
    return HRRT_OK;
}

extern "C" int hrrt_forward_proj3d(hrrt_ctx *ctx, const float *d_image, float *d_prj, int view, void *stream) {
    if (!ctx || !d_image || !d_prj) return fail(HRRT_ERR_ARG, "null argument");
    if (view < 0 || view >= ctx->geo.views / 2) return fail(HRRT_ERR_ARG, "view %d out of range", view);
    int r;
    if ((r = ctx->ensure_pair_scratch())) return r;
    hrrt::PairBatch b{view, 0, 1};
    HIP_TRY(hrrt::launch_forward(ctx->dg, b, d_image, ctx->pc1, hrrt::FpIO(), hrrt::FP_PLAIN, S(stream)));
    HIP_TRY(hrrt::launch_paircol_to_estimate(ctx->dg, view, ctx->pc1, d_prj, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_forward_proj3d_noalloc(hrrt_ctx *ctx, const float *d_image, float *d_prj, int view,
                                           void *stream) {
    if (!ctx || !d_image || !d_prj) return fail(HRRT_ERR_ARG, "null argument");
    if (view < 0 || view >= ctx->geo.views / 2) return fail(HRRT_ERR_ARG, "view %d out of range", view);
    if (ctx->col_floats() > (size_t)HRRT_NOALLOC_SCRATCH_FLOATS)
        return fail(HRRT_ERR_ARG, "a pair-column of this geometry (%zu floats) exceeds the %d-float scratch",
                    ctx->col_floats(), HRRT_NOALLOC_SCRATCH_FLOATS);
    if (!ctx->noalloc_scratch)
        return fail(HRRT_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(ctx->noalloc_err));
    hrrt::PairBatch b{view, 0, 1};
    HIP_TRY(hrrt::launch_forward(ctx->dg, b, d_image, ctx->noalloc_scratch, hrrt::FpIO(), hrrt::FP_PLAIN, S(stream)));
    HIP_TRY(hrrt::launch_paircol_to_estimate(ctx->dg, view, ctx->noalloc_scratch, d_prj, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_back_proj3d(hrrt_ctx *ctx, float *d_prj, float *d_ima, int view, void *stream) {
    if (!ctx || !d_prj || !d_ima) return fail(HRRT_ERR_ARG, "null argument");
    if (view < 0 || view >= ctx->geo.views / 2) return fail(HRRT_ERR_ARG, "view %d out of range", view);
    int r;
    if ((r = ctx->ensure_pair_scratch())) return r;
    hrrt::PairBatch b{view, 0, 1};
    HIP_TRY(hrrt::launch_estimate_to_paircol(ctx->dg, d_prj, ctx->pc1, S(stream)));
    HIP_TRY(hrrt::launch_backproject(ctx->dg, b, ctx->pc1, 0, 0, 0, 0, ctx->bufA, ctx->bufB, S(stream)));
    HIP_TRY(hrrt::launch_rotate_add(ctx->dg, view, ctx->bufA, ctx->bufB, d_ima, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_back_proj3d_incremental(hrrt_ctx *ctx, float *d_prj, float *d_ima, int view, void *stream) {
    if (!ctx || !d_prj || !d_ima) return fail(HRRT_ERR_ARG, "null argument");
    if (view < 0 || view >= ctx->geo.views / 2) return fail(HRRT_ERR_ARG, "view %d out of range", view);
    if (!ctx->dg.inc_ok) return fail(HRRT_ERR_UNSUPPORTED, "incremental back projection is not available for this geometry");
    int r;
    if ((r = ctx->ensure_pair_scratch())) return r;
    hrrt::PairBatch b{view, 0, 1};
    HIP_TRY(hrrt::launch_estimate_to_paircol(ctx->dg, d_prj, ctx->pc1, S(stream)));
    HIP_TRY(hrrt::launch_backproject_incremental(ctx->dg, b, ctx->pc1, 0, 0, ctx->bufA, ctx->bufB, S(stream)));
    HIP_TRY(hrrt::launch_rotate_add(ctx->dg, view, ctx->bufA, ctx->bufB, d_ima, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_back_proj3d_rotframe(hrrt_ctx *ctx, float *d_prj, float *d_imagebuf, int view, int incremental,
                                         void *stream) {
    if (!ctx || !d_prj || !d_imagebuf) return fail(HRRT_ERR_ARG, "null argument");
    if (view < 0 || view >= ctx->geo.views / 2) return fail(HRRT_ERR_ARG, "view %d out of range", view);
    if (incremental && !ctx->dg.inc_ok)
        return fail(HRRT_ERR_UNSUPPORTED, "incremental back projection is not available for this geometry");
    int r;
    if ((r = ctx->ensure_pair_scratch())) return r;
    hrrt::PairBatch b{view, 0, 1};
    HIP_TRY(hrrt::launch_estimate_to_paircol(ctx->dg, d_prj, ctx->pc1, S(stream)));
    if (incremental) HIP_TRY(hrrt::launch_backproject_incremental(ctx->dg, b, ctx->pc1, 0, 0, ctx->bufA, ctx->bufB, S(stream)));
    else HIP_TRY(hrrt::launch_backproject(ctx->dg, b, ctx->pc1, 0, 0, 0, 0, ctx->bufA, ctx->bufB, S(stream)));
    HIP_TRY(hrrt::launch_halves_sum(ctx->dg, ctx->bufA, ctx->bufB, d_imagebuf, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_rotate_add(hrrt_ctx *ctx, const float *d_imagebuf, float *d_ima, int view, void *stream) {
    if (!ctx || !d_imagebuf || !d_ima) return fail(HRRT_ERR_ARG, "null argument");
    if (view < 0 || view >= ctx->geo.views / 2) return fail(HRRT_ERR_ARG, "view %d out of range", view);
    int r;
    if ((r = ctx->zalloc((void **)&ctx->rot_zero, sizeof(float) * ctx->vol_floats()))) return r;
    HIP_TRY(hrrt::launch_rotate_add(ctx->dg, view, d_imagebuf, ctx->rot_zero, d_ima, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_set_backproject_schedule(hrrt_ctx *ctx, int nthreads) {
    if (!ctx) return fail(HRRT_ERR_ARG, "null context");
    if (nthreads < 0) return fail(HRRT_ERR_ARG, "nthreads must be >= 0");
    if (nthreads > 0 && !ctx->dg.inc_ok)
        return fail(HRRT_ERR_UNSUPPORTED, "incremental back projection is not available for this geometry");
    ctx->bp_threads = nthreads;
    ctx->drop_graphs();
    return HRRT_OK;
}

// make1d_psf_filter psf.cpp:20-70: the sum of two Gaussians evaluated in double, stored in float
// and normalised in float; fwhm1 == 0 selects the HRRT (model 999) defaults
static int psf_kernel(float fwhm1, float fwhm2, float ratio, float pixel_size, std::vector<float> &k) {
    if (fwhm1 == 0.0f) {
        fwhm1 = (float)(0.9 * sqrt(8.0 * log(2.0)));
        fwhm2 = (float)(2.5 * sqrt(8.0 * log(2.0)));
        ratio = (float)0.05;
    }
    if (!(fwhm1 > 0.0f)) return -1;
    const bool two = fwhm2 > 0.0f && ratio > 0.0f;
    const int size = 2 * ((int)(1.5 * (two ? fwhm2 : fwhm1) * sqrt(log(10.0) / log(2.)) / pixel_size / 2.0)) + 1;
    if (size < 1 || size > 1023) return -1;
    k.assign(size, 0.0f);
    const float sigma1 = (float)(fwhm1 / (sqrt(8.0 * log(2.0)) * pixel_size));
    const float sigma2 = (float)(fwhm2 / (sqrt(8.0 * log(2.0)) * pixel_size));
    const int bin0 = (size - 1) / 2;
    float norm = 0.0f;
    for (int bin = 0; bin < size; bin++) {
        k[bin] = (float)exp(-0.5 * (pow((double)(bin - bin0) / sigma1, 2.0)));
        if (two) k[bin] = (float)(k[bin] + ratio * exp(-0.5 * (pow((double)(bin - bin0) / sigma2, (double)2.0))));
        norm += k[bin];
    }
    for (int bin = 0; bin < size; bin++) k[bin] /= norm;
    return size;
}

extern "C" int hrrt_set_psf(hrrt_ctx *ctx, int enable, float fwhm1, float fwhm2, float ratio) {
    if (!ctx) return fail(HRRT_ERR_ARG, "null context");
    ctx->drop_graphs();
    if (!enable) {
        ctx->psf_on = false;
        return HRRT_OK;
    }
    const hrrt::Geometry &G = ctx->geo;
    std::vector<float> k[3];
    const float px[3] = {G.x_size, G.y_size, G.z_size};
    for (int a = 0; a < 3; a++)
        if (psf_kernel(fwhm1, fwhm2, ratio, px[a], k[a]) < 1)
            return fail(HRRT_ERR_ARG, "invalid PSF parameters (fwhm1 %g, fwhm2 %g, ratio %g)", fwhm1, fwhm2, ratio);
    // the z pass reads the z kernel at the y kernel's indices (psf.cpp:356-374)
    if (k[2].size() < k[1].size())
        return fail(HRRT_ERR_UNSUPPORTED, "z kernel (%zu taps) shorter than the y kernel (%zu taps)", k[2].size(),
                    k[1].size());
    const int hy = ((int)k[1].size() - 1) / 2;
    std::vector<float> normz(G.z4, 0.0f);
    for (int mask = -hy; mask < hy; mask++)
        for (int z = 0; z < G.z4; z++) {
            const int m2 = z + mask;
            if (m2 < G.z_min || m2 >= G.z_max) continue;
            normz[z] += k[2][mask + hy];
        }
    for (int z = 0; z < G.z4; z++)
        if (normz[z] > 0) normz[z] = 1 / normz[z];
    std::vector<float> packed(k[0]);
    packed.insert(packed.end(), k[1].begin(), k[1].end());
    packed.insert(packed.end(), normz.begin(), normz.end());
    if (ctx->d_psf) (void)hipFree(ctx->d_psf);
    ctx->d_psf = nullptr;
    if (hipMalloc((void **)&ctx->d_psf, sizeof(float) * packed.size()) != hipSuccess)
        return fail(HRRT_ERR_NOMEM, "hipMalloc of the PSF kernels failed");
    HIP_TRY(hipMemcpy(ctx->d_psf, packed.data(), sizeof(float) * packed.size(), hipMemcpyHostToDevice));
    for (int a = 0; a < 3; a++) ctx->psf_k[a] = k[a];
    ctx->psf_normz = normz;
    ctx->psf_on = true;
    return HRRT_OK;
}

extern "C" int hrrt_get_psf_kernel(const hrrt_ctx *ctx, int axis, float *taps, int *size) {
    if (!ctx || !size) return fail(HRRT_ERR_ARG, "null argument");
    if (axis < 0 || axis > 2) return fail(HRRT_ERR_ARG, "axis %d out of range", axis);
    if (!ctx->psf_on) return fail(HRRT_ERR_ARG, "the PSF model is not enabled");
    const std::vector<float> &k = ctx->psf_k[axis];
    if (taps) std::memcpy(taps, k.data(), sizeof(float) * k.size());
    *size = (int)k.size();
    return HRRT_OK;
}

extern "C" int hrrt_convolve3d(hrrt_ctx *ctx, const float *d_in, float *d_out, void *stream) {
    if (!ctx || !d_in || !d_out) return fail(HRRT_ERR_ARG, "null argument");
    if (!ctx->psf_on) return fail(HRRT_ERR_ARG, "the PSF model is not enabled (hrrt_set_psf)");
    int r;
    if ((r = ctx->ensure_psf_scratch())) return r;
    HIP_TRY(ctx->convolve3d(d_in, d_out, S(stream)));
    return HRRT_OK;
}

// ---------------------------------------------------------------------------
// Image and sensitivity files (SaveImage S3:4016-4096, write_flat_image FIO:300-333,
// Init_Image -I S3:3970-3993, write_norm / read_norm FIO:360-427, write_image_header
// write_image_header.cpp:67-319)
// ---------------------------------------------------------------------------
extern "C" int hrrt_image_to_flat(hrrt_ctx *ctx, const float *d_image, float *d_flat, int flip, void *stream) {
    if (!ctx || !d_image || !d_flat) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_flat_image(ctx->dg, d_image, d_flat, 1, flip ? 1 : 0, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_image_from_flat(hrrt_ctx *ctx, const float *d_flat, float *d_image, int flip, void *stream) {
    if (!ctx || !d_image || !d_flat) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_flat_image(ctx->dg, d_flat, d_image, 0, flip ? 1 : 0, S(stream)));
    return HRRT_OK;
}

static size_t flat_floats(const hrrt_ctx *ctx) { return (size_t)ctx->geo.z_pixels * ctx->geo.N * ctx->geo.N; }

static const char *base_name(const char *path) {
    const char *p = std::strrchr(path, '/');
    return p ? p + 1 : path;
}

// scatter fraction from the scatter file's Interfile header (<name>.h33), 0 when absent
static float scatter_fraction(const char *scatter_file) {
    std::string hdr(scatter_file);
    const size_t dot = hdr.rfind('.');
    if (dot != std::string::npos && hdr.find('/', dot) == std::string::npos) hdr.resize(dot);
    hdr += ".h33";
    FILE *fp = std::fopen(hdr.c_str(), "rt");
    if (!fp) return 0.0f;
    char line[1024];
    float scf = 0.0f;
    while (std::fgets(line, sizeof line, fp)) {
        if (!std::strstr(line, "scatter fraction")) continue;
        const char *p = std::strstr(line, ":=");
        if (p && std::sscanf(p + 2, "%g", &scf) == 1) break;
    }
    std::fclose(fp);
    return scf;
}

static int write_header(const hrrt_ctx *ctx, const char *image_path, const hrrt_image_info *info) {
    const hrrt::Geometry &G = ctx->geo;
    std::string name(base_name(image_path)), ext;
    const size_t dot = name.rfind('.');
    if (dot != std::string::npos) {
        ext = name.substr(dot);
        name.resize(dot);
    }
    const std::string hdr = std::string(image_path) + ".hdr";
    FILE *out = std::fopen(hdr.c_str(), "wt");
    if (!out) return fail(HRRT_ERR_ARG, "cannot open image header file '%s'", hdr.c_str());
    FILE *sino = nullptr;      // lines of the sinogram's header are carried forward
    const char *sino_file = (info && info->true_file && *info->true_file) ? info->true_file
                            : (info && info->prompt_file && *info->prompt_file) ? info->prompt_file
                                                                                : nullptr;
    if (sino_file) sino = std::fopen((std::string(sino_file) + ".hdr").c_str(), "r");
    std::fprintf(out, "!INTERFILE\n");
    std::fprintf(out, "name of data file := %s%s\n", name.c_str(), ext.c_str());
    std::fprintf(out, "data format := image\n");
    std::fprintf(out, "number format := %s\n", "float");
    std::fprintf(out, "number of bytes per pixel := %d\n", 4);
    std::fprintf(out, "number of dimensions := 3\n");
    std::fprintf(out, "matrix size [1] := %d\n", G.N);
    std::fprintf(out, "matrix size [2] := %d\n", G.N);
    std::fprintf(out, "matrix size [3] := %d\n", G.z_pixels);
    std::fprintf(out, "scaling factor (mm/pixel) [1] := %f\n", G.x_size);
    std::fprintf(out, "scaling factor (mm/pixel) [2] := %f\n", G.y_size);
    std::fprintf(out, "scaling factor (mm/pixel) [3] := %f\n", G.z_size);
    static const char *methods[] = {"UW", "AW", "ANW", "OP"};
    const int rt = info ? info->recontype : 3;
    if (rt >= 0 && rt <= 3)
        std::fprintf(out, "reconstruction method := inki-2006-may-OSEM3D-%s%s\n", methods[rt], ctx->psf_on ? "-PSF" : "");
    std::fprintf(out, "reconstruction version := %s\n", (info && info->sw_version) ? info->sw_version : "");
    std::fprintf(out, "reconstruction build ID := %s\n", (info && info->sw_build_id) ? info->sw_build_id : "");
    if (info) {
        const struct { const char *key, *file; } used[] = {
            {"true", info->true_file},   {"prompt", info->prompt_file}, {"delayed", info->delayed_file},
            {"norm", info->norm_file},   {"atten", info->atten_file},   {"scatter", info->scatter_file}};
        for (const auto &u : used) {
            if (!u.file || !*u.file) continue;
            std::fprintf(out, "%s file used := %s\n", u.key, base_name(u.file));
            if (u.file == info->scatter_file) {
                const float scf = scatter_fraction(u.file);
                if (scf > 0) std::fprintf(out, "scatter fraction := %g\n", scf);
            }
        }
    }
    std::fprintf(out, "number of iterations := %d\n", info ? info->iterations : 0);
    std::fprintf(out, "number of subsets := %d\n", G.p.subsets);
    if (sino) {
        static const char *skip[] = {"name of data file",    "data format",       "number format",
                                     "number of bytes per",  "number of dimensions", "matrix size",
                                     "scaling factor",       "reconstruction method", "true file used",
                                     "prompt file used",     "delayed file used", "norm file used",
                                     "atten file used",      "scatter file used", "number of iterations",
                                     "number of subsets",    "CPSCC-OSEM3D build", "CPSCC-Server build"};
        char line[1024];
        while (std::fgets(line, sizeof line, sino)) {
            if (!std::strncmp(line, "!INTERFILE", 10) || !std::strncmp(line, "!Interfile", 10) ||
                !std::strncmp(line, "!interfile", 10))
                continue;
            bool drop = false;
            for (const char *k : skip) drop = drop || std::strstr(line, k) != nullptr;
            if (!drop) std::fprintf(out, "%s", line);
        }
        std::fclose(sino);
    }
    std::fclose(out);
    return HRRT_OK;
}

extern "C" int hrrt_write_image(hrrt_ctx *ctx, const float *d_image, const char *path, int flip,
                                const hrrt_image_info *info, void *stream) {
    if (!ctx || !d_image || !path) return fail(HRRT_ERR_ARG, "null argument");
    int r;
    if ((r = ctx->zalloc((void **)&ctx->flat, sizeof(float) * flat_floats(ctx)))) return r;
    std::vector<float> host(flat_floats(ctx));
    HIP_TRY(hrrt::launch_flat_image(ctx->dg, d_image, ctx->flat, 1, flip ? 1 : 0, S(stream)));
    HIP_TRY(hipMemcpyAsync(host.data(), ctx->flat, sizeof(float) * host.size(), hipMemcpyDeviceToHost, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(HRRT_ERR_ARG, "cannot open '%s'", path);
    const size_t plane = (size_t)ctx->geo.N * ctx->geo.N;
    for (int z = 0; z < ctx->geo.z_pixels; z++)     // plane by plane, as write_flat_image does
        if (std::fwrite(host.data() + z * plane, sizeof(float), plane, f) != plane) {
            std::fclose(f);
            return fail(HRRT_ERR_ARG, "write error on '%s'", path);
        }
    std::fclose(f);
    return write_header(ctx, path, info);
}

extern "C" int hrrt_read_image(hrrt_ctx *ctx, const char *path, float *d_image, int flip, void *stream) {
    if (!ctx || !d_image || !path) return fail(HRRT_ERR_ARG, "null argument");
    int r;
    if ((r = ctx->zalloc((void **)&ctx->flat, sizeof(float) * flat_floats(ctx)))) return r;
    std::vector<float> host(flat_floats(ctx));
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(HRRT_ERR_ARG, "cannot open initial image '%s'", path);
    const size_t got = std::fread(host.data(), sizeof(float), host.size(), f);
    std::fclose(f);
    if (got != host.size()) return fail(HRRT_ERR_ARG, "'%s' holds %zu floats, %zu expected", path, got, host.size());
    HIP_TRY(hipMemcpyAsync(ctx->flat, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice, S(stream)));
    HIP_TRY(hrrt::launch_flat_image(ctx->dg, ctx->flat, d_image, 0, flip ? 1 : 0, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_write_normfac(hrrt_ctx *ctx, const float *d_normfac, const char *path, void *stream) {
    if (!ctx || !d_normfac || !path) return fail(HRRT_ERR_ARG, "null argument");
    const size_t n = ctx->vol_floats();
    std::vector<float> host(n);
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(HRRT_ERR_ARG, "cannot open '%s'", path);
    for (int s = 0; s < ctx->geo.p.subsets; s++) {   // subset after subset, [X][Y][Z4] each
        if (hipMemcpyAsync(host.data(), d_normfac + s * n, sizeof(float) * n, hipMemcpyDeviceToHost, S(stream)) !=
                hipSuccess ||
            hipStreamSynchronize(S(stream)) != hipSuccess) {
            std::fclose(f);
            return fail(HRRT_ERR_HIP, "copy of sensitivity subset %d failed", s);
        }
        if (std::fwrite(host.data(), sizeof(float), n, f) != n) {
            std::fclose(f);
            return fail(HRRT_ERR_ARG, "write error on '%s'", path);
        }
    }
    std::fclose(f);
    return HRRT_OK;
}

extern "C" int hrrt_read_normfac(hrrt_ctx *ctx, const char *path, float *d_normfac, void *stream) {
    if (!ctx || !d_normfac || !path) return fail(HRRT_ERR_ARG, "null argument");
    const size_t n = ctx->vol_floats();
    std::vector<float> host(n);
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(HRRT_ERR_ARG, "cannot open '%s'", path);
    for (int s = 0; s < ctx->geo.p.subsets; s++) {
        if (std::fread(host.data(), sizeof(float), n, f) != n) {
            std::fclose(f);
            return fail(HRRT_ERR_ARG, "'%s' ends before subset %d", path, s);
        }
        if (hipMemcpyAsync(d_normfac + s * n, host.data(), sizeof(float) * n, hipMemcpyHostToDevice, S(stream)) !=
                hipSuccess ||
            hipStreamSynchronize(S(stream)) != hipSuccess) {
            std::fclose(f);
            return fail(HRRT_ERR_HIP, "copy of sensitivity subset %d failed", s);
        }
    }
    std::fclose(f);
    return HRRT_OK;
}

extern "C" int hrrt_get_backproject_schedule(const hrrt_ctx *ctx, int *nthreads) {
    if (!ctx || !nthreads) return fail(HRRT_ERR_ARG, "null argument");
    *nthreads = ctx->bp_threads;
    return HRRT_OK;
}

extern "C" int hrrt_update_estimate_w3(hrrt_ctx *ctx, float *d_est, const float *d_add, const int16_t *d_prompt,
                                       void *stream) {
    if (!ctx || !d_est || !d_add || !d_prompt) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_ratio_w3(ctx->dg, d_est, d_add, d_prompt, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_update_estimate_w012(hrrt_ctx *ctx, float *d_est, const float *d_data, void *stream) {
    if (!ctx || !d_est || !d_data) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_ratio_w012(ctx->dg, d_est, d_data, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_cal_updateimage(hrrt_ctx *ctx, float *d_image, float *d_correction, const float *d_normfac,
                                    void *stream) {
    if (!ctx || !d_image || !d_correction || !d_normfac) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_update_image(ctx->dg, d_image, d_correction, d_normfac, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_normfac_calculate(hrrt_ctx *ctx, const float *d_sens, float *d_out, void *stream) {
    if (!ctx || !d_sens || !d_out) return fail(HRRT_ERR_ARG, "null argument");
    HIP_TRY(hrrt::launch_normfac_invert(ctx->dg, d_sens, d_out, S(stream)));
    return HRRT_OK;
}

extern "C" int hrrt_forward_proj3d_host(hrrt_ctx *ctx, const float *image, float *prj, int view) {
    if (!ctx || !image || !prj) return fail(HRRT_ERR_ARG, "null argument");
    int r;
    const size_t ni = ctx->vol_floats() * sizeof(float);
    const size_t np = (size_t)2 * ctx->geo.segmentsize * ctx->geo.N * sizeof(float);
    if ((r = ctx->zalloc((void **)&ctx->h_img, ni)) || (r = ctx->zalloc((void **)&ctx->h_prj, np))) return r;
    HIP_TRY(hipMemcpy(ctx->h_img, image, ni, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->h_prj, prj, np, hipMemcpyHostToDevice));
    if ((r = hrrt_forward_proj3d(ctx, ctx->h_img, ctx->h_prj, view, nullptr))) return r;
    HIP_TRY(hipMemcpy(prj, ctx->h_prj, np, hipMemcpyDeviceToHost));
    return HRRT_OK;
}

extern "C" int hrrt_back_proj3d_host(hrrt_ctx *ctx, float *prj, float *ima, int view) {
    if (!ctx || !prj || !ima) return fail(HRRT_ERR_ARG, "null argument");
    int r;
    const size_t ni = ctx->vol_floats() * sizeof(float);
    const size_t np = (size_t)2 * ctx->geo.segmentsize * ctx->geo.N * sizeof(float);
    if ((r = ctx->zalloc((void **)&ctx->h_img, ni)) || (r = ctx->zalloc((void **)&ctx->h_prj, np))) return r;
    HIP_TRY(hipMemcpy(ctx->h_img, ima, ni, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->h_prj, prj, np, hipMemcpyHostToDevice));
    if ((r = hrrt_back_proj3d(ctx, ctx->h_prj, ctx->h_img, view, nullptr))) return r;
    HIP_TRY(hipMemcpy(prj, ctx->h_prj, np, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ima, ctx->h_img, ni, hipMemcpyDeviceToHost));
    return HRRT_OK;
}

extern "C" int hrrt_back_proj3d_rotframe_host(hrrt_ctx *ctx, float *prj, float *imagebuf, int view, int incremental) {
    if (!ctx || !prj || !imagebuf) return fail(HRRT_ERR_ARG, "null argument");
    const size_t ni = sizeof(float) * ctx->vol_floats(), np = sizeof(float) * 2 * (size_t)ctx->geo.segmentsize * ctx->geo.N;
    int r;
    if ((r = ctx->zalloc((void **)&ctx->h_img, ni)) || (r = ctx->zalloc((void **)&ctx->h_prj, np))) return r;
    HIP_TRY(hipMemcpy(ctx->h_img, imagebuf, ni, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->h_prj, prj, np, hipMemcpyHostToDevice));
    if ((r = hrrt_back_proj3d_rotframe(ctx, ctx->h_prj, ctx->h_img, view, incremental, nullptr))) return r;
    HIP_TRY(hipMemcpy(prj, ctx->h_prj, np, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(imagebuf, ctx->h_img, ni, hipMemcpyDeviceToHost));
    return HRRT_OK;
}

extern "C" int hrrt_rotate_add_host(hrrt_ctx *ctx, const float *imagebuf, float *ima, int view) {
    if (!ctx || !imagebuf || !ima) return fail(HRRT_ERR_ARG, "null argument");
    const size_t ni = sizeof(float) * ctx->vol_floats();
    int r;
    if ((r = ctx->zalloc((void **)&ctx->h_img, ni)) || (r = ctx->zalloc((void **)&ctx->h_buf, ni))) return r;
    HIP_TRY(hipMemcpy(ctx->h_buf, imagebuf, ni, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->h_img, ima, ni, hipMemcpyHostToDevice));
    if ((r = hrrt_rotate_add(ctx, ctx->h_buf, ctx->h_img, view, nullptr))) return r;
    HIP_TRY(hipMemcpy(ima, ctx->h_img, ni, hipMemcpyDeviceToHost));
    return HRRT_OK;
}
