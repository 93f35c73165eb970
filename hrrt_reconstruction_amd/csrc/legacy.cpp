// The reference projector's own entry points (hrrt_osem3d.h:197-211) with their C++
// signatures, as marshalling shims over the C ABI of libhrrt_osem.so: a driver written
// against hrrt_osem3dv_sbrt_span3.cpp links this library in its place and keeps its
// float*** images and float** estimates in host memory.
//
//   dependencies / free_dependencies     PJ:199-601 / the matching frees
//   forward_proj3d2, forward_proj3d_thread1   PJ:1271, 1373 (both: hrrt_forward_proj3d)
//   back_proj3d2                         PJ:1457 (direct projector, rotated into ima)
//   back_proj3d_thread1                  PJ:1622 (incremental projector into imagebuf, no rotation)
//   rotate_image_buf                     PJ:1701 (imagebuf[i] rotated by view[i] into ima)
//
// The reference reads its scanner and option globals inside dependencies(); this library
// cannot see a driver's globals, so the driver passes them first through
// hrrt_legacy_set_params (declared in include/hrrt_legacy.h).
//
// Images are gathered column by column (ima[x][y] -> Z4 floats), as the reference allocates
// each [x][y] column on its own (alloc_imagexyzf_thread S3:2628-2663); estimates row by row
// (prj[r] -> N floats, 2*segmentsize rows).  Every call is synchronous and calls are
// serialised by one mutex: the reference calls the *_thread1 variants from several threads
// at once (pt_osem3d_proj_thread1 S3:1910-1938), each with its own imagebuf and prjbuf, and
// those stay independent because every call copies its buffers in and out.
//
// The projectors read the reference's norm_zero_mask global (columns whose normalisation is
// zero in every plane, set by prepare_norm S3:897-909); the driver hands it over with
// hrrt_legacy_set_norm_zero_mask after prepare_norm.  Until then every column counts.
//
// Returns 1 on success as the reference does; 0 with a message on stderr on failure.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "hrrt_legacy.h"
#include "hrrt_osem.h"

namespace {

struct Legacy {
    std::mutex mu;
    hrrt_params params{};
    bool have_params = false;
    hrrt_ctx *ctx = nullptr;
    hrrt_dims dims{};
    float *d_ima = nullptr, *d_buf = nullptr, *d_prj = nullptr;
    std::vector<float> h_ima, h_buf, h_prj;
};

Legacy &state() {
    static Legacy s;
    return s;
}

int fail_msg(const char *where, const char *what) {
    std::fprintf(stderr, "%s: %s\n", where, what);
    return 0;
}

int fail_hip(const char *where, hipError_t e) { return fail_msg(where, hipGetErrorString(e)); }

int fail_ctx(const char *where) { return fail_msg(where, hrrt_last_error()); }

void release(Legacy &L) {
    if (L.d_ima) (void)hipFree(L.d_ima);
    if (L.d_buf) (void)hipFree(L.d_buf);
    if (L.d_prj) (void)hipFree(L.d_prj);
    L.d_ima = L.d_buf = L.d_prj = nullptr;
    if (L.ctx) hrrt_destroy(L.ctx);
    L.ctx = nullptr;
    L.h_ima.clear();
    L.h_buf.clear();
    L.h_prj.clear();
    L.h_ima.shrink_to_fit();
    L.h_buf.shrink_to_fit();
    L.h_prj.shrink_to_fit();
}

size_t image_floats(const Legacy &L) { return (size_t)L.dims.image_floats; }
size_t prj_rows(const Legacy &L) { return (size_t)2 * L.dims.segmentsize; }
size_t prj_floats(const Legacy &L) { return prj_rows(L) * L.dims.N; }

// ima[x][y][0..Z4) <-> flat [X][Y][Z4]
void gather_image(const Legacy &L, float ***ima, float *flat) {
    const int N = L.dims.N, Z4 = L.dims.z4;
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++) std::memcpy(flat + ((size_t)x * N + y) * Z4, ima[x][y], sizeof(float) * Z4);
}

void scatter_image(const Legacy &L, const float *flat, float ***ima) {
    const int N = L.dims.N, Z4 = L.dims.z4;
    for (int x = 0; x < N; x++)
        for (int y = 0; y < N; y++) std::memcpy(ima[x][y], flat + ((size_t)x * N + y) * Z4, sizeof(float) * Z4);
}

void gather_prj(const Legacy &L, float **prj, float *flat) {
    const size_t N = L.dims.N;
    for (size_t r = 0; r < prj_rows(L); r++) std::memcpy(flat + r * N, prj[r], sizeof(float) * N);
}

void scatter_prj(const Legacy &L, const float *flat, float **prj) {
    const size_t N = L.dims.N;
    for (size_t r = 0; r < prj_rows(L); r++) std::memcpy(prj[r], flat + r * N, sizeof(float) * N);
}

#define LTRY(where, expr)                              \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return fail_hip(where, e_); \
    } while (0)

int check_view(const Legacy &L, const char *where, int view) {
    if (!L.ctx) return fail_msg(where, "dependencies() has not been called");
    if (view < 0 || view >= L.dims.views / 2) return fail_msg(where, "view out of range");
    return 1;
}

int put_prj(Legacy &L, const char *where, float **prj) {
    gather_prj(L, prj, L.h_prj.data());
    LTRY(where, hipMemcpy(L.d_prj, L.h_prj.data(), sizeof(float) * prj_floats(L), hipMemcpyHostToDevice));
    return 1;
}

int get_prj(Legacy &L, const char *where, float **prj) {
    LTRY(where, hipMemcpy(L.h_prj.data(), L.d_prj, sizeof(float) * prj_floats(L), hipMemcpyDeviceToHost));
    scatter_prj(L, L.h_prj.data(), prj);
    return 1;
}

int put_image(Legacy &L, const char *where, float ***ima, std::vector<float> &h, float *d) {
    gather_image(L, ima, h.data());
    LTRY(where, hipMemcpy(d, h.data(), sizeof(float) * image_floats(L), hipMemcpyHostToDevice));
    return 1;
}

int get_image(Legacy &L, const char *where, float ***ima, std::vector<float> &h, const float *d) {
    LTRY(where, hipMemcpy(h.data(), d, sizeof(float) * image_floats(L), hipMemcpyDeviceToHost));
    scatter_image(L, h.data(), ima);
    return 1;
}

int forward(const char *where, float ***ima, float **prj, int view) {
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    if (!ima || !prj) return fail_msg(where, "null argument");
    if (!check_view(L, where, view)) return 0;
    // bins the projector does not write keep the caller's values, as in the reference
    if (!put_image(L, where, ima, L.h_ima, L.d_ima) || !put_prj(L, where, prj)) return 0;
    if (hrrt_forward_proj3d(L.ctx, L.d_ima, L.d_prj, view, nullptr) != HRRT_OK) return fail_ctx(where);
    return get_prj(L, where, prj);
}

}  // namespace

extern "C" int hrrt_legacy_set_params(const hrrt_params *params) {
    if (!params) return fail_msg("hrrt_legacy_set_params", "null argument");
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    L.params = *params;
    L.have_params = true;
    return 1;
}

extern "C" hrrt_ctx *hrrt_legacy_context(void) {
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    return L.ctx;
}

extern "C" int hrrt_legacy_set_norm_zero_mask(short **mask) {
    const char *where = "hrrt_legacy_set_norm_zero_mask";
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    if (!mask) return fail_msg(where, "null argument");
    if (!L.ctx) return fail_msg(where, "dependencies() has not been called");
    const int rows = L.dims.views / 2, cols = L.dims.N / 2 + 1;
    std::vector<int8_t> m((size_t)rows * cols);
    for (int v = 0; v < rows; v++) {
        if (!mask[v]) return fail_msg(where, "null mask row");
        for (int xr = 0; xr < cols; xr++) m[(size_t)v * cols + xr] = mask[v][xr] != 0;
    }
    if (hrrt_set_norm_zero_mask(L.ctx, m.data()) != HRRT_OK) return fail_ctx(where);
    return 1;
}

int dependencies(int nprojs, int nviews, int verbose) {
    (void)verbose;
    const char *where = "dependencies";
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    if (!L.have_params) return fail_msg(where, "hrrt_legacy_set_params() has not been called");
    release(L);
    if (hrrt_create(&L.params, &L.ctx) != HRRT_OK) {
        L.ctx = nullptr;
        return fail_ctx(where);
    }
    if (hrrt_get_dims(L.ctx, &L.dims) != HRRT_OK) {
        release(L);
        return fail_ctx(where);
    }
    if (nprojs != L.dims.N || nviews != L.dims.views) {
        std::fprintf(stderr, "%s: nprojs %d, nviews %d do not match the parameters (%d, %d)\n", where, nprojs,
                     nviews, L.dims.N, L.dims.views);
        release(L);
        return 0;
    }
    hipError_t e;
    if ((e = hipMalloc((void **)&L.d_ima, sizeof(float) * image_floats(L))) != hipSuccess ||
        (e = hipMalloc((void **)&L.d_buf, sizeof(float) * image_floats(L))) != hipSuccess ||
        (e = hipMalloc((void **)&L.d_prj, sizeof(float) * prj_floats(L))) != hipSuccess) {
        release(L);
        return fail_hip(where, e);
    }
    L.h_ima.resize(image_floats(L));
    L.h_buf.resize(image_floats(L));
    L.h_prj.resize(prj_floats(L));
    return 1;
}

void free_dependencies() {
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    release(L);
}

int forward_proj3d2(float ***ima, float **prj, int view, int numthread, float ***imagebuf, float ***prjbuf) {
    (void)numthread, (void)imagebuf, (void)prjbuf;
    return forward("forward_proj3d2", ima, prj, view);
}

int forward_proj3d_thread1(float ***ima, float **prj, int view, int numthread, float ***imagebuf, float ***prjbuf) {
    (void)numthread, (void)imagebuf, (void)prjbuf;
    return forward("forward_proj3d_thread1", ima, prj, view);
}

int back_proj3d2(float **prj, float ***ima, int view, int numthread, float ***imagebuf, float ***prjbuf) {
    (void)numthread, (void)imagebuf, (void)prjbuf;
    const char *where = "back_proj3d2";
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    if (!ima || !prj) return fail_msg(where, "null argument");
    if (!check_view(L, where, view)) return 0;
    if (!put_image(L, where, ima, L.h_ima, L.d_ima) || !put_prj(L, where, prj)) return 0;
    if (hrrt_back_proj3d(L.ctx, L.d_prj, L.d_ima, view, nullptr) != HRRT_OK) return fail_ctx(where);
    return get_prj(L, where, prj) && get_image(L, where, ima, L.h_ima, L.d_ima);
}

int back_proj3d_thread1(float **prj, float ***ima, int view, int numthread, float ***imagebuf, float ***prjbuf) {
    (void)ima, (void)numthread, (void)prjbuf;
    const char *where = "back_proj3d_thread1";
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    if (!imagebuf || !prj) return fail_msg(where, "null argument");
    if (!check_view(L, where, view)) return 0;
    // FOV columns are overwritten; the rest of imagebuf keeps the caller's values
    if (!put_image(L, where, imagebuf, L.h_buf, L.d_buf) || !put_prj(L, where, prj)) return 0;
    if (hrrt_back_proj3d_rotframe(L.ctx, L.d_prj, L.d_buf, view, 1, nullptr) != HRRT_OK) return fail_ctx(where);
    return get_prj(L, where, prj) && get_image(L, where, imagebuf, L.h_buf, L.d_buf);
}

int rotate_image_buf(float ****imagebuf, float ***ima, int numthread, int *view) {
    const char *where = "rotate_image_buf";
    Legacy &L = state();
    std::lock_guard<std::mutex> lock(L.mu);
    if (!imagebuf || !ima || !view || numthread < 0) return fail_msg(where, "invalid argument");
    if (!L.ctx) return fail_msg(where, "dependencies() has not been called");
    for (int i = 0; i < numthread; i++)
        if (!imagebuf[i] || view[i] < 0 || view[i] >= L.dims.views / 2) return fail_msg(where, "invalid buffer or view");
    if (!put_image(L, where, ima, L.h_ima, L.d_ima)) return 0;
    for (int i = 0; i < numthread; i++) {       // in buffer order, as the reference accumulates
        if (!put_image(L, where, imagebuf[i], L.h_buf, L.d_buf)) return 0;
        if (hrrt_rotate_add(L.ctx, L.d_buf, L.d_ima, view[i], nullptr) != HRRT_OK) return fail_ctx(where);
    }
    return get_image(L, where, ima, L.h_ima, L.d_ima);
}
