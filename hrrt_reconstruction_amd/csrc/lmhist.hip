// lmhist.hip -- list-mode histogramming (lmhistogram_mp) on gfx950, and its C ABI
// (include/hrrt_lmhist.h).
//
// The reference walks each packet of PACKET_SIZE 64-bit words sequentially: a word pair is an
// event or a tag and advances the walk by two words, an out-of-sync word advances it by one
// (rebin_packet histogram_mp.cpp:1053-1060).  The walk is split here into 2048-word chunks,
// one wave each:
//   lm_summary_kernel  bitmasks of the out-of-sync words of each parity in the chunk, and from
//                      them the parity the walk leaves the chunk with, for either entry parity
//   lm_resolve_kernel  per packet, the entry parity of every chunk (the walk restarts at word 0
//                      of each packet)
//   lm_window_kernel   the frame's start and end time tags (first in walk order)
//   lm_hist_kernel     every word pair the walk visits inside the frame: tags into the
//                      statistics, events decoded, rebinned through the module-pair LUT with
//                      the reference's float operations, and added with integer atomics into
//                      the sinogram counts and the coincidence maps
//   lm_finalize_kernel counts -> int16 bins (wrapping) and saturating uint8 delayed bins
// Integer sums do not depend on the order of the adds, so the result equals the reference's
// sequential histogram exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/hrrt_lmhist.h"
#include "delays.h"
#include "kernels.h"

using namespace hrrt_dl;

namespace {

constexpr long long PACKET = HRRT_LM_PACKET_WORDS;
constexpr int CHUNK = 2048;                          // words per chunk, one wave
constexpr int CHUNKS_PER_PACKET = (int)(PACKET / CHUNK);
constexpr int WAVES = 4;                             // chunks per workgroup
constexpr int NBLOCKS = HRRT_LM_NBLOCKS;
constexpr unsigned long long NONE = ~0ull;

// statistics scratch (u64 words); each counter on its own 128-byte line, since every wave adds
// into all of them
constexpr int ST_LINE = 16;
enum { ST_PROMPTS = 0, ST_RANDOMS = ST_LINE, ST_EVENTS = 2 * ST_LINE, ST_TAGS = 3 * ST_LINE,
       ST_TIME_TAGS = 4 * ST_LINE, ST_REJECTED = 5 * ST_LINE, ST_TX = 6 * ST_LINE, ST_LAST_TIME = 7 * ST_LINE,
       ST_START = 8 * ST_LINE, ST_END = 9 * ST_LINE, ST_BLOCKS = 10 * ST_LINE };
constexpr int ST_WORDS = ST_BLOCKS + 3 * NBLOCKS;    // then block totals, sample counts, last (packed)

// EWTYPES geometry_info.h:75 packed 2 bits per entry: 0 prompt, 1 delayed, 2 tag, 3 out of sync
constexpr unsigned ewtypes_packed() {
    constexpr int t[16] = {3, 3, 1, 0, 3, 3, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3};
    unsigned v = 0;
    for (int i = 0; i < 16; i++) v |= (unsigned)t[i] << (2 * i);
    return v;
}
constexpr unsigned EWT = ewtypes_packed();

// HRRT_MPAIRS geometry_info.h:68-73 on the device
__constant__ int c_mpairs[NMP + 1][2] = {{-1, -1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {0, 6}, {1, 3}, {1, 4},
                                         {1, 5},   {1, 6}, {1, 7}, {2, 4}, {2, 5}, {2, 6}, {2, 7}, {3, 5},
                                         {3, 6},   {3, 7}, {4, 6}, {4, 7}, {5, 7}};

__device__ inline int word_type(unsigned ew1, unsigned ew2) {
    const unsigned idx = ((ew2 & 0xc0000000u) >> 30) | ((ew1 & 0xc0000000u) >> 28);
    return (int)((EWT >> (2 * idx)) & 3u);
}

struct LmArgs {
    const unsigned *w;
    long long nwords;
    int mode, start_ms, end_ms, ignore_border, doi;
    const int4 *sol;        // [21][144][144] {nsino, bits(d), bits(z), 0}
    const float2 *z;        // [104] {ring z, ring z in planes + 0.5}
    const int2 *seg;        // [128] by 2*(int)seg + (seg < 0) + 64: {first | last << 16, offset}
    int nsino;
    long long S;            // nplanes * nsino
    int *cnt;               // [2][S]
    unsigned *coin_p, *coin_d;
    unsigned *chunks;       // per chunk: exit parity for entry 0 (bit 0) and 1 (bit 1), entry parity (bit 2)
    unsigned long long *st;
};

__device__ inline long long packet_end(long long q, long long nwords) {
    const long long pe = (q / PACKET + 1) * PACKET;
    return pe < nwords ? pe : nwords;
}

// bits 0, 2, 4, ... of x -> bits 0, 1, 2, ...
__device__ inline unsigned even_bits(unsigned long long x) {
    x &= 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x >> 4)) & 0x00ff00ff00ff00ffull;
    x = (x | (x >> 8)) & 0x0000ffff0000ffffull;
    x = (x | (x >> 16)) & 0x00000000ffffffffull;
    return (unsigned)x;
}

// mask[parity * 32 + i] bit j: word cs + 2*(32 i + j) + parity is out of sync and inside its
// packet's walk (it has a successor in the packet).  One wave; cs is a multiple of CHUNK.
__device__ void chunk_masks(const unsigned *__restrict__ w, long long cs, long long nwords, unsigned *mask, int lane) {
    const long long pend = packet_end(cs, nwords);
    for (int i = 0; i < CHUNK / 64; i++) {
        const long long q = cs + i * 64 + lane;
        bool t3 = false;
        if (q + 1 < pend) t3 = word_type(w[q], w[q + 1]) == 3;
        const unsigned long long b = __ballot(t3);
        if (lane == 0) {
            mask[i] = even_bits(b);
            mask[32 + i] = even_bits(b >> 1);
        }
    }
}

// The walk through one chunk from relative position `entry` (0 or 1): the relative positions
// of the out-of-sync words it steps on (sw, may be null) and the parity it leaves with.
__device__ int chunk_walk(const unsigned *mask, int entry, unsigned short *sw, int &nsw) {
    int p = entry;
    nsw = 0;
    while (p < CHUNK) {
        const int par = p & 1, k = p >> 1;
        int f = -1;
        for (int wi = k >> 5; wi < 32; wi++) {
            unsigned m = mask[par * 32 + wi];
            if (wi == (k >> 5)) m &= ~0u << (k & 31);
            if (m) {
                f = (wi << 5) + __ffs(m) - 1;
                break;
            }
        }
        if (f < 0) break;
        const int pos = 2 * f + par;
        if (sw) sw[nsw] = (unsigned short)pos;
        nsw++;
        p = pos + 1;
    }
    return p & 1;
}

__global__ __launch_bounds__(WAVES * 64) void lm_summary_kernel(LmArgs a, long long nchunks) {
    __shared__ unsigned masks[WAVES][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long c = (long long)blockIdx.x * WAVES + wv;
    if (c >= nchunks) return;
    chunk_masks(a.w, c * CHUNK, a.nwords, masks[wv], lane);
    if (lane == 0) {
        int n;
        const int e0 = chunk_walk(masks[wv], 0, nullptr, n);
        const int e1 = chunk_walk(masks[wv], 1, nullptr, n);
        a.chunks[c] = (unsigned)(e0 | (e1 << 1));
    }
}

// one workgroup per packet: the chunk summaries staged in LDS, walked in order by one thread
// (the walk enters the packet's first chunk at word 0)
__global__ __launch_bounds__(256) void lm_resolve_kernel(unsigned *chunks, long long nchunks) {
    __shared__ unsigned sum[CHUNKS_PER_PACKET];
    const long long c0 = (long long)blockIdx.x * CHUNKS_PER_PACKET;
    const int n = (int)(c0 + CHUNKS_PER_PACKET < nchunks ? CHUNKS_PER_PACKET : nchunks - c0);
    for (int i = threadIdx.x; i < n; i += 256) sum[i] = chunks[c0 + i] & 3u;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned e = 0;
        for (int i = 0; i < n; i++) {
            const unsigned v = sum[i];
            sum[i] = v | (e << 2);
            e = (v >> e) & 1u;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) chunks[c0 + i] = sum[i];
}

// Per wave: masks and the switch list of the walk through chunk c; returns the entry parity.
__device__ int chunk_visit_setup(const LmArgs &a, long long c, unsigned *mask, unsigned short *sw, int &nsw,
                                 int lane) {
    chunk_masks(a.w, c * CHUNK, a.nwords, mask, lane);
    const int entry = (int)((a.chunks[c] >> 2) & 1u);
    if (lane == 0) {
        int n;
        chunk_walk(mask, entry, sw, n);
        mask[64] = (unsigned)n;
    }
    __builtin_amdgcn_wave_barrier();
    nsw = (int)mask[64];
    return entry;
}

// Does the walk start a word pair at relative position r?
__device__ inline bool visited(int r, int entry, const unsigned short *sw, int nsw) {
    int lo = 0, hi = nsw;                  // switches before r
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sw[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    const int start = lo == 0 ? entry : sw[lo - 1] + 1;
    if (lo < nsw && sw[lo] == r) return false;   // an out-of-sync word the walk steps on
    return r >= start && ((r - start) & 1) == 0;
}

template <int PASS>
__global__ __launch_bounds__(WAVES * 64) void lm_window_kernel(LmArgs a, long long nchunks) {
    __shared__ unsigned masks[WAVES][65];
    __shared__ unsigned short sws[WAVES][CHUNK];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long c = (long long)blockIdx.x * WAVES + wv;
    if (c >= nchunks) return;
    // pass 1 looks for the end tag from the start tag on (the start tag itself may end the frame)
    long long from = 0;
    if (PASS == 1 && a.start_ms > 0) {
        if (a.st[ST_START] == NONE) return;
        from = (long long)a.st[ST_START];
    }
    int nsw;
    const int entry = chunk_visit_setup(a, c, masks[wv], sws[wv], nsw, lane);
    const long long cs = c * CHUNK, pend = packet_end(cs, a.nwords);
    const int bound = PASS == 0 ? a.start_ms : a.end_ms;
    for (int i = 0; i < CHUNK / 64; i++) {
        const int r = i * 64 + lane;
        const long long q = cs + r;
        if (q + 1 >= pend || q < from) continue;
        if (!visited(r, entry, sws[wv], nsw)) continue;
        const unsigned ew1 = a.w[q], ew2 = a.w[q + 1];
        if (word_type(ew1, ew2) != 2) continue;
        const unsigned tag = (ew1 & 0xffffu) | ((ew2 & 0xffffu) << 16);
        if ((tag & 0xE0000000u) != 0x80000000u) continue;
        if ((int)(tag & 0x3fffffffu) >= bound) atomicMin(&a.st[PASS == 0 ? ST_START : ST_END], (unsigned long long)q);
    }
}

// rebin_event LM_Rebinner_mp.cpp:188-219 on the device tables: sinogram address or -1
__device__ inline long long rebin(const LmArgs &a, int mp, int alayer, int ax, int ay, int blayer, int bx, int by) {
    const int4 e = a.sol[((mp * NXD) + ax + NX * alayer) * NXD + bx + NX * blayer];
    if (e.x < 0) return -1;
    const float sd = __int_as_float(e.y), sz = __int_as_float(e.z);
    const float2 za = a.z[ay];
    const float dz = a.z[by].x - za.x;
    const int plane = (int)(za.y + sz * dz);
    const float s = 0.5f + dz * sd;
    const int si = (int)s;
    const int2 ent = a.seg[64 + min(max(2 * si + (s < 0.0f ? 1 : 0), -64), 63)];
    if (plane < (ent.x & 0xffff) || plane > (ent.x >> 16)) return -1;
    return (long long)(plane + ent.y) * a.nsino + e.x;
}

// sum of a lane counter over the wave, added once by lane 0
__device__ inline void wave_flush(unsigned long long *dst, unsigned v, int lane) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

__global__ __launch_bounds__(WAVES * 64) void lm_hist_kernel(LmArgs a, long long nchunks) {
    __shared__ unsigned masks[WAVES][65];
    __shared__ unsigned short sws[WAVES][CHUNK];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long c = (long long)blockIdx.x * WAVES + wv;
    if (c >= nchunks) return;
    // the frame: from its start tag (or word 0) through its end tag (or the last word)
    const unsigned long long s0 = a.st[ST_START], s1 = a.st[ST_END];
    if (a.start_ms > 0 && s0 == NONE) return;
    const long long lo = a.start_ms > 0 ? (long long)s0 : 0;
    const long long hi = (a.end_ms > 0 && s1 != NONE) ? (long long)s1 + 1 : a.nwords;
    const long long cs = c * CHUNK, pend = packet_end(cs, a.nwords);
    if (cs + CHUNK <= lo || cs >= hi) return;
    int nsw;
    const int entry = chunk_visit_setup(a, c, masks[wv], sws[wv], nsw, lane);
    const int xhigh = a.ignore_border ? NX - 1 : NX;
    unsigned n_tags = 0, n_time = 0, n_events = 0, n_rej = 0, n_tx = 0, n_prompts = 0, n_randoms = 0;
    unsigned long long last_time = 0;
    for (int i = 0; i < CHUNK / 64; i++) {
        const int r = i * 64 + lane;
        const long long q = cs + r;
        bool live = q + 1 < pend && q >= lo && q < hi && visited(r, entry, sws[wv], nsw);
        unsigned ew1 = 0, ew2 = 0;
        int type = 3;
        if (live) {
            ew1 = a.w[q];
            ew2 = a.w[q + 1];
            type = word_type(ew1, ew2);
        }
        // tags (process_tagword histogram_mp.cpp:225-295)
        const bool is_tag = live && type == 2;
        const unsigned tag = (ew1 & 0xffffu) | ((ew2 & 0xffffu) << 16);
        const bool time_tag = is_tag && (tag & 0xE0000000u) == 0x80000000u;
        n_tags += is_tag;
        n_time += time_tag;
        if (time_tag) last_time = ((unsigned long long)q << 30) | (tag & 0x3fffffffu);   // q grows with i
        if (is_tag && (tag & 0xE0000000u) == 0xA0000000u) {
            const int block = (int)((tag & 0x1ff80000u) >> 19);
            const unsigned v = tag & 0x0007ffffu;
            if (block < NBLOCKS) {
                atomicAdd(&a.st[ST_BLOCKS + block], (unsigned long long)v);
                atomicAdd(&a.st[ST_BLOCKS + NBLOCKS + block], 1ull);
                atomicMax(&a.st[ST_BLOCKS + 2 * NBLOCKS + block], ((unsigned long long)q << 19) | v);
            }
        }
        // events (rebin_packet histogram_mp.cpp:1111-1227)
        const bool is_event = live && type < 2;
        n_events += is_event;
        const int mpe = (int)(((ew1 & 0x00070000u) >> 16) | ((ew2 & 0x00070000u) >> 13));
        const int ax = (int)(ew1 & 0xff), ay = (int)((ew1 & 0xff00) >> 8);
        const int bx = (int)(ew2 & 0xff), by = (int)((ew2 & 0xff00) >> 8);
        const int doia = a.doi ? (int)((ew1 & 0x01C00000u) >> 22) : 1;
        const int doib = a.doi ? (int)((ew2 & 0x01C00000u) >> 22) : 1;
        const bool bad = is_event && (mpe < 1 || mpe > NMP || ax >= NX || ay >= NY || bx >= NX || by >= NY);
        n_rej += bad;
        const bool ok = is_event && !bad;
        const bool tx = ok && (doia == 7 || doib == 7);
        n_tx += tx;
        const bool em = ok && !tx && doia <= 1 && doib <= 1;
        long long addr = -1;
        if (em) {
            if (a.ignore_border && (ax * bx == 0 || ax == xhigh || bx == xhigh)) addr = -1;
            else if (a.mode == 2 && type != 0) addr = -1;
            else addr = rebin(a, mpe, doia, ax, ay, doib, bx, by);
            unsigned *coin = type == 0 ? a.coin_p : a.coin_d;
            if (coin) {
                const int ahead = c_mpairs[mpe][0], bhead = c_mpairs[mpe][1];
                atomicAdd(&coin[doia * XYH + ahead * NX + ax + ay * HX], 1u);
                atomicAdd(&coin[doib * XYH + bhead * NX + bx + by * HX], 1u);
            }
        }
        n_prompts += addr >= 0 && type == 0;
        n_randoms += addr >= 0 && type == 1;
        if (addr >= 0) {
            if (a.mode == 0) atomicAdd(&a.cnt[addr], type == 0 ? 1 : -1);
            else if (a.mode == 2 || type == 0) atomicAdd(&a.cnt[addr], 1);
            else atomicAdd(&a.cnt[a.S + addr], 1);
        }
    }
    wave_flush(&a.st[ST_TAGS], n_tags, lane);
    wave_flush(&a.st[ST_TIME_TAGS], n_time, lane);
    wave_flush(&a.st[ST_EVENTS], n_events, lane);
    wave_flush(&a.st[ST_REJECTED], n_rej, lane);
    wave_flush(&a.st[ST_TX], n_tx, lane);
    wave_flush(&a.st[ST_PROMPTS], n_prompts, lane);
    wave_flush(&a.st[ST_RANDOMS], n_randoms, lane);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(last_time, o);
        last_time = other > last_time ? other : last_time;
    }
    if (lane == 0 && last_time) atomicMax(&a.st[ST_LAST_TIME], last_time);
}

__global__ void lm_finalize_kernel(const int *__restrict__ cnt, long long S, int mode, int16_t *__restrict__ sino,
                                   uint8_t *__restrict__ delayed) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    sino[i] = (int16_t)(uint16_t)(unsigned)cnt[i];
    if (mode == 1) {
        const int d = cnt[S + i];
        if (delayed) delayed[i] = (uint8_t)min(d, 255);
        else sino[S + i] = (int16_t)(uint16_t)(unsigned)d;
    }
}

// Eight bins per thread: two 16-byte count loads, one 16-byte sinogram store (and one
// 8-byte delayed store).  Needs S % 8 == 0 and 16-byte aligned outputs (8 for the bytes).
__device__ __forceinline__ unsigned pack16(int a, int b) { return ((unsigned)a & 0xffffu) | ((unsigned)b << 16); }

__device__ __forceinline__ unsigned pack8(int a, int b, int c, int d) {
    return (unsigned)min(a, 255) | ((unsigned)min(b, 255) << 8) | ((unsigned)min(c, 255) << 16) |
           ((unsigned)min(d, 255) << 24);
}

__global__ void lm_finalize8_kernel(const int4 *__restrict__ cnt, long long S8, int mode, uint4 *__restrict__ sino,
                                    uint2 *__restrict__ delayed) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S8) return;
    const int4 a = cnt[2 * i], b = cnt[2 * i + 1];
    sino[i] = make_uint4(pack16(a.x, a.y), pack16(a.z, a.w), pack16(b.x, b.y), pack16(b.z, b.w));
    if (mode == 1) {
        const int4 c = cnt[2 * (S8 + i)], d = cnt[2 * (S8 + i) + 1];
        if (delayed) delayed[i] = make_uint2(pack8(c.x, c.y, c.z, c.w), pack8(d.x, d.y, d.z, d.w));
        else sino[S8 + i] = make_uint4(pack16(c.x, c.y), pack16(c.z, c.w), pack16(d.x, d.y), pack16(d.z, d.w));
    }
}

__global__ void lm_stats_init_kernel(unsigned long long *st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ST_WORDS) st[i] = (i == ST_START || i == ST_END) ? NONE : 0ull;
}

__global__ void lm_zero_kernel(unsigned *p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// The module-pair LUT by (module pair, A crystal, B crystal) on the device, and the counts.
int ensure_lm(hrrt_delays &c, long long nchunks) {
    int r = upload(c);
    if (r != HRRT_OK) return r;
    int dev = -1;
    DTRY(hipGetDevice(&dev));
    if (dev != c.device) return hrrt_dl::fail(HRRT_ERR_ARG, "context tables live on another device");
    if (!c.d_sol) {
        std::vector<int4> sol(c.lut.size());
        for (size_t i = 0; i < c.lut.size(); i++) {
            int4 e;
            e.x = c.lut[i].nsino;
            memcpy(&e.y, &c.lut[i].d, 4);
            memcpy(&e.z, &c.lut[i].z, 4);
            e.w = 0;
            sol[i] = e;
        }
        DTRY(hipMalloc(&c.d_sol, sizeof(int4) * sol.size()));
        DTRY(hipMemcpy(c.d_sol, sol.data(), sizeof(int4) * sol.size(), hipMemcpyHostToDevice));
    }
    if (!c.d_lm_counts)
        DTRY(hipMalloc(&c.d_lm_counts, sizeof(int) * 2 * (size_t)c.nplanes * (size_t)c.nsino));
    if (!c.d_lm_stats) DTRY(hipMalloc(&c.d_lm_stats, sizeof(unsigned long long) * ST_WORDS));
    if (nchunks > c.lm_chunk_cap) {
        if (c.d_lm_chunks) DTRY(hipFree(c.d_lm_chunks));
        c.d_lm_chunks = nullptr;
        DTRY(hipMalloc(&c.d_lm_chunks, sizeof(unsigned) * (size_t)nchunks));
        c.lm_chunk_cap = nchunks;
    }
    return HRRT_OK;
}

}  // namespace

extern "C" int hrrt_lm_histogram(hrrt_delays *c, const uint32_t *d_words, long long nwords, const hrrt_lm_options *opt,
                                 int16_t *d_sino, uint8_t *d_delayed, uint32_t *d_coin_p, uint32_t *d_coin_d,
                                 hrrt_lm_stats *stats, void *stream) {
    if (!c || !opt || !d_sino || (nwords > 0 && !d_words) || nwords < 0) return hrrt_dl::fail(HRRT_ERR_ARG, "null argument");
    if (opt->mode == 7) return hrrt_dl::fail(HRRT_ERR_UNSUPPORTED, "transmission histogramming is not supported");
    if (opt->mode < 0 || opt->mode > 2) return hrrt_dl::fail(HRRT_ERR_ARG, "mode must be 0, 1 or 2");
    if (nwords >= (1ll << 34)) return hrrt_dl::fail(HRRT_ERR_ARG, "stream longer than 2^34 words");
    hipStream_t s = (hipStream_t)stream;
    const long long nchunks = (nwords + CHUNK - 1) / CHUNK;
    int r = ensure_lm(*c, std::max(nchunks, 1ll));
    if (r != HRRT_OK) return r;
    const long long S = (long long)c->nplanes * c->nsino;
    LmArgs a;
    a.w = d_words;
    a.nwords = nwords;
    a.mode = opt->mode;
    a.start_ms = opt->start_ms;
    a.end_ms = opt->end_ms;
    a.ignore_border = opt->ignore_border != 0;
    a.doi = opt->no_doi == 0;
    a.sol = c->d_sol;
    a.z = c->d_z;
    a.seg = c->d_seg;
    a.nsino = c->nsino;
    a.S = S;
    a.cnt = c->d_lm_counts;
    a.coin_p = d_coin_p;
    a.coin_d = d_coin_d;
    a.chunks = c->d_lm_chunks;
    a.st = c->d_lm_stats;
    lm_stats_init_kernel<<<(ST_WORDS + 255) / 256, 256, 0, s>>>(a.st);
    lm_zero_kernel<<<2048, 256, 0, s>>>((unsigned *)a.cnt, (opt->mode == 1 ? 2 : 1) * S);
    if (d_coin_p) lm_zero_kernel<<<128, 256, 0, s>>>(d_coin_p, NCRYS);
    if (d_coin_d) lm_zero_kernel<<<128, 256, 0, s>>>(d_coin_d, NCRYS);
    DTRY(hipGetLastError());
    if (nchunks > 0) {
        const unsigned grid = (unsigned)((nchunks + WAVES - 1) / WAVES);
        const long long npackets = (nchunks + CHUNKS_PER_PACKET - 1) / CHUNKS_PER_PACKET;
        lm_summary_kernel<<<grid, WAVES * 64, 0, s>>>(a, nchunks);
        lm_resolve_kernel<<<(unsigned)npackets, 256, 0, s>>>(a.chunks, nchunks);
        if (opt->start_ms > 0) lm_window_kernel<0><<<grid, WAVES * 64, 0, s>>>(a, nchunks);
        if (opt->end_ms > 0) lm_window_kernel<1><<<grid, WAVES * 64, 0, s>>>(a, nchunks);
        lm_hist_kernel<<<grid, WAVES * 64, 0, s>>>(a, nchunks);
    }
    if (S % 8 == 0 && (uintptr_t)d_sino % 16 == 0 && (uintptr_t)d_delayed % 8 == 0)
        lm_finalize8_kernel<<<(unsigned)((S / 8 + 255) / 256), 256, 0, s>>>(
            (const int4 *)a.cnt, S / 8, opt->mode, (uint4 *)d_sino, (uint2 *)d_delayed);
    else
        lm_finalize_kernel<<<(unsigned)((S + 255) / 256), 256, 0, s>>>(a.cnt, S, opt->mode, d_sino, d_delayed);
    DTRY(hipGetLastError());
    if (stats) {
        std::vector<unsigned long long> h(ST_WORDS);
        DTRY(hipMemcpyAsync(h.data(), a.st, sizeof(unsigned long long) * ST_WORDS, hipMemcpyDeviceToHost, s));
        DTRY(hipStreamSynchronize(s));
        memset(stats, 0, sizeof(*stats));
        stats->prompts = (long long)h[ST_PROMPTS];
        stats->randoms = (long long)h[ST_RANDOMS];
        stats->events = (long long)h[ST_EVENTS];
        stats->tags = (long long)h[ST_TAGS];
        stats->time_tags = (long long)h[ST_TIME_TAGS];
        stats->rejected = (long long)h[ST_REJECTED];
        stats->transmission = (long long)h[ST_TX];
        const bool never = opt->start_ms > 0 && h[ST_START] == NONE;
        stats->first_word = never ? -1 : (opt->start_ms > 0 ? (long long)h[ST_START] : 0);
        stats->end_word = never ? -1 : ((opt->end_ms > 0 && h[ST_END] != NONE) ? (long long)h[ST_END] + 2 : nwords);
        stats->last_time_ms = h[ST_TIME_TAGS] ? (int)(h[ST_LAST_TIME] & 0x3fffffffu) : -1;
        for (int b = 0; b < NBLOCKS; b++) {
            stats->block_total[b] = (long long)h[ST_BLOCKS + b];
            stats->block_nsamples[b] = (int)h[ST_BLOCKS + NBLOCKS + b];
            stats->block_last[b] = (int)(h[ST_BLOCKS + 2 * NBLOCKS + b] & 0x7ffffu);
        }
    }
    return HRRT_OK;
}
