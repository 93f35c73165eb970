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
//                      statistics, events decoded and rebinned through the module-pair LUT with
//                      the reference's float operations; each binned event written as a key
//                      2*bin + delayed into its chunk's CHUNK/2 key slots (the unused slots hold
//                      NOKEY), and each event's two crystals as 16-bit entries into the chunk's
//                      per-head segments
//   radix sort         the keys by 8192-bin tile (rocPRIM, key bits 14 and up; NOKEY sorts last)
//   lm_tiles_kernel    the first key of every tile in the sorted keys
//   lm_tile_kernel     per tile: its keys counted in LDS, then the tile's int16 bins (wrapping)
//                      and saturating uint8 delayed bins written in order, 16 per thread
//   lm_coin_kernel     per head and range of chunks: the head's crystal entries counted in LDS
//   lm_coin_sum_kernel the ranges' counts summed into the coincidence maps
// No global atomics touch a sinogram bin or a crystal counter: scattered device-scope atomics
// into a 1.9 GB int32 count array and the 0.5 MB maps took 5.6 ms of a 6.8 ms call for 40 M
// events at span 3, against 1.9 ms for the whole call this way.
// Integer counts do not depend on the order of the events, so the result equals the
// reference's sequential histogram exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/hrrt_lmhist.h"
#include "delays.h"
#include "kernels.h"

using namespace hrrt_dl;

namespace {

constexpr long long PACKET = HRRT_LM_PACKET_WORDS;
constexpr int CHUNK = 2048;                          // words per chunk, one wave
constexpr int TILE_BITS = 13, TILE = 1 << TILE_BITS;  // sinogram bins per histogram tile
constexpr unsigned NOKEY = 0xffffffffu;               // an unused key slot
constexpr int COIN_HEAD = 2 * NDOIS * NY * NX;        // coincidence counters of one head: [map][layer][ring][x]
constexpr int COIN_RANGES = 32;                       // chunk ranges of the coincidence-map pass
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

// HRRT_MPAIRS geometry_info.h:68-73, packed 3 bits per module pair (A heads, B heads)
constexpr int MPAIRS[NMP + 1][2] = {{-1, -1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {0, 6}, {1, 3}, {1, 4},
                                    {1, 5},   {1, 6}, {1, 7}, {2, 4}, {2, 5}, {2, 6}, {2, 7}, {3, 5},
                                    {3, 6},   {3, 7}, {4, 6}, {4, 7}, {5, 7}};
constexpr unsigned long long mp_pack(int side) {
    unsigned long long v = 0;
    for (int m = 0; m <= NMP; m++) v |= (unsigned long long)(MPAIRS[m][side] & 7) << (3 * m);
    return v;
}
constexpr unsigned long long MP_A = mp_pack(0), MP_B = mp_pack(1);

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
    unsigned *keys;         // [nchunks][CHUNK/2] binned events 2*bin + (delayed), then NOKEY
    unsigned *coin_p, *coin_d;
    unsigned short *cseg;   // [nchunks][NHEADS][CHUNK/2] crystal entries (index in COIN_HEAD) by head
    unsigned *ccnt;         // [nchunks][NHEADS] entries per chunk and head
    unsigned *chunks;       // per chunk: exit parity for entry 0 (bit 0) and 1 (bit 1), entry parity (bit 2)
    unsigned *cmask;        // [nchunks][64] out-of-sync masks of each chunk (chunk_mask_word)
    bool w16;               // the stream is 16-byte aligned
    unsigned long long *st;
};

__device__ inline long long packet_end(long long q, long long nwords) {
    const long long pe = (q / PACKET + 1) * PACKET;
    return pe < nwords ? pe : nwords;
}

// bits 0..15 of x -> bits 0, 2, 4, ..., 30
__device__ inline unsigned spread16(unsigned x) {
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// Mask word `lane` of chunk cs (a multiple of CHUNK): mask[parity * 32 + i] bit j says word
// cs + 2*(32 i + j) + parity is out of sync and inside its packet's walk (it has a successor
// in the packet).  One wave; each lane reads four consecutive words per 256-word block (one
// 16-byte load when the stream is 16-byte aligned), the fifth from its neighbour.
__device__ unsigned chunk_mask_word(const unsigned *__restrict__ w, long long cs, long long nwords, bool w16, int lane) {
    const long long pend = packet_end(cs, nwords);
    const int par = lane >> 5, mi = lane & 3, kk = (lane & 31) >> 2;
    unsigned mine = 0;
#pragma unroll
    for (int k = 0; k < CHUNK / 256; k++) {
        const long long q0 = cs + 256 * k + 4 * lane;
        unsigned v[5];
        if (w16 && q0 + 3 < nwords) {
            const uint4 x = *reinterpret_cast<const uint4 *>(w + q0);
            v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = q0 + j < nwords ? w[q0 + j] : 0u;
        }
        v[4] = __shfl_down(v[0], 1);
        if (lane == 63) v[4] = q0 + 4 < pend ? w[q0 + 4] : 0u;
        unsigned long long b[4];
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = __ballot(q0 + j + 1 < pend && word_type(v[j], v[j + 1]) == 3);
        // relative position 256 k + 4 l + j: parity j & 1, bit 2 (l mod 16) + (j >> 1) of mask word 4 k + l / 16
        const unsigned long long lo = par ? b[1] : b[0], hi = par ? b[3] : b[2];
        const unsigned word = spread16((unsigned)(lo >> (16 * mi))) | (spread16((unsigned)(hi >> (16 * mi))) << 1);
        if (kk == k) mine = word;
    }
    return mine;
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
    const unsigned m = chunk_mask_word(a.w, c * CHUNK, a.nwords, a.w16, lane);
    masks[wv][lane] = m;
    a.cmask[c * 64 + lane] = m;
    __builtin_amdgcn_wave_barrier();
    // the walks from entry 0 (lane 0) and entry 1 (lane 32) side by side
    int e = 0;
    if ((lane & 31) == 0) {
        int n;
        e = chunk_walk(masks[wv], lane >> 5, nullptr, n);
    }
    const int e1 = __shfl(e, 32);
    if (lane == 0) a.chunks[c] = (unsigned)(e | (e1 << 1));
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
    mask[lane] = a.cmask[c * 64 + lane];
    __builtin_amdgcn_wave_barrier();
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

// rebin_event LM_Rebinner_mp.cpp:188-219 from the event's LUT entry e (gathered from a.sol) and
// the ring and segment tables in LDS: sinogram address or -1
__device__ inline long long rebin(int4 e, int ay, int by, const float2 *zs, const int2 *segs, int nsino) {
    if (e.x < 0) return -1;
    const float sd = __int_as_float(e.y), sz = __int_as_float(e.z);
    const float2 za = zs[ay];
    const float dz = zs[by].x - za.x;
    const int plane = (int)(za.y + sz * dz);
    const float s = 0.5f + dz * sd;
    const int si = (int)s;
    const int2 ent = segs[64 + min(max(2 * si + (s < 0.0f ? 1 : 0), -64), 63)];
    if (plane < (ent.x & 0xffff) || plane > (ent.x >> 16)) return -1;
    return (long long)(plane + ent.y) * nsino + e.x;
}

// sum of a lane counter over the wave, added once by lane 0
__device__ inline void wave_flush(unsigned long long *dst, unsigned v, int lane) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

__global__ __launch_bounds__(WAVES * 64) void lm_hist_kernel(LmArgs a, long long nchunks) {
    __shared__ unsigned masks[WAVES][65];
    __shared__ unsigned short sws[WAVES][CHUNK];
    __shared__ float2 zs[NY];
    __shared__ int2 segs[128];
    for (int j = threadIdx.x; j < NY; j += blockDim.x) zs[j] = a.z[j];
    for (int j = threadIdx.x; j < 128; j += blockDim.x) segs[j] = a.seg[j];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long c = (long long)blockIdx.x * WAVES + wv;
    if (c >= nchunks) return;
    // a chunk holds at most CHUNK/2 events (the walk steps two words past an event)
    unsigned *keys = a.keys + c * (CHUNK / 2);
    auto no_keys_from = [&](unsigned j0) {
        for (unsigned j = j0 + lane; j < CHUNK / 2; j += 64) keys[j] = NOKEY;
    };
    // the frame: from its start tag (or word 0) through its end tag (or the last word)
    const unsigned long long s0 = a.st[ST_START], s1 = a.st[ST_END];
    const long long lo = a.start_ms > 0 ? (long long)s0 : 0;
    const long long hi = (a.end_ms > 0 && s1 != NONE) ? (long long)s1 + 1 : a.nwords;
    const long long cs = c * CHUNK, pend = packet_end(cs, a.nwords);
    if ((a.start_ms > 0 && s0 == NONE) || cs + CHUNK <= lo || cs >= hi) {
        no_keys_from(0);
        if (a.cseg && lane < NHEADS) a.ccnt[c * NHEADS + lane] = 0;
        return;
    }
    unsigned short *cseg = a.cseg + c * (NHEADS * (CHUNK / 2));
    unsigned cc[NHEADS];   // entries per head so far (wave-uniform)
#pragma unroll
    for (int h = 0; h < NHEADS; h++) cc[h] = 0;
    int nsw;
    const int entry = chunk_visit_setup(a, c, masks[wv], sws[wv], nsw, lane);
    const int xhigh = a.ignore_border ? NX - 1 : NX;
    unsigned n_tags = 0, n_time = 0, n_events = 0, n_rej = 0, n_tx = 0, n_prompts = 0, n_randoms = 0;
    unsigned nk = 0;   // keys of the chunk so far (wave-uniform)
    unsigned long long last_time = 0;

    // event fields (rebin_packet histogram_mp.cpp:1111-1227)
    struct Ev {
        int type, mpe, ax, ay, bx, by, doia, doib;
        bool is_event, bad, tx, em, binnable;
    };
    auto decode = [&](unsigned ew1, unsigned ew2, bool live) __attribute__((always_inline)) {
        Ev v;
        v.type = live ? word_type(ew1, ew2) : 3;
        v.is_event = v.type < 2;
        v.mpe = (int)(((ew1 & 0x00070000u) >> 16) | ((ew2 & 0x00070000u) >> 13));
        v.ax = (int)(ew1 & 0xff), v.ay = (int)((ew1 & 0xff00) >> 8);
        v.bx = (int)(ew2 & 0xff), v.by = (int)((ew2 & 0xff00) >> 8);
        v.doia = a.doi ? (int)((ew1 & 0x01C00000u) >> 22) : 1;
        v.doib = a.doi ? (int)((ew2 & 0x01C00000u) >> 22) : 1;
        v.bad = v.is_event && (v.mpe < 1 || v.mpe > NMP || v.ax >= NX || v.ay >= NY || v.bx >= NX || v.by >= NY);
        const bool ok = v.is_event && !v.bad;
        v.tx = ok && (v.doia == 7 || v.doib == 7);
        v.em = ok && !v.tx && v.doia <= 1 && v.doib <= 1;
        v.binnable = v.em && !(a.ignore_border && (v.ax * v.bx == 0 || v.ax == xhigh || v.bx == xhigh)) &&
                     !(a.mode == 2 && v.type != 0);
        return v;
    };
    auto step = [&](long long q, unsigned ew1, unsigned ew2, const Ev &v, int4 e) __attribute__((always_inline)) {
        // tags (process_tagword histogram_mp.cpp:225-295)
        const bool is_tag = v.type == 2;
        const unsigned tag = (ew1 & 0xffffu) | ((ew2 & 0xffffu) << 16);
        const bool time_tag = is_tag && (tag & 0xE0000000u) == 0x80000000u;
        n_tags += is_tag;
        n_time += time_tag;
        if (time_tag) last_time = ((unsigned long long)q << 30) | (tag & 0x3fffffffu);   // q grows along the lane
        if (is_tag && (tag & 0xE0000000u) == 0xA0000000u) {
            const int block = (int)((tag & 0x1ff80000u) >> 19);
            const unsigned val = tag & 0x0007ffffu;
            if (block < NBLOCKS) {
                atomicAdd(&a.st[ST_BLOCKS + block], (unsigned long long)val);
                atomicAdd(&a.st[ST_BLOCKS + NBLOCKS + block], 1ull);
                atomicMax(&a.st[ST_BLOCKS + 2 * NBLOCKS + block], ((unsigned long long)q << 19) | val);
            }
        }
        n_events += v.is_event;
        n_rej += v.bad;
        n_tx += v.tx;
        const long long addr = v.binnable ? rebin(e, v.ay, v.by, zs, segs, a.nsino) : -1;
        // coincidence maps (histogram_mp.cpp:1218-1222, ring stride of histogram_mp_fromarchive.cpp:1092):
        // the event's two crystals, each an entry [map][layer][ring][x] in its head's segment of the
        // chunk (the two heads of a module pair differ)
        if (a.cseg) {
            const bool cn = v.em && (v.type == 0 ? a.coin_p : a.coin_d) != nullptr;
            const int ahead = cn ? (int)((MP_A >> (3 * v.mpe)) & 7) : -1;
            const int bhead = cn ? (int)((MP_B >> (3 * v.mpe)) & 7) : -1;
            const int mo = v.type == 0 ? 0 : NDOIS * NY * NX;
            const unsigned short ka = (unsigned short)(mo + (v.doia * NY + v.ay) * NX + v.ax);
            const unsigned short kb = (unsigned short)(mo + (v.doib * NY + v.by) * NX + v.bx);
#pragma unroll
            for (int h = 0; h < NHEADS; h++) {
                const bool on = ahead == h || bhead == h;
                const unsigned long long bh = __ballot(on);
                if (on)
                    cseg[h * (CHUNK / 2) + cc[h] +
                         __builtin_amdgcn_mbcnt_hi((unsigned)(bh >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bh, 0u))] =
                        ahead == h ? ka : kb;
                cc[h] += (unsigned)__popcll(bh);
            }
        }
        n_prompts += addr >= 0 && v.type == 0;
        n_randoms += addr >= 0 && v.type == 1;
        // trues mode counts prompt - delayed, mode 1 the two apart, mode 2 prompts only (the
        // delayed events have no bin there): every mode keeps the delayed flag in bit 0
        const bool has = addr >= 0;
        const unsigned long long bal = __ballot(has);
        if (has) {
            const unsigned pre = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
            keys[nk + pre] = (unsigned)(2 * addr + (v.type != 0 ? 1 : 0));
        }
        nk += (unsigned)__popcll(bal);
    };
    // The walk visits one position of each aligned word pair (2k, 2k + 1): the one of its
    // parity, entry parity flipped by every out-of-sync word before 2k; an out-of-sync word is
    // visited but not decoded, and when it is the pair's even word the walk goes on to the odd
    // one.  Each lane takes a pair per row; rows of 64 pairs in groups of G: the next group's
    // words are loaded and this group's LUT entries gathered before any of the group is processed.
    constexpr int G = 4, NIT = CHUNK / 128;
    static_assert(NIT % G == 0, "whole groups");
    const unsigned short *sw = sws[wv];
    auto load_words = [&](int i0, unsigned *x1, unsigned *x2, int *pp, bool *l) __attribute__((always_inline)) {
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int x = 2 * ((i0 + g) * 64 + lane);
            int n = 0, m = nsw;   // out-of-sync words visited before x
            while (n < m) {
                const int mid = (n + m) >> 1;
                if (sw[mid] < x) n = mid + 1;
                else m = mid;
            }
            const int par = (entry + n) & 1;
            int p = x + par;
            if (n < nsw && sw[n] == p) p = (par == 0 && !(n + 1 < nsw && sw[n + 1] == x + 1)) ? x + 1 : -1;
            const long long q = cs + p;
            pp[g] = p;
            l[g] = p >= 0 && q + 1 < pend && q >= lo && q < hi;
            x1[g] = l[g] ? a.w[q] : 0u;
            x2[g] = l[g] ? a.w[q + 1] : 0u;
        }
    };
    unsigned w1[G], w2[G];
    int pos[G];
    bool lv[G];
    load_words(0, w1, w2, pos, lv);
    for (int i0 = 0; i0 < NIT; i0 += G) {
        unsigned n1[G], n2[G];
        int np[G];
        bool nl[G];
#pragma unroll
        for (int g = 0; g < G; g++) n1[g] = n2[g] = 0u, np[g] = -1, nl[g] = false;
        if (i0 + G < NIT) load_words(i0 + G, n1, n2, np, nl);
        Ev v[G];
        int4 e[G];
#pragma unroll
        for (int g = 0; g < G; g++) {
            v[g] = decode(w1[g], w2[g], lv[g]);
            e[g] = make_int4(-1, 0, 0, 0);
            if (v[g].binnable)
                e[g] = a.sol[((v[g].mpe * NXD) + v[g].ax + NX * v[g].doia) * NXD + v[g].bx + NX * v[g].doib];
        }
#pragma unroll
        for (int g = 0; g < G; g++) step(cs + pos[g], w1[g], w2[g], v[g], e[g]);
#pragma unroll
        for (int g = 0; g < G; g++) w1[g] = n1[g], w2[g] = n2[g], pos[g] = np[g], lv[g] = nl[g];
    }
    no_keys_from(nk);
    if (a.cseg) {
        unsigned n = 0;
#pragma unroll
        for (int h = 0; h < NHEADS; h++) n = lane == h ? cc[h] : n;
        if (lane < NHEADS) a.ccnt[c * NHEADS + lane] = n;
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

// tiles[t] = index of the first sorted key of tile t or later, for t in [0, ntiles] (NOKEY is
// later than every tile): a binary search per tile.  The sort orders key bits TILE_BITS + 1 and
// up, so key < t << (TILE_BITS + 1) is monotone along the sorted keys.
__global__ void lm_tiles_kernel(const unsigned *__restrict__ keys, long long n, long long ntiles,
                                unsigned *__restrict__ tiles) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    const unsigned v = (unsigned)(t << (TILE_BITS + 1));
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    tiles[t] = (unsigned)lo;
}

constexpr int TILE_THREADS = 512, TILE_PER_THREAD = TILE / TILE_THREADS;   // 16 bins per thread
static_assert(TILE_PER_THREAD == 16, "the tile write-out packs 16 bins per thread");

__device__ inline unsigned pack16(int lo, int hi) { return ((unsigned)lo & 0xffffu) | ((unsigned)hi << 16); }
__device__ inline unsigned pack8(const int *v) {
    return (unsigned)min(v[0], 255) | (unsigned)min(v[1], 255) << 8 | (unsigned)min(v[2], 255) << 16 |
           (unsigned)min(v[3], 255) << 24;
}

// Tiles of TILE bins, grid-strided: a tile's keys counted per bin in LDS, then its bins in order,
// 16 per thread: mode 0 int16(prompts - delayed), mode 2 int16(prompts), mode 1 (BOTH: the two
// counted apart) int16(prompts) and min(delayed, 255) as bytes, or int16 delayed at sino[S + bin]
// without a byte output.  int16 values wrap as the reference's short counters do.  vec: the
// outputs take 16-byte stores (sino, and delayed or sino + S, 16-byte aligned).
template <bool BOTH>
__global__ __launch_bounds__(TILE_THREADS) void lm_tile_kernel(const unsigned *__restrict__ keys,
                                                               const unsigned *__restrict__ tiles, long long ntiles,
                                                               long long S, int mode, int16_t *__restrict__ sino,
                                                               uint8_t *__restrict__ delayed, int vec) {
    constexpr int NC = BOTH ? 2 : 1;
    __shared__ int4 cnt4[NC * TILE / 4];
    int *cnt = reinterpret_cast<int *>(cnt4);
    const int j0 = threadIdx.x * TILE_PER_THREAD;
    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        for (int j = threadIdx.x; j < NC * TILE / 4; j += TILE_THREADS) cnt4[j] = make_int4(0, 0, 0, 0);
        __syncthreads();
        const unsigned a = tiles[t], b = tiles[t + 1];
        for (unsigned i = a + threadIdx.x; i < b; i += TILE_THREADS) {
            const unsigned k = keys[i];
            const int bin = (int)((k >> 1) & (TILE - 1));
            if (BOTH) atomicAdd(&cnt[(int)(k & 1) * TILE + bin], 1);
            else atomicAdd(&cnt[bin], (mode == 0 && (k & 1)) ? -1 : 1);
        }
        __syncthreads();
        const long long b0 = t * TILE + j0;
        int p[TILE_PER_THREAD], d[TILE_PER_THREAD];
#pragma unroll
        for (int q = 0; q < TILE_PER_THREAD / 4; q++) {
            const int4 v = cnt4[(j0 >> 2) + q];
            p[4 * q] = v.x, p[4 * q + 1] = v.y, p[4 * q + 2] = v.z, p[4 * q + 3] = v.w;
            if (BOTH) {
                const int4 w = cnt4[((TILE + j0) >> 2) + q];
                d[4 * q] = w.x, d[4 * q + 1] = w.y, d[4 * q + 2] = w.z, d[4 * q + 3] = w.w;
            }
        }
        if (vec && b0 + TILE_PER_THREAD <= S) {
            uint4 *o = reinterpret_cast<uint4 *>(sino + b0);
            o[0] = make_uint4(pack16(p[0], p[1]), pack16(p[2], p[3]), pack16(p[4], p[5]), pack16(p[6], p[7]));
            o[1] = make_uint4(pack16(p[8], p[9]), pack16(p[10], p[11]), pack16(p[12], p[13]), pack16(p[14], p[15]));
            if (BOTH) {
                if (delayed) {
                    *reinterpret_cast<uint4 *>(delayed + b0) = make_uint4(pack8(d), pack8(d + 4), pack8(d + 8), pack8(d + 12));
                } else {
                    uint4 *o2 = reinterpret_cast<uint4 *>(sino + S + b0);
                    o2[0] = make_uint4(pack16(d[0], d[1]), pack16(d[2], d[3]), pack16(d[4], d[5]), pack16(d[6], d[7]));
                    o2[1] = make_uint4(pack16(d[8], d[9]), pack16(d[10], d[11]), pack16(d[12], d[13]), pack16(d[14], d[15]));
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < TILE_PER_THREAD; q++) {
                const long long i = b0 + q;
                if (i < S) {
                    sino[i] = (int16_t)(uint16_t)(unsigned)p[q];
                    if (BOTH) {
                        if (delayed) delayed[i] = (uint8_t)min(d[q], 255);
                        else sino[S + i] = (int16_t)(uint16_t)(unsigned)d[q];
                    }
                }
            }
        }
        __syncthreads();   // the counters are cleared for the next tile
    }
}

// Coincidence maps, pass 1: per (chunk range, head) the head's entries of the range's chunks
// counted in LDS, written as the range's partial counts.  Each wave takes one chunk at a time and
// loads its first 256 entries together with the chunk's count.
__global__ __launch_bounds__(1024) void lm_coin_kernel(const unsigned short *__restrict__ cseg,
                                                       const unsigned *__restrict__ ccnt, long long nchunks,
                                                       unsigned *__restrict__ part) {
    __shared__ unsigned cnt[COIN_HEAD];
    const int r = blockIdx.x, h = blockIdx.y;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j = threadIdx.x; j < COIN_HEAD; j += blockDim.x) cnt[j] = 0;
    __syncthreads();
    const long long per = (nchunks + COIN_RANGES - 1) / COIN_RANGES;
    const long long c0 = r * per, c1 = min(nchunks, c0 + per);
    for (long long c = c0 + wv; c < c1; c += nw) {
        const unsigned short *e = cseg + (c * NHEADS + h) * (CHUNK / 2);
        const unsigned n = ccnt[c * NHEADS + h];
        unsigned short v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = e[k * 64 + lane];   // inside the chunk's 1024 slots
#pragma unroll
        for (int k = 0; k < 4; k++)
            if ((unsigned)(k * 64 + lane) < n) atomicAdd(&cnt[v[k]], 1u);
        for (unsigned j = 256 + lane; j < n; j += 64) atomicAdd(&cnt[e[j]], 1u);
    }
    __syncthreads();
    unsigned *o = part + ((long long)r * NHEADS + h) * COIN_HEAD;
    for (int j = threadIdx.x; j < COIN_HEAD; j += blockDim.x) o[j] = cnt[j];
}

// Coincidence maps, pass 2: the ranges' partial counts summed into the .ch halves
// [layer][ring][head][x] (a NULL map is skipped).
__global__ void lm_coin_sum_kernel(const unsigned *__restrict__ part, unsigned *__restrict__ coin_p,
                                   unsigned *__restrict__ coin_d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * NCRYS) return;
    const int map = i / NCRYS, k = i % NCRYS;
    unsigned *out = map ? coin_d : coin_p;
    if (!out) return;
    const int x = k % NX, head = (k / NX) % NHEADS, ring = (k / HX) % NY, layer = k / XYH;
    const int local = ((map * NDOIS + layer) * NY + ring) * NX + x;
    unsigned sum = 0;
    for (int r = 0; r < COIN_RANGES; r++) sum += part[((long long)r * NHEADS + head) * COIN_HEAD + local];
    out[k] = sum;
}

__global__ void lm_stats_init_kernel(unsigned long long *st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ST_WORDS) st[i] = (i == ST_START || i == ST_END) ? NONE : 0ull;
}

// The module-pair LUT by (module pair, A crystal, B crystal) on the device, and the counts.
int ensure_lm(hrrt_delays &c, long long nchunks, bool coin) {
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
    if (!c.d_lm_stats) DTRY(hipMalloc(&c.d_lm_stats, sizeof(unsigned long long) * ST_WORDS));
    if (!c.d_lm_tiles) {
        const long long S = (long long)c.nplanes * c.nsino;
        DTRY(hipMalloc(&c.d_lm_tiles, sizeof(unsigned) * (size_t)((S + TILE - 1) / TILE + 1)));
    }
    const long long keys = nchunks * (CHUNK / 2);
    if (keys > c.lm_key_cap) {
        if (c.d_lm_keys) DTRY(hipFree(c.d_lm_keys));
        if (c.d_lm_sort_tmp) DTRY(hipFree(c.d_lm_sort_tmp));
        c.d_lm_keys = nullptr;
        c.d_lm_sort_tmp = nullptr;
        c.lm_key_cap = 0;
        DTRY(hipMalloc(&c.d_lm_keys, sizeof(unsigned) * 2 * (size_t)keys));
        size_t bytes = 0;
        DTRY(rocprim::radix_sort_keys(nullptr, bytes, c.d_lm_keys, c.d_lm_keys + keys, (size_t)keys, 0, 32));
        DTRY(hipMalloc(&c.d_lm_sort_tmp, bytes));
        c.lm_sort_bytes = bytes;
        c.lm_key_cap = keys;
    }
    if (coin && !c.d_lm_coin_part)
        DTRY(hipMalloc(&c.d_lm_coin_part, sizeof(unsigned) * (size_t)COIN_RANGES * NHEADS * COIN_HEAD));
    if (coin && nchunks > c.lm_coin_cap) {
        if (c.d_lm_coin_seg) DTRY(hipFree(c.d_lm_coin_seg));
        if (c.d_lm_coin_cnt) DTRY(hipFree(c.d_lm_coin_cnt));
        c.d_lm_coin_seg = nullptr;
        c.d_lm_coin_cnt = nullptr;
        c.lm_coin_cap = 0;
        DTRY(hipMalloc(&c.d_lm_coin_seg, sizeof(unsigned short) * (size_t)nchunks * NHEADS * (CHUNK / 2)));
        DTRY(hipMalloc(&c.d_lm_coin_cnt, sizeof(unsigned) * (size_t)nchunks * NHEADS));
        c.lm_coin_cap = nchunks;
    }
    if (nchunks > c.lm_chunk_cap) {
        if (c.d_lm_chunks) DTRY(hipFree(c.d_lm_chunks));
        c.d_lm_chunks = nullptr;
        DTRY(hipMalloc(&c.d_lm_chunks, sizeof(unsigned) * 65 * (size_t)nchunks));   // summaries, then masks
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
    if ((long long)c->nplanes * c->nsino > (1ll << 30)) return hrrt_dl::fail(HRRT_ERR_ARG, "sinogram larger than 2^30 bins");
    // key slots (one per two words) are indexed in 32 bits
    if (nwords > (1ll << 33) - (1ll << 22)) return hrrt_dl::fail(HRRT_ERR_ARG, "stream longer than 2^33 - 2^22 words");
    hipStream_t s = (hipStream_t)stream;
    const long long nchunks = (nwords + CHUNK - 1) / CHUNK;
    const bool coin = d_coin_p || d_coin_d;
    int r = ensure_lm(*c, std::max(nchunks, 1ll), coin);
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
    a.keys = c->d_lm_keys;
    a.coin_p = d_coin_p;
    a.coin_d = d_coin_d;
    a.cseg = coin ? c->d_lm_coin_seg : nullptr;
    a.ccnt = coin ? c->d_lm_coin_cnt : nullptr;
    a.chunks = c->d_lm_chunks;
    a.cmask = c->d_lm_chunks + c->lm_chunk_cap;
    a.w16 = (uintptr_t)d_words % 16 == 0;
    a.st = c->d_lm_stats;
    lm_stats_init_kernel<<<(ST_WORDS + 255) / 256, 256, 0, s>>>(a.st);
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
    DTRY(hipGetLastError());
    // the key slots by tile: the tile bits (key bits 13 and up) sorted, through enough bits that
    // NOKEY's exceed every tile's
    const long long ntiles = (S + TILE - 1) / TILE;
    const long long nk = nchunks * (CHUNK / 2);
    unsigned *sorted = c->d_lm_keys + c->lm_key_cap;
    if (nk > 0) {
        unsigned end_bit = TILE_BITS + 1;
        while ((ntiles >> (end_bit - TILE_BITS - 1)) != 0) end_bit++;
        size_t bytes = 0;
        DTRY(rocprim::radix_sort_keys(nullptr, bytes, c->d_lm_keys, sorted, (size_t)nk, TILE_BITS + 1, end_bit, s));
        if (bytes > c->lm_sort_bytes)
            return hrrt_dl::fail(HRRT_ERR_HIP, "radix sort scratch of " + std::to_string(bytes) + " bytes exceeds its buffer");
        DTRY(rocprim::radix_sort_keys(c->d_lm_sort_tmp, bytes, c->d_lm_keys, sorted, (size_t)nk, TILE_BITS + 1, end_bit, s));
    }
    lm_tiles_kernel<<<(unsigned)((ntiles + 1 + 255) / 256), 256, 0, s>>>(sorted, nk, ntiles, c->d_lm_tiles);
    const unsigned tgrid = (unsigned)std::min(ntiles, 1024ll);   // every workgroup resident (4 per CU)
    const bool both = opt->mode == 1;
    const int vec = (uintptr_t)d_sino % 16 == 0 &&
                    (!both || (d_delayed ? (uintptr_t)d_delayed % 16 == 0 : S % 8 == 0));
    if (both)
        lm_tile_kernel<true><<<tgrid, TILE_THREADS, 0, s>>>(sorted, c->d_lm_tiles, ntiles, S, opt->mode, d_sino, d_delayed, vec);
    else
        lm_tile_kernel<false><<<tgrid, TILE_THREADS, 0, s>>>(sorted, c->d_lm_tiles, ntiles, S, opt->mode, d_sino, d_delayed, vec);
    if (coin) {
        lm_coin_kernel<<<dim3(COIN_RANGES, NHEADS), 1024, 0, s>>>(a.cseg, a.ccnt, nchunks, c->d_lm_coin_part);
        lm_coin_sum_kernel<<<(2 * NCRYS + 255) / 256, 256, 0, s>>>(c->d_lm_coin_part, d_coin_p, d_coin_d);
    }
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
