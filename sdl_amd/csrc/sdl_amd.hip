// MI355X (gfx950) implementation of the SDL module-map chain behind the C-ABI
// of include/sdl_amd.h.
//
// Data layout (all flat SoA in HBM, one "slot" per (event, sensor module)):
//   hits        x y z idx (input) + rt r3 phi slot, and the two strip-edge
//               points of endcap-2S hits (Event.cc:132-138)
//   slots       hit range (counting sort by slot), MD/SG/TP/TL ranges
//   objects     dense arrays ordered by their inner object: MDs by lower slot,
//               SGs by inner MD, TPs and TLs by inner SG, TCs by tracklet.
// Each combinatorial stage runs one wavefront per work item (lower slot for
// MD, inner MD for SG, inner SG for TL); the 64 lanes evaluate candidates in
// the reference's loop order and survivors are compacted with a wave ballot
// into a slab sized by the candidate count, then repacked densely after an
// exclusive scan.  TP, TC-from-triplets and TC-from-tracklets are integer joins
// over those dense arrays (see the comments at each kernel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <iterator>
#include <map>
#include <sstream>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/sdl_amd.h"
#include "sdl_cuts.h"
#include "sdl_mlp.h"
#include "sdl_module.h"

using namespace sdl;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_OK(call)                                                                          \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) return fail(SDL_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kNVariantsTL = 3;  // tracklet algorithm variants with device kernels: BBBB, BBEE, EEEE
// Tracklet algorithms (the TLAlgo values of Algo.h:21-26)
enum { TLA_DEFAULT = 0, TLA_NM1 = 1, TLA_ALLCOMB = 2 };

// Set while a chain is being captured into a HIP graph: an allocation is not
// allowed then, so a buffer that must grow ends the capture (the chain then
// runs uncaptured once, which grows it).
thread_local bool g_capturing = false;
thread_local bool g_captureBlocked = false;
// Bumped by every (re)allocation on any thread, so a captured graph knows its
// buffers moved (a handle may be driven from several host threads in turn).
std::atomic<unsigned long long> g_allocEpoch{0};
// The stream of the handle whose call runs on this thread (set by the C-ABI
// entry points, OnStream): buffers grow in that stream's order (hipFreeAsync,
// hipMallocAsync), so one handle's growth never waits for another handle's
// work.  Without one (geometry loading) allocation is synchronous.
thread_local hipStream_t g_allocStream = nullptr;
struct OnStream {
    hipStream_t prev;
    explicit OnStream(hipStream_t s) : prev(g_allocStream) { g_allocStream = s; }
    ~OnStream() { g_allocStream = prev; }
};

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    int ensure(size_t n) {
        if (n <= cap && p) return 0;
        if (g_capturing) {
            g_captureBlocked = true;
            return fail(SDL_ERR_STATE, "allocation during graph capture");
        }
        ++g_allocEpoch;
        size_t want = std::max<size_t>(n, 1);
        want = std::max(want, cap + cap / 2);
        // queued work may still read the old buffer: free it in stream order
        if (p) {
            if (g_allocStream) {
                (void)hipFreeAsync(p, g_allocStream);
            } else {
                (void)hipDeviceSynchronize();
                (void)hipFree(p);
            }
        }
        p = nullptr;
        cap = 0;
        const hipError_t e = g_allocStream ? hipMallocAsync((void**)&p, want * sizeof(T), g_allocStream)
                                           : hipMalloc(&p, want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(SDL_ERR_HIP, "device allocation failed");
        }
        cap = want;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// ------------------------------------------------------------------ run totals
// Totals of one run, written on the device by the scans (and read back by the
// host once per call).  Kernels take their item counts from here, so a whole
// chain can be enqueued without a host round trip.  Buffers sized ahead of
// time are checked against the totals by the kernel that would fill them
// (Guard): on overflow it raises TOT_ABORT and writes nothing, every later
// kernel of the call then sees zero items, and the host runs the call again
// with buffers sized from the totals.
enum {
    TOT_MDCAP = 0, TOT_MD, TOT_SGCAP, TOT_SG, TOT_TP, TOT_TLCAP, TOT_TL, TOT_TPE, TOT_TCP, TOT_TCL,
    TOT_TLC0, TOT_TLC1, TOT_TLC2,  // second-stage tracklet candidates per variant
    TOT_TLP0, TOT_TLP1, TOT_TLP2,  // first-stage tracklet pairs per variant
    TOT_TLPAIRS, TOT_TLCANDS,      // all first-stage pairs, all second-stage candidates
    TOT_ABORT, TOT_UNKNOWN, TOT_N
};

// Number of work items: fixed by the host (idx < 0) or a run total.
struct Count {
    long long host;
    const long long* tot;
    int idx;
    __device__ __forceinline__ long long get() const {
        if (tot[TOT_ABORT]) return 0;
        return idx < 0 ? host : tot[idx];
    }
};
// A buffer of `cap` entries about to receive tot[idx] entries.
struct Guard {
    long long* tot;
    int idx;
    long long cap;
    int stage;
    // true (and the run aborted) when the total does not fit; uniform over the grid
    __device__ __forceinline__ bool over() const {
        if (tot[TOT_ABORT]) return true;
        const long long t = tot[idx];
        if (t >= 0 && t <= cap) return false;
        if (threadIdx.x == 0) tot[TOT_ABORT] = 1 + stage;
        return true;
    }
};

// Largest k in [0, n) with off[k] <= j, for off[0] <= j < off[n].
__device__ __forceinline__ long long rangeOf(const long long* off, long long n, long long j) {
    long long lo = 0, hi = n;
    while (hi - lo > 1) {
        const long long mid = (lo + hi) >> 1;
        if (off[mid] <= j) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------ scans
// Exclusive scan of n counts into n+1 offsets (out[n] = total) in one launch:
// chained scan with decoupled look-back.  Tiles take their index from a ticket
// counter, so every predecessor of a tile has started when the tile looks
// back, and a predecessor publishes its aggregate before it waits on anything:
// every wait ends.  A tile's state word packs
// [epoch:20][flag:2][value:42]; the epoch changes after every scan (the last
// tile to finish advances it and re-arms the ticket), so the state array needs
// no clearing between scans.  Totals of 2^42 or more are reported as -1.
// A state word carries its own value, so publishing and polling are relaxed
// device-scope atomics: no fence (which would write back or invalidate the
// XCD's L2) is needed.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr unsigned long long kScanValMask = (1ULL << 42) - 1;
constexpr unsigned kScanEpochMax = (1u << 20) - 1;
enum : unsigned { SCAN_AGG = 1, SCAN_INCL = 2 };

struct ScanCtl {
    unsigned long long ticket;
    unsigned long long done;
    unsigned long long epoch;
};

__device__ __forceinline__ unsigned long long scanPack(unsigned epoch, unsigned flag, long long v) {
    const unsigned long long val = v < 0 || (unsigned long long)v >= kScanValMask ? kScanValMask : (unsigned long long)v;
    return ((unsigned long long)epoch << 44) | ((unsigned long long)flag << 42) | val;
}

typedef long long ll2v __attribute__((ext_vector_type(2)));

// The kScanItems items a thread owns, zeros past n; whole groups of aligned
// items are read with vector loads.
template <class TIn>
__device__ __forceinline__ void loadItems(const TIn* in, long long b, long long n, bool vec, long long (&v)[kScanItems]) {
    if (vec && b + kScanItems <= n) {
        if constexpr (sizeof(TIn) == 4) {
            const int4 a = *reinterpret_cast<const int4*>(in + b);
            const int4 c = *reinterpret_cast<const int4*>(in + b + 4);
            const int w[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) v[k] = (long long)TIn(w[k]);
            return;
        } else if constexpr (sizeof(TIn) == 1) {
            const unsigned long long w = *reinterpret_cast<const unsigned long long*>(in + b);
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) v[k] = (long long)TIn((unsigned char)(w >> (8 * k)));
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = b + k < n ? (long long)in[b + k] : 0;
}

// Writes the exclusive scan of a thread's items v (zeros past n), starting at r.
template <class TIn>
__device__ __forceinline__ void writeTile(long long* out, long long b, long long n, bool vec,
                                          const long long (&v)[kScanItems], long long r) {
    if (vec && b + kScanItems <= n) {
        ll2v o[kScanItems / 2];
#pragma unroll
        for (int k = 0; k < kScanItems; k += 2) {
            o[k / 2].x = r;
            r += v[k];
            o[k / 2].y = r;
            r += v[k + 1];
        }
#pragma unroll
        for (int k = 0; k < kScanItems / 2; ++k) reinterpret_cast<ll2v*>(out + b)[k] = o[k];
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (b + k < n) out[b + k] = r;
            r += v[k];
        }
    }
}

// Block-wide inclusive scan of one value per thread: returns (exclusive prefix
// of the thread, block total).
__device__ __forceinline__ void blockScan(long long s, long long* waveSum, long long& excl, long long& total) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    long long x = s;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const long long y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) waveSum[wave] = x;
    __syncthreads();
    long long wPrefix = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) {
        if (w < wave) wPrefix += waveSum[w];
        total += waveSum[w];
    }
    excl = wPrefix + x - s;
}

// Each tile scans kScanTile consecutive items held in registers: it sums them,
// publishes the sum, looks back over its predecessors' words for its prefix
// and writes the offsets.  Every tile of the launch takes part (the ticket
// protocol needs all of them); tiles past the end of the input scan zeros.
// Used for at most kWave tiles: one look-back window then covers every
// predecessor (each further window costs a device-coherent round trip).
template <class TIn>
__global__ void __launch_bounds__(kBlock) k_scan(const TIn* in, long long* out, Count cn, int nTiles,
                                                 unsigned long long* state, unsigned long long stateCap,
                                                 ScanCtl* ctl, long long* tot, int totIdx) {
    __shared__ long long waveSum[kWavesPerBlock];
    __shared__ long long prefixSh;
    __shared__ int tileSh;
    __shared__ unsigned epochSh;
    if (threadIdx.x == 0) {
        tileSh = int(__hip_atomic_fetch_add(&ctl->ticket, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        epochSh = unsigned(__hip_atomic_load(&ctl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    __syncthreads();
    const int tile = tileSh;
    const unsigned epoch = epochSh;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    long long n = cn.get();
    if (n > (long long)nTiles * kScanTile) {  // launched for fewer items than the total: abort the run
        if (threadIdx.x == 0) tot[TOT_ABORT] = 1 + 100;
        n = 0;
    }
    const bool vec = ((unsigned long long)in & (sizeof(TIn) * kScanItems - 1)) == 0 &&
                     ((unsigned long long)out & 15) == 0;
    const long long b = (long long)tile * kScanTile + (long long)threadIdx.x * kScanItems;
    long long v[kScanItems];
    loadItems(in, b, n, vec, v);
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    long long excl0, agg;
    blockScan(s, waveSum, excl0, agg);
    if (wave == 0) {
        long long excl = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(&state[0], scanPack(epoch, SCAN_INCL, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(&state[tile], scanPack(epoch, SCAN_AGG, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int j = tile - 1;
            for (;;) {
                const int idx = j - lane;
                unsigned long long w = 0;
                bool ready = idx < 0;
                for (;;) {
                    if (!ready) {
                        w = __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ready = unsigned(w >> 44) == epoch && ((w >> 42) & 3u) != 0;
                    }
                    if (__all(ready)) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                const bool incl = idx >= 0 && ((w >> 42) & 3u) == SCAN_INCL;
                const long long val = idx >= 0 ? (long long)(w & kScanValMask) : 0;
                const unsigned long long inclMask = __ballot(incl);
                // lanes up to and including the nearest inclusive predecessor
                const int stop = inclMask ? __builtin_ctzll(inclMask) : kWave - 1;
                long long pv = lane <= stop ? val : 0;
#pragma unroll
                for (int off = 1; off < kWave; off <<= 1) pv += __shfl_xor(pv, off);
                excl += pv;
                if (inclMask) break;
                j -= kWave;
            }
            if (lane == 0)
                __hip_atomic_store(&state[tile], scanPack(epoch, SCAN_INCL, excl + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) prefixSh = excl;
    }
    __syncthreads();
    const long long prefix = prefixSh;
    writeTile<TIn>(out, b, n, vec, v, prefix + excl0);
    if (tile == nTiles - 1 && threadIdx.x == 0) {
        const long long total = prefix + agg;
        const long long t = (unsigned long long)prefix >= kScanValMask || (unsigned long long)total >= kScanValMask ? -1 : total;
        out[n] = t;
        if (totIdx >= 0) tot[totIdx] = t;
    }
    // the last tile to finish re-arms the ticket and advances the epoch
    __shared__ bool lastSh;
    if (threadIdx.x == 0)
        lastSh = __hip_atomic_fetch_add(&ctl->done, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                 (unsigned long long)(nTiles - 1);
    __syncthreads();
    if (!lastSh) return;
    if (epoch == kScanEpochMax) {
        for (unsigned long long i = threadIdx.x; i < stateCap; i += blockDim.x) state[i] = 0;
    }
    if (threadIdx.x == 0) {
        ctl->ticket = 0;
        ctl->done = 0;
        ctl->epoch = epoch == kScanEpochMax ? 1 : epoch + 1;
    }
}

// Large scans, pass 1: the sum of every tile of kScanTile items.
template <class TIn>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(const TIn* in, Count cn, long long* tileSum) {
    __shared__ long long waveSum[kWavesPerBlock];
    const long long n = cn.get();
    const bool vec = ((unsigned long long)in & (sizeof(TIn) * kScanItems - 1)) == 0;
    const long long b = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    long long v[kScanItems];
    loadItems(in, b, n, vec, v);
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & (kWave - 1)) == 0) waveSum[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) t += waveSum[w];
        tileSum[blockIdx.x] = t;
    }
}

// Large scans, pass 2: each tile takes its prefix from tileOff (the scanned
// tile sums) or, without it, sums its predecessors' tile sums itself (used up
// to kScanSelfTiles tiles: a few thousand values, read from cache), then scans
// its items; the last tile writes the total to out[n] and tot[totIdx].
constexpr int kScanSelfTiles = 4096;
template <class TIn>
__global__ void __launch_bounds__(kBlock) k_scan_apply(const TIn* in, long long* out, Count cn, int nTiles,
                                                       const long long* tileSum, const long long* tileOff,
                                                       long long* tot, int totIdx) {
    __shared__ long long waveSum[kWavesPerBlock];
    __shared__ long long prefixSh;
    const int tile = blockIdx.x;
    long long n = cn.get();
    const bool over = n > (long long)nTiles * kScanTile;  // launched for fewer items than the total: abort the run
    if (over) n = 0;
    long long prefix;
    if (tileOff) {
        prefix = tileOff[tile];
    } else {
        long long p = 0;
        for (int t = threadIdx.x; t < tile; t += blockDim.x) p += tileSum[t];
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) p += __shfl_xor(p, off);
        if ((threadIdx.x & (kWave - 1)) == 0) waveSum[threadIdx.x / kWave] = p;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long t = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) t += waveSum[w];
            prefixSh = t;
        }
        __syncthreads();
        prefix = prefixSh;
    }
    const bool vec = ((unsigned long long)in & (sizeof(TIn) * kScanItems - 1)) == 0 &&
                     ((unsigned long long)out & 15) == 0;
    const long long b = (long long)tile * kScanTile + (long long)threadIdx.x * kScanItems;
    long long v[kScanItems];
    loadItems(in, b, n, vec, v);
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    long long excl, agg;
    blockScan(s, waveSum, excl, agg);
    writeTile<TIn>(out, b, n, vec, v, prefix + excl);
    if (tile == nTiles - 1 && threadIdx.x == 0) {
        if (over) tot[TOT_ABORT] = 1 + 100;
        const long long total = prefix + agg;
        const long long t = (unsigned long long)total >= kScanValMask ? -1 : total;
        out[n] = t;
        if (totIdx >= 0) tot[totIdx] = t;
    }
}

// Zero-fill (a kernel rather than a memset, so that captured graphs replay it reliably).
template <class T>
__global__ void k_zero(T* p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

// Zero-fill of up to kZeroArrays arrays in one launch (sizes in 32-bit words).
constexpr int kZeroArrays = 4;
struct ZeroSet {
    unsigned* p[kZeroArrays];
    long long end[kZeroArrays];  // running word totals
    int n;
};

__global__ void k_zero_many(ZeroSet z) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long start = 0;
    for (int a = 0; a < z.n; ++a) {
        if (i < z.end[a]) {
            z.p[a][i - start] = 0u;
            return;
        }
        start = z.end[a];
    }
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long n, unsigned long long v) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct ScanScratch {
    unsigned long long* state = nullptr;
    unsigned long long cap = 0;
    ScanCtl* ctl = nullptr;
    static constexpr int kLevels = 3;  // 2048^3 items
    DBuf<long long> tileSum[kLevels], tileOff[kLevels];
    ~ScanScratch() {
        if (state) (void)hipFree(state);
        if (ctl) (void)hipFree(ctl);
    }
    // Makes room for nTiles tile states; new memory starts at epoch 0 (never current).
    int ensure(unsigned long long nTiles, hipStream_t st) {
        if (g_capturing && (!ctl || nTiles > cap)) {
            g_captureBlocked = true;
            return fail(SDL_ERR_STATE, "allocation during graph capture");
        }
        if (!ctl || nTiles > cap) ++g_allocEpoch;
        if (!ctl) {
            if (hipMalloc(&ctl, sizeof(ScanCtl)) != hipSuccess) return fail(SDL_ERR_HIP, "hipMalloc failed");
            const ScanCtl init{0, 0, 1};
            if (hipMemcpyAsync(ctl, &init, sizeof(init), hipMemcpyHostToDevice, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return fail(SDL_ERR_HIP, "scan state init failed");
        }
        if (nTiles <= cap) return 0;
        const unsigned long long want = std::max<unsigned long long>(nTiles, cap * 2);
        if (state) (void)hipFreeAsync(state, st);
        state = nullptr;
        cap = 0;
        if (hipMallocAsync((void**)&state, want * sizeof(unsigned long long), st) != hipSuccess) {
            state = nullptr;
            return fail(SDL_ERR_HIP, "device allocation failed");
        }
        hipLaunchKernelGGL(k_fill_u64, dim3(unsigned((want + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, state, want, 0ULL);
        cap = want;
        return 0;
    }
};

// out must hold n+1 entries; out[n] (and tot[totIdx], if totIdx >= 0) receives
// the total (-1 if it does not fit 42 bits).  nMax bounds the item count on
// the host (it sizes the launch); the count itself may live on the device.
// Up to kWave tiles: one single-pass launch.  More: tile sums, then the tiles'
// offsets (from the sums directly, or by scanning the sums at the next level),
// then the items.
template <class TIn>
int scanExclusive(const TIn* in, long long* out, Count cn, long long nMax, ScanScratch& sc, hipStream_t st,
                  long long* tot, int totIdx, int level = 0) {
    const long long nTiles = std::max(1LL, (nMax + kScanTile - 1) / kScanTile);
    if (nTiles <= kWave) {
        if (int r = sc.ensure((unsigned long long)nTiles, st)) return r;
        hipLaunchKernelGGL(k_scan<TIn>, dim3(unsigned(nTiles)), dim3(kBlock), 0, st, in, out, cn, int(nTiles), sc.state,
                           sc.cap, sc.ctl, tot, totIdx);
        return 0;
    }
    if (nTiles > (1LL << 30) || level >= ScanScratch::kLevels) return fail(SDL_ERR_CAPACITY, "scan too large");
    DBuf<long long>& sum = sc.tileSum[level];
    DBuf<long long>& off = sc.tileOff[level];
    if (int r = sum.ensure(nTiles)) return r;
    hipLaunchKernelGGL(k_scan_reduce<TIn>, dim3(unsigned(nTiles)), dim3(kBlock), 0, st, in, cn, sum.p);
    const long long* tileOff = nullptr;
    if (nTiles > kScanSelfTiles) {
        if (int r = off.ensure(nTiles + 1)) return r;
        if (int r = scanExclusive<long long>(sum.p, off.p, Count{nTiles, tot, -1}, nTiles, sc, st, tot, -1, level + 1))
            return r;
        tileOff = off.p;
    }
    hipLaunchKernelGGL(k_scan_apply<TIn>, dim3(unsigned(nTiles)), dim3(kBlock), 0, st, in, out, cn, int(nTiles),
                       sum.p, tileOff, tot, totIdx);
    return 0;
}

// ------------------------------------------------------------------ hits
// detId -> module index as an open-addressing hash table (linear probing, a
// power-of-two size of at least 4 NM slots, empty slots holding kNoDet): a hit
// finds its module in one or two loads instead of a binary search over the
// sorted detIds (15 dependent loads for 27 K modules).
constexpr uint32_t kNoDet = 0xFFFFFFFFu;
__host__ __device__ inline uint32_t detHash(uint32_t d, int bits) { return (d * 2654435761u) >> (32 - bits); }

__global__ void k_hit_prep(const float* x, const float* y, const float* z, const unsigned* detId,
                           const long long* evOff, int E, const uint2* detTable, int hashBits, int NM,
                           const ModDev* mods,
                           float* rt, float* r3, float* phi, float* alpha, int* slot, float4* eHi, float4* eLo,
                           int* slotCount, long long* tot) {
    const long long nh = evOff[E];  // hits of the batch (uploaded with the event offsets)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nh) return;
    // event of the hit
    int lo = 0, hi = E;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (evOff[mid] <= i) lo = mid; else hi = mid;
    }
    const int e = lo;
    // module of the hit (the table always has empty slots, so every probe ends)
    const uint32_t d = detId[i];
    const uint32_t mask = (1u << hashBits) - 1u;
    int a = -1;
    for (uint32_t h = detHash(d, hashBits); d != kNoDet; h = (h + 1u) & mask) {
        const uint2 t = detTable[h];
        if (t.x == d) {
            a = int(t.y);
            break;
        }
        if (t.x == kNoDet) break;
    }
    if (a < 0 || a >= NM) {
        atomicAdd((unsigned long long*)&tot[TOT_UNKNOWN], 1ULL);
        slot[i] = -1;
        return;
    }
    const float xx = x[i], yy = y[i], zz = z[i];
    const float rtv = rtOf(xx, yy);
    rt[i] = rtv;
    r3[i] = r3Of(xx, yy, zz);
    phi[i] = phiOf(xx, yy);
    alpha[i] = asinF(fmin2(rtv * k2Rinv() / kPtCut, kSinAlphaMax));
    const int s = e * NM + a;
    slot[i] = s;
    const ModDev m = mods[a];
    if (m.subdet == SUB_ENDCAP && m.type == TYPE_2S) {
        HitView h{xx, yy, zz, 0, 0, 0, 0};
        HitView hh = edgeOf(h, m.centroidPhi, true);
        HitView hl = edgeOf(h, m.centroidPhi, false);
        eHi[i] = make_float4(hh.x, hh.y, hh.rt, hh.phi);
        eLo[i] = make_float4(hl.x, hl.y, hl.rt, hl.phi);
    }
    atomicAdd(&slotCount[s], 1);
}

__global__ void k_hit_scatter(const long long* evOff, int E, const int* slot, const long long* slotStart, int* slotFill,
                              int* sortedHit) {
    const long long nh = evOff[E];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nh) return;
    const int s = slot[i];
    if (s < 0) return;
    const int pos = atomicAdd(&slotFill[s], 1);
    sortedHit[slotStart[s] + pos] = i;
}

// Sorts each segment ascending: restores insertion order inside each slot
// (Module::addHit appends, Module.cc:128-137) and makes atomically scattered
// lists deterministic.  kSortLanes lanes per segment (segments hold a few
// elements on average); up to 4 * kSortLanes elements are ranked in registers,
// longer segments are sorted by the group's first lane.  The hit and SG lists
// use 16 lanes; the tracklet lists (two or three per inner SG on average at
// PU200 density) use 8, where fewer idle lanes outweigh the rare list longer
// than 32 that one lane sorts.
constexpr int kSortLanesWide = 16, kSortLanesTL = 8;
template <int kSortLanes>
__global__ void __launch_bounds__(kBlock) k_sort_segments(Count cseg, const long long* start, int* vals) {
    const int nseg = int(cseg.get());
    const int gl = threadIdx.x & (kSortLanes - 1);
    const long long sg = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kSortLanes;
    if (sg >= nseg) return;  // uniform over the group
    const long long a = start[sg], b = start[sg + 1];
    const int n = int(b - a);
    if (n <= 1) return;
    if (n <= 4 * kSortLanes) {  // every lane of the group takes part in the shuffles below
        int v[4], rank[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int idx = gl + k * kSortLanes;
            v[k] = idx < n ? vals[a + idx] : INT_MAX;
        }
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            const int m = min(n - kb * kSortLanes, kSortLanes);
            for (int src = 0; src < m; ++src) {
                const int o = __shfl(v[kb], src, kSortLanes);
                const int j = kb * kSortLanes + src;
#pragma unroll
                for (int k = 0; k < 4; ++k) rank[k] += (o < v[k]) || (o == v[k] && j < gl + k * kSortLanes);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (gl + k * kSortLanes < n) vals[a + rank[k]] = v[k];
        return;
    }
    if (gl != 0) return;
    for (long long i = a + 1; i < b; ++i) {
        int v = vals[i];
        long long j = i - 1;
        while (j >= a && vals[j] > v) {
            vals[j + 1] = vals[j];
            --j;
        }
        vals[j + 1] = v;
    }
}

__device__ __forceinline__ HitView loadHit(int h, const float* x, const float* y, const float* z, const float* rt,
                                           const float* r3, const float* phi, const float* alpha) {
    return HitView{x[h], y[h], z[h], rt[h], r3[h], phi[h], alpha[h]};
}

// ------------------------------------------------------------------ MD
// Candidate count per lower slot: lower hits x upper hits (Event.cc:257-305).
// (also clears the slot's MD count, which the MD kernel accumulates)
__global__ void k_md_cap(int NS, int NM, const ModDev* mods, const long long* slotHitStart, int* mdCap, int* mdCount) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= NS) return;
    mdCount[s] = 0;
    const int m = s % NM;
    const ModDev md = mods[m];
    int cap = 0;
    if (md.isLower && md.partner >= 0) {
        const int ps = s - m + md.partner;
        const long long nLo = slotHitStart[s + 1] - slotHitStart[s];
        const long long nUp = slotHitStart[ps + 1] - slotHitStart[ps];
        cap = int(nLo * nUp);
    }
    mdCap[s] = cap;
}

// Pseudo mini-doublets (Event.cc:307-430): every hit on the anchor side (the
// lower sensor, or the upper one when the lower is the strip sensor of a PS
// module) paired with the first hit of the other sensor; one candidate each,
// all accepted (AllComb, MiniDoublet.cc:223-226).
__device__ __forceinline__ bool pseudoLoopsLower(const ModDev& md) { return !(md.type == TYPE_PS && md.layerType != LT_PIXEL); }

__global__ void k_pmd_cap(int NS, int NM, const ModDev* mods, const long long* slotHitStart, int* mdCap,
                          int* mdCount) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= NS) return;
    mdCount[s] = 0;
    const int m = s % NM;
    const ModDev md = mods[m];
    int cap = 0;
    if (md.isLower && md.partner >= 0) {
        const int ps = s - m + md.partner;
        const long long nLo = slotHitStart[s + 1] - slotHitStart[s];
        const long long nUp = slotHitStart[ps + 1] - slotHitStart[ps];
        cap = int(pseudoLoopsLower(md) ? (nUp > 0 ? nLo : 0) : (nLo > 0 ? nUp : 0));
    }
    mdCap[s] = cap;
}

// Records of hits, mini-doublets and segments (32 bytes at most) are gathered
// with 32-bit byte offsets: the load then takes the array base from scalar
// registers and a 32-bit lane offset, where a 64-bit index costs a sign
// extension, a shift and a 64-bit add per gather.  The sizing steps refuse a
// call with more than kMaxObjects32 hits, mini-doublets or segments.
constexpr long long kMaxObjects32 = (1LL << 32) / 32;
template <class T>
__device__ __forceinline__ T ld32(const T* base, int i) {
    static_assert(sizeof(T) <= 32, "record size");
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + unsigned(i) * unsigned(sizeof(T)));
}

struct HitsDev {
    const float *x, *y, *z, *rt, *r3, *phi, *alpha;
    const int* idx;
    const float4 *eHi, *eLo;
};

// One thread per candidate (lower hit x upper hit of a lower slot, in the
// reference's loop order): the MD algorithm, its cached values at the
// candidate's index, a pass flag, and a count per slot.
__global__ void __launch_bounds__(kBlock)
k_md(Count cCand, int NS, int NM, const ModDev* mods, HitsDev H, const long long* slotHitStart, const int* sortedHit,
     const long long* mdCapOff, int2* slabHits, float4* slabF, int* slabSlot, unsigned char* flag, int* mdCount,
     int allComb, Guard g) {
    if (g.over()) return;
    const long long nCand = cCand.get();
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nCand) return;
    const int item = int(rangeOf(mdCapOff, NS, j));
    const int m = item % NM;
    const ModDev md = mods[m];
    const int ps = item - m + md.partner;
    const long long upStart = slotHitStart[ps];
    const int nUp = int(slotHitStart[ps + 1] - upStart);
    const int k = int(j - mdCapOff[item]);
    const int i = k / nUp, jj = k - i * nUp;
    const int lo = sortedHit[slotHitStart[item] + i];
    const int up = sortedHit[upStart + jj];
    const HitView hl = loadHit(lo, H.x, H.y, H.z, H.rt, H.r3, H.phi, H.alpha);
    const HitView hu = loadHit(up, H.x, H.y, H.z, H.rt, H.r3, H.phi, H.alpha);
    // AllComb_MDAlgo (MiniDoublet.cc:223-226): every candidate, the cached values
    // left at the constructor's 0 (MiniDoublet.cc:41-51)
    MDOut o{0.f, 0.f, 0.f, 0.f};
    const bool pass = allComb || mdAlgo(hl, hu, md, o);
    flag[j] = pass ? 1 : 0;
    if (pass) {
        slabHits[j] = make_int2(lo, up);
        slabF[j] = make_float4(o.dz, o.shiftedDz, o.dphi, o.dphichange);
        slabSlot[j] = item;
        atomicAdd(&mdCount[item], 1);
    }
}

// One thread per pseudo MD, written in the slab format of k_md with the
// cached values 0 (MiniDoublet.cc:41-51).
__global__ void k_pmd(Count cCand, int NS, int NM, const ModDev* mods, const long long* slotHitStart,
                      const int* sortedHit, const long long* mdCapOff, int2* slabHits, float4* slabF, int* slabSlot,
                      unsigned char* flag, int* mdCount, Guard g) {
    if (g.over()) return;
    const long long nCand = cCand.get();
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nCand) return;
    const int item = int(rangeOf(mdCapOff, NS, j));
    const int m = item % NM;
    const ModDev md = mods[m];
    const int ps = item - m + md.partner;
    const int k = int(j - mdCapOff[item]);
    const bool loopLower = pseudoLoopsLower(md);
    const int lo = sortedHit[slotHitStart[item] + (loopLower ? k : 0)];
    const int up = sortedHit[slotHitStart[ps] + (loopLower ? 0 : k)];
    flag[j] = 1;
    slabHits[j] = make_int2(lo, up);
    slabF[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    slabSlot[j] = item;
    atomicAdd(&mdCount[item], 1);
}

struct MDDev {
    int *lo, *up, *anchor, *slot;
    float4* a;   // anchor x, y, z, rt
    float4* b;   // anchor r3, phi; dz, shiftedDz
    float2* c;   // dphi, dphichange
    float4* p;   // alpha1 = asin(min(rt k2, .95)), tanf(alpha1), tan(alpha1)/alpha1 (double), asin(min(rt k2 / 3, .95))
    double* q;   // mulsFactor(z, rt) of the anchor
};

// Dense repack; the anchor is the pixel hit of a PS module, else the lower hit
// (MiniDoublet.cc:60-81).
// One thread per candidate: survivors to their dense position (the exclusive
// scan of the pass flags, i.e. slot order then the reference's loop order).
__global__ void k_md_repack(Count cCand, const ModDev* mods, int NM, HitsDev H, const unsigned char* flag,
                            const long long* pos, const int2* slabHits, const float4* slabF, const int* slabSlot,
                            MDDev D, Guard g) {
    if (g.over()) return;
    const long long nCand = cCand.get();
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nCand || !flag[j]) return;
    const long long o = pos[j];
    const int2 h = slabHits[j];
    const float4 f = slabF[j];
    const int item = slabSlot[j];
    const ModDev md = mods[item % NM];
    const bool upAnchor = md.type == TYPE_PS && md.layerType != LT_PIXEL;
    const int an = upAnchor ? h.y : h.x;
    D.lo[o] = h.x;
    D.up[o] = h.y;
    D.anchor[o] = an;
    D.slot[o] = item;
    const float rta = H.rt[an];
    const float a1 = H.alpha[an];
    D.a[o] = make_float4(H.x[an], H.y[an], H.z[an], rta);
    D.b[o] = make_float4(H.r3[an], H.phi[an], f.x, f.y);
    D.c[o] = make_float2(f.z, f.w);
    D.p[o] = make_float4(a1, tanF(a1), float(tan(double(a1)) / double(a1)),
                         asinF(fmin2(rta * k2Rinv() / 3.0f, kSinAlphaMax)));
    D.q[o] = mulsFactor(H.z[an], rta);
}

__device__ __forceinline__ MDView loadMD(int i, const MDDev& D) {
    const float4 a = D.a[i], b = D.b[i];
    const float2 c = D.c[i];
    const float4 p = D.p[i];
    MDView v;
    v.a = HitView{a.x, a.y, a.z, a.w, b.x, b.y, p.x};
    v.dz = b.z;
    v.shiftedDz = b.w;
    v.dphi = c.x;
    v.dphichange = c.y;
    v.alpha1 = p.x;
    v.tanAlpha1 = p.y;
    v.mod = 0;
    return v;
}
__device__ __forceinline__ HitView loadAnchor(int i, const MDDev& D) {
    const float4 a = ld32(D.a, i), b = ld32(D.b, i);
    return HitView{a.x, a.y, a.z, a.w, b.x, b.y, 0.f};
}

// ------------------------------------------------------------------ SG
// SG candidates of one inner MD: MDs of every connected lower module (Event.cc:457-522).
// kSgCapLanes lanes per MD, one connection each per round.
constexpr int kSgCapLanes = 8;
// (also clears the MD's counts of the SGs ending at it, accumulated later)
__global__ void k_sg_cap(Count cMD, int NM, const int* mdSlot, const int* connStart, const int* connList,
                         const int* mdCount, int* sgCap, int* sgEndCount, int* sgEndFill) {
    const int nMD = int(cMD.get());
    const int gl = threadIdx.x & (kSgCapLanes - 1);
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kSgCapLanes;
    if (i >= nMD) return;  // uniform over the group
    const int s = mdSlot[i];
    const int m = s % NM, eb = s - m;
    int c = 0;
    for (int k = connStart[m] + gl; k < connStart[m + 1]; k += kSgCapLanes) c += mdCount[eb + connList[k]];
#pragma unroll
    for (int off = 1; off < kSgCapLanes; off <<= 1) c += __shfl_xor(c, off, kSgCapLanes);
    if (gl == 0) {
        sgCap[i] = c;
        sgEndCount[i] = 0;
        sgEndFill[i] = 0;
    }
}

struct SGF {
    float v[6];  // dphi, dphimin, dphimax, dphichange, dphichangemin, dphichangemax
};

// Layer bin of a module: barrel layers 0..5, endcap layers 6..10 (the layers
// of Event.h's per-layer counters and of SDL::Layer).
__host__ __device__ __forceinline__ int layerBin(const ModDev& m) {
    return m.subdet == SUB_BARREL ? m.layer - 1 : 6 + m.layer - 1;
}

constexpr int kNCounters = 10;
constexpr int kNBins = 11;

// Flattened walk over the connected modules of one work item, W lanes per
// item: returns the connection entry holding candidate k and the index inside
// it.  pre holds the group's inclusive prefix of per-entry counts (entry j in
// lane j of the group, 1 <= nEntries <= W), so the entries with pre <= k are a
// prefix [0, e): a binary search of log2(W) shuffles finds e.  Must be called
// by all W lanes of the group (the shuffles read every lane of it).
template <int W>
__device__ __forceinline__ int findEntry(int k, int pre, int nEntries, int& inner) {
    int e = 0;
#pragma unroll
    for (int step = W / 2; step >= 1; step >>= 1) {
        const int j = e + step - 1;
        const int pj = __shfl(pre, min(j, nEntries - 1), W);
        if (j < nEntries && pj <= k) e += step;
    }
    const int prev = __shfl(pre, max(e - 1, 0), W);
    inner = k - (e > 0 ? prev : 0);
    return e;
}

// One SG candidate (inner MD `in` on module mi, outer MD mdOut on module mO):
// the segment algorithm with the endcap-2S strip edges of the outer anchor.
__device__ __forceinline__ bool sgCandidate(const MDView& in, const ModDev& mi, int mdOut, const ModDev& mO,
                                            bool connFound, const HitsDev& H, const MDDev& D, SGOut& o,
                                            float* rec = nullptr) {
    const MDView out = loadMD(mdOut, D);
    HitView hiE = out.a, loE = out.a;
    if (mO.subdet == SUB_ENDCAP && mO.type == TYPE_2S) {
        const int an = D.anchor[mdOut];
        const float4 a = H.eHi[an], b = H.eLo[an];
        hiE.phi = a.w;
        loE.phi = b.w;
    }
    return sgAlgo(in, out, mi, mO, connFound, hiE, loE, o, rec);
}

// Appends the passing SG candidates of a group of W lanes (the wave, or one
// half of it), in lane order, to the inner MD's slab; gl = lane in the group.
template <int W = kWave>
__device__ __forceinline__ void appendSG(bool pass, int lane, int& base, long long off, int mdOut, const SGOut& o,
                                         int* slabOut, SGF* slabF) {
    const int gl = lane & (W - 1);
    unsigned long long mask = __ballot(pass) >> (lane - gl);
    if constexpr (W < kWave) mask &= (1ull << W) - 1ull;
    if (pass) {
        const int pos = base + __popcll(mask & ((1ull << gl) - 1ull));
        slabOut[off + pos] = mdOut;
        SGF f;
        f.v[0] = o.dphi; f.v[1] = o.dphimin; f.v[2] = o.dphimax;
        f.v[3] = o.dphichange; f.v[4] = o.dphichangemin; f.v[5] = o.dphichangemax;
        slabF[off + pos] = f;
    }
    base += __popcll(mask);
}

// One group of kSgLanes lanes per inner MD (two MDs per wave: half of the MDs
// of a PU200 event have fewer than 30 candidates, so a wave per MD left about
// 40 % of its lanes idle); the lanes walk the flattened (connected module,
// outer MD) space.
constexpr int kSgLanes = 32;
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) k_sg(Count cMD, int NM, const ModDev* mods, HitsDev H, MDDev D,
                                               const int* connStart, const int* connList, const long long* mdStart,
                                               const int* mdCount, const int* sgCap, const long long* sgCapOff,
                                               int* slabOut, SGF* slabF, int* sgCount, int allComb, Guard g) {
    if (g.over()) return;
    const int nMD = int(cMD.get());
    const int lane = threadIdx.x & (kWave - 1);
    const int gl = lane & (kSgLanes - 1);
    const int item = int(((long long)blockIdx.x * blockDim.x + threadIdx.x) / kSgLanes);
    if (item >= nMD) return;  // uniform over the group
    const int cap = sgCap[item];
    if (cap == 0) {
        if (gl == 0) sgCount[item] = 0;
        return;
    }
    const int sIn = D.slot[item];
    const int mIn = sIn % NM, eb = sIn - mIn;
    const ModDev mi = mods[mIn];
    const MDView in = loadMD(item, D);
    const int c0 = connStart[mIn], nc = connStart[mIn + 1] - c0;
    const long long off = sgCapOff[item];
    int base = 0;
    // connection entries are processed in chunks of kSgLanes (one per lane)
    for (int cb = 0; cb < nc; cb += kSgLanes) {
        const int ne = min(kSgLanes, nc - cb);
        int cnt = 0;
        if (gl < ne) cnt = mdCount[eb + connList[c0 + cb + gl]];
        // inclusive prefix over the group's lanes
        int pre = cnt;
        for (int o = 1; o < kSgLanes; o <<= 1) {
            const int t = __shfl_up(pre, o, kSgLanes);
            if (gl >= o) pre += t;
        }
        const int total = __shfl(pre, kSgLanes - 1, kSgLanes);
        for (int k0 = 0; k0 < total; k0 += kSgLanes) {
            const int k = k0 + gl;
            bool pass = false;
            SGOut o{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // AllComb leaves the constructor's values (Segment.cc:35-56)
            int mdOut = 0;
            int inner;
            const int e = findEntry<kSgLanes>(k, pre, ne, inner);
            if (k < total) {
                const int modOut = connList[c0 + cb + e];
                mdOut = int(mdStart[eb + modOut]) + inner;
                pass = allComb || sgCandidate(in, mi, mdOut, mods[modOut], true, H, D, o);
            }
            appendSG<kSgLanes>(pass, lane, base, off, mdOut, o, slabOut, slabF);
        }
    }
    if (gl == 0) sgCount[item] = base;
}

// ------------------------------------------------------- layer-pair form
// createSegments / createTracklets / createTrackCandidates (Event.cc:432-443,
// 599-609, 755-872): every object of an inner layer against every object of an
// outer layer, for fixed lists of layer pairs (Layer.cc:3-52).  Layer bins:
// barrel layers 1-6 = 0-5, endcap layers 1-5 = 6-10.
//
// Objects of one layer of one event: the geometry lists the modules by layer
// bin (layerMods, with layerModStart[bin] .. [bin + 1]); a per-slot object array
// (start) then gives each module's range, and a per-event list of the objects
// of every layer is filled in that order (layerOff[e * NM + position] = offset).

// createSegments' layer pairs (inner bin, outer bin), Layer.cc:35-52.
constexpr int kNSgLayerPairs = 16;
__constant__ int kSgLayerPairs[kNSgLayerPairs][2] = {
    {0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5},  // barrel 1-2 ... 5-6
    {0, 6}, {1, 6}, {2, 6}, {3, 6}, {4, 6},  // barrel 1 ... 5 - endcap 1
    {6, 7}, {7, 8}, {7, 9}, {7, 10}, {8, 9}, {9, 10}};  // endcap 1-2, 2-3, 2-4, 2-5, 3-4, 4-5

__global__ void k_layer_count(int E, int NM, const int* layerMods, const long long* start, int* cnt) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)E * NM) return;
    const long long e = i / NM;
    const int m = layerMods[i % NM];
    cnt[i] = int(start[e * NM + m + 1] - start[e * NM + m]);
}

__global__ void k_layer_fill(int E, int NM, const int* layerMods, const long long* start, const long long* off,
                             int* list) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)E * NM) return;
    const long long e = i / NM;
    const int m = layerMods[i % NM];
    const long long a = start[e * NM + m], n = start[e * NM + m + 1] - a, d = off[i];
    for (long long k = 0; k < n; ++k) list[d + k] = int(a + k);
}

// [a, b): the objects of layer bin L of event e in a list filled by k_layer_fill.
__device__ __forceinline__ void layerRange(const long long* off, const int* layerModStart, int NM, int e, int L,
                                           long long& a, long long& b) {
    a = off[(long long)e * NM + layerModStart[L]];
    b = off[(long long)e * NM + layerModStart[L + 1]];
}

// Per inner MD: SG candidates of the layer pairs that start at its layer.
__global__ void k_sgl_cap(Count cMD, int NM, const ModDev* mods, const int* mdSlot, const long long* layerOff,
                          const int* layerModStart, int* sgCap, int* sgEndCount, int* sgEndFill) {
    const int nMD = int(cMD.get());
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nMD) return;
    const int slot = mdSlot[j];
    const int e = slot / NM, L = layerBin(mods[slot % NM]);
    long long c = 0;
    for (int p = 0; p < kNSgLayerPairs; ++p)
        if (kSgLayerPairs[p][0] == L) {
            long long a, b;
            layerRange(layerOff, layerModStart, NM, e, kSgLayerPairs[p][1], a, b);
            c += b - a;
        }
    sgCap[j] = int(c);
    sgEndCount[j] = 0;  // counts of the SGs ending at the MD, accumulated later
    sgEndFill[j] = 0;
}

// One wave per inner MD: the SG algorithm against every MD of each outer layer
// (Event.cc:793-821), survivors appended to the MD's slab as in k_sg.  The
// endcap algorithm's connection test (Segment.cc:627-634) is evaluated here,
// since layer pairs need not be connected modules.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_sgl(Count cMD, int NM, const ModDev* mods, HitsDev H, MDDev D, const int* connStart, const int* connList,
      const long long* layerOff, const int* layerModStart, const int* layerMd, const int* sgCap,
      const long long* sgCapOff, int* slabOut, SGF* slabF, int* sgCount, int allComb, Guard g) {
    if (g.over()) return;
    const int nMD = int(cMD.get());
    const int item = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (item >= nMD) return;
    if (sgCap[item] == 0) {
        if (lane == 0) sgCount[item] = 0;
        return;
    }
    const int sIn = D.slot[item];
    const int mIn = sIn % NM, e = sIn / NM;
    const ModDev mi = mods[mIn];
    const int L = layerBin(mi);
    const MDView in = loadMD(item, D);
    const int c0 = connStart[mIn], c1 = connStart[mIn + 1];
    const long long off = sgCapOff[item];
    int base = 0;
    for (int p = 0; p < kNSgLayerPairs; ++p) {
        if (kSgLayerPairs[p][0] != L) continue;
        long long a, b;
        layerRange(layerOff, layerModStart, NM, e, kSgLayerPairs[p][1], a, b);
        for (long long k0 = a; k0 < b; k0 += kWave) {
            const long long k = k0 + lane;
            bool pass = false;
            SGOut o{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // AllComb leaves the constructor's values
            int mdOut = 0;
            if (k < b) {
                mdOut = layerMd[k];
                const int modOut = D.slot[mdOut] % NM;
                const ModDev mO = mods[modOut];
                bool conn = false;
                if (mO.subdet != SUB_BARREL && !allComb)
                    for (int q = c0; q < c1; ++q) conn = conn || connList[q] == modOut;
                pass = allComb || sgCandidate(in, mi, mdOut, mO, conn, H, D, o);
            }
            appendSG(pass, lane, base, off, mdOut, o, slabOut, slabF);
        }
    }
    if (lane == 0) sgCount[item] = base;
}

struct SGDev {
    int *in, *out;
    SGF* f;
    float4* x;    // rt(out - in), out.rt - in.rt, deltaPhi(out, out - in), deltaPhi(in, out)   (anchors)
    float2* e;    // deltaPhi(edge, edge - in) for the high / low strip edges of an endcap-2S outer anchor
    int* flags;   // 1 inner barrel, 2 outer barrel, 4 inner PS, 8 outer PS, 16 outer endcap 2S
    int* key;     // counter bin of the inner MD's lower module: event * kNBins + layer bin
};

// Per-segment quantities that the tracklet algebra reads for either role; also
// the SGs ending at each MD (sgEndCount, zeroed before) and the triplets of each
// inner SG (tpCount: the SGs that start at its outer MD, Triplet.cc:50-122).
__global__ void k_sg_extra(Count cSG, int NM, const ModDev* mods, HitsDev H, MDDev D, SGDev S,
                           const long long* sgStart, int* sgEndCount, int* tpCount) {
    const int nSG = int(cSG.get());
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nSG) return;
    const int mi = S.in[s], mo = S.out[s];
    const HitView a = loadAnchor(mi, D), b = loadAnchor(mo, D);
    const ModDev Mi = mods[D.slot[mi] % NM], Mo = mods[D.slot[mo] % NM];
    const float dx = b.x - a.x, dy = b.y - a.y;
    S.x[s] = make_float4(rtOf(dx, dy), b.rt - a.rt, phiMpiPi(phiOf(dx, dy) - b.phi), dPhi(a.phi, b.phi));
    const bool ec2s = Mo.subdet == SUB_ENDCAP && Mo.type == TYPE_2S;
    S.flags[s] = (Mi.subdet == SUB_BARREL ? 1 : 0) | (Mo.subdet == SUB_BARREL ? 2 : 0) | (Mi.type == TYPE_PS ? 4 : 0) |
                 (Mo.type == TYPE_PS ? 8 : 0) | (ec2s ? 16 : 0);
    S.key[s] = (D.slot[mi] / NM) * kNBins + layerBin(Mi);
    atomicAdd(&sgEndCount[mo], 1);
    tpCount[s] = int(sgStart[mo + 1] - sgStart[mo]);
    float2 e = make_float2(0.f, 0.f);
    if (ec2s) {
        const int an = D.anchor[mo];
        const float4 hi = H.eHi[an], lo = H.eLo[an];
        e.x = phiMpiPi(phiOf(hi.x - a.x, hi.y - a.y) - hi.w);
        e.y = phiMpiPi(phiOf(lo.x - a.x, lo.y - a.y) - lo.w);
    }
    S.e[s] = e;
}

// kRepackLanes lanes per inner MD (a PU200 MD stores about four SGs).
constexpr int kRepackLanes = 8;
__global__ void k_sg_repack(Count cMD, const long long* sgCapOff, const int* sgCount, const long long* sgStart,
                            const int* slabOut, const SGF* slabF, SGDev S, Guard g) {
    if (g.over()) return;
    const int nMD = int(cMD.get());
    const int item = int(((long long)blockIdx.x * blockDim.x + threadIdx.x) / kRepackLanes);
    const int gl = threadIdx.x & (kRepackLanes - 1);
    if (item >= nMD) return;
    const int n = sgCount[item];
    const long long src = sgCapOff[item], dst = sgStart[item];
    for (int k = gl; k < n; k += kRepackLanes) {
        S.in[dst + k] = item;
        S.out[dst + k] = slabOut[src + k];
        S.f[dst + k] = slabF[src + k];
    }
}

// Per-slot kernels over the MDs of a slot: kSlotLanes lanes per slot, one MD
// each per round; a ballot over the group keeps MD order.
constexpr int kSlotLanes = 8;
__device__ __forceinline__ unsigned long long slotGroupMask(bool p) {
    const int lane = threadIdx.x & (kWave - 1), gl = threadIdx.x & (kSlotLanes - 1);
    return (__ballot(p) >> (lane - gl)) & ((1ull << kSlotLanes) - 1ull);
}

// Per-slot ranges of objects ordered by an inner MD: first = start[mdStart[slot]].
// Per slot: where its SGs start, and how many of its MDs start at least one SG
// (the tracklet pairs' outer MDs).
__global__ void k_slot_range(Count cNS, const long long* mdStart, const long long* sgStart, long long* slotSgStart,
                             int* slotMdSg) {
    const int NS = int(cNS.get());
    const int gl = threadIdx.x & (kSlotLanes - 1);
    const long long s = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kSlotLanes;
    if (s > NS) return;  // uniform over the group
    if (gl == 0) slotSgStart[s] = sgStart[mdStart[s]];
    if (s == NS) return;
    int c = 0;
    const long long m0 = mdStart[s], m1 = mdStart[s + 1];
    for (long long mb = m0; mb < m1; mb += kSlotLanes) {
        const long long m = mb + gl;
        c += __popcll(slotGroupMask(m < m1 && sgStart[m + 1] > sgStart[m]));
    }
    if (gl == 0) slotMdSg[s] = c;
}

// SGs ending at each MD (reverse index, used by the joins)
__global__ void k_scatter_by(Count cn, const int* key, const long long* start, int* fill, int* list) {
    const int n = int(cn.get());
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = key[i];
    const int pos = atomicAdd(&fill[k], 1);
    list[start[k] + pos] = i;
}

// ------------------------------------------------------------------ TP
// Triplet = two SGs sharing the middle MD (Triplet.cc:50-122, Segment.cc:861-866).
// The inner SG s ends at MD m; the outer SGs stored at m's module with inner MD m
// are exactly the dense range of SGs whose inner MD is m.
// tpKey: the counter key of the TP's inner SG (its layer bin), which the TC and
// counter kernels would otherwise reach through tpIn.
// kTpLanes lanes per inner SG write its triplets.
constexpr int kTpLanes = 4;
__global__ void k_tp_write(Count cSG, const int* sgOut, const long long* sgStart, const long long* tpStart,
                           const int* sgKey, int* tpIn, int* tpOutArr, int* tpKey, Guard g) {
    if (g.over()) return;
    const int nSG = int(cSG.get());
    const int gl = threadIdx.x & (kTpLanes - 1);
    const long long sl = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kTpLanes;
    if (sl >= nSG) return;
    const int s = int(sl);
    const int m = sgOut[s], key = sgKey[s];
    const long long a = sgStart[m], n = sgStart[m + 1] - a, d = tpStart[s];
    for (long long k = gl; k < n; k += kTpLanes) {
        tpIn[d + k] = s;
        tpOutArr[d + k] = int(a + k);
        tpKey[d + k] = key;
    }
}

// ------------------------------------------------------------------ AllComb joins
// The AllComb algorithms accept every candidate (Triplet.cc:45-48,
// Tracklet.cc:139-142, TrackCandidate.cc:138-141).  For these three stages the
// candidates of inner object i are the objects stored at the modules connected
// to the module of a key MD of i, in connection-list order and, inside a
// module, in storage order (the reference's loops):
//   triplets (AC_TP):  inner SG s, the SGs of the modules connected to the
//                      module of s's inner MD                  Event.cc:539-597
//   tracklets (AC_TL): inner SG s, the SGs of the modules connected to the
//                      module of s's outer MD                  Event.cc:627-693
//   TCs from tracklets (AC_TCL): inner TL l, the TLs of the modules connected to
//                      the module of its inner SG's outer MD   Event.cc:999-1062
// objStart holds the per-slot object ranges (SGs or TLs).
enum { AC_TP = 0, AC_TL = 1, AC_TCL = 2 };

struct AcIn {
    int kind, NM;
    const int *sgIn, *sgOut, *tlIn, *mdSlot, *connStart, *connList;
    const long long* objStart;
    __device__ __forceinline__ int keySlot(int i) const {
        const int md = kind == AC_TP ? sgIn[i] : (kind == AC_TL ? sgOut[i] : sgOut[tlIn[i]]);
        return mdSlot[md];
    }
};

// Candidates (= objects) per inner object; cap (optional) the same as 64-bit.
__global__ void k_ac_count(Count cIn, AcIn A, int* cnt, long long* cap) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cIn.get()) return;
    const int b = A.keySlot(int(i));
    const int m = b % A.NM, eb = b - m;
    long long c = 0;
    for (int k = A.connStart[m]; k < A.connStart[m + 1]; ++k) {
        const int t = eb + A.connList[k];
        c += A.objStart[t + 1] - A.objStart[t];
    }
    cnt[i] = int(c);
    if (cap) cap[i] = c;
}

// The objects (i, o) of inner object i at start[i]: into a and b (triplets,
// tracklets; key[i] also into bKey when given) or as pairs (TCs).
__global__ void k_ac_write(Count cIn, AcIn A, const long long* start, int* outA, int* outB, const int* key,
                           int* outKey, int2* outPair, Guard g) {
    if (g.over()) return;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cIn.get()) return;
    const int b = A.keySlot(int(i));
    const int m = b % A.NM, eb = b - m;
    long long d = start[i];
    for (int k = A.connStart[m]; k < A.connStart[m + 1]; ++k) {
        const int t = eb + A.connList[k];
        for (long long o = A.objStart[t]; o < A.objStart[t + 1]; ++o, ++d) {
            if (outPair) {
                outPair[d] = make_int2(int(i), int(o));
            } else {
                outA[d] = int(i);
                outB[d] = int(o);
                if (outKey) outKey[d] = key[i];
            }
        }
    }
}

// Layer-pair TCs with AllComb (Event.cc:840-872): every TL of barrel layer 1
// against every TL of barrel layer 3 of its event; layerTl / layerTlOff are the
// per-event layer lists of the tracklets.
__global__ void k_tcll_ac_write(Count cTL, int NM, const ModDev* mods, const int* tlIn, const int* sgIn,
                                const int* mdSlot, const long long* layerTlOff, const int* layerModStart,
                                const int* layerTl, const long long* start, int2* tc, Guard g) {
    if (g.over()) return;
    const int nTL = int(cTL.get());
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nTL) return;
    const int slot = mdSlot[sgIn[tlIn[l]]];
    if (layerBin(mods[slot % NM]) != 0) return;
    const int e = slot / NM;
    const long long a = layerTlOff[(long long)e * NM + layerModStart[2]];
    const long long b = layerTlOff[(long long)e * NM + layerModStart[3]];
    long long d = start[l];
    for (long long k = a; k < b; ++k) tc[d++] = make_int2(l, layerTl[k]);
}

// ------------------------------------------------------------------ TL
// TL candidates of one inner SG: SGs stored at every module connected to the
// inner SG's outer-MD module (Event.cc:627-693).  The tracklet algorithm runs
// as a cascade of dense passes:
//  1. pairs: (inner SG, MD of a connected module that starts an SG), listed per
//     algorithm variant (BBBB, BBEE, EEEE; uniform over a connected module);
//  2. k_tl_stage1<V>, one thread per pair: the first-stage cuts, which involve
//     only the inner SG and the outer SG's inner MD (Tracklet.cc:902-1153,
//     1540-1771, 2018-2270), once per pair instead of once per candidate;
//  3. k_tl_expand<V>, one thread per surviving pair: its outer SGs, densely;
//  4. k_tl_stage2<V>, one thread per candidate: the remaining cuts;
//  5. survivors grouped by inner SG, sorted by outer SG inside each group.
__device__ __forceinline__ int tlModuleVariant(int innerFlags, const ModDev& outLoModule) {
    return tlVariant(innerFlags & 1, innerFlags & 2, outLoModule.subdet == SUB_BARREL);
}

// Per inner SG: TL candidates (for the counters) and pairs per variant.
// kCapLanes lanes per SG, one connection each per round, summed by shuffles.
constexpr int kCapLanes = 4;
__global__ void k_tl_cap(Count cSG, int NM, const ModDev* mods, const int* sgOut, const int* sgFlags,
                         const int* mdSlot, const int* connStart, const int* connList, const long long* slotSgStart,
                         const int* slotMdSg, long long* tlCap, int* nPair, long long stride, int* tlCount,
                         int* tlFill) {
    const int nSG = int(cSG.get());
    const int gl = threadIdx.x & (kCapLanes - 1);
    const long long s = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kCapLanes;
    if (s >= stride) return;  // uniform over the group
    if (s >= nSG) {  // the three per-variant rows are scanned as one array
        if (gl < kNVariantsTL) nPair[gl * stride + s] = 0;
        return;
    }
    const int b = mdSlot[sgOut[s]];
    const int m = b % NM, eb = b - m;
    const int fl = sgFlags[s];
    long long c = 0;
    int np[kNVariantsTL] = {0, 0, 0};
    for (int k = connStart[m] + gl; k < connStart[m + 1]; k += kCapLanes) {
        const int t = eb + connList[k];
        c += slotSgStart[t + 1] - slotSgStart[t];
        const int v = tlModuleVariant(fl, mods[connList[k]]);
        if (v < kNVariantsTL) np[v] += slotMdSg[t];
    }
#pragma unroll
    for (int off = 1; off < kCapLanes; off <<= 1) {
        c += __shfl_xor(c, off, kCapLanes);
#pragma unroll
        for (int v = 0; v < kNVariantsTL; ++v) np[v] += __shfl_xor(np[v], off, kCapLanes);
    }
    if (gl == 0) {
        tlCap[s] = c;
        for (int v = 0; v < kNVariantsTL; ++v) nPair[v * stride + s] = np[v];
        tlCount[s] = 0;  // accumulated by the second stage
        tlFill[s] = 0;   // and by the scatter
    }
}

// The items of one variant in a list ordered by variant: [b, e) with
// b = off1[v * stride] (then off2[b] when off2 is set), e likewise for v + 1.
// A launch sized for `cap` items aborts the run if the variant has more.
struct VarRange {
    const long long* off1;
    const long long* off2;
    long long stride;
    int v;
    long long cap;
    long long* tot;
    __device__ __forceinline__ bool get(long long& b, long long& e) const {
        b = e = 0;
        if (tot[TOT_ABORT]) return false;
        long long p0 = off1[v * stride], p1 = off1[(v + 1) * stride];
        if (off2) {
            p0 = off2[p0];
            p1 = off2[p1];
        }
        if (p1 - p0 > cap) {
            if (threadIdx.x == 0) tot[TOT_ABORT] = 1 + SDL_STAGE_TL;
            return false;
        }
        b = p0;
        e = p1;
        return true;
    }
};

// The MDs of each slot that start at least one SG, in MD order.
__global__ void k_slot_md_sg_list(Count cNS, const long long* mdStart, const long long* sgStart,
                                  const long long* listStart, int* list) {
    const int NS = int(cNS.get());
    const int gl = threadIdx.x & (kSlotLanes - 1);
    const long long t = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kSlotLanes;
    if (t >= NS) return;  // uniform over the group
    long long d = listStart[t];
    const long long m0 = mdStart[t], m1 = mdStart[t + 1];
    for (long long mb = m0; mb < m1; mb += kSlotLanes) {  // the same trips for the whole group
        const long long m = mb + gl;
        const bool has = m < m1 && sgStart[m + 1] > sgStart[m];
        const unsigned long long g = slotGroupMask(has);
        if (has) list[d + __popcll(g & ((1ull << gl) - 1ull))] = int(m);
        d += __popcll(g);
    }
}

// A group of kEnumLanes lanes per inner SG writes its pairs, in connected-module
// and MD order, into one list ordered by variant (offsets pairOff[v * stride + s]).
// Lane k of the group takes connected module k (in rounds of kEnumLanes); one
// scan of the per-variant counts packed into 21-bit fields places every
// module's run, and the lanes then write the round's pairs contiguously.
constexpr int kEnumLanes = 8;
__global__ void __launch_bounds__(kBlock) k_tl_enum(Count cSG, int NM, const ModDev* mods, const int* sgOut,
                                                    const int* sgFlags, const int* mdSlot, const int* connStart,
                                                    const int* connList, const long long* listStart, const int* list,
                                                    const long long* pairOff, long long stride, int2* pair, Guard g) {
    if (g.over()) return;
    const int nSG = int(cSG.get());
    const int gl = threadIdx.x & (kEnumLanes - 1);
    const long long s = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kEnumLanes;
    if (s >= nSG) return;  // uniform over the group
    const int b = mdSlot[sgOut[s]];
    const int m = b % NM, eb = b - m;
    const int fl = sgFlags[s];
    long long pos[kNVariantsTL];
    for (int v = 0; v < kNVariantsTL; ++v) pos[v] = pairOff[v * stride + s];
    const int k0 = connStart[m], k1 = connStart[m + 1];
    for (int kc = k0; kc < k1; kc += kEnumLanes) {
        const int k = kc + gl;
        int v = kNVariantsTL, cnt = 0;
        long long ls = 0;
        if (k < k1) {
            v = tlModuleVariant(fl, mods[connList[k]]);
            if (v < kNVariantsTL) {
                const int t = eb + connList[k];
                ls = listStart[t];
                cnt = int(listStart[t + 1] - ls);
            }
        }
        // inclusive scan of the packed per-variant counts over the group
        const unsigned long long mine = v < kNVariantsTL ? (unsigned long long)cnt << (21 * v) : 0ULL;
        unsigned long long inc = mine;
#pragma unroll
        for (int off = 1; off < kEnumLanes; off <<= 1) {
            const unsigned long long y = __shfl_up(inc, off, kEnumLanes);
            if (gl >= off) inc += y;
        }
        const unsigned long long exc = inc - mine;
        const unsigned long long tot = __shfl(inc, kEnumLanes - 1, kEnumLanes);
        auto field = [](unsigned long long w, int f) { return (long long)((w >> (21 * f)) & 0x1FFFFFULL); };
        // this lane's run: destination and flat position among the round's pairs
        const long long dst = v < kNVariantsTL ? pos[v] + field(exc, v) : 0;
        const int flat = int(field(exc, 0) + field(exc, 1) + field(exc, 2));
        const int total = int(field(tot, 0) + field(tot, 1) + field(tot, 2));
        // every lane of the group takes part in each step (the shuffles read
        // all of the group's lanes)
        for (int j0 = 0; j0 < total; j0 += kEnumLanes) {
            const int j = j0 + gl;
            // the last lane whose run starts at or before j owns pair j
            int lo = 0;
#pragma unroll
            for (int step = kEnumLanes / 2; step >= 1; step >>= 1) {
                const int f = __shfl(flat, lo + step, kEnumLanes);
                if (f <= j) lo += step;
            }
            const int q = j - __shfl(flat, lo, kEnumLanes);
            const long long d = __shfl(dst, lo, kEnumLanes);
            const long long l0 = __shfl(ls, lo, kEnumLanes);
            if (j < total) pair[d + q] = make_int2(int(s), list[l0 + q]);
        }
        for (int vv = 0; vv < kNVariantsTL; ++vv) pos[vv] += field(tot, vv);
    }
}

__device__ __forceinline__ TLOutLo loadOutLo(int mLo, int flags, const MDDev& D) {
    TLOutLo L;
    L.a = loadAnchor(mLo, D);
    const float4 p = ld32(D.p, mLo);
    L.alpha1 = p.x;
    L.dzScale = p.z;
    L.alphaReg3 = p.w;
    L.barrel = flags & 1;
    L.ps = flags & 4;
    return L;
}

__device__ __forceinline__ TLInner loadInner(int s, const MDDev& D, const SGDev& S) {
    TLInner I;
    const int mi = ld32(S.in, s), mo = ld32(S.out, s);
    I.lo = loadAnchor(mi, D);
    I.up = loadAnchor(mo, D);
    const SGF fi = ld32(S.f, s);
    I.dphichange = fi.v[3];
    I.dphichangemin = fi.v[4];
    I.dphichangemax = fi.v[5];
    const float4 xi = ld32(S.x, s);
    I.dr = xi.x;
    I.d = xi.y;
    I.alphaReg3Lo = ld32(D.p, mi).w;
    I.sqrtR3Rt = ld32(D.q, mi);
    const int fl = ld32(S.flags, s);
    I.bLo = fl & 1; I.bUp = fl & 2; I.psLo = fl & 4; I.psUp = fl & 8;
    return I;
}

__device__ __forceinline__ TLOutUp loadOutUp(int so, int flo, HitsDev H, const MDDev& D, const SGDev& S) {
    TLOutUp U;
    const int mUp = ld32(S.out, so);
    U.a = loadAnchor(mUp, D);
    const SGF sf = ld32(S.f, so);
    U.dphimin = sf.v[1]; U.dphimax = sf.v[2];
    U.dphichange = sf.v[3]; U.dphichangemin = sf.v[4]; U.dphichangemax = sf.v[5];
    const float4 sx = ld32(S.x, so);
    U.dr = sx.x; U.d = sx.y; U.aOU = sx.z; U.dPhiPos = sx.w;
    U.ec2s = flo & 16;
    U.hiX = U.hiY = U.hiRt = U.hiPhi = U.loX = U.loY = U.loRt = U.loPhi = U.aOUhi = U.aOUlo = 0.f;
    if (U.ec2s) {
        const int an = ld32(D.anchor, mUp);
        const float4 hi = ld32(H.eHi, an), lo = ld32(H.eLo, an);
        const float2 ee = ld32(S.e, so);
        U.hiX = hi.x; U.hiY = hi.y; U.hiRt = hi.z; U.hiPhi = hi.w;
        U.loX = lo.x; U.loY = lo.y; U.loRt = lo.z; U.loPhi = lo.w;
        U.aOUhi = ee.x; U.aOUlo = ee.y;
    }
    return U;
}

// The three variants of a tracklet stage in one launch: blocks [0, nb[0]) run
// variant 0, the next nb[1] blocks variant 1, the rest variant 2 (nb[v] blocks
// cover variant v's capacity), so the variants' tails and heads overlap instead
// of the chip draining between three launches.
struct VarRanges {
    VarRange r[kNVariantsTL];
    int nb[kNVariantsTL];
};
__device__ __forceinline__ int variantOfBlock(const VarRanges& R, int& blk) {
    blk = blockIdx.x;
    for (int v = 0; v < kNVariantsTL - 1; ++v) {
        if (blk < R.nb[v]) return v;
        blk -= R.nb[v];
    }
    return kNVariantsTL - 1;
}

// First stage, one thread per pair: carry = (sdlCut, thetaMuls, sin of the
// midpoint deltaPhi); nOut = number of outer SGs to try (0 if the pair failed).
template <int V>
__device__ __forceinline__ void tlStage1Item(const VarRange& R, int blk, MDDev D, SGDev S, const long long* sgStart,
                                             const int2* pair, float4* carry, int* nOut) {
    long long b0, e0;
    R.get(b0, e0);
    const long long i = b0 + (long long)blk * blockDim.x + threadIdx.x;
    if (i >= e0) return;
    const int2 pr = pair[i];
    const int sgA = int(ld32(sgStart, pr.y));
    const TLInner I = loadInner(pr.x, D, S);
    const TLOutLo L = loadOutLo(pr.y, ld32(S.flags, sgA), D);
    TLCarry c{0.f, 0.f, 0.f};
    const bool ok = tlStage1(V, I, L, c);
    if (ok) carry[i] = make_float4(c.sdlCut, c.thetaMuls, c.sinDPhiMid, 0.f);  // read for surviving pairs only
    nOut[i] = ok ? int(ld32(sgStart, pr.y + 1)) - sgA : 0;
}
__global__ void __launch_bounds__(kBlock) k_tl_stage1(VarRanges R, MDDev D, SGDev S, const long long* sgStart,
                                                      const int2* pair, float4* carry, int* nOut) {
    int blk;
    const int v = variantOfBlock(R, blk);
    if (v == TL_BBBB) tlStage1Item<TL_BBBB>(R.r[v], blk, D, S, sgStart, pair, carry, nOut);
    else if (v == TL_BBEE) tlStage1Item<TL_BBEE>(R.r[v], blk, D, S, sgStart, pair, carry, nOut);
    else tlStage1Item<TL_EEEE>(R.r[v], blk, D, S, sgStart, pair, carry, nOut);
}

// ---- learned proposal (pairs that passed the first stage are scored)
// Features of a pair that passed the first stage: the stored records of the
// inner SG and the outer SG's inner MD, and where the pair lies in the
// first-stage windows.
__device__ __forceinline__ void tlPairFeatures(const TLInner& I, const TLOutLo& L, const TLCarry& c,
                                               const TLMargins& m, int variant, float* f) {
    const float drIn = fmaxf(I.up.rt - I.lo.rt, 1e-3f);
    const float drOut = fmaxf(L.a.rt - I.lo.rt, 1e-3f);
    const float pointIn = (I.up.z - I.lo.z) / drIn;
    const float pointOut = (L.a.z - I.lo.z) / drOut;
    f[0] = I.dphichange * 10.f;
    f[1] = (I.dphichangemax - I.dphichangemin) * 10.f;
    f[2] = I.dr * 0.05f;
    f[3] = I.d * 0.05f;
    f[4] = I.lo.rt * 0.01f;
    f[5] = I.lo.z * 0.01f;
    f[6] = I.up.rt * 0.01f;
    f[7] = I.up.z * 0.01f;
    f[8] = L.a.rt * 0.01f;
    f[9] = L.a.z * 0.01f;
    f[10] = L.alpha1;
    f[11] = I.alphaReg3Lo;
    f[12] = L.alphaReg3;
    f[13] = phiMpiPi(L.a.phi - I.up.phi) * 10.f;
    f[14] = fminf(fmaxf(pointOut - pointIn, -5.f), 5.f);
    f[15] = m.w1 - 0.5f;
    f[16] = m.w2 - 0.5f;
    f[17] = m.phi;
    f[18] = c.sdlCut * 10.f;
    f[19] = I.psLo ? 1.f : 0.f;
    f[20] = I.psUp ? 1.f : 0.f;
    f[21] = L.ps ? 1.f : 0.f;
    f[22] = variant == TL_BBEE ? 1.f : 0.f;
    f[23] = variant == TL_EEEE ? 1.f : 0.f;
}

// First stage of pair i with the proposal's features (ok = passed the cuts).
template <int V>
__device__ __forceinline__ bool stage1Features(long long i, const int2* pair, const long long* sgStart, MDDev D,
                                               SGDev S, TLCarry& c, int& nOuter, float* f) {
    const int2 pr = pair[i];
    const int sgA = int(ld32(sgStart, pr.y));
    const TLInner I = loadInner(pr.x, D, S);
    const TLOutLo L = loadOutLo(pr.y, ld32(S.flags, sgA), D);
    TLMargins m{0.f, 0.f, 0.f};
    c = TLCarry{0.f, 0.f, 0.f};
    const bool ok = tlStage1(V, I, L, c, &m);
    nOuter = int(ld32(sgStart, pr.y + 1)) - sgA;
    if (ok) tlPairFeatures(I, L, c, m, V, f);
    return ok;
}

// First stage with the learned proposal: one wave per tile of 64 pairs,
// kScoreWaves waves per block (the block loads the network's LDS operands once).
// Each lane runs the first-stage cuts for its pair; the pairs that pass write
// their BF16 feature rows to the wave's LDS tile (ballot compaction) and are
// scored 16 rows at a time (one MFMA column block) with the BF16 network.  A
// pair keeps its outer SGs (nOut) only if its logit reaches the threshold;
// pairs that fail the cuts get nOut = 0.  scoredPerWave[wave] = the number of
// pairs the wave scored (a diagnostic; no atomics on a shared counter).  The
// waves per block are a launch parameter (dynamic LDS: kScoredLdsPerWave bytes
// per wave).
constexpr int kScoredLdsPerWave = 64 * 4 * 16 + 64 * 4 + 64 * 4;  // rows, pair indices, outer SG counts
template <int V>
__global__ void __launch_bounds__(1024)
k_tl_stage1_scored(VarRange R, MDDev D, SGDev S, const long long* sgStart, const int2* pairAll, const mlp::Packed* P,
                   float logitCut, float4* carryAll, int* nOutAll, int* scoredPerWave) {
    __shared__ __align__(16) mlp::Vecs Vs;
    extern __shared__ __align__(16) char dynLds[];
    mlp::loadVecs(P, Vs);  // includes a block barrier
    long long b0, e0;
    R.get(b0, e0);
    const long long n = e0 - b0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nWaves = blockDim.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    mlp::short8v* rows = reinterpret_cast<mlp::short8v*>(dynLds) + w * 64 * 4;
    int* qPair = reinterpret_cast<int*>(dynLds + nWaves * 64 * 4 * 16) + w * 64;
    int* qOut = reinterpret_cast<int*>(dynLds + nWaves * (64 * 4 * 16 + 64 * 4)) + w * 64;
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x);  // this lane's pair
    bool ok = false;
    int nOuter = 0;
    if (i < n) {
        TLCarry cr;
        float f[mlp::kIn];
        ok = stage1Features<V>(i, pairAll + b0, sgStart, D, S, cr, nOuter, f);
        if (ok) {
            carryAll[b0 + i] = make_float4(cr.sdlCut, cr.thetaMuls, cr.sinDPhiMid, 0.f);
            const unsigned long long m = __ballot(1);
            const int pos = __popcll(m & ((1ull << lane) - 1));
            mlp::short8v v[4];
            mlp::featureRow(f, v);
#pragma unroll
            for (int k = 0; k < 4; ++k) rows[pos * 4 + k] = v[k];
            qPair[pos] = int(i);
            qOut[pos] = nOuter;
        } else {
            nOutAll[b0 + i] = 0;
        }
    }
    const int q = __popcll(__ballot(ok));  // rows queued by this wave
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (lane == 0 && wave * 64 < n) scoredPerWave[wave] = q;
    if (q == 0) return;  // uniform over the wave (no block barrier follows)
    mlp::waveSync();
    mlp::Frags F;
    mlp::loadFrags(P, lane, F);
    for (int base = 0; base < q; base += 16) {  // rows past q hold stale data: they only affect their own logits
        const int r = base + c;
        const float z = mlp::logit16(__builtin_bit_cast(mlp::bf16x8, rows[r * 4 + g]), lane, F, Vs);
        if (g == 0 && r < q) nOutAll[b0 + qPair[r]] = z >= logitCut ? qOut[r] : 0;
    }
}

// Training rows of the plain chain: for every pair, its features and label
// (1: the pair yields at least one tracklet, 0: it passed the first stage but
// yields none, -1: it failed the first stage and is never scored).
template <int V>
__global__ void k_tl_pair_rows(VarRange R, MDDev D, SGDev S, const long long* sgStart, const int2* pair,
                               const long long* candOff, const unsigned char* flag, float* X, float* Y) {
    long long b0, e0;
    R.get(b0, e0);
    const long long i = b0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= e0) return;
    float f[mlp::kIn];
    TLCarry c;
    int nOuter;
    const bool ok = stage1Features<V>(i, pair, sgStart, D, S, c, nOuter, f);
    for (int k = 0; k < mlp::kIn; ++k) X[i * mlp::kIn + k] = ok ? f[k] : 0.f;
    bool any = false;
    for (long long cc = candOff[i]; cc < candOff[i + 1] && !any; ++cc) any = flag[cc] != 0;
    Y[i] = ok ? (any ? 1.f : 0.f) : -1.f;
}

// Candidates (pair, outer SG) of the surviving pairs, one thread per pair.
__global__ void k_tl_expand(Count cP, const long long* sgStart, const int2* pair, const int* nOut,
                            const long long* candOff, int2* cand, Guard g) {
    if (g.over()) return;
    const long long n = cP.get();
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = nOut[i];
    if (c == 0) return;
    const int a = int(sgStart[pair[i].y]);
    const long long d = candOff[i];
    for (int q = 0; q < c; ++q) cand[d + q] = make_int2(int(i), a + q);
}

// Second stage, one thread per candidate; survivors are counted per inner SG.
template <int V>
__device__ __forceinline__ void tlStage2Item(const VarRange& R, int blk, const HitsDev& H, const MDDev& D,
                                             const SGDev& S, const int2* cand, const int2* pair, const float4* carry,
                                             unsigned char* flag, int* tlCount, int nm1) {
    long long b0, e0;
    R.get(b0, e0);
    const long long i = b0 + (long long)blk * blockDim.x + threadIdx.x;
    if (i >= e0) return;
    const int2 c = cand[i];
    const int2 pr = pair[c.x];
    const float4 cf = carry[c.x];
    const int so = c.y;
    const int flo = ld32(S.flags, so);
    const TLInner I = loadInner(pr.x, D, S);
    const TLOutLo L = loadOutLo(pr.y, flo, D);
    const TLOutUp U = loadOutUp(so, flo, H, D, S);
    const TLCarry cc{cf.x, cf.y, cf.z};
    const bool pass = tlStage2(V, (flo & 2) != 0, I, L, U, cc, nullptr, nm1 != 0);
    flag[i] = pass ? 1 : 0;
    if (pass) atomicAdd(&tlCount[pr.x], 1);
}
__global__ void __launch_bounds__(kBlock) k_tl_stage2(VarRanges R, HitsDev H, MDDev D, SGDev S, const int2* cand,
                                                      const int2* pair, const float4* carry, unsigned char* flag,
                                                      int* tlCount, int nm1) {
    int blk;
    const int v = variantOfBlock(R, blk);
    if (v == TL_BBBB) tlStage2Item<TL_BBBB>(R.r[v], blk, H, D, S, cand, pair, carry, flag, tlCount, nm1);
    else if (v == TL_BBEE) tlStage2Item<TL_BBEE>(R.r[v], blk, H, D, S, cand, pair, carry, flag, tlCount, nm1);
    else tlStage2Item<TL_EEEE>(R.r[v], blk, H, D, S, cand, pair, carry, flag, tlCount, nm1);
}


// Survivors to their inner SG's range (the order inside a range is fixed
// afterwards by sorting on the outer SG); four candidates per thread, whose
// flags are read as one word (flag arrays are padded to a multiple of four).
__global__ void k_tl_scatter(Count cC, const int2* cand, const int2* pair, const unsigned char* flag,
                             const long long* tlStart, int* tlFill, int* tlIn, int* tlOutArr, Guard g) {
    if (g.over()) return;
    const long long n = cC.get();
    const long long i0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= n) return;
    const unsigned word = *reinterpret_cast<const unsigned*>(flag + i0);
    if (word == 0) return;
    for (int q = 0; q < 4; ++q) {
        const long long i = i0 + q;
        if (i >= n || !((word >> (8 * q)) & 0xffu)) continue;
        const int2 c = cand[i];
        const int s = pair[c.x].x;
        const long long d = tlStart[s] + atomicAdd(&tlFill[s], 1);
        tlIn[d] = s;
        tlOutArr[d] = c.y;
    }
}

// ------------------------------------------------------------------ TC
// TC from tracklets (Event.cc:999-1062, TrackCandidate.cc:143-226): the dR cut
// never rejects, so the pair passes iff the outer SG of the inner TL is the
// inner SG of the outer TL; those TLs are the dense range of TLs of that SG.
__global__ void k_tcl_count(Count cTL, const int* tlOutArr, const long long* tlStart, int* cnt) {
    const int nTL = int(cTL.get());
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nTL) return;
    const int s = tlOutArr[l];
    cnt[l] = int(tlStart[s + 1] - tlStart[s]);
}
// TCs of inner TL l: the cnt[l] TLs that start at l's outer SG.
__global__ void k_tcl_write(Count cTL, const int* tlOutArr, const long long* tlStart, const int* cnt,
                            const long long* start, int2* tc, Guard g) {
    if (g.over()) return;
    const int nTL = int(cTL.get());
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nTL) return;
    const long long a = tlStart[tlOutArr[l]], n = cnt[l], d = start[l];
    for (long long k = 0; k < n; ++k) tc[d + k] = make_int2(l, int(a + k));
}

// Layer-pair tracklets (Event.cc:599-609, 755-790; tracklet layer pairs barrel
// 1-3, 2-4, 3-5, Layer.cc:3-33): per inner SG of barrel layers 1-3, every SG of
// the layer two further out (tlCap, also the candidate counter of the inner
// layer, Event.cc:780).
__global__ void k_tll_cap(Count cSG, int NM, const ModDev* mods, const int* sgIn, const int* mdSlot,
                          const long long* layerOff, const int* layerModStart, long long* tlCap, int* capInt) {
    const int nSG = int(cSG.get());
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nSG) return;
    const int slot = mdSlot[sgIn[s]];
    const int e = slot / NM, L = layerBin(mods[slot % NM]);
    long long c = 0;
    if (L <= 2) {
        long long a, b;
        layerRange(layerOff, layerModStart, NM, e, L + 2, a, b);
        c = b - a;
    }
    tlCap[s] = c;
    capInt[s] = int(c);
}

// One wave per inner SG: both tracklet stages against every SG of the outer
// layer (whose inner MDs are barrel modules, so the algorithm variant is the
// inner SG's); survivors appended to the SG's slab in candidate order.
__global__ void __launch_bounds__(kBlock) k_tll(Count cSG, int NM, const ModDev* mods, HitsDev H, MDDev D, SGDev S,
                                                const long long* layerOff, const int* layerModStart, const int* layerSg,
                                                const int* capInt, const long long* capOff, int* slab, int* tlCount,
                                                int algo, Guard g) {
    if (g.over()) return;
    const int nSG = int(cSG.get());
    const int item = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (item >= nSG) return;
    if (capInt[item] == 0) {
        if (lane == 0) tlCount[item] = 0;
        return;
    }
    const int slot = D.slot[S.in[item]];
    const int e = slot / NM, L = layerBin(mods[slot % NM]);
    long long a, b;
    layerRange(layerOff, layerModStart, NM, e, L + 2, a, b);
    const TLInner I = loadInner(item, D, S);
    const int fl = S.flags[item];
    const int V = tlVariant(fl & 1, fl & 2, true);
    const long long off = capOff[item];
    int base = 0;
    for (long long k0 = a; k0 < b; k0 += kWave) {
        const long long k = k0 + lane;
        bool pass = false;
        int so = 0;
        if (k < b) so = layerSg[k];
        if (k < b && algo == TLA_ALLCOMB) {
            pass = true;
        } else if (k < b && V < kNVariantsTL) {
            const int flo = S.flags[so];
            const TLOutLo Lo = loadOutLo(S.in[so], flo, D);
            TLCarry c{0.f, 0.f, 0.f};
            if (tlStage1(V, I, Lo, c)) {
                const TLOutUp U = loadOutUp(so, flo, H, D, S);
                pass = tlStage2(V, (flo & 2) != 0, I, Lo, U, c, nullptr, algo == TLA_NM1);
            }
        }
        const unsigned long long mask = __ballot(pass);
        if (pass) slab[off + base + __popcll(mask & ((1ull << lane) - 1ull))] = so;
        base += __popcll(mask);
    }
    if (lane == 0) tlCount[item] = base;
}

// Tracklets of inner SG s, in slab order, at tlStart[s].
__global__ void k_tll_repack(Count cSG, const long long* capOff, const long long* tlStart, const int* slab, int* tlIn,
                             int* tlOut, Guard g) {
    if (g.over()) return;
    const int nSG = int(cSG.get());
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nSG) return;
    const long long d = tlStart[s], n = tlStart[s + 1] - d, o = capOff[s];
    for (long long k = 0; k < n; ++k) {
        tlIn[d + k] = s;
        tlOut[d + k] = slab[o + k];
    }
}

// Tracklets via navigation (Event.cc:696-751): inner SG s1, every SG s2 that
// starts at s1's outer MD, every SG s3 that starts at s2's outer MD; the
// candidate is (s1, s3).  Per s1: the number of candidates (also its counter).
__global__ void k_tln_cap(Count cSG, const int* sgOut, const long long* sgStart, long long* tlCap, int* capInt) {
    const int nSG = int(cSG.get());
    const int s1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (s1 >= nSG) return;
    const int m1 = sgOut[s1];
    long long c = 0;
    for (long long s2 = sgStart[m1]; s2 < sgStart[m1 + 1]; ++s2) {
        const int m2 = sgOut[s2];
        c += sgStart[m2 + 1] - sgStart[m2];
    }
    tlCap[s1] = c;
    capInt[s1] = int(c);
}

// One wave per inner SG: the lanes take the candidates of each s2 in turn; both
// tracklet stages with the variant of each candidate's anchors.
__global__ void __launch_bounds__(kBlock) k_tln(Count cSG, int NM, const ModDev* mods, HitsDev H, MDDev D, SGDev S,
                                                const long long* sgStart, const int* capInt, const long long* capOff,
                                                int* slab, int* tlCount, int algo, Guard g) {
    if (g.over()) return;
    const int nSG = int(cSG.get());
    const int item = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (item >= nSG) return;
    if (capInt[item] == 0) {
        if (lane == 0) tlCount[item] = 0;
        return;
    }
    const TLInner I = loadInner(item, D, S);
    const int fl = S.flags[item];
    const int m1 = S.out[item];
    const long long off = capOff[item];
    int base = 0;
    for (long long s2 = sgStart[m1]; s2 < sgStart[m1 + 1]; ++s2) {
        const int m2 = S.out[s2];
        const long long a = sgStart[m2], b = sgStart[m2 + 1];
        for (long long k0 = a; k0 < b; k0 += kWave) {
            const long long s3 = k0 + lane;
            bool pass = false;
            if (s3 < b && algo == TLA_ALLCOMB) {
                pass = true;
            } else if (s3 < b) {
                const int flo = S.flags[s3];
                const int mLo = S.in[s3];
                const int V = tlModuleVariant(fl, mods[D.slot[mLo] % NM]);
                if (V < kNVariantsTL) {
                    const TLOutLo Lo = loadOutLo(mLo, flo, D);
                    TLCarry c{0.f, 0.f, 0.f};
                    if (tlStage1(V, I, Lo, c)) {
                        const TLOutUp U = loadOutUp(int(s3), flo, H, D, S);
                        pass = tlStage2(V, (flo & 2) != 0, I, Lo, U, c, nullptr, algo == TLA_NM1);
                    }
                }
            }
            const unsigned long long mask = __ballot(pass);
            if (pass) slab[off + base + __popcll(mask & ((1ull << lane) - 1ull))] = int(s3);
            base += __popcll(mask);
        }
    }
    if (lane == 0) tlCount[item] = base;
}

// TLs per (event, layer bin of the inner SG): the layer lists of tracklets.
__global__ void k_tl_per_layer(Count cTL, int NM, const ModDev* mods, const int* tlIn, const int* sgIn,
                               const int* mdSlot, int* perLayer) {
    const int nTL = int(cTL.get());
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nTL) return;
    const int slot = mdSlot[sgIn[tlIn[l]]];
    atomicAdd(&perLayer[(slot / NM) * kNBins + layerBin(mods[slot % NM])], 1);
}

// Layer-pair track candidates (Event.cc:823-872: barrel layer 1 against barrel
// layer 3 only): inner TL l of layer bin 0 pairs with every TL of bin 2; the
// default algorithm passes iff l's outer SG is the other TL's inner SG
// (TrackCandidate.cc:143-226, A-13), i.e. the TLs that start at l's outer SG
// when that SG lies in bin 2.  cand[l] = the candidates counted for l.
__global__ void k_tcll_count(Count cTL, int NM, const ModDev* mods, const int* tlIn, const int* tlOut,
                             const int* sgIn, const int* mdSlot, const long long* tlStart, const int* perLayer,
                             int allComb, int* cnt, long long* cand) {
    const int nTL = int(cTL.get());
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nTL) return;
    const int slot = mdSlot[sgIn[tlIn[l]]];
    const int e = slot / NM;
    int c = 0;
    long long k = 0;
    if (layerBin(mods[slot % NM]) == 0) {
        k = perLayer[e * kNBins + 2];
        const int so = tlOut[l];
        const int slotO = mdSlot[sgIn[so]];
        if (allComb) c = int(k);  // AllComb_TCAlgo: every candidate
        else if (layerBin(mods[slotO % NM]) == 2) c = int(tlStart[so + 1] - tlStart[so]);
    }
    cnt[l] = c;
    cand[l] = k;
}

// TC from triplets (Event.cc:889-982).  The candidate (inner TP t ending at MD C,
// outer TP t' = (s3, s4) starting at MD D) needs the bridge SG(C, D) to pass the
// SG algorithm and the tracklet (bridge, s4) to pass the TL algorithm.  Both are
// pure functions of their inputs that the SG and TL stages already evaluated on
// exactly these candidates (D's module is connected to C's because s3..s4 and the
// bridge are module-map objects), so: for every tracklet (b, s4) with b = C->D,
// the TCs are {TPs ending at C} x {TP(s3, s4)} when the SG s3 = D->s4.inner exists.
//
// The TPs ending at each MD are listed once.  Entry k of the SG-end lists is the
// pair (MD m, SG s2 ending at m); it owns the TPs (s1, s2) for the SGs s1 ending
// at s2's inner MD, so the list of MD m is the range [tpeOff[sgEndStart[m]],
// tpeOff[sgEndStart[m+1]]) of the scan tpeOff of the per-entry counts.
__global__ void k_tpe_count(Count cSG, const int* sgEndList, const long long* sgEndStart, const int* sgIn, int* cnt) {
    const int nSG = int(cSG.get());
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nSG) return;
    const int mx = sgIn[sgEndList[k]];
    cnt[k] = int(sgEndStart[mx + 1] - sgEndStart[mx]);
}
// kTpeLanes lanes per listed SG write its TPs (nTP in total; the two totals
// are equal, every TP (s1, s2) is listed once).
constexpr int kTpeLanes = 8;
__global__ void k_tpe_write(Count cTP, Count cSG, const long long* tpeOff, const int* sgEndList,
                            const long long* sgEndStart, const int* sgIn, const long long* sgStart,
                            const long long* tpStart, int* tpEndList, Guard g) {
    if (g.over()) return;
    const long long nTP = cTP.get();
    const int nSG = int(cSG.get());
    const int gl = threadIdx.x & (kTpeLanes - 1);
    const long long k = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kTpeLanes;
    if (k >= nSG) return;
    const long long a = tpeOff[k], e = min(tpeOff[k + 1], nTP);
    if (a + gl >= e) return;
    const int s2 = sgEndList[k];
    const int mx = sgIn[s2];
    const long long b1 = sgEndStart[mx], d = s2 - sgStart[mx];
    for (long long j = a + gl; j < e; j += kTpeLanes) tpEndList[j] = int(tpStart[sgEndList[b1 + (j - a)]] + d);
}
// Per tracklet: the bridge test (s3 exists), the number of TCs and where their
// inner TPs are listed; info = (first listed inner TP, outer TP).  tlDefault
// (after DefaultNm1 tracklets, else null): the tracklets the default algorithm passes.
__global__ void k_tcp_count(Count cTL, const int* tlIn, const int* tlOutArr, const int* sgIn, const int* sgOut,
                            const long long* sgStart, const long long* sgEndStart, const long long* tpeOff,
                            const long long* tpStart, const unsigned char* tlDefault, int2* info, int* cnt) {
    const int nTL = int(cTL.get());
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nTL) return;
    if (tlDefault && !tlDefault[l]) {  // a DefaultNm1 tracklet the default algorithm rejects bridges nothing
        cnt[l] = 0;
        info[l] = make_int2(0, -1);
        return;
    }
    const int b = tlIn[l], s4 = tlOutArr[l];
    const int mdC = sgIn[b], mdD = sgOut[b], mdE = sgIn[s4];
    int s3 = -1;
    for (long long k = sgStart[mdD]; k < sgStart[mdD + 1]; ++k)
        if (sgOut[k] == mdE) {
            s3 = int(k);
            break;
        }
    const long long a = tpeOff[sgEndStart[mdC]], e = tpeOff[sgEndStart[mdC + 1]];
    cnt[l] = s3 >= 0 ? int(e - a) : 0;
    info[l] = make_int2(int(a), s3 >= 0 ? int(tpStart[s3] + (s4 - sgStart[mdE])) : -1);
}

// ------------------------------------------------------------------ counters

// All per-layer counters of the stages in `stages` (bit s = stage s), one
// grid-stride pass over slots, MDs, SGs, TPs and TLs (Event.cc:1169-1344).  A
// candidate or object counts at the lower module of its innermost MD (TCs
// from triplets are counted where they are written).  Bins are pre-aggregated
// in LDS, one global atomic per touched bin and block.
constexpr int kMaxLdsBins = 4096;
// The per-layer counters are kept in kCounterCopies copies, each block adding
// into copy (block index mod kCounterCopies) and the readers summing them:
// device-scope atomics on one address serialise (about 10 ns each), so a
// single copy makes the per-block flushes of a kernel queue behind each other.
constexpr int kCounterCopies = 16;

struct CounterIn {
    int NS, NM, E, stages;
    const ModDev* mods;
    const int *connStart, *connList;
    const int *mdCap, *mdCount;                              // per slot
    const int *mdSlot, *sgCap, *sgCount;                     // per MD
    const int *sgIn, *sgOut, *tpCount, *tlCount;             // per SG
    const long long* tlCap;                                  // per SG
    const long long *slotSgStart, *slotTpStart, *slotTlStart;
    const int *tpIn, *tpOut;                                 // per TP
    const int *tlIn, *tclCount;                              // per TL
    const long long* tcCand;  // per TL: TC candidates of the layer-pair form (null: the module-map count)
    long long* connSum;       // per slot, [3][NS]: SGs, TPs, TLs on the modules connected to the slot's module
    const int* sgKey;         // per SG: event * kNBins + layer bin of its inner MD's lower module
    const int* tpKey;         // per TP: sgKey of its inner SG
};

__device__ __forceinline__ long long connected(const CounterIn& C, int slot, const long long* slotStart) {
    const int m = slot % C.NM, eb = slot - m;
    long long c = 0;
    for (int k = C.connStart[m]; k < C.connStart[m + 1]; ++k) {
        const int t = eb + C.connList[k];
        c += slotStart[t + 1] - slotStart[t];
    }
    return c;
}

// Objects on the modules connected to each slot's module, for the candidate
// counters that count them per object (SGs for TP candidates, TPs and TLs for TC
// candidates): one walk of the connection list per slot instead of one per
// object.  kConnLanes lanes per slot take one connection each and sum by shuffles.
constexpr int kConnLanes = 16;
__global__ void __launch_bounds__(kBlock) k_conn_sums(CounterIn C, const long long* tot) {
    const int gl = threadIdx.x & (kConnLanes - 1);
    const int slot = (blockIdx.x * blockDim.x + threadIdx.x) / kConnLanes;
    if (slot >= C.NS || tot[TOT_ABORT] != 0) return;  // uniform over the group
    const int m = slot % C.NM, eb = slot - m;
    const bool sg = C.stages & 4, tp = C.stages & 16, tl = (C.stages & 32) && !C.tcCand;
    long long a = 0, b = 0, c = 0;
    for (int k = C.connStart[m] + gl; k < C.connStart[m + 1]; k += kConnLanes) {
        const int t = eb + C.connList[k];
        if (sg) a += C.slotSgStart[t + 1] - C.slotSgStart[t];
        if (tp) b += C.slotTpStart[t + 1] - C.slotTpStart[t];
        if (tl) c += C.slotTlStart[t + 1] - C.slotTlStart[t];
    }
#pragma unroll
    for (int off = 1; off < kConnLanes; off <<= 1) {
        a += __shfl_xor(a, off, kConnLanes);
        b += __shfl_xor(b, off, kConnLanes);
        c += __shfl_xor(c, off, kConnLanes);
    }
    if (gl != 0) return;
    if (sg) C.connSum[slot] = a;
    if (tp) C.connSum[C.NS + slot] = b;
    if (tl) C.connSum[2 * C.NS + slot] = c;
}

// Adds v (0: nothing) to bins[key] for every lane of the wave, one atomic per
// distinct key (lanes of a wave mostly share one).  Every lane must be active.
__device__ __forceinline__ void waveBinAdd(unsigned long long* bins, int key, long long v) {
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long rem = __ballot(v != 0);
    while (rem) {
        const int leader = __builtin_ctzll(rem);
        const int k = __shfl(key, leader);
        const bool mine = ((rem >> lane) & 1ULL) && key == k;
        long long x = mine ? v : 0;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) x += __shfl_xor(x, off);
        if (lane == leader) atomicAdd(&bins[k], (unsigned long long)x);
        rem &= ~__ballot(mine);
    }
}

__global__ void __launch_bounds__(kBlock) k_counters(CounterIn C, const long long* tot, unsigned long long* counters) {
    extern __shared__ unsigned long long ldsBins[];  // nb entries when they fit kMaxLdsBins (else none)
    const int nb = C.E * kNCounters * kNBins;
    unsigned long long* copy = counters + size_t(blockIdx.x % kCounterCopies) * nb;
    const bool lds = nb <= kMaxLdsBins;
    if (lds)
        for (int b = threadIdx.x; b < nb; b += blockDim.x) ldsBins[b] = 0;
    __syncthreads();
    const bool aborted = tot[TOT_ABORT] != 0;
    const long long nSlot = aborted || !(C.stages & 1) ? 0 : C.NS;
    const long long nMd = aborted || !(C.stages & 2) ? 0 : tot[TOT_MD];
    const long long nSg = aborted || !(C.stages & (4 | 8)) ? 0 : tot[TOT_SG];
    const long long nTp = aborted || !(C.stages & 16) ? 0 : tot[TOT_TP];
    const long long nTl = aborted || !(C.stages & 32) ? 0 : tot[TOT_TL];
    const long long n = nSlot + nMd + nSg + nTp + nTl;
    const long long stride = (long long)gridDim.x * blockDim.x;
    // block-uniform loop: every lane reaches the wave-wide adds
    for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += stride) {
        long long j = base + threadIdx.x;
        int key = 0;  // event * kNBins + layer bin
        int w[4] = {0, 0, 0, 0};
        long long v[4] = {0, 0, 0, 0};
        if (j < n) {
            if (j < nSlot) {  // MD candidates / objects per lower slot (Event.cc:257-305)
                key = int(j / C.NM) * kNBins + layerBin(C.mods[j % C.NM]);
                w[0] = 0; v[0] = C.mdCap[j];
                w[1] = 5; v[1] = C.mdCount[j];
            } else if ((j -= nSlot) < nMd) {  // SG candidates / objects per inner MD (Event.cc:457-522)
                const int slot = C.mdSlot[j];
                key = (slot / C.NM) * kNBins + layerBin(C.mods[slot % C.NM]);
                w[0] = 1; v[0] = C.sgCap[j];
                w[1] = 6; v[1] = C.sgCount[j];
            } else if ((j -= nMd) < nSg) {  // TP (Event.cc:539-597) and TL (Event.cc:627-693) per inner SG
                key = C.sgKey[j];
                if (C.stages & 4) {
                    w[0] = 2; v[0] = C.connSum[C.mdSlot[C.sgIn[j]]];
                    w[1] = 7; v[1] = C.tpCount[j];
                }
                if (C.stages & 8) {
                    w[2] = 3; v[2] = C.tlCap[j];
                    w[3] = 8; v[3] = C.tlCount[j];
                }
            } else if ((j -= nSg) < nTp) {  // TC candidates per inner TP: TPs at modules connected to its outermost MD (Event.cc:889-982)
                key = C.tpKey[j];
                w[0] = 4; v[0] = C.connSum[C.NS + C.mdSlot[C.sgOut[C.tpOut[j]]]];
            } else {  // TC candidates and objects per inner TL (Event.cc:999-1062; layer pairs: Event.cc:840-872)
                j -= nTp;
                const int s = C.tlIn[j];
                key = C.sgKey[s];
                w[0] = 4; v[0] = C.tcCand ? C.tcCand[j] : C.connSum[2 * C.NS + C.mdSlot[C.sgOut[s]]];
                w[1] = 9; v[1] = C.tclCount[j];
            }
        }
        const int key0 = (key / kNBins) * kNCounters * kNBins + key % kNBins;
        // one atomic per distinct bin and wave (lanes of a wave mostly share a bin)
        unsigned long long* dst = lds ? ldsBins : copy;
#pragma unroll
        for (int c = 0; c < 4; ++c) waveBinAdd(dst, key0 + w[c] * kNBins, v[c]);
    }
    __syncthreads();
    if (lds)
        for (int b = threadIdx.x; b < nb; b += blockDim.x)
            if (ldsBins[b]) atomicAdd(&copy[b], ldsBins[b]);
}

// Adds 1 to bins[key] for every active lane: one LDS atomic per distinct key of
// the wave instead of one per lane (lanes of a wave mostly share a key).
__device__ __forceinline__ void binCount(unsigned long long* bins, int key) {
    const int lane = threadIdx.x & (kWave - 1);
    bool done = false;
    for (;;) {
        const unsigned long long rem = __ballot(!done);
        if (!rem) break;
        const int leader = __builtin_ctzll(rem);
        const int k = __shfl(key, leader);
        const bool mine = !done && key == k;
        const unsigned long long m = __ballot(mine);
        if (lane == leader) atomicAdd(&bins[k], (unsigned long long)__popcll(m));
        done = done || mine;
    }
}

// TCs from triplets, kTcpLanes lanes per tracklet (grid-stride over tracklets):
// the lanes copy the tracklet's inner TPs.  Each TC counts at the module of its
// inner TP (Event.cc:893-897,965) in a per-block LDS histogram.
constexpr int kTcpLanes = 16;
__global__ void __launch_bounds__(kBlock) k_tcp_write(Count cTL, const long long* start, const int2* info,
                                                      const int* tpEndList, const int* tpKey, int E,
                                                      int2* tc, unsigned long long* counters, Guard g) {
    if (g.over()) return;
    const int nTL = int(cTL.get());
    extern __shared__ unsigned long long bins[];  // nb entries when they fit kMaxLdsBins (else none)
    const int nb = E * kNBins;
    counters += size_t(blockIdx.x % kCounterCopies) * E * kNCounters * kNBins;
    const bool lds = nb <= kMaxLdsBins;
    if (lds)
        for (int b = threadIdx.x; b < nb; b += blockDim.x) bins[b] = 0;
    __syncthreads();
    const int gl = threadIdx.x & (kTcpLanes - 1);
    const int g0 = (blockIdx.x * blockDim.x + threadIdx.x) / kTcpLanes, ng = gridDim.x * blockDim.x / kTcpLanes;
    for (int l = g0; l < nTL; l += ng) {
        const long long a = start[l], n = start[l + 1] - a;
        if (n == 0) continue;
        const int2 f = info[l];
        for (long long q = gl; q < n; q += kTcpLanes) {
            const int t = tpEndList[f.x + q];
            tc[a + q] = make_int2(t, f.y);
            const int key = tpKey[t];  // the TC's inner TP's inner SG
            if (lds) binCount(bins, key);
            else atomicAdd(&counters[((key / kNBins) * kNCounters + 9) * kNBins + key % kNBins], 1ULL);
        }
    }
    __syncthreads();
    if (lds)
        for (int b = threadIdx.x; b < nb; b += blockDim.x)
            if (bins[b]) atomicAdd(&counters[((b / kNBins) * kNCounters + 9) * kNBins + (b % kNBins)], bins[b]);
}


// Per-slot object starts for objects ordered by inner SG: slotStart[s] = start[slotSgStart[s]]
// With pairOff set (the module-map tracklet stage), the thread of slot NS also
// records the pairs and candidates per variant for the next call's size
// predictions.
__global__ void k_slot_range_sg(Count cNS, const long long* slotSgStart, const long long* objStartBySg,
                                long long* slotObjStart, const long long* pairOff = nullptr,
                                const long long* candOff = nullptr, long long stride = 0, long long* tot = nullptr) {
    const int NS = int(cNS.get());
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > NS) return;
    slotObjStart[s] = objStartBySg[slotSgStart[s]];
    if (s != NS || !pairOff || tot[TOT_ABORT]) return;
    for (int v = 0; v < kNVariantsTL; ++v) {
        const long long p0 = pairOff[v * stride], p1 = pairOff[(v + 1) * stride];
        tot[TOT_TLP0 + v] = p1 - p0;
        tot[TOT_TLC0 + v] = candOff[p1] - candOff[p0];
    }
}

// ------------------------------------------------------------------ export
__global__ void k_export_md(long long a, long long n, const int* lo, const int* up, const int* idx, int* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[2 * i] = idx[lo[a + i]];
    out[2 * i + 1] = idx[up[a + i]];
}
__device__ __forceinline__ void putSG(int s, const int* sgIn, const int* sgOut, const int* lo, const int* up,
                                      const int* idx, int* o) {
    const int a = sgIn[s], b = sgOut[s];
    o[0] = idx[lo[a]]; o[1] = idx[up[a]]; o[2] = idx[lo[b]]; o[3] = idx[up[b]];
}
// kind: 1 SG, 2 TP/TL pair of SGs, 3 TC of TP pairs, 4 TC of TL pairs
// ------------------------------------------------------------------ per-object values
// The reference's typed getters and reco variables of stored MDs, SGs and TLs
// (MiniDoublet.h, Segment.h, Tracklet.h; layouts kMDVars, kSGVars, kTLVars in
// sdl_cuts.h): every object of [a, a + n) is evaluated once more with the
// algorithms' recording on; one thread per object.  Pseudo MDs (no cuts run)
// keep their zero values and have no miniCut (NaN).
__global__ void k_vars_md(long long a, long long n, int NM, const ModDev* mods, HitsDev H, MDDev D, bool pseudo,
                          float* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long m = a + i;
    float rec[kMDVars] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (pseudo) {
        rec[6] = __builtin_nanf("");
    } else {
        const HitView hl = loadHit(D.lo[m], H.x, H.y, H.z, H.rt, H.r3, H.phi, H.alpha);
        const HitView hu = loadHit(D.up[m], H.x, H.y, H.z, H.rt, H.r3, H.phi, H.alpha);
        MDOut o;
        mdAlgo(hl, hu, mods[D.slot[m] % NM], o, rec);
    }
    for (int k = 0; k < kMDVars; ++k) out[i * kMDVars + k] = rec[k];
}

__global__ void k_vars_sg(long long a, long long n, int NM, const ModDev* mods, HitsDev H, MDDev D, SGDev S,
                          float* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long s = a + i;
    const int mi = S.in[s], mo = S.out[s];
    float rec[kSGVars];
    for (int k = 0; k < kSGVars; ++k) rec[k] = 0.f;
    SGOut o;
    // the connection test only decides pass/fail before any value is computed
    sgCandidate(loadMD(mi, D), mods[D.slot[mi] % NM], mo, mods[D.slot[mo] % NM], true, H, D, o, rec);
    for (int k = 0; k < kSGVars; ++k) out[i * kSGVars + k] = rec[k];
}

__global__ void k_vars_tl(long long a, long long n, int NM, const ModDev* mods, HitsDev H, MDDev D, SGDev S,
                          const int* tlIn, const int* tlOutArr, float* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = tlIn[a + i], so = tlOutArr[a + i];
    const int mdOutLo = S.in[so];
    const int flo = S.flags[so];
    const int V = tlModuleVariant(S.flags[s], mods[D.slot[mdOutLo] % NM]);
    // written in place: the typed values start at the constructor's 0, the map
    // at its reset (Tracklet.cc:152-248), the _v2 algorithm's own keys absent
    float* rec = out + i * kTLVars;
    float* tk = rec + kTLTyped;
    for (int k = 0; k < kTLTyped; ++k) rec[k] = 0.f;
    for (int k = 0; k < kTLKeys; ++k) tk[k] = -999.f;
    const int v2Keys[10] = {TK_alpha1GeV_OutLo, TK_drt_InSeg, TK_drt_tl_axis, TK_rtRatio_OutLoInLo, TK_rt_InLo,
                            TK_rt_OutLo, TK_z_InLo, TK_z_OutLo, TK_zpitch_InLo, TK_zpitch_OutLo};
    for (int k = 0; k < 10; ++k) tk[v2Keys[k]] = __builtin_nanf("");
    tk[TK_algo] = V == TL_BBBB ? 1.f : (V == TL_BBEE ? 2.f : 3.f);  // Tracklet.cc:256-293
    const TLInner I = loadInner(s, D, S);
    const TLOutLo L = loadOutLo(mdOutLo, flo, D);  // the flags' inner-MD bits are those of every SG at mdOutLo
    const TLOutUp U = loadOutUp(so, flo, H, D, S);
    TLCarry c{0.f, 0.f, 0.f};
    // a stored TL passed; under DefaultNm1 it may fail the waived cut, so the values are read with it waived
    if (tlStage1(V, I, L, c, nullptr, tk) && tlStage2(V, (flo & 2) != 0, I, L, U, c, rec, true) && V != TL_BBEE) {
        // the anchors' module detIds: the lower sensor's, or its partner's when the anchor is the upper hit
        const int md[4] = {S.in[s], S.out[s], mdOutLo, S.out[so]};
        uint32_t detId[4];
        for (int k = 0; k < 4; ++k) {
            const ModDev& lm = mods[D.slot[md[k]] % NM];
            detId[k] = D.anchor[md[k]] == D.lo[md[k]] ? lm.detId : (lm.partner >= 0 ? mods[lm.partner].detId : 0u);
        }
        tlCommonKeys(I.lo, I.up, L.a, U.a, detId, tk);
    }
}

// Under DefaultNm1_TLAlgo the stored tracklets include those that fail the
// waived dBeta cut, but TCs from triplets re-run the default tracklet
// algorithm on their bridges (Event.cc:940-956).  The TC join therefore reads,
// per stored tracklet, whether the default algorithm passes it: both cut
// stages again, nothing waived.
__global__ void k_tl_default(Count cTL, int NM, const ModDev* mods, HitsDev H, MDDev D, SGDev S, const int* tlIn,
                             const int* tlOutArr, unsigned char* def) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cTL.get()) return;
    const int s = tlIn[i], so = tlOutArr[i];
    const int mdOutLo = S.in[so];
    const int flo = S.flags[so];
    const int V = tlModuleVariant(S.flags[s], mods[D.slot[mdOutLo] % NM]);
    const TLInner I = loadInner(s, D, S);
    const TLOutLo L = loadOutLo(mdOutLo, flo, D);
    const TLOutUp U = loadOutUp(so, flo, H, D, S);
    TLCarry c{0.f, 0.f, 0.f};
    def[i] = (tlStage1(V, I, L, c) && tlStage2(V, (flo & 2) != 0, I, L, U, c, nullptr, false)) ? 1 : 0;
}

// TrackCandidate::runTrackCandidateDefaultAlgo's recorded values (TrackCandidate.cc:183-207)
// of TCs [a, a + n) made from tracklets: innerR through the inner TL's first
// three anchors, outerR through the outer TL's last three, dR = (innerR - outerR) / innerR.
__global__ void k_vars_tcl(long long a, long long n, MDDev D, SGDev S, const int* tlIn, const int* tlOutArr,
                           const int2* tcl, float* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 tc = tcl[a + i];
    const int sIn1 = tlIn[tc.x], sOut1 = tlOutArr[tc.x], sIn2 = tlIn[tc.y], sOut2 = tlOutArr[tc.y];
    const float4 iA = D.a[S.in[sIn1]], iB = D.a[S.out[sIn1]], iC = D.a[S.in[sOut1]];
    const float4 oA = D.a[S.out[sIn2]], oB = D.a[S.in[sOut2]], oC = D.a[S.out[sOut2]];
    const float innerR = radiusFromA(iA.x, iA.y, iB.x, iB.y, iC.x, iC.y);
    const float outerR = radiusFromA(oA.x, oA.y, oB.x, oB.y, oC.x, oC.y);
    out[i * kTCLVars] = (innerR - outerR) / innerR;
    out[i * kTCLVars + 1] = innerR;
}

__global__ void k_export(long long a, long long n, int kind, const int* pA, const int* pB, const int2* pC,
                         const int* tpIn, const int* tpOutArr, const int* tlIn, const int* tlOutArr, const int* sgIn,
                         const int* sgOut, const int* lo, const int* up, const int* idx, int* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long o = a + i;
    if (kind == 1) {
        putSG(int(o), sgIn, sgOut, lo, up, idx, out + 4 * i);
    } else if (kind == 2) {
        putSG(pA[o], sgIn, sgOut, lo, up, idx, out + 8 * i);
        putSG(pB[o], sgIn, sgOut, lo, up, idx, out + 8 * i + 4);
    } else {
        const int2 c = pC[o];
        const int* in = kind == 3 ? tpIn : tlIn;
        const int* ou = kind == 3 ? tpOutArr : tlOutArr;
        putSG(in[c.x], sgIn, sgOut, lo, up, idx, out + 16 * i);
        putSG(ou[c.x], sgIn, sgOut, lo, up, idx, out + 16 * i + 4);
        putSG(in[c.y], sgIn, sgOut, lo, up, idx, out + 16 * i + 8);
        putSG(ou[c.y], sgIn, sgOut, lo, up, idx, out + 16 * i + 12);
    }
}

// ------------------------------------------------------------------ host geometry
// The module table entry of a detId: the decode of sdl_module.h plus the
// per-module constants of the mini-doublet algebra.
ModDev decodeModule(uint32_t d) {
    const ModuleId id = decodeDetId(d);
    ModDev m{};
    m.detId = d;
    m.partner = -1;
    m.subdet = uint8_t(id.subdet);
    m.side = uint8_t(id.side);
    m.layer = uint8_t(id.layer);
    m.ring = uint8_t(id.ring);
    m.rod = uint8_t(id.rod);
    m.isLower = id.isLower;
    m.type = id.type;
    m.layerType = id.layerType;
    // MiniDoublet::isTighterTiltedModules / moduleGapSize (MiniDoublet.cc:1242-1332)
    bool tighter = false;
    if (m.subdet == SUB_BARREL && m.side != SIDE_CENTER) {
        tighter = m.layer == 3 || (m.side == SIDE_NEGZ && m.layer == 2 && m.rod > 5) ||
                  (m.side == SIDE_POSZ && m.layer == 2 && m.rod < 8) ||
                  (m.side == SIDE_NEGZ && m.layer == 1 && m.rod > 9) || (m.side == SIDE_POSZ && m.layer == 1 && m.rod < 4);
    }
    const float flat[6] = {0.26f, 0.16f, 0.16f, 0.18f, 0.18f, 0.18f};
    const int iL = m.layer - 1;
    if (m.subdet == SUB_BARREL && m.side == SIDE_CENTER) m.gap = (iL >= 0 && iL < 6) ? flat[iL] : 0.f;
    else if (tighter) m.gap = 0.26f;
    else if (m.subdet == SUB_ENDCAP) {
        const int j = m.ring - 1;
        if (iL == 0 || iL == 1) m.gap = j < 10 ? 0.4f : 0.18f;
        else if (iL == 2 || iL == 3) m.gap = j < 8 ? 0.4f : 0.18f;
        else m.gap = j < 9 ? 0.4f : 0.18f;
    } else m.gap = 0.4f;
    return m;
}

uint32_t partnerOf(uint32_t d) { return decodeDetId(d).partner; }

}  // namespace

struct sdl_geom {
    int device = 0;
    int NM = 0;
    std::vector<uint32_t> detIds;
    std::vector<ModDev> mods;
    std::vector<int> connStart, connList;
    uint2* dDetTable = nullptr;  // detId -> module hash table (detHash)
    int hashBits = 0;
    ModDev* dMods = nullptr;
    int *dConnStart = nullptr, *dConnList = nullptr;
    // modules ordered by layer bin (then index), and where each bin starts
    int *dLayerMods = nullptr, *dLayerModStart = nullptr;
    int nCU = 256;  // the device's compute units (bounded grids)
};

struct sdl_mlp {
    int device = 0;
    hipStream_t st = nullptr;
    float* dParams = nullptr;   // float32 master copy
    float* dGrad = nullptr;
    float* dM = nullptr;        // Adam moments
    float* dV = nullptr;
    float* dLoss = nullptr;
    sdl::mlp::Packed* dPacked = nullptr;  // BF16 fragments for inference
    long long adamSteps = 0;
};

// Launch shapes of the kernels whose best shape depends on the batch and the
// device: block sizes of the thread-per-item kernels, the wave-per-MD segment
// kernel's block, and the waves per block of the learned first stage.
enum { TUNE_MD_BLOCK = 0, TUNE_SG_BLOCK, TUNE_TL1_BLOCK, TUNE_TL2_BLOCK, TUNE_TL_EXPAND_BLOCK, TUNE_SCORED_WAVES, TUNE_N };
struct Tuning {
    int v[TUNE_N] = {256, 256, 256, 256, 256, 8};
    bool operator==(const Tuning& o) const {
        for (int k = 0; k < TUNE_N; ++k)
            if (v[k] != o.v[k]) return false;
        return true;
    }
};
const char* const kTuneNames[TUNE_N] = {"md_block", "sg_block", "tl1_block", "tl2_block", "tl_expand_block",
                                        "scored_waves"};
// Candidate values the autotuner tries for each knob (0 ends a list).
const int kTuneValues[TUNE_N][5] = {{64, 128, 256, 0}, {64, 128, 256, 0}, {64, 128, 256, 0}, {64, 128, 256, 0},
                                    {64, 128, 256, 0}, {2, 4, 8, 0}};

struct sdl_event {
    sdl_geom* g = nullptr;
    hipStream_t st = nullptr;
    // host staging (add_hits)
    std::vector<float> hx, hy, hz;
    std::vector<int> hidx;
    std::vector<unsigned> hdet;
    std::vector<long long> hoff{0};
    bool hostStaged = true;
    // batch description
    int E = 0;
    long long NH = 0;
    std::vector<long long> evOff;
    const float *px = nullptr, *py = nullptr, *pz = nullptr;
    const int* pidx = nullptr;
    const unsigned* pdet = nullptr;
    int stage = -1;   // last completed stage
    bool hitsReady = false;
    long long nMD = 0, nSG = 0, nTP = 0, nTL = 0, nTCP = 0, nTCL = 0;
    // run totals on the device; sizes of the current call (exact or predicted)
    DBuf<long long> tot;
    bool spec = false;
    long long need[TOT_N] = {};
    // totals of the last completed call, to predict buffer sizes from
    bool haveHistory = false;
    long long lastTot[TOT_N] = {};
    long long lastNH = 0;
    unsigned ranMask = 0;  // stages run on the current batch (bit = stage)
    // launch shapes (chosen by sdl_event_autotune; defaults below)
    Tuning tune;
    // captured chain (speculative mode)
    bool useGraphs = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    struct {
        int E = -1;
        const void *px = nullptr, *py = nullptr, *pz = nullptr, *pidx = nullptr, *pdet = nullptr;
        unsigned long long allocEpoch = 0;
        int headroomFrom = 0;
        double headroomFactor = 0;
        const void* proposal = nullptr;
        float proposalLogit = 0.f;
        Tuning tune;
    } graphKey;
    // prediction headroom: factor for the totals of stages >= headroomFrom
    int headroomFrom = SDL_STAGE_TCL + 1;
    double headroomFactor = 1.25;
    // device buffers
    DBuf<float> bx, by, bz;
    DBuf<int> bidx;
    DBuf<unsigned> bdet;
    DBuf<long long> dEvOff;
    // the offsets dEvOff holds, and the allocation epoch when they were sent
    std::vector<long long> evOffSent;
    unsigned long long evOffSentEpoch = ~0ULL;
    DBuf<float> rt, r3, phi, halpha;
    DBuf<int> hslot;
    DBuf<float4> eHi, eLo;
    DBuf<int> slotHitCount, slotFill, sortedHit;
    DBuf<long long> slotHitStart;
    DBuf<int> mdCap, mdCount;
    DBuf<long long> mdCapOff, mdStart;
    DBuf<int2> mdSlabHits;
    DBuf<unsigned char> mdFlag;
    DBuf<int> mdSlabSlot;
    DBuf<long long> mdPos;
    DBuf<float4> mdSlabF;
    DBuf<int> mdLo, mdUp, mdAnchor, mdSlot;
    DBuf<float4> mdA, mdB, mdP;
    DBuf<double> mdQ;
    DBuf<float2> mdC;
    DBuf<int> sgCap, sgCount, sgSlabOut, sgIn, sgOut;
    DBuf<long long> sgCapOff, sgStart, slotSgStart;
    DBuf<SGF> sgSlabF, sgF;
    DBuf<float4> sgX;
    DBuf<float2> sgE;
    DBuf<int> sgFlags, sgKey;
    DBuf<int> sgEndCount, sgEndFill, sgEndList;
    DBuf<long long> sgEndStart;
    DBuf<int> tpCount, tpIn, tpOut, tpKey, tpeCount, tpEndList;
    int tlNm1 = 0;  // the tracklet stage runs DefaultNm1_TLAlgo (dBeta cut waived)
    int tlAlgo = TLA_DEFAULT;  // algorithm of the last tracklet stage (TLAlgo value)
    int curAlgo = 0;           // algorithm of the stage being run by sdl_event_run (Algo.h value)
    unsigned acMask = 0;       // stages of the current batch run with an AllComb algorithm (bit = stage)
    bool forceExact = false;   // size every buffer exactly (calls with a non-default algorithm)
    DBuf<long long> layerTlOff;  // per-event layer lists of the tracklets (AllComb layer-pair TCs)
    DBuf<int> layerTl;
    DBuf<unsigned char> tlDefault;  // after DefaultNm1 tracklets: those the default algorithm passes
    DBuf<long long> tpeOff;
    DBuf<long long> tpStart, slotTpStart;
    DBuf<long long> tlCap, tlCapOff, tlStart, slotTlStart;
    DBuf<int> tlCount, tlIn, tlOut;
    DBuf<int> slotMdSg, slotMdSgList, tlNP, tlFill;
    DBuf<long long> slotMdSgStart;
    // tracklet pairs and candidates, one list each, ordered by variant
    DBuf<long long> tlPOff, tlCOff;
    DBuf<int2> tlPair, tlCand;
    DBuf<float4> tlCarry;
    DBuf<int> tlNOut;
    DBuf<unsigned char> tlFlag;
    // layer-pair stages: per-event layer lists of MDs and SGs, TL candidate slabs
    DBuf<int> layerCnt, layerMd, layerSg, tlCapInt, tllSlab, tlPerLayer;
    DBuf<long long> connSum;  // per slot [3][NS], for the counters
    DBuf<long long> layerMdOff, layerSgOff, tcCand;
    DBuf<int> tlScored;  // learned proposal: pairs scored per wave, per variant (diagnostic)
    long long tlScoredStride = 0;
    // learned proposal (null: the plain chain)
    sdl_mlp* proposal = nullptr;
    float proposalLogit = 0.f;
    // run statistics: stage runs, graph replays, graph captures, exact reruns
    long long stats[4] = {0, 0, 0, 0};
    DBuf<int> tcCount;
    DBuf<int2> tcpInfo;
    DBuf<long long> tcpStart, tclStart;
    DBuf<int2> tcp, tcl;
    DBuf<unsigned long long> counters;
    DBuf<long long> tmp64;
    DBuf<int> exportBuf;
    DBuf<float> exportVars;
    ScanScratch scan;
    std::vector<unsigned long long> hostCounters;
    // optional per-kernel timing of the combinatorial kernels (prep, MD, SG, TL)
    bool timing = false;
    hipEvent_t tev[8][2] = {};
    double kernelMs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

enum { T_PREP = 0, T_MD = 1, T_SG = 2, T_TL = 3, T_NTIMERS = 4 };

void timerBegin(sdl_event* ev, int k) {
    if (!ev->timing) return;
    if (!ev->tev[k][0]) {
        (void)hipEventCreate(&ev->tev[k][0]);
        (void)hipEventCreate(&ev->tev[k][1]);
    }
    (void)hipEventRecord(ev->tev[k][0], ev->st);
}
void timerEnd(sdl_event* ev, int k) {
    if (!ev->timing) return;
    (void)hipEventRecord(ev->tev[k][1], ev->st);
}
void timerCollect(sdl_event* ev, int k) {
    if (!ev->timing || !ev->tev[k][0]) return;
    float ms = 0.f;
    if (hipEventSynchronize(ev->tev[k][1]) == hipSuccess && hipEventElapsedTime(&ms, ev->tev[k][0], ev->tev[k][1]) == hipSuccess)
        ev->kernelMs[k] = ms;
}

long long readLL(const long long* d, hipStream_t st) {
    long long v = 0;
    if (hipMemcpyAsync(&v, d, sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    return v;
}

unsigned gridFor(long long n, int per) { return unsigned((n + per - 1) / per); }

// The three geometry tables as the reference's singletons hold them:
// TiltedGeometry drdz and slope (TiltedGeometry.cc:36-39), EndcapGeometry's
// lower slope and column 7, read as the centroid phi (EndcapGeometry.cc:48-58),
// and the connection lists in file order (ModuleConnectionMap.cc:36-44).
struct GeomTables {
    std::unordered_map<uint32_t, float> drdz, slope, ecSlope, ecPhi;
    std::vector<std::pair<uint32_t, std::vector<uint32_t>>> conn;
};

int parseGeomFiles(const char* tilted, const char* endcap, const char* map, GeomTables& T) {
    std::ifstream ft(tilted), fe(endcap), fm(map);
    if (!ft) return fail(SDL_ERR_IO, std::string("cannot open ") + tilted);
    if (!fe) return fail(SDL_ERR_IO, std::string("cannot open ") + endcap);
    if (!fm) return fail(SDL_ERR_IO, std::string("cannot open ") + map);
    std::string line;
    while (std::getline(ft, line)) {
        std::stringstream ss(line);
        unsigned d; float a, b;
        if (ss >> d >> a >> b) { T.drdz[d] = a; T.slope[d] = b; }
    }
    while (std::getline(fe, line)) {
        std::stringstream ss(line);
        unsigned d; float avgr2, yl, sl, yh, sh, c7, c8, cz;
        if (ss >> d >> avgr2 >> yl >> sl >> yh >> sh >> c7 >> c8 >> cz) { T.ecSlope[d] = sl; T.ecPhi[d] = c7; }
    }
    while (std::getline(fm, line)) {
        std::stringstream ss(line);
        unsigned d; int n;
        if (!(ss >> d >> n)) continue;
        std::vector<uint32_t> v;
        for (int i = 0; i < n; ++i) {
            unsigned c;
            if (!(ss >> c)) break;
            v.push_back(c);
        }
        T.conn.push_back({d, v});
    }
    return 0;
}

// The device-ready module table and CSR connection map of a set of tables.
int buildGeom(const GeomTables& T, sdl_geom& G) {
    std::unordered_set<uint32_t> ids;
    for (auto& kv : T.conn) {
        std::unordered_set<uint32_t> seen;
        for (uint32_t c : kv.second) {
            if (!seen.insert(c).second)
                return fail(SDL_ERR_DUPLICATE_CONNECTION,
                            "connection list of " + std::to_string(kv.first) + " repeats " + std::to_string(c));
            ids.insert(c);
        }
        ids.insert(kv.first);
    }
    for (auto& kv : T.drdz) ids.insert(kv.first);
    for (auto& kv : T.ecSlope) ids.insert(kv.first);
    std::vector<uint32_t> base(ids.begin(), ids.end());
    for (uint32_t d : base) ids.insert(partnerOf(d));
    G.detIds.assign(ids.begin(), ids.end());
    std::sort(G.detIds.begin(), G.detIds.end());
    G.NM = int(G.detIds.size());
    std::unordered_map<uint32_t, int> index;
    for (int i = 0; i < G.NM; ++i) index[G.detIds[i]] = i;
    auto get = [](const std::unordered_map<uint32_t, float>& m, uint32_t k) {
        auto it = m.find(k);
        return it == m.end() ? 0.f : it->second;  // a missing key reads 0 (TiltedGeometry.cc:45,50)
    };
    G.mods.resize(G.NM);
    for (int i = 0; i < G.NM; ++i) {
        ModDev m = decodeModule(G.detIds[i]);
        const uint32_t p = partnerOf(G.detIds[i]);
        m.partner = index.count(p) ? index[p] : -1;
        const uint32_t strip = m.layerType == LT_PIXEL ? p : G.detIds[i];
        m.drdzStrip = get(T.drdz, strip);
        m.slopeStrip = m.subdet == SUB_ENDCAP ? get(T.ecSlope, strip) : get(T.slope, strip);
        m.centroidPhi = get(T.ecPhi, G.detIds[i]);
        G.mods[i] = m;
    }
    std::vector<std::vector<int>> adj(G.NM);
    for (auto& kv : T.conn) {
        std::vector<int>& a = adj[index[kv.first]];
        a.clear();  // a later line for the same detId replaces the earlier one (ModuleConnectionMap.cc:45)
        for (uint32_t c : kv.second) a.push_back(index[c]);
    }
    G.connStart.assign(G.NM + 1, 0);
    for (int i = 0; i < G.NM; ++i) G.connStart[i + 1] = G.connStart[i] + int(adj[i].size());
    G.connList.clear();
    for (int i = 0; i < G.NM; ++i) G.connList.insert(G.connList.end(), adj[i].begin(), adj[i].end());
    return 0;
}

int loadGeomFiles(const char* tilted, const char* endcap, const char* map, sdl_geom& G) {
    GeomTables T;
    if (int r = parseGeomFiles(tilted, endcap, map, T)) return r;
    return buildGeom(T, G);
}

int ensureHits(sdl_event* ev) {
    const long long NH = ev->NH;
    const int NS = ev->E * ev->g->NM;
    if (NH > kMaxObjects32) return fail(SDL_ERR_ARG, "more than 134,217,728 hits in one call: split the batch");
    if (int r = ev->rt.ensure(NH)) return r;
    if (int r = ev->r3.ensure(NH)) return r;
    if (int r = ev->phi.ensure(NH)) return r;
    if (int r = ev->halpha.ensure(NH)) return r;
    if (int r = ev->hslot.ensure(NH)) return r;
    if (int r = ev->eHi.ensure(NH)) return r;
    if (int r = ev->eLo.ensure(NH)) return r;
    if (int r = ev->sortedHit.ensure(NH)) return r;
    if (int r = ev->slotHitCount.ensure(NS)) return r;
    if (int r = ev->slotFill.ensure(NS)) return r;
    if (int r = ev->slotHitStart.ensure(NS + 1)) return r;
    if (int r = ev->counters.ensure(size_t(kCounterCopies) * ev->E * kNCounters * kNBins)) return r;
    return 0;
}

HitsDev hitsView(sdl_event* ev) {
    return HitsDev{ev->px, ev->py, ev->pz, ev->rt.p, ev->r3.p, ev->phi.p, ev->halpha.p, ev->pidx, ev->eHi.p, ev->eLo.p};
}
MDDev mdView(sdl_event* ev) {
    return MDDev{ev->mdLo.p, ev->mdUp.p, ev->mdAnchor.p, ev->mdSlot.p, ev->mdA.p, ev->mdB.p, ev->mdC.p, ev->mdP.p,
                 ev->mdQ.p};
}
SGDev sgView(sdl_event* ev) {
    return SGDev{ev->sgIn.p, ev->sgOut.p, ev->sgF.p, ev->sgX.p, ev->sgE.p, ev->sgFlags.p, ev->sgKey.p};
}

Count hostCount(sdl_event* ev, long long n) { return Count{n, ev->tot.p, -1}; }
Count devCount(sdl_event* ev, int k) { return Count{0, ev->tot.p, k}; }
Guard guardOf(sdl_event* ev, int k, long long cap, int stage) { return Guard{ev->tot.p, k, cap, stage}; }

// Host-side size of run total k for this call: read back (exact mode) or
// predicted from the last call, scaled by the hit count, with 25 % headroom
// (speculative mode; kernels check it and the call runs again on overflow).
long long sizeOf(sdl_event* ev, int k) {
    long long v;
    if (!ev->spec) {
        v = readLL(ev->tot.p + k, ev->st);
    } else {
        static const int stageOf[TOT_N] = {SDL_STAGE_MD, SDL_STAGE_MD, SDL_STAGE_SG, SDL_STAGE_SG, SDL_STAGE_TP,
                                           SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TCP, SDL_STAGE_TCP, SDL_STAGE_TCL,
                                           SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL,
                                           SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL, 0, 0};
        const double scale = ev->lastNH > 0 ? double(ev->NH) / double(ev->lastNH) : 1.0;
        const bool custom = stageOf[k] >= ev->headroomFrom;
        const double f = custom ? ev->headroomFactor : 1.25;
        v = (long long)(double(ev->lastTot[k]) * scale * f) + (custom && f < 1.0 ? 0 : 1024);
    }
    ev->need[k] = v;
    return v;
}


// Hit ingestion: module lookup, derived quantities, counting sort by slot.
// Hits on modules unknown to the geometry are counted (TOT_UNKNOWN) and left out.
template <class T>
void zeroAsync(T* p, long long n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_zero<T>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, p, n);
}

// One launch zeroing several arrays: zeroMany(st, {span(p, n), ...}).
struct ZeroSpan {
    void* p;
    long long bytes;
};

template <class T>
ZeroSpan span(T* p, long long n) {
    static_assert(sizeof(T) % 4 == 0, "zeroMany works in 32-bit words");
    return ZeroSpan{p, n * (long long)sizeof(T)};
}

void zeroMany(hipStream_t st, std::initializer_list<ZeroSpan> spans) {
    ZeroSet z{};
    long long total = 0;
    for (const ZeroSpan& sp : spans) {
        if (sp.bytes <= 0) continue;
        if (z.n == kZeroArrays) {
            hipLaunchKernelGGL(k_zero_many, dim3(gridFor(total, kBlock)), dim3(kBlock), 0, st, z);
            z = ZeroSet{};
            total = 0;
        }
        total += sp.bytes / 4;
        z.p[z.n] = static_cast<unsigned*>(sp.p);
        z.end[z.n++] = total;
    }
    if (z.n) hipLaunchKernelGGL(k_zero_many, dim3(gridFor(total, kBlock)), dim3(kBlock), 0, st, z);
}

// Uploads the batch's event offsets (outside any captured graph: the host array
// is read when the copy runs).  Skipped when the device copy already holds
// them: the same values were sent and no buffer has been allocated since (an
// allocation anywhere advances the epoch, so a buffer freed and allocated
// again at the same address is never taken for the old one).
int uploadOffsets(sdl_event* ev) {
    if (int r = ev->dEvOff.ensure(ev->E + 1)) return r;
    if (ev->evOffSentEpoch == g_allocEpoch && ev->evOffSent == ev->evOff) return 0;
    HIP_OK(hipMemcpyAsync(ev->dEvOff.p, ev->evOff.data(), sizeof(long long) * (ev->E + 1), hipMemcpyHostToDevice, ev->st));
    ev->evOffSent = ev->evOff;
    ev->evOffSentEpoch = g_allocEpoch;
    return 0;
}

int runPrep(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const int NS = ev->E * G->NM;
    ev->ranMask = 0;
    if (int r = ensureHits(ev)) return r;
    zeroMany(st, {span(ev->tot.p, TOT_N), span(ev->slotHitCount.p, NS), span(ev->slotFill.p, NS),
                  span(ev->counters.p, (long long)kCounterCopies * ev->E * kNCounters * kNBins)});
    // launches cover the hit buffers' capacity; the kernels read the hit count
    const long long nhGrid = (long long)ev->rt.cap;
    timerBegin(ev, T_PREP);
    hipLaunchKernelGGL(k_hit_prep, dim3(gridFor(nhGrid, kBlock)), dim3(kBlock), 0, st, ev->px, ev->py, ev->pz, ev->pdet,
                       ev->dEvOff.p, ev->E, G->dDetTable, G->hashBits, G->NM, G->dMods, ev->rt.p, ev->r3.p, ev->phi.p, ev->halpha.p,
                       ev->hslot.p, ev->eHi.p, ev->eLo.p, ev->slotHitCount.p, ev->tot.p);
    timerEnd(ev, T_PREP);
    if (int r = scanExclusive<int>(ev->slotHitCount.p, ev->slotHitStart.p, hostCount(ev, NS), NS, ev->scan, st,
                                   ev->tot.p, -1))
        return r;
    hipLaunchKernelGGL(k_hit_scatter, dim3(gridFor(nhGrid, kBlock)), dim3(kBlock), 0, st, ev->dEvOff.p, ev->E,
                       ev->hslot.p, ev->slotHitStart.p, ev->slotFill.p, ev->sortedHit.p);
    hipLaunchKernelGGL(k_sort_segments<kSortLanesWide>, dim3(gridFor((long long)NS * kSortLanesWide, kBlock)), dim3(kBlock), 0, st,
                       hostCount(ev, NS),
                       ev->slotHitStart.p, ev->sortedHit.p);
    ev->hitsReady = true;
    return 0;
}

// Mini-doublets: the default algorithm over every (lower hit, upper hit) pair
// (createMiniDoublets), or with pseudo = true the pseudo mini-doublets
// (createPseudoMiniDoubletsFromAnchorModule); stored alike.
int runMD(sdl_event* ev, bool pseudo) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const int NS = ev->E * G->NM;
    if (int r = ev->mdCap.ensure(NS)) return r;
    if (int r = ev->mdCount.ensure(NS)) return r;
    if (int r = ev->mdCapOff.ensure(NS + 1)) return r;
    if (int r = ev->mdStart.ensure(NS + 1)) return r;
    if (pseudo)
        hipLaunchKernelGGL(k_pmd_cap, dim3(gridFor(NS, kBlock)), dim3(kBlock), 0, st, NS, G->NM, G->dMods,
                           ev->slotHitStart.p, ev->mdCap.p, ev->mdCount.p);
    else
        hipLaunchKernelGGL(k_md_cap, dim3(gridFor(NS, kBlock)), dim3(kBlock), 0, st, NS, G->NM, G->dMods,
                           ev->slotHitStart.p, ev->mdCap.p, ev->mdCount.p);
    if (int r = scanExclusive<int>(ev->mdCap.p, ev->mdCapOff.p, hostCount(ev, NS), NS, ev->scan, st, ev->tot.p,
                                   TOT_MDCAP))
        return r;
    const long long slab = sizeOf(ev, TOT_MDCAP);
    if (slab < 0) return fail(SDL_ERR_HIP, "readback failed");
    if (int r = ev->mdSlabHits.ensure(slab)) return r;
    if (int r = ev->mdSlabF.ensure(slab)) return r;
    if (int r = ev->mdFlag.ensure(std::max(slab, 1LL))) return r;
    if (int r = ev->mdSlabSlot.ensure(std::max(slab, 1LL))) return r;
    if (int r = ev->mdPos.ensure(std::max(slab, 1LL) + 1)) return r;
    HitsDev H = hitsView(ev);
    timerBegin(ev, T_MD);
    if (slab > 0 && pseudo)
        hipLaunchKernelGGL(k_pmd, dim3(gridFor(slab, kBlock)), dim3(kBlock), 0, st, devCount(ev, TOT_MDCAP), NS, G->NM,
                           G->dMods, ev->slotHitStart.p, ev->sortedHit.p, ev->mdCapOff.p, ev->mdSlabHits.p,
                           ev->mdSlabF.p, ev->mdSlabSlot.p, ev->mdFlag.p, ev->mdCount.p,
                           guardOf(ev, TOT_MDCAP, slab, SDL_STAGE_MD));
    if (slab > 0 && !pseudo)
        hipLaunchKernelGGL(k_md, dim3(gridFor(slab, ev->tune.v[TUNE_MD_BLOCK])), dim3(ev->tune.v[TUNE_MD_BLOCK]), 0, st,
                           devCount(ev, TOT_MDCAP), NS, G->NM,
                           G->dMods, H, ev->slotHitStart.p, ev->sortedHit.p, ev->mdCapOff.p, ev->mdSlabHits.p,
                           ev->mdSlabF.p, ev->mdSlabSlot.p, ev->mdFlag.p, ev->mdCount.p, ev->curAlgo == 1 ? 1 : 0,
                           guardOf(ev, TOT_MDCAP, slab, SDL_STAGE_MD));
    if (int r = scanExclusive<unsigned char>(ev->mdFlag.p, ev->mdPos.p, devCount(ev, TOT_MDCAP), slab, ev->scan, st,
                                             ev->tot.p, -1))
        return r;
    timerEnd(ev, T_MD);
    if (int r = scanExclusive<int>(ev->mdCount.p, ev->mdStart.p, hostCount(ev, NS), NS, ev->scan, st, ev->tot.p, TOT_MD))
        return r;
    const long long nMD = sizeOf(ev, TOT_MD);
    if (nMD < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nMD = nMD;
    const long long n = std::max(nMD, 1LL);
    if (n > kMaxObjects32) return fail(SDL_ERR_ARG, "more than 134,217,728 mini-doublets in one call: split the batch");
    for (DBuf<int>* b : {&ev->mdLo, &ev->mdUp, &ev->mdAnchor, &ev->mdSlot})
        if (int r = b->ensure(n)) return r;
    for (DBuf<float4>* b : {&ev->mdA, &ev->mdB, &ev->mdP})
        if (int r = b->ensure(n)) return r;
    if (int r = ev->mdC.ensure(n)) return r;
    if (int r = ev->mdQ.ensure(n)) return r;
    if (slab > 0)
        hipLaunchKernelGGL(k_md_repack, dim3(gridFor(slab, kBlock)), dim3(kBlock), 0, st, devCount(ev, TOT_MDCAP),
                           G->dMods, G->NM, H, ev->mdFlag.p, ev->mdPos.p, ev->mdSlabHits.p, ev->mdSlabF.p,
                           ev->mdSlabSlot.p, mdView(ev), guardOf(ev, TOT_MD, nMD, SDL_STAGE_MD));
    return 0;
}

int fillLayerList(sdl_event* ev, const long long* start, DBuf<long long>& off, DBuf<int>& list, long long nObj);

// Segments: module-map candidates (createSegmentsWithModuleMap), or with
// layers = true the layer-pair candidates (createSegments); the objects are
// stored alike, by inner MD.
int runSG(sdl_event* ev, bool layers) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const int NS = ev->E * G->NM;
    const long long nMD = ev->nMD;
    const long long nm = std::max(nMD, 1LL);
    if (layers)
        if (int r = fillLayerList(ev, ev->mdStart.p, ev->layerMdOff, ev->layerMd, nMD)) return r;
    for (DBuf<int>* b : {&ev->sgCap, &ev->sgCount, &ev->sgEndCount, &ev->sgEndFill})
        if (int r = b->ensure(nm)) return r;
    for (DBuf<long long>* b : {&ev->sgCapOff, &ev->sgStart, &ev->sgEndStart})
        if (int r = b->ensure(nm + 1)) return r;
    if (int r = ev->slotSgStart.ensure(NS + 1)) return r;
    if (int r = ev->slotMdSg.ensure(NS)) return r;
    // reverse index (SGs by outer MD): cleared by the capacity kernel, counted by
    // k_sg_extra, filled after its scan
    const Count cMD = devCount(ev, TOT_MD);
    if (nMD > 0 && !layers)
        hipLaunchKernelGGL(k_sg_cap, dim3(gridFor(nMD * kSgCapLanes, kBlock)), dim3(kBlock), 0, st, cMD, G->NM,
                           ev->mdSlot.p,
                           G->dConnStart, G->dConnList, ev->mdCount.p, ev->sgCap.p, ev->sgEndCount.p,
                           ev->sgEndFill.p);
    if (nMD > 0 && layers)
        hipLaunchKernelGGL(k_sgl_cap, dim3(gridFor(nMD, kBlock)), dim3(kBlock), 0, st, cMD, G->NM, G->dMods,
                           ev->mdSlot.p, ev->layerMdOff.p, G->dLayerModStart, ev->sgCap.p, ev->sgEndCount.p,
                           ev->sgEndFill.p);
    if (int r = scanExclusive<int>(ev->sgCap.p, ev->sgCapOff.p, cMD, nMD, ev->scan, st, ev->tot.p, TOT_SGCAP)) return r;
    const long long slab = sizeOf(ev, TOT_SGCAP);
    if (slab < 0) return fail(SDL_ERR_HIP, "readback failed");
    if (int r = ev->sgSlabOut.ensure(slab)) return r;
    if (int r = ev->sgSlabF.ensure(slab)) return r;
    timerBegin(ev, T_SG);
    if (nMD > 0 && !layers)
        hipLaunchKernelGGL(k_sg, dim3(gridFor(nMD, ev->tune.v[TUNE_SG_BLOCK] / kSgLanes)), dim3(ev->tune.v[TUNE_SG_BLOCK]), 0,
                           st, cMD, G->NM, G->dMods,
                           hitsView(ev), mdView(ev), G->dConnStart, G->dConnList, ev->mdStart.p, ev->mdCount.p,
                           ev->sgCap.p, ev->sgCapOff.p, ev->sgSlabOut.p, ev->sgSlabF.p, ev->sgCount.p,
                           ev->curAlgo == 1 ? 1 : 0, guardOf(ev, TOT_SGCAP, slab, SDL_STAGE_SG));
    if (nMD > 0 && layers)
        hipLaunchKernelGGL(k_sgl, dim3(gridFor(nMD, kWavesPerBlock)), dim3(kBlock), 0, st, cMD, G->NM, G->dMods,
                           hitsView(ev), mdView(ev), G->dConnStart, G->dConnList, ev->layerMdOff.p, G->dLayerModStart,
                           ev->layerMd.p, ev->sgCap.p, ev->sgCapOff.p, ev->sgSlabOut.p, ev->sgSlabF.p, ev->sgCount.p,
                           ev->curAlgo == 1 ? 1 : 0, guardOf(ev, TOT_SGCAP, slab, SDL_STAGE_SG));
    timerEnd(ev, T_SG);
    if (int r = scanExclusive<int>(ev->sgCount.p, ev->sgStart.p, cMD, nMD, ev->scan, st, ev->tot.p, TOT_SG)) return r;
    const long long nSG = sizeOf(ev, TOT_SG);
    if (nSG < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nSG = nSG;
    const long long ns = std::max(nSG, 1LL);
    if (ns > kMaxObjects32) return fail(SDL_ERR_ARG, "more than 134,217,728 segments in one call: split the batch");
    for (DBuf<int>* b : {&ev->sgIn, &ev->sgOut, &ev->sgFlags, &ev->sgKey, &ev->sgEndList, &ev->tpCount})
        if (int r = b->ensure(ns)) return r;
    if (int r = ev->sgF.ensure(ns)) return r;
    if (int r = ev->sgX.ensure(ns)) return r;
    if (int r = ev->sgE.ensure(ns)) return r;
    const Count cSG = devCount(ev, TOT_SG);
    if (nMD > 0)
        hipLaunchKernelGGL(k_sg_repack, dim3(gridFor(nMD * kRepackLanes, kBlock)), dim3(kBlock), 0, st, cMD,
                           ev->sgCapOff.p, ev->sgCount.p, ev->sgStart.p, ev->sgSlabOut.p, ev->sgSlabF.p, sgView(ev),
                           guardOf(ev, TOT_SG, nSG, SDL_STAGE_SG));
    if (nSG > 0)
        hipLaunchKernelGGL(k_sg_extra, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, G->NM, G->dMods,
                           hitsView(ev), mdView(ev), sgView(ev), ev->sgStart.p, ev->sgEndCount.p, ev->tpCount.p);
    hipLaunchKernelGGL(k_slot_range, dim3(gridFor((NS + 1LL) * kSlotLanes, kBlock)), dim3(kBlock), 0, st, hostCount(ev, NS),
                       ev->mdStart.p, ev->sgStart.p, ev->slotSgStart.p, ev->slotMdSg.p);
    if (int r = scanExclusive<int>(ev->sgEndCount.p, ev->sgEndStart.p, cMD, nMD, ev->scan, st, ev->tot.p, -1)) return r;
    if (nSG > 0) {
        hipLaunchKernelGGL(k_scatter_by, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, ev->sgOut.p,
                           ev->sgEndStart.p, ev->sgEndFill.p, ev->sgEndList.p);
        hipLaunchKernelGGL(k_sort_segments<kSortLanesWide>, dim3(gridFor(nMD * kSortLanesWide, kBlock)), dim3(kBlock), 0, st,
                           cMD,
                           ev->sgEndStart.p, ev->sgEndList.p);
    }
    return 0;
}


int runTP(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const long long nSG = ev->nSG;
    const long long ns = std::max(nSG, 1LL);
    const int NS = ev->E * G->NM;
    if (int r = ev->tpCount.ensure(ns)) return r;
    if (int r = ev->tpStart.ensure(ns + 1)) return r;
    if (int r = ev->slotTpStart.ensure(NS + 1)) return r;
    const Count cSG = devCount(ev, TOT_SG);
    const bool allComb = ev->curAlgo == 1;  // AllComb_TPAlgo: every candidate (Triplet.cc:45-48)
    const AcIn A{AC_TP, G->NM, ev->sgIn.p, ev->sgOut.p, nullptr, ev->mdSlot.p, G->dConnStart, G->dConnList,
                 ev->slotSgStart.p};
    // tpCount was written by the SG stage (k_sg_extra) for the default join
    if (allComb && nSG > 0)
        hipLaunchKernelGGL(k_ac_count, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, A, ev->tpCount.p,
                           (long long*)nullptr);
    if (int r = scanExclusive<int>(ev->tpCount.p, ev->tpStart.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TP)) return r;
    const long long nTP = sizeOf(ev, TOT_TP);
    if (nTP < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTP = nTP;
    if (int r = ev->tpIn.ensure(std::max(nTP, 1LL))) return r;
    if (int r = ev->tpOut.ensure(std::max(nTP, 1LL))) return r;
    if (int r = ev->tpKey.ensure(std::max(nTP, 1LL))) return r;
    if (nSG > 0 && !allComb)
        hipLaunchKernelGGL(k_tp_write, dim3(gridFor(nSG * kTpLanes, kBlock)), dim3(kBlock), 0, st, cSG, ev->sgOut.p,
                           ev->sgStart.p, ev->tpStart.p, ev->sgKey.p, ev->tpIn.p, ev->tpOut.p, ev->tpKey.p,
                           guardOf(ev, TOT_TP, nTP, SDL_STAGE_TP));
    if (nSG > 0 && allComb)
        hipLaunchKernelGGL(k_ac_write, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, A, ev->tpStart.p,
                           ev->tpIn.p, ev->tpOut.p, ev->sgKey.p, ev->tpKey.p, (int2*)nullptr,
                           guardOf(ev, TOT_TP, nTP, SDL_STAGE_TP));
    hipLaunchKernelGGL(k_slot_range_sg, dim3(gridFor(NS + 1, kBlock)), dim3(kBlock), 0, st, hostCount(ev, NS),
                       ev->slotSgStart.p, ev->tpStart.p, ev->slotTpStart.p);
    return 0;
}

// Host-side item count of variant v in a variant-ordered list: read back
// (exact mode) or predicted (speculative; the kernels check it).
long long variantSize(sdl_event* ev, int totK, const long long* off1, const long long* off2, long long stride, int v) {
    if (ev->spec) return sizeOf(ev, totK);
    long long p0 = readLL(off1 + v * stride, ev->st), p1 = readLL(off1 + (v + 1) * stride, ev->st);
    if (p0 < 0 || p1 < 0) return -1;
    if (off2) {
        p0 = readLL(off2 + p0, ev->st);
        p1 = readLL(off2 + p1, ev->st);
    }
    return p1 - p0;
}

// Module-map tracklets with AllComb_TLAlgo (Tracklet.cc:139-142): every
// candidate (Event.cc:627-693), also those whose subdet combination the default
// algorithm rejects outright; stored by inner SG like the default ones.
int runTLAllComb(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const long long nSG = ev->nSG;
    const long long ns = std::max(nSG, 1LL);
    const int NS = ev->E * G->NM;
    if (int r = ev->tlCap.ensure(ns)) return r;
    if (int r = ev->tlCount.ensure(ns)) return r;
    if (int r = ev->tlStart.ensure(ns + 1)) return r;
    if (int r = ev->slotTlStart.ensure(NS + 1)) return r;
    const Count cSG = devCount(ev, TOT_SG);
    const AcIn A{AC_TL, G->NM, ev->sgIn.p, ev->sgOut.p, nullptr, ev->mdSlot.p, G->dConnStart, G->dConnList,
                 ev->slotSgStart.p};
    timerBegin(ev, T_TL);
    if (nSG > 0)
        hipLaunchKernelGGL(k_ac_count, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, A, ev->tlCount.p,
                           ev->tlCap.p);
    if (int r = scanExclusive<int>(ev->tlCount.p, ev->tlStart.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TL)) return r;
    const long long nTL = sizeOf(ev, TOT_TL);
    if (nTL < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTL = nTL;
    if (int r = ev->tlIn.ensure(std::max(nTL, 1LL))) return r;
    if (int r = ev->tlOut.ensure(std::max(nTL, 1LL))) return r;
    if (nSG > 0)
        hipLaunchKernelGGL(k_ac_write, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, A, ev->tlStart.p,
                           ev->tlIn.p, ev->tlOut.p, (const int*)nullptr, (int*)nullptr, (int2*)nullptr,
                           guardOf(ev, TOT_TL, nTL, SDL_STAGE_TL));
    timerEnd(ev, T_TL);
    hipLaunchKernelGGL(k_slot_range_sg, dim3(gridFor(NS + 1, kBlock)), dim3(kBlock), 0, st, hostCount(ev, NS),
                       ev->slotSgStart.p, ev->tlStart.p, ev->slotTlStart.p);
    return 0;
}

int runTL(sdl_event* ev) {
    if (ev->tlAlgo == TLA_ALLCOMB) return runTLAllComb(ev);
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const long long nSG = ev->nSG;
    const long long ns = std::max(nSG, 1LL);
    const int NS = ev->E * G->NM;
    if (int r = ev->tlCap.ensure(ns)) return r;
    if (int r = ev->tlCount.ensure(ns)) return r;
    if (int r = ev->tlFill.ensure(ns)) return r;
    if (int r = ev->tlStart.ensure(ns + 1)) return r;
    if (int r = ev->slotTlStart.ensure(NS + 1)) return r;
    if (int r = ev->tlNP.ensure(kNVariantsTL * ns)) return r;
    if (int r = ev->tlPOff.ensure(kNVariantsTL * ns + 1)) return r;
    const Count cSG = devCount(ev, TOT_SG);
    timerBegin(ev, T_TL);
    // slotMdSg was written by the SG stage (k_slot_range)
    if (int r = ev->slotMdSgStart.ensure(NS + 1)) return r;
    if (int r = ev->slotMdSgList.ensure(std::max(ev->nMD, 1LL))) return r;
    if (int r = scanExclusive<int>(ev->slotMdSg.p, ev->slotMdSgStart.p, hostCount(ev, NS), NS, ev->scan, st, ev->tot.p,
                                   -1))
        return r;
    hipLaunchKernelGGL(k_slot_md_sg_list, dim3(gridFor((long long)NS * kSlotLanes, kBlock)), dim3(kBlock), 0, st,
                       hostCount(ev, NS),
                       ev->mdStart.p, ev->sgStart.p, ev->slotMdSgStart.p, ev->slotMdSgList.p);
    hipLaunchKernelGGL(k_tl_cap, dim3(gridFor(ns * kCapLanes, kBlock)), dim3(kBlock), 0, st, cSG, G->NM, G->dMods,
                       ev->sgOut.p,
                       ev->sgFlags.p, ev->mdSlot.p, G->dConnStart, G->dConnList, ev->slotSgStart.p, ev->slotMdSg.p,
                       ev->tlCap.p, ev->tlNP.p, ns, ev->tlCount.p, ev->tlFill.p);
    // pairs: one list, variant-major
    if (int r = scanExclusive<int>(ev->tlNP.p, ev->tlPOff.p, hostCount(ev, kNVariantsTL * ns), kNVariantsTL * ns,
                                   ev->scan, st, ev->tot.p, TOT_TLPAIRS))
        return r;
    const long long nPairs = sizeOf(ev, TOT_TLPAIRS);
    if (nPairs < 0) return fail(SDL_ERR_HIP, "readback failed");
    const long long npc = std::max(nPairs, 1LL);
    if (int r = ev->tlPair.ensure(npc)) return r;
    if (int r = ev->tlCarry.ensure(npc)) return r;
    if (int r = ev->tlNOut.ensure(npc)) return r;
    if (int r = ev->tlCOff.ensure(npc + 1)) return r;
    if (nSG > 0)
        hipLaunchKernelGGL(k_tl_enum, dim3(gridFor(nSG * kEnumLanes, kBlock)), dim3(kBlock), 0, st, cSG, G->NM, G->dMods,
                           ev->sgOut.p, ev->sgFlags.p, ev->mdSlot.p, G->dConnStart, G->dConnList,
                           ev->slotMdSgStart.p, ev->slotMdSgList.p, ev->tlPOff.p, ns, ev->tlPair.p,
                           guardOf(ev, TOT_TLPAIRS, nPairs, SDL_STAGE_TL));
    long long np[kNVariantsTL];
    for (int v = 0; v < kNVariantsTL; ++v) {
        np[v] = variantSize(ev, TOT_TLP0 + v, ev->tlPOff.p, nullptr, ns, v);
        if (np[v] < 0) return fail(SDL_ERR_HIP, "readback failed");
    }
    // first stage per pair; with the learned proposal, the pairs that pass are listed and scored
    const MDDev D = mdView(ev);
    const SGDev S = sgView(ev);
    // the proposal was trained on the default chain: DefaultNm1 tracklets run without it
    const bool learned = ev->proposal != nullptr && ev->tlAlgo == TLA_DEFAULT;
    const long long waveStride = (nPairs + 63) / 64 + 1;  // per-variant region of the per-wave scored counts
    ev->tlScoredStride = waveStride;
    if (learned)
        if (int r = ev->tlScored.ensure(kNVariantsTL * waveStride)) return r;
    const int b1 = ev->tune.v[TUNE_TL1_BLOCK];
    VarRanges R1{};
    long long nb1 = 0;
    for (int v = 0; v < kNVariantsTL; ++v) {
        R1.r[v] = VarRange{ev->tlPOff.p, nullptr, ns, v, np[v], ev->tot.p};
        // a speculative launch gives every variant a block, so that a variant
        // predicted empty still checks its true count against the capacity
        R1.nb[v] = std::max(int(gridFor(np[v], b1)), ev->spec ? 1 : 0);
        nb1 += R1.nb[v];
    }
    if (!learned && nb1 > 0)
        hipLaunchKernelGGL(k_tl_stage1, dim3(unsigned(nb1)), dim3(b1), 0, st, R1, D, S, ev->sgStart.p, ev->tlPair.p,
                           ev->tlCarry.p, ev->tlNOut.p);
    for (int v = 0; learned && v < kNVariantsTL; ++v) {
        if (R1.nb[v] == 0) continue;
        const VarRange R = R1.r[v];
        {
            const int sw = ev->tune.v[TUNE_SCORED_WAVES];
            const dim3 sgrid(std::max(gridFor(np[v], sw * kWave), 1u));
            const size_t lds = size_t(sw) * kScoredLdsPerWave;
            const mlp::Packed* P = ev->proposal->dPacked;
            const float cut = ev->proposalLogit;
            if (v == 0)
                hipLaunchKernelGGL(k_tl_stage1_scored<TL_BBBB>, sgrid, dim3(sw * kWave), lds, st, R, D, S, ev->sgStart.p,
                                   ev->tlPair.p, P, cut, ev->tlCarry.p, ev->tlNOut.p, ev->tlScored.p + 0 * waveStride);
            else if (v == 1)
                hipLaunchKernelGGL(k_tl_stage1_scored<TL_BBEE>, sgrid, dim3(sw * kWave), lds, st, R, D, S, ev->sgStart.p,
                                   ev->tlPair.p, P, cut, ev->tlCarry.p, ev->tlNOut.p, ev->tlScored.p + 1 * waveStride);
            else
                hipLaunchKernelGGL(k_tl_stage1_scored<TL_EEEE>, sgrid, dim3(sw * kWave), lds, st, R, D, S, ev->sgStart.p,
                                   ev->tlPair.p, P, cut, ev->tlCarry.p, ev->tlNOut.p, ev->tlScored.p + 2 * waveStride);
        }
    }
    // candidates of the surviving pairs: one list, variant-major
    const Count cPairs = devCount(ev, TOT_TLPAIRS);
    if (int r = scanExclusive<int>(ev->tlNOut.p, ev->tlCOff.p, cPairs, nPairs, ev->scan, st, ev->tot.p, TOT_TLCANDS))
        return r;
    const long long nCands = sizeOf(ev, TOT_TLCANDS);
    if (nCands < 0) return fail(SDL_ERR_HIP, "readback failed");
    if (int r = ev->tlCand.ensure(std::max(nCands, 1LL))) return r;
    if (int r = ev->tlFlag.ensure((std::max(nCands, 1LL) + 3) / 4 * 4)) return r;
    if (nPairs > 0)
        hipLaunchKernelGGL(k_tl_expand, dim3(gridFor(nPairs, ev->tune.v[TUNE_TL_EXPAND_BLOCK])),
                           dim3(ev->tune.v[TUNE_TL_EXPAND_BLOCK]), 0, st, cPairs, ev->sgStart.p,
                           ev->tlPair.p, ev->tlNOut.p, ev->tlCOff.p, ev->tlCand.p,
                           guardOf(ev, TOT_TLCANDS, nCands, SDL_STAGE_TL));
    long long nc[kNVariantsTL];
    for (int v = 0; v < kNVariantsTL; ++v) {
        nc[v] = variantSize(ev, TOT_TLC0 + v, ev->tlPOff.p, ev->tlCOff.p, ns, v);
        if (nc[v] < 0) return fail(SDL_ERR_HIP, "readback failed");
    }
    // second stage per candidate (tlCount and tlFill were cleared by k_tl_cap)
    const HitsDev H = hitsView(ev);
    const int b2 = ev->tune.v[TUNE_TL2_BLOCK];
    VarRanges R2{};
    long long nb2 = 0;
    for (int v = 0; v < kNVariantsTL; ++v) {
        R2.r[v] = VarRange{ev->tlPOff.p, ev->tlCOff.p, ns, v, nc[v], ev->tot.p};
        R2.nb[v] = std::max(int(gridFor(nc[v], b2)), ev->spec ? 1 : 0);
        nb2 += R2.nb[v];
    }
    if (nb2 > 0)
        hipLaunchKernelGGL(k_tl_stage2, dim3(unsigned(nb2)), dim3(b2), 0, st, R2, H, D, S, ev->tlCand.p, ev->tlPair.p,
                           ev->tlCarry.p, ev->tlFlag.p, ev->tlCount.p, ev->tlNm1);
    if (int r = scanExclusive<int>(ev->tlCount.p, ev->tlStart.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TL)) return r;
    const long long nTL = sizeOf(ev, TOT_TL);
    if (nTL < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTL = nTL;
    if (int r = ev->tlIn.ensure(std::max(nTL, 1LL))) return r;
    if (int r = ev->tlOut.ensure(std::max(nTL, 1LL))) return r;
    if (nCands > 0)
        hipLaunchKernelGGL(k_tl_scatter, dim3(gridFor((nCands + 3) / 4, kBlock)), dim3(kBlock), 0, st,
                           devCount(ev, TOT_TLCANDS), ev->tlCand.p, ev->tlPair.p, ev->tlFlag.p, ev->tlStart.p,
                           ev->tlFill.p, ev->tlIn.p, ev->tlOut.p, guardOf(ev, TOT_TL, nTL, SDL_STAGE_TL));
    if (nSG > 0)
        hipLaunchKernelGGL(k_sort_segments<kSortLanesTL>, dim3(gridFor(nSG * kSortLanesTL, kBlock)), dim3(kBlock), 0, st,
                           cSG,
                           ev->tlStart.p, ev->tlOut.p);
    timerEnd(ev, T_TL);
    hipLaunchKernelGGL(k_slot_range_sg, dim3(gridFor(NS + 1, kBlock)), dim3(kBlock), 0, st, hostCount(ev, NS),
                       ev->slotSgStart.p, ev->tlStart.p, ev->slotTlStart.p, ev->tlPOff.p, ev->tlCOff.p, ns, ev->tot.p);
    return 0;
}

int runTCP(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const long long nTL = ev->nTL, nSG = ev->nSG, nTP = ev->nTP;
    const long long nl = std::max(nTL, 1LL);
    if (int r = ev->tpeCount.ensure(std::max(nSG, 1LL))) return r;
    if (int r = ev->tpeOff.ensure(std::max(nSG, 1LL) + 1)) return r;
    if (int r = ev->tpEndList.ensure(std::max(nTP, 1LL))) return r;
    if (int r = ev->tcCount.ensure(nl)) return r;
    if (int r = ev->tcpInfo.ensure(nl)) return r;
    if (int r = ev->tcpStart.ensure(nl + 1)) return r;
    const Count cSG = devCount(ev, TOT_SG), cTP = devCount(ev, TOT_TP), cTL = devCount(ev, TOT_TL);
    if (nSG > 0)
        hipLaunchKernelGGL(k_tpe_count, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, ev->sgEndList.p,
                           ev->sgEndStart.p, ev->sgIn.p, ev->tpeCount.p);
    if (int r = scanExclusive<int>(ev->tpeCount.p, ev->tpeOff.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TPE)) return r;
    if (nTP > 0 && nSG > 0)
        hipLaunchKernelGGL(k_tpe_write, dim3(gridFor(nSG * kTpeLanes, kBlock)), dim3(kBlock), 0, st, cTP, cSG, ev->tpeOff.p,
                           ev->sgEndList.p, ev->sgEndStart.p, ev->sgIn.p, ev->sgStart.p, ev->tpStart.p, ev->tpEndList.p,
                           guardOf(ev, TOT_TPE, nTP, SDL_STAGE_TCP));
    const unsigned char* tlDefault = nullptr;
    if (ev->tlAlgo != TLA_DEFAULT) {
        if (int r = ev->tlDefault.ensure(nl)) return r;
        tlDefault = ev->tlDefault.p;
        if (nTL > 0)
            hipLaunchKernelGGL(k_tl_default, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, G->NM, G->dMods,
                               hitsView(ev), mdView(ev), sgView(ev), ev->tlIn.p, ev->tlOut.p, ev->tlDefault.p);
    }
    if (nTL > 0)
        hipLaunchKernelGGL(k_tcp_count, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, ev->tlIn.p,
                           ev->tlOut.p, ev->sgIn.p, ev->sgOut.p, ev->sgStart.p, ev->sgEndStart.p, ev->tpeOff.p,
                           ev->tpStart.p, tlDefault, ev->tcpInfo.p, ev->tcCount.p);
    if (int r = scanExclusive<int>(ev->tcCount.p, ev->tcpStart.p, cTL, nTL, ev->scan, st, ev->tot.p, TOT_TCP)) return r;
    const long long nTCP = sizeOf(ev, TOT_TCP);
    if (nTCP < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTCP = nTCP;
    if (int r = ev->tcp.ensure(std::max(nTCP, 1LL))) return r;
    if (nTCP > 0)
        hipLaunchKernelGGL(k_tcp_write, dim3(std::min(gridFor(nTL, kBlock / kTcpLanes), 65536u)), dim3(kBlock), ev->E * kNBins <= kMaxLdsBins ? size_t(ev->E) * kNBins * 8 : 0, st, cTL,
                           ev->tcpStart.p, ev->tcpInfo.p, ev->tpEndList.p, ev->tpKey.p, ev->E,
                           ev->tcp.p, ev->counters.p, guardOf(ev, TOT_TCP, nTCP, SDL_STAGE_TCP));
    return 0;
}

int runTCL(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const long long nTL = ev->nTL;
    const long long nl = std::max(nTL, 1LL);
    if (int r = ev->tcCount.ensure(nl)) return r;
    if (int r = ev->tclStart.ensure(nl + 1)) return r;
    const Count cTL = devCount(ev, TOT_TL);
    // AllComb_TCAlgo (TrackCandidate.cc:138-141): every candidate
    const bool allComb = ev->curAlgo == 1;
    const AcIn A{AC_TCL, G->NM, ev->sgIn.p, ev->sgOut.p, ev->tlIn.p, ev->mdSlot.p, G->dConnStart, G->dConnList,
                 ev->slotTlStart.p};
    if (nTL > 0 && !allComb)
        hipLaunchKernelGGL(k_tcl_count, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, ev->tlOut.p,
                           ev->tlStart.p, ev->tcCount.p);
    if (nTL > 0 && allComb)
        hipLaunchKernelGGL(k_ac_count, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, A, ev->tcCount.p,
                           (long long*)nullptr);
    if (int r = scanExclusive<int>(ev->tcCount.p, ev->tclStart.p, cTL, nTL, ev->scan, st, ev->tot.p, TOT_TCL)) return r;
    const long long nTCL = sizeOf(ev, TOT_TCL);
    if (nTCL < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTCL = nTCL;
    if (int r = ev->tcl.ensure(std::max(nTCL, 1LL))) return r;
    if (nTL > 0 && !allComb)
        hipLaunchKernelGGL(k_tcl_write, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, ev->tlOut.p,
                           ev->tlStart.p, ev->tcCount.p, ev->tclStart.p, ev->tcl.p,
                           guardOf(ev, TOT_TCL, nTCL, SDL_STAGE_TCL));
    if (nTL > 0 && allComb)
        hipLaunchKernelGGL(k_ac_write, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, A, ev->tclStart.p,
                           (int*)nullptr, (int*)nullptr, (const int*)nullptr, (int*)nullptr, ev->tcl.p,
                           guardOf(ev, TOT_TCL, nTCL, SDL_STAGE_TCL));
    return 0;
}

// ------------------------------------------------------- layer-pair stages
// The objects of every layer of every event, from a per-slot range array.
int fillLayerList(sdl_event* ev, const long long* start, DBuf<long long>& off, DBuf<int>& list, long long nObj) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const long long NS = (long long)ev->E * G->NM;
    if (int r = ev->layerCnt.ensure(NS)) return r;
    if (int r = off.ensure(NS + 1)) return r;
    if (int r = list.ensure(std::max(nObj, 1LL))) return r;
    hipLaunchKernelGGL(k_layer_count, dim3(gridFor(NS, kBlock)), dim3(kBlock), 0, st, ev->E, G->NM, G->dLayerMods,
                       start, ev->layerCnt.p);
    if (int r = scanExclusive<int>(ev->layerCnt.p, off.p, hostCount(ev, NS), NS, ev->scan, st, ev->tot.p, -1)) return r;
    hipLaunchKernelGGL(k_layer_fill, dim3(gridFor(NS, kBlock)), dim3(kBlock), 0, st, ev->E, G->NM, G->dLayerMods,
                       start, off.p, list.p);
    return 0;
}

int runTLLayers(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const int NS = ev->E * G->NM;
    const long long nSG = ev->nSG;
    const long long ns = std::max(nSG, 1LL);
    if (int r = fillLayerList(ev, ev->slotSgStart.p, ev->layerSgOff, ev->layerSg, nSG)) return r;
    if (int r = ev->tlCap.ensure(ns)) return r;
    if (int r = ev->tlCapInt.ensure(ns)) return r;
    if (int r = ev->tlCapOff.ensure(ns + 1)) return r;
    if (int r = ev->tlCount.ensure(ns)) return r;
    if (int r = ev->tlStart.ensure(ns + 1)) return r;
    if (int r = ev->slotTlStart.ensure(NS + 1)) return r;
    const Count cSG = devCount(ev, TOT_SG);
    if (nSG > 0)
        hipLaunchKernelGGL(k_tll_cap, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, G->NM, G->dMods,
                           ev->sgIn.p, ev->mdSlot.p, ev->layerSgOff.p, G->dLayerModStart, ev->tlCap.p, ev->tlCapInt.p);
    if (int r = scanExclusive<int>(ev->tlCapInt.p, ev->tlCapOff.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TLCANDS))
        return r;
    const long long slab = sizeOf(ev, TOT_TLCANDS);
    if (slab < 0) return fail(SDL_ERR_HIP, "readback failed");
    if (int r = ev->tllSlab.ensure(std::max(slab, 1LL))) return r;
    if (nSG > 0)
        hipLaunchKernelGGL(k_tll, dim3(gridFor(nSG, kWavesPerBlock)), dim3(kBlock), 0, st, cSG, G->NM, G->dMods,
                           hitsView(ev), mdView(ev), sgView(ev), ev->layerSgOff.p, G->dLayerModStart, ev->layerSg.p,
                           ev->tlCapInt.p, ev->tlCapOff.p, ev->tllSlab.p, ev->tlCount.p, ev->tlAlgo,
                           guardOf(ev, TOT_TLCANDS, slab, SDL_STAGE_TL));
    if (int r = scanExclusive<int>(ev->tlCount.p, ev->tlStart.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TL)) return r;
    const long long nTL = sizeOf(ev, TOT_TL);
    if (nTL < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTL = nTL;
    if (int r = ev->tlIn.ensure(std::max(nTL, 1LL))) return r;
    if (int r = ev->tlOut.ensure(std::max(nTL, 1LL))) return r;
    if (nSG > 0)
        hipLaunchKernelGGL(k_tll_repack, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, ev->tlCapOff.p,
                           ev->tlStart.p, ev->tllSlab.p, ev->tlIn.p, ev->tlOut.p, guardOf(ev, TOT_TL, nTL, SDL_STAGE_TL));
    hipLaunchKernelGGL(k_slot_range_sg, dim3(gridFor(NS + 1, kBlock)), dim3(kBlock), 0, st, hostCount(ev, NS),
                       ev->slotSgStart.p, ev->tlStart.p, ev->slotTlStart.p);
    return 0;
}

int runTLNavigation(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const int NS = ev->E * G->NM;
    const long long nSG = ev->nSG;
    const long long ns = std::max(nSG, 1LL);
    if (int r = ev->tlCap.ensure(ns)) return r;
    if (int r = ev->tlCapInt.ensure(ns)) return r;
    if (int r = ev->tlCapOff.ensure(ns + 1)) return r;
    if (int r = ev->tlCount.ensure(ns)) return r;
    if (int r = ev->tlStart.ensure(ns + 1)) return r;
    if (int r = ev->slotTlStart.ensure(NS + 1)) return r;
    const Count cSG = devCount(ev, TOT_SG);
    if (nSG > 0)
        hipLaunchKernelGGL(k_tln_cap, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, ev->sgOut.p, ev->sgStart.p,
                           ev->tlCap.p, ev->tlCapInt.p);
    if (int r = scanExclusive<int>(ev->tlCapInt.p, ev->tlCapOff.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TLCANDS))
        return r;
    const long long slab = sizeOf(ev, TOT_TLCANDS);
    if (slab < 0) return fail(SDL_ERR_HIP, "readback failed");
    if (int r = ev->tllSlab.ensure(std::max(slab, 1LL))) return r;
    if (nSG > 0)
        hipLaunchKernelGGL(k_tln, dim3(gridFor(nSG, kWavesPerBlock)), dim3(kBlock), 0, st, cSG, G->NM, G->dMods,
                           hitsView(ev), mdView(ev), sgView(ev), ev->sgStart.p, ev->tlCapInt.p, ev->tlCapOff.p,
                           ev->tllSlab.p, ev->tlCount.p, ev->tlAlgo, guardOf(ev, TOT_TLCANDS, slab, SDL_STAGE_TL));
    if (int r = scanExclusive<int>(ev->tlCount.p, ev->tlStart.p, cSG, nSG, ev->scan, st, ev->tot.p, TOT_TL)) return r;
    const long long nTL = sizeOf(ev, TOT_TL);
    if (nTL < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTL = nTL;
    if (int r = ev->tlIn.ensure(std::max(nTL, 1LL))) return r;
    if (int r = ev->tlOut.ensure(std::max(nTL, 1LL))) return r;
    if (nSG > 0)
        hipLaunchKernelGGL(k_tll_repack, dim3(gridFor(nSG, kBlock)), dim3(kBlock), 0, st, cSG, ev->tlCapOff.p,
                           ev->tlStart.p, ev->tllSlab.p, ev->tlIn.p, ev->tlOut.p, guardOf(ev, TOT_TL, nTL, SDL_STAGE_TL));
    hipLaunchKernelGGL(k_slot_range_sg, dim3(gridFor(NS + 1, kBlock)), dim3(kBlock), 0, st, hostCount(ev, NS),
                       ev->slotSgStart.p, ev->tlStart.p, ev->slotTlStart.p);
    return 0;
}

int runTCLayers(sdl_event* ev) {
    sdl_geom* G = ev->g;
    hipStream_t st = ev->st;
    const long long nTL = ev->nTL;
    const long long nl = std::max(nTL, 1LL);
    if (int r = ev->tcCount.ensure(nl)) return r;
    if (int r = ev->tcCand.ensure(nl)) return r;
    if (int r = ev->tclStart.ensure(nl + 1)) return r;
    if (int r = ev->tcpStart.ensure(nl + 1)) return r;
    if (int r = ev->tlPerLayer.ensure((long long)ev->E * kNBins)) return r;
    const bool allComb = ev->curAlgo == 1;  // AllComb_TCAlgo (TrackCandidate.cc:138-141)
    // no TCs from triplets in this form
    zeroMany(st, {span(ev->tcpStart.p, nl + 1), span(ev->tot.p + TOT_TCP, 1),
                  span(ev->tlPerLayer.p, (long long)ev->E * kNBins)});
    ev->nTCP = 0;
    const Count cTL = devCount(ev, TOT_TL);
    if (nTL > 0) {
        hipLaunchKernelGGL(k_tl_per_layer, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, G->NM, G->dMods,
                           ev->tlIn.p, ev->sgIn.p, ev->mdSlot.p, ev->tlPerLayer.p);
        hipLaunchKernelGGL(k_tcll_count, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, G->NM, G->dMods,
                           ev->tlIn.p, ev->tlOut.p, ev->sgIn.p, ev->mdSlot.p, ev->tlStart.p, ev->tlPerLayer.p,
                           allComb ? 1 : 0, ev->tcCount.p, ev->tcCand.p);
    }
    if (int r = scanExclusive<int>(ev->tcCount.p, ev->tclStart.p, cTL, nTL, ev->scan, st, ev->tot.p, TOT_TCL)) return r;
    const long long nTCL = sizeOf(ev, TOT_TCL);
    if (nTCL < 0) return fail(SDL_ERR_HIP, "readback failed");
    ev->nTCL = nTCL;
    if (int r = ev->tcl.ensure(std::max(nTCL, 1LL))) return r;
    if (nTL > 0 && !allComb)
        hipLaunchKernelGGL(k_tcl_write, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, ev->tlOut.p,
                           ev->tlStart.p, ev->tcCount.p, ev->tclStart.p, ev->tcl.p,
                           guardOf(ev, TOT_TCL, nTCL, SDL_STAGE_TCL));
    if (nTL > 0 && allComb) {
        if (int r = fillLayerList(ev, ev->slotTlStart.p, ev->layerTlOff, ev->layerTl, nTL)) return r;
        hipLaunchKernelGGL(k_tcll_ac_write, dim3(gridFor(nTL, kBlock)), dim3(kBlock), 0, st, cTL, G->NM, G->dMods,
                           ev->tlIn.p, ev->sgIn.p, ev->mdSlot.p, ev->layerTlOff.p, G->dLayerModStart, ev->layerTl.p,
                           ev->tclStart.p, ev->tcl.p, guardOf(ev, TOT_TCL, nTCL, SDL_STAGE_TCL));
    }
    return 0;
}

int runStage(sdl_event* ev, int stage) {
    switch (stage) {
        case SDL_STAGE_MD: return runMD(ev, false);
        case SDL_STAGE_MD_PSEUDO: return runMD(ev, true);
        case SDL_STAGE_SG: return runSG(ev, false);
        case SDL_STAGE_TP: return runTP(ev);
        case SDL_STAGE_TL: return runTL(ev);
        case SDL_STAGE_TCP: return runTCP(ev);
        case SDL_STAGE_TCL: return runTCL(ev);
        case SDL_STAGE_SG_LAYERS: return runSG(ev, true);
        case SDL_STAGE_TL_LAYERS: return runTLLayers(ev);
        case SDL_STAGE_TC_LAYERS: return runTCLayers(ev);
        case SDL_STAGE_TL_NAV: return runTLNavigation(ev);
        default: return fail(SDL_ERR_ARG, "bad stage");
    }
}

// Per-layer counters of stages s0..s1 (one pass, after the stages ran).
int launchCounters(sdl_event* ev, int s0, int s1) {
    sdl_geom* G = ev->g;
    CounterIn C;
    C.NS = ev->E * G->NM;
    C.NM = G->NM;
    C.E = ev->E;
    C.stages = 0;
    C.tcCand = nullptr;
    for (int s = s0; s <= s1; ++s) {
        if (s == SDL_STAGE_SG_LAYERS) continue;  // createSegmentsFromTwoLayers counts nothing (Event.cc:793-821)
        if (s == SDL_STAGE_MD_PSEUDO) {  // counted per lower module (Event.cc:354-364)
            C.stages |= 1 << SDL_STAGE_MD;
            continue;
        }
        if (s == SDL_STAGE_TL_LAYERS || s == SDL_STAGE_TL_NAV) {  // Event.cc:778-786, 720-734
            C.stages |= 1 << SDL_STAGE_TL;
        } else if (s == SDL_STAGE_TC_LAYERS) {    // counted on the inner layer (Event.cc:859-867)
            C.stages |= 1 << SDL_STAGE_TCL;
            C.tcCand = ev->tcCand.p;
        } else {
            C.stages |= 1 << s;
        }
    }
    C.mods = G->dMods;
    C.connStart = G->dConnStart;
    C.connList = G->dConnList;
    C.mdCap = ev->mdCap.p; C.mdCount = ev->mdCount.p;
    C.mdSlot = ev->mdSlot.p; C.sgCap = ev->sgCap.p; C.sgCount = ev->sgCount.p;
    C.sgIn = ev->sgIn.p; C.sgOut = ev->sgOut.p; C.tpCount = ev->tpCount.p; C.tlCount = ev->tlCount.p;
    C.tlCap = ev->tlCap.p;
    C.slotSgStart = ev->slotSgStart.p; C.slotTpStart = ev->slotTpStart.p; C.slotTlStart = ev->slotTlStart.p;
    C.tpIn = ev->tpIn.p; C.tpOut = ev->tpOut.p;
    C.tlIn = ev->tlIn.p; C.tclCount = ev->tcCount.p;
    if (int r = ev->connSum.ensure(3LL * C.NS)) return r;
    C.connSum = ev->connSum.p;
    C.sgKey = ev->sgKey.p;
    C.tpKey = ev->tpKey.p;
    if (C.stages & (4 | 16 | 32))
        hipLaunchKernelGGL(k_conn_sums, dim3(gridFor((long long)C.NS * kConnLanes, kBlock)), dim3(kBlock), 0, ev->st,
                           C, ev->tot.p);
    // one thread per slot and object (host counts bound the device totals; the
    // kernel's grid-stride loop covers any excess)
    long long n = 0;
    if (C.stages & 1) n += C.NS;
    if (C.stages & 2) n += ev->nMD;
    if (C.stages & (4 | 8)) n += ev->nSG;
    if (C.stages & 16) n += ev->nTP;
    if (C.stages & 32) n += ev->nTL;
    // with LDS bins, a bounded grid: every block flushes its bins with global atomics
    const int nb = C.E * kNCounters * kNBins;
    const bool lds = nb <= kMaxLdsBins;
    const long long cap = lds ? 4LL * G->nCU : 65536;
    const unsigned grid = unsigned(std::min<long long>(std::max<long long>(gridFor(n, kBlock), 1), cap));
    hipLaunchKernelGGL(k_counters, dim3(grid), dim3(kBlock), lds ? size_t(nb) * 8 : 0, ev->st, C, ev->tot.p,
                       ev->counters.p);
    return 0;
}

// Enqueues hit preparation (if pending), stages s0..s1 and their counters.
int enqueueStages(sdl_event* ev, int s0, int s1) {
    if (!ev->hitsReady)
        if (int r = runPrep(ev)) return r;
    for (int s = s0; s <= s1; ++s)
        if (int r = runStage(ev, s)) return r;
    return launchCounters(ev, s0, s1);
}

bool sameGraphKey(const sdl_event* ev) {
    const auto& k = ev->graphKey;
    return k.E == ev->E && k.px == ev->px && k.py == ev->py && k.pz == ev->pz && k.pidx == ev->pidx &&
           k.pdet == ev->pdet && k.allocEpoch == g_allocEpoch && k.headroomFrom == ev->headroomFrom &&
           k.headroomFactor == ev->headroomFactor && k.proposal == ev->proposal &&
           k.proposalLogit == ev->proposalLogit && k.tune == ev->tune;
}

void dropGraph(sdl_event* ev) {
    if (ev->graphExec) (void)hipGraphExecDestroy(ev->graphExec);
    if (ev->graph) (void)hipGraphDestroy(ev->graph);
    ev->graphExec = nullptr;
    ev->graph = nullptr;
}

// Speculative whole chain as a HIP graph: replayed while the batch shape, the
// input pointers and every buffer are unchanged, captured again otherwise.
// Returns 1 if the chain was launched, 0 if it must run uncaptured.
int launchChainGraph(sdl_event* ev) {
    hipStream_t st = ev->st;
    if (ev->graphExec && sameGraphKey(ev)) {
        HIP_OK(hipGraphLaunch(ev->graphExec, st));
        ev->hitsReady = true;
        ++ev->stats[1];
        return 1;
    }
    dropGraph(ev);
    HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    g_capturing = true;
    g_captureBlocked = false;
    const int r = enqueueStages(ev, SDL_STAGE_MD, SDL_STAGE_TCL);
    g_capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    ev->hitsReady = false;
    if (r != 0 || e != hipSuccess || !graph || g_captureBlocked) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return 0;
    }
    if (hipGraphInstantiate(&ev->graphExec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        ev->graphExec = nullptr;
        (void)hipGetLastError();
        return 0;
    }
    ev->graph = graph;
    auto& k = ev->graphKey;
    k.E = ev->E; k.px = ev->px; k.py = ev->py; k.pz = ev->pz; k.pidx = ev->pidx; k.pdet = ev->pdet;
    k.allocEpoch = g_allocEpoch; k.headroomFrom = ev->headroomFrom; k.headroomFactor = ev->headroomFactor;
    k.proposal = ev->proposal; k.proposalLogit = ev->proposalLogit; k.tune = ev->tune;
    HIP_OK(hipGraphLaunch(ev->graphExec, st));
    ev->hitsReady = true;
    ++ev->stats[2];
    return 1;
}

// Runs stages s0..s1 of the batch with one host round trip at the end when the
// last call's totals predict the buffer sizes (as a replayed graph for a whole
// chain); otherwise (first call, or a prediction that fell short) each stage
// reads back the totals it sizes from.
int runStages(sdl_event* ev, int s0, int s1) {
    hipStream_t st = ev->st;
    if (int r = ev->tot.ensure(TOT_N)) return r;
    if (!ev->hitsReady) {
        if (int r = uploadOffsets(ev)) return r;
    } else {
        zeroAsync(ev->tot.p + TOT_ABORT, 1, st);
    }
    // the layer-pair stages and calls with a non-default algorithm are sized exactly
    ev->spec = ev->haveHistory && s0 < SDL_STAGE_SG_LAYERS && !ev->forceExact;
    bool launched = false;
    if (ev->spec && ev->useGraphs && !ev->timing && !ev->hitsReady && s0 == SDL_STAGE_MD && s1 == SDL_STAGE_TCL) {
        const int r = launchChainGraph(ev);
        if (r < 0) return r;
        launched = r == 1;
    }
    if (!launched)
        if (int r = enqueueStages(ev, s0, s1)) return r;
    ++ev->stats[0];
    long long t[TOT_N];
    HIP_OK(hipMemcpyAsync(t, ev->tot.p, sizeof(t), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (t[TOT_UNKNOWN] > 0)
        return fail(SDL_ERR_UNKNOWN_DETID, std::to_string(t[TOT_UNKNOWN]) + " hits have a detId unknown to the geometry");
    if (t[TOT_ABORT] != 0) {
        if (!ev->spec) return fail(SDL_ERR_CAPACITY, "device capacity check failed on an exactly sized run");
        // rerun from the stage that overflowed, sized exactly; a captured graph
        // holds the capacities that fell short, so the next call captures anew
        // (from the totals of this run) instead of replaying it into another rerun
        const int from = t[TOT_ABORT] > SDL_STAGE_TCL + 1 ? s0 : int(t[TOT_ABORT] - 1);
        if (launched) dropGraph(ev);
        ++ev->stats[3];
        ev->spec = false;
        zeroAsync(ev->tot.p + TOT_ABORT, 1, st);
        for (int s = std::max(from, s0); s <= s1; ++s)
            if (int r = runStage(ev, s)) return r;
        if (int r = launchCounters(ev, s0, s1)) return r;
        HIP_OK(hipMemcpyAsync(t, ev->tot.p, sizeof(t), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (t[TOT_ABORT] != 0) return fail(SDL_ERR_CAPACITY, "device capacity check failed on an exactly sized run");
    }
    ev->nMD = t[TOT_MD]; ev->nSG = t[TOT_SG]; ev->nTP = t[TOT_TP];
    ev->nTL = t[TOT_TL]; ev->nTCP = t[TOT_TCP]; ev->nTCL = t[TOT_TCL];
    for (int s = s0; s <= s1; ++s) ev->ranMask |= 1u << s;
    // remember the sizes of the stages that ran (the prediction for the next call)
    static const int stageOfTot[TOT_ABORT] = {SDL_STAGE_MD, SDL_STAGE_MD, SDL_STAGE_SG, SDL_STAGE_SG, SDL_STAGE_TP,
                                              SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TCP, SDL_STAGE_TCP, SDL_STAGE_TCL,
                                              SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL,
                                              SDL_STAGE_TL, SDL_STAGE_TL, SDL_STAGE_TL};
    for (int k = 0; k < TOT_ABORT; ++k)
        if (stageOfTot[k] >= s0 && stageOfTot[k] <= s1) ev->lastTot[k] = t[k];
    if (s1 == SDL_STAGE_TCL) {
        ev->lastNH = ev->NH;
        ev->haveHistory = true;
    }
    return 0;
}

int stageRange(sdl_event* ev, int e, int stage, long long& a, long long& n) {
    // objects of event e are contiguous: compute first/last via the slot ranges
    const int NM = ev->g->NM;
    const long long s0 = (long long)e * NM, s1 = (long long)(e + 1) * NM;
    auto rd = [&](const long long* p, long long i) { return readLL(p + i, ev->st); };
    long long lo = 0, hi = 0;
    switch (stage) {
        case SDL_STAGE_MD: lo = rd(ev->mdStart.p, s0); hi = rd(ev->mdStart.p, s1); break;
        case SDL_STAGE_SG: lo = rd(ev->slotSgStart.p, s0); hi = rd(ev->slotSgStart.p, s1); break;
        case SDL_STAGE_TP: lo = rd(ev->slotTpStart.p, s0); hi = rd(ev->slotTpStart.p, s1); break;
        case SDL_STAGE_TL: lo = rd(ev->slotTlStart.p, s0); hi = rd(ev->slotTlStart.p, s1); break;
        case SDL_STAGE_TCP:
        case SDL_STAGE_TCL: {
            const long long l0 = rd(ev->slotTlStart.p, s0), l1 = rd(ev->slotTlStart.p, s1);
            const long long* st = stage == SDL_STAGE_TCP ? ev->tcpStart.p : ev->tclStart.p;
            lo = rd(st, l0);
            hi = rd(st, l1);
            break;
        }
        default: return fail(SDL_ERR_ARG, "bad stage");
    }
    if (lo < 0 || hi < 0) return fail(SDL_ERR_HIP, "readback failed");
    a = lo;
    n = hi - lo;
    return 0;
}

int beginBatch(sdl_event* ev) {
    if (!ev->hostStaged) return 0;
    // upload host-staged hits
    ev->E = int(ev->hoff.size()) - 1;
    ev->evOff = ev->hoff;
    ev->NH = ev->evOff.back();
    const long long n = std::max(ev->NH, 1LL);
    if (int r = ev->bx.ensure(n)) return r;
    if (int r = ev->by.ensure(n)) return r;
    if (int r = ev->bz.ensure(n)) return r;
    if (int r = ev->bidx.ensure(n)) return r;
    if (int r = ev->bdet.ensure(n)) return r;
    if (ev->NH > 0) {
        HIP_OK(hipMemcpyAsync(ev->bx.p, ev->hx.data(), sizeof(float) * ev->NH, hipMemcpyHostToDevice, ev->st));
        HIP_OK(hipMemcpyAsync(ev->by.p, ev->hy.data(), sizeof(float) * ev->NH, hipMemcpyHostToDevice, ev->st));
        HIP_OK(hipMemcpyAsync(ev->bz.p, ev->hz.data(), sizeof(float) * ev->NH, hipMemcpyHostToDevice, ev->st));
        HIP_OK(hipMemcpyAsync(ev->bidx.p, ev->hidx.data(), sizeof(int) * ev->NH, hipMemcpyHostToDevice, ev->st));
        HIP_OK(hipMemcpyAsync(ev->bdet.p, ev->hdet.data(), sizeof(unsigned) * ev->NH, hipMemcpyHostToDevice, ev->st));
    }
    ev->px = ev->bx.p; ev->py = ev->by.p; ev->pz = ev->bz.p; ev->pidx = ev->bidx.p; ev->pdet = ev->bdet.p;
    return 0;
}

// float -> bf16 bits, round to nearest even (NaN kept quiet)
short bf16Host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return short((u >> 16) | 0x40u);
    const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return short(r >> 16);
}

// Parameters (layout of sdl_mlp.h) -> MFMA operand fragments in lane order.
void packMlp(const float* P, sdl::mlp::Packed& K) {
    using namespace sdl::mlp;
    for (int t = 0; t < kTiles; ++t)
        for (int l = 0; l < 64; ++l) {
            const int row = t * 16 + (l % 16), g = l / 16;  // A operand row: output neuron
            for (int i = 0; i < 8; ++i) {
                const int k = 8 * g + i;  // layer 1: K slot = input feature (24: the bias, then zeros)
                K.a1[t][l][i] = bf16Host(k < kIn ? P[kW1 + row * kIn + k] : (k == kIn ? P[kB1 + row] : 0.f));
                for (int u = 0; u < kSteps; ++u) {
                    const int h = hiddenOfSlot(u, g, i);
                    K.a2[t][u][l][i] = bf16Host(P[kW2 + row * kH + h]);
                    K.a3[t][u][l][i] = bf16Host(P[kW3 + row * kH + h]);
                }
            }
        }
    for (int k = 0; k < kH; ++k) {
        K.b2[k] = P[kB2 + k];
        K.b3[k] = P[kB3 + k];
        K.w4[k] = P[kW4 + k];
    }
    K.b4 = P[kB4];
}

int repackMlp(sdl_mlp* m) {
    std::vector<float> p(sdl::mlp::kNParams);
    HIP_OK(hipMemcpyAsync(p.data(), m->dParams, sizeof(float) * p.size(), hipMemcpyDeviceToHost, m->st));
    HIP_OK(hipStreamSynchronize(m->st));
    auto* K = new sdl::mlp::Packed();
    packMlp(p.data(), *K);
    const hipError_t e = hipMemcpyAsync(m->dPacked, K, sizeof(*K), hipMemcpyHostToDevice, m->st);
    const hipError_t e2 = hipStreamSynchronize(m->st);
    delete K;
    if (e != hipSuccess || e2 != hipSuccess) return fail(SDL_ERR_HIP, "MLP upload failed");
    return 0;
}

// ---- geometry: device upload and the binary cache
int uploadGeom(sdl_geom* G) {
    (void)hipGetDevice(&G->device);
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, G->device) == hipSuccess && cu > 0) G->nCU = cu;
    auto up = [&](void** d, const void* h, size_t bytes) -> int {
        if (hipMalloc(d, std::max<size_t>(bytes, 4)) != hipSuccess) return fail(SDL_ERR_HIP, "hipMalloc failed");
        if (bytes && hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return fail(SDL_ERR_HIP, "hipMemcpy failed");
        return 0;
    };
    std::vector<int> layerMods(G->NM), layerModStart(kNBins + 1, 0);
    for (int m = 0; m < G->NM; ++m) {
        const int b = layerBin(G->mods[m]);
        if (b < 0 || b >= kNBins) return fail(SDL_ERR_IO, "module with a layer outside barrel 1-6 / endcap 1-5");
        ++layerModStart[b + 1];
    }
    for (int b = 0; b < kNBins; ++b) layerModStart[b + 1] += layerModStart[b];
    {
        std::vector<int> fill(layerModStart.begin(), layerModStart.end() - 1);
        for (int m = 0; m < G->NM; ++m) layerMods[fill[layerBin(G->mods[m])]++] = m;
    }
    int bits = 4;
    while ((1LL << bits) < 4LL * G->NM) ++bits;
    std::vector<uint2> table(size_t(1) << bits, make_uint2(kNoDet, 0u));
    for (int m = 0; m < G->NM; ++m) {
        if (G->detIds[m] == kNoDet) return fail(SDL_ERR_IO, "module detId 0xFFFFFFFF (reserved)");
        uint32_t h = detHash(G->detIds[m], bits);
        while (table[h].x != kNoDet) h = (h + 1u) & uint32_t(table.size() - 1);
        table[h] = make_uint2(G->detIds[m], uint32_t(m));
    }
    G->hashBits = bits;
    int r = up((void**)&G->dDetTable, table.data(), sizeof(uint2) * table.size());
    if (!r) r = up((void**)&G->dMods, G->mods.data(), sizeof(ModDev) * G->mods.size());
    if (!r) r = up((void**)&G->dConnStart, G->connStart.data(), sizeof(int) * G->connStart.size());
    if (!r) r = up((void**)&G->dConnList, G->connList.data(), sizeof(int) * G->connList.size());
    if (!r) r = up((void**)&G->dLayerMods, layerMods.data(), sizeof(int) * layerMods.size());
    if (!r) r = up((void**)&G->dLayerModStart, layerModStart.data(), sizeof(int) * layerModStart.size());
    return r;
}

// Binary cache: the parsed geometry exactly as the device holds it.
//   "SDLGEOM1" | u32 version | u32 sizeof(ModDev) | i32 NM | i64 number of connections
//   | u32 detIds[NM] (sorted) | ModDev mods[NM] | i32 connStart[NM + 1] | i32 connList[]
//   | u64 FNV-1a hash of everything before it
constexpr char kGeomMagic[8] = {'S', 'D', 'L', 'G', 'E', 'O', 'M', '1'};
constexpr uint32_t kGeomVersion = 1;

uint64_t fnv1a(const std::string& b) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : b) {
        h ^= ch;
        h *= 1099511628211ull;
    }
    return h;
}

template <class T>
void putRaw(std::string& b, const T* p, size_t n) {
    b.append(reinterpret_cast<const char*>(p), sizeof(T) * n);
}

int writeGeomBinary(const sdl_geom& G, const char* path) {
    std::string b(kGeomMagic, 8);
    const uint32_t ver = kGeomVersion, msz = sizeof(ModDev);
    const int32_t nm = G.NM;
    const int64_t nc = (int64_t)G.connList.size();
    putRaw(b, &ver, 1);
    putRaw(b, &msz, 1);
    putRaw(b, &nm, 1);
    putRaw(b, &nc, 1);
    putRaw(b, G.detIds.data(), G.detIds.size());
    putRaw(b, G.mods.data(), G.mods.size());
    putRaw(b, G.connStart.data(), G.connStart.size());
    putRaw(b, G.connList.data(), G.connList.size());
    const uint64_t h = fnv1a(b);
    putRaw(b, &h, 1);
    const std::string tmp = std::string(path) + ".tmp";
    std::ofstream f(tmp, std::ios::binary);
    if (!f || !f.write(b.data(), (std::streamsize)b.size())) return fail(SDL_ERR_IO, std::string("cannot write ") + tmp);
    f.close();
    if (std::rename(tmp.c_str(), path) != 0) return fail(SDL_ERR_IO, std::string("cannot rename to ") + path);
    return 0;
}

int readGeomBinary(const char* path, sdl_geom& G) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(SDL_ERR_IO, std::string("cannot open ") + path);
    const std::string b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const size_t head = 8 + 4 + 4 + 4 + 8;
    auto bad = [&](const char* why) { return fail(SDL_ERR_IO, std::string(path) + ": not a valid geometry cache (" + why + ")"); };
    if (b.size() < head + 8 || std::memcmp(b.data(), kGeomMagic, 8) != 0) return bad("magic");
    uint32_t ver, msz;
    int32_t nm;
    int64_t nc;
    std::memcpy(&ver, b.data() + 8, 4);
    std::memcpy(&msz, b.data() + 12, 4);
    std::memcpy(&nm, b.data() + 16, 4);
    std::memcpy(&nc, b.data() + 20, 8);
    if (ver != kGeomVersion) return bad("version");
    if (msz != sizeof(ModDev)) return bad("module record size");
    if (nm < 0 || nc < 0 || nc > (int64_t(1) << 31)) return bad("sizes");
    const size_t need = head + 4 * size_t(nm) + sizeof(ModDev) * size_t(nm) + 4 * (size_t(nm) + 1) + 4 * size_t(nc) + 8;
    if (b.size() != need) return bad("length");
    uint64_t h;
    std::memcpy(&h, b.data() + need - 8, 8);
    if (fnv1a(b.substr(0, need - 8)) != h) return bad("checksum");
    size_t o = head;
    G.NM = nm;
    G.detIds.resize(nm);
    G.mods.resize(nm);
    G.connStart.resize(size_t(nm) + 1);
    G.connList.resize(size_t(nc));
    std::memcpy(G.detIds.data(), b.data() + o, 4 * size_t(nm)); o += 4 * size_t(nm);
    std::memcpy(G.mods.data(), b.data() + o, sizeof(ModDev) * size_t(nm)); o += sizeof(ModDev) * size_t(nm);
    std::memcpy(G.connStart.data(), b.data() + o, 4 * (size_t(nm) + 1)); o += 4 * (size_t(nm) + 1);
    std::memcpy(G.connList.data(), b.data() + o, 4 * size_t(nc));
    // structure: sorted detIds, a monotone CSR covering the list, module indices in range
    for (int k = 1; k < nm; ++k)
        if (G.detIds[k - 1] >= G.detIds[k]) return bad("detIds not sorted");
    if (G.connStart[0] != 0 || G.connStart[nm] != nc) return bad("connection offsets");
    for (int k = 0; k < nm; ++k)
        if (G.connStart[k] > G.connStart[k + 1]) return bad("connection offsets");
    for (int64_t k = 0; k < nc; ++k)
        if (G.connList[k] < 0 || G.connList[k] >= nm) return bad("connection target");
    return 0;
}

// The per-layer counters of all events, summed over the counter copies.
int readCounters(sdl_event* ev, std::vector<unsigned long long>& all) {
    const size_t per = size_t(ev->E) * kNCounters * kNBins;
    std::vector<unsigned long long> copies(per * kCounterCopies);
    HIP_OK(hipMemcpyAsync(copies.data(), ev->counters.p, sizeof(unsigned long long) * copies.size(),
                          hipMemcpyDeviceToHost, ev->st));
    HIP_OK(hipStreamSynchronize(ev->st));
    all.assign(per, 0ULL);
    for (int c = 0; c < kCounterCopies; ++c)
        for (size_t i = 0; i < per; ++i) all[i] += copies[size_t(c) * per + i];
    return 0;
}

}  // namespace

// ====================================================================== C-ABI
extern "C" {

const char* sdl_last_error(void) { return g_err.c_str(); }

int sdl_geom_load(const char* tiltedPath, const char* endcapPath, const char* mapPath, sdl_geom** out) {
    if (!tiltedPath || !endcapPath || !mapPath || !out) return fail(SDL_ERR_ARG, "null argument");
    auto* G = new sdl_geom();
    if (int r = loadGeomFiles(tiltedPath, endcapPath, mapPath, *G)) { delete G; return r; }
    if (int r = uploadGeom(G)) { sdl_geom_free(G); return r; }
    *out = G;
    return 0;
}

int sdl_geom_load_tables(int nT, const unsigned* tDet, const float* tDrdz, const float* tSlope, int nE,
                         const unsigned* eDet, const float* eCols, int nM, const unsigned* mDet, const int* mCount,
                         const unsigned* mList, sdl_geom** out) {
    if (!out || nT < 0 || nE < 0 || nM < 0 || (nT && (!tDet || !tDrdz || !tSlope)) || (nE && (!eDet || !eCols)) ||
        (nM && (!mDet || !mCount)))
        return fail(SDL_ERR_ARG, "bad arguments");
    GeomTables T;
    for (int i = 0; i < nT; ++i) { T.drdz[tDet[i]] = tDrdz[i]; T.slope[tDet[i]] = tSlope[i]; }
    // the eight columns after the detId (EndcapGeometry.cc:48): avgr2 yl sl yh sh c7 c8 cz
    for (int i = 0; i < nE; ++i) { T.ecSlope[eDet[i]] = eCols[8 * i + 2]; T.ecPhi[eDet[i]] = eCols[8 * i + 5]; }
    long long k = 0;
    for (int i = 0; i < nM; ++i) {
        if (mCount[i] < 0 || (mCount[i] && !mList)) return fail(SDL_ERR_ARG, "bad connection count");
        T.conn.push_back({mDet[i], std::vector<uint32_t>(mList + k, mList + k + mCount[i])});
        k += mCount[i];
    }
    auto* G = new sdl_geom();
    if (int r = buildGeom(T, *G)) { delete G; return r; }
    if (int r = uploadGeom(G)) { sdl_geom_free(G); return r; }
    *out = G;
    return 0;
}

int sdl_module_decode(unsigned int detId, unsigned int out[11]) {
    if (!out) return fail(SDL_ERR_ARG, "null argument");
    const ModuleId m = decodeDetId(detId);
    const unsigned v[11] = {m.subdet, m.side, m.layer, m.rod, m.ring, m.module, m.isLower, m.isInverted,
                            m.partner, m.type, m.layerType};
    std::copy(v, v + 11, out);
    return 0;
}

int sdl_map_build(long long nTracks, const long long* trackOffsets, const int* station, const unsigned* detId,
                  const char* mapPath, int merge) {
    if (nTracks < 0 || (nTracks && (!trackOffsets || !station || !detId)) || !mapPath)
        return fail(SDL_ERR_ARG, "bad arguments");
    // the existing file as ModuleConnectionMap::load reads it (ModuleConnectionMap.cc:18-47):
    // one list per detId, a repeated detId keeps its last line
    std::map<uint32_t, std::vector<uint32_t>> conn;
    if (merge) {
        std::ifstream fm(mapPath);
        std::string line;
        while (fm && std::getline(fm, line)) {
            std::stringstream ss(line);
            unsigned d; int n;
            if (!(ss >> d >> n)) continue;
            std::vector<uint32_t> v;
            for (int i = 0; i < n; ++i) {
                unsigned c;
                if (!(ss >> c)) break;
                v.push_back(c);
            }
            conn[d] = std::move(v);
        }
    }
    std::map<uint32_t, std::vector<uint32_t>> built;  // this call's connections
    for (long long t = 0; t < nTracks; ++t) {
        const long long a = trackOffsets[t], b = trackOffsets[t + 1];
        if (b < a) return fail(SDL_ERR_ARG, "track offsets must not decrease");
        // stations in order; every lower module of one station connects to every
        // lower module of the next (an upper sensor stands for its lower partner)
        std::vector<uint32_t> prev, cur;
        int curStation = 0;
        for (long long k = a; k <= b; ++k) {
            if (k == b || (k > a && station[k] != curStation)) {
                if (!cur.empty()) {
                    for (uint32_t p : prev)
                        for (uint32_t c : cur) built[p].push_back(c);
                    prev.swap(cur);
                    cur.clear();
                }
                if (k == b) break;
            }
            if (k > a && station[k] < curStation) return fail(SDL_ERR_ARG, "stations of a track must not decrease");
            curStation = station[k];
            const ModuleId m = decodeDetId(detId[k]);
            const uint32_t lower = m.isLower ? detId[k] : m.partner;
            if (std::find(cur.begin(), cur.end(), lower) == cur.end()) cur.push_back(lower);
        }
    }
    // merged as ModuleConnectionMap::add merges a file (ModuleConnectionMap.cc:49-85):
    // a detId with new connections gets the sorted union without duplicates,
    // every other list stays as read
    for (auto& kv : built) {
        std::vector<uint32_t>& v = conn[kv.first];
        v.insert(v.end(), kv.second.begin(), kv.second.end());
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::ofstream fo(mapPath);
    if (!fo) return fail(SDL_ERR_IO, std::string("cannot write ") + mapPath);
    for (auto& kv : conn) {
        const std::vector<uint32_t>& v = kv.second;
        fo << kv.first << ' ' << v.size();
        for (uint32_t c : v) fo << ' ' << c;
        fo << '\n';
    }
    if (!fo) return fail(SDL_ERR_IO, std::string("cannot write ") + mapPath);
    return 0;
}

int sdl_geom_convert(const char* tiltedPath, const char* endcapPath, const char* mapPath, const char* binaryPath) {
    if (!tiltedPath || !endcapPath || !mapPath || !binaryPath) return fail(SDL_ERR_ARG, "null argument");
    sdl_geom G;
    if (int r = loadGeomFiles(tiltedPath, endcapPath, mapPath, G)) return r;
    return writeGeomBinary(G, binaryPath);
}

int sdl_geom_save(const sdl_geom* g, const char* binaryPath) {
    if (!g || !binaryPath) return fail(SDL_ERR_ARG, "null argument");
    return writeGeomBinary(*g, binaryPath);
}

int sdl_geom_check_binary(const char* binaryPath, int* numModules, long long* numConnections) {
    if (!binaryPath) return fail(SDL_ERR_ARG, "null argument");
    sdl_geom G;
    if (int r = readGeomBinary(binaryPath, G)) return r;
    if (numModules) *numModules = G.NM;
    if (numConnections) *numConnections = (long long)G.connList.size();
    return 0;
}

int sdl_geom_load_binary(const char* binaryPath, sdl_geom** out) {
    if (!binaryPath || !out) return fail(SDL_ERR_ARG, "null argument");
    auto* G = new sdl_geom();
    if (int r = readGeomBinary(binaryPath, *G)) { delete G; return r; }
    if (int r = uploadGeom(G)) { sdl_geom_free(G); return r; }
    *out = G;
    return 0;
}

void sdl_geom_free(sdl_geom* g) {
    if (!g) return;
    if (g->dDetTable) (void)hipFree(g->dDetTable);
    if (g->dMods) (void)hipFree(g->dMods);
    if (g->dConnStart) (void)hipFree(g->dConnStart);
    if (g->dConnList) (void)hipFree(g->dConnList);
    if (g->dLayerMods) (void)hipFree(g->dLayerMods);
    if (g->dLayerModStart) (void)hipFree(g->dLayerModStart);
    delete g;
}

int sdl_geom_num_modules(const sdl_geom* g) { return g ? g->NM : 0; }

int sdl_event_create(sdl_geom* g, sdl_event** out) {
    if (!g || !out) return fail(SDL_ERR_ARG, "null argument");
    auto* ev = new sdl_event();
    ev->g = g;
    if (hipSetDevice(g->device) != hipSuccess || hipStreamCreateWithFlags(&ev->st, hipStreamNonBlocking) != hipSuccess) {
        delete ev;
        return fail(SDL_ERR_HIP, "stream creation failed");
    }
    *out = ev;
    return 0;
}

void sdl_event_destroy(sdl_event* ev) {
    if (!ev) return;
    (void)hipStreamSynchronize(ev->st);
    // device buffers are released by their destructors
    dropGraph(ev);
    for (auto& t : ev->tev)
        for (auto& x : t)
            if (x) (void)hipEventDestroy(x);
    (void)hipStreamDestroy(ev->st);
    delete ev;
}

int sdl_event_reset(sdl_event* ev) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev) return fail(SDL_ERR_ARG, "null event");
    ev->hx.clear(); ev->hy.clear(); ev->hz.clear(); ev->hidx.clear(); ev->hdet.clear();
    ev->hoff.assign(1, 0);
    ev->hostStaged = true;
    ev->E = 0; ev->NH = 0;
    ev->stage = -1;
    ev->hitsReady = false;
    ev->nMD = ev->nSG = ev->nTP = ev->nTL = ev->nTCP = ev->nTCL = 0;
    return 0;
}

int sdl_event_add_hits(sdl_event* ev, int n, const float* x, const float* y, const float* z, const int* idx,
                       const unsigned int* detId) {
    if (!ev || n < 0 || (n > 0 && (!x || !y || !z || !idx || !detId))) return fail(SDL_ERR_ARG, "bad arguments");
    if (!ev->hostStaged) sdl_event_reset(ev);
    ev->hx.insert(ev->hx.end(), x, x + n);
    ev->hy.insert(ev->hy.end(), y, y + n);
    ev->hz.insert(ev->hz.end(), z, z + n);
    ev->hidx.insert(ev->hidx.end(), idx, idx + n);
    ev->hdet.insert(ev->hdet.end(), detId, detId + n);
    if (ev->hoff.size() == 1) ev->hoff.push_back(0);
    ev->hoff.back() += n;
    ev->stage = -1;
    ev->hitsReady = false;
    return 0;
}

int sdl_event_next(sdl_event* ev) {
    if (!ev) return fail(SDL_ERR_ARG, "null event");
    if (ev->hoff.size() == 1) ev->hoff.push_back(0);
    ev->hoff.push_back(ev->hoff.back());
    return 0;
}

int sdl_event_set_hits_device(sdl_event* ev, int nevents, const long long* offsets, const float* dx, const float* dy,
                              const float* dz, const int* didx, const unsigned int* ddetId) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || nevents <= 0 || !offsets) return fail(SDL_ERR_ARG, "bad arguments");
    ev->hostStaged = false;
    ev->E = nevents;
    ev->evOff.assign(offsets, offsets + nevents + 1);
    ev->NH = ev->evOff.back();
    ev->px = dx; ev->py = dy; ev->pz = dz; ev->pidx = didx; ev->pdet = ddetId;
    ev->stage = -1;
    ev->hitsReady = false;
    return 0;
}

int sdl_event_run(sdl_event* ev, int stage, int algo) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev) return fail(SDL_ERR_ARG, "null event");
    if (stage < SDL_STAGE_MD || stage > SDL_STAGE_MD_PSEUDO) return fail(SDL_ERR_ARG, "bad stage");
    const bool tlStage = stage == SDL_STAGE_TL || stage == SDL_STAGE_TL_LAYERS || stage == SDL_STAGE_TL_NAV;
    // algorithms, the Algo.h values: 0 Default and 1 AllComb, for the tracklet
    // stages 1 DefaultNm1 and 2 AllComb.  TCs from triplets ignore theirs (the
    // bridges always run the default algorithms, Event.cc:874-982) and pseudo
    // mini-doublets always run AllComb (Event.cc:353).
    if (algo < 0 || algo > (tlStage ? 2 : 1)) return fail(SDL_ERR_ARG, "unknown algorithm value for this stage");
    // each stage reads the objects of the stage before it: module-map chain MD,
    // SG, TP, TL, TCP, TCL; layer-pair stages after MD (segments), after either
    // segment stage (tracklets) and after either tracklet stage (track
    // candidates).  The module-map joins (TP, TCP, TCL) assume module-map
    // inputs, so they do not follow a layer-pair stage.
    const int prev = ev->stage == SDL_STAGE_MD_PSEUDO ? SDL_STAGE_MD : ev->stage;  // both MD forms are alike
    bool ok = stage == prev + 1 && stage <= SDL_STAGE_TCL;
    if (stage == SDL_STAGE_MD_PSEUDO) ok = ev->stage == -1;
    if (stage == SDL_STAGE_SG_LAYERS) ok = prev == SDL_STAGE_MD;
    if (stage == SDL_STAGE_TL_LAYERS) ok = ev->stage == SDL_STAGE_SG || ev->stage == SDL_STAGE_SG_LAYERS;
    if (stage == SDL_STAGE_TL_NAV)
        ok = ev->stage == SDL_STAGE_SG || ev->stage == SDL_STAGE_SG_LAYERS || ev->stage == SDL_STAGE_TP;
    if (stage == SDL_STAGE_TC_LAYERS)
        ok = ev->stage == SDL_STAGE_TL || ev->stage == SDL_STAGE_TL_LAYERS || ev->stage == SDL_STAGE_TL_NAV;
    // navigation tracklets over module-map segments keep the TCL join exact
    if (stage == SDL_STAGE_TCL && ev->stage == SDL_STAGE_TL_NAV) ok = !(ev->ranMask & (1u << SDL_STAGE_SG_LAYERS));
    // TCs from tracklets (a join over the stored tracklets) may also follow the
    // module-map tracklets directly
    if (stage == SDL_STAGE_TCL && ev->stage == SDL_STAGE_TL) ok = true;
    if (!ok) return fail(SDL_ERR_STATE, "stages must run in order");
    // TCs from triplets run the default segment and tracklet algorithms on their
    // bridges (Event.cc:929-957).  The device reads those decisions from the
    // stored segments and tracklets: after default or DefaultNm1 / AllComb
    // tracklets it re-evaluates the stored tracklets, but AllComb segments or
    // triplets leave nothing it could read them from.
    if (stage == SDL_STAGE_TCP && (ev->acMask & ((1u << SDL_STAGE_SG) | (1u << SDL_STAGE_TP))))
        return fail(SDL_ERR_STATE, "TCs from triplets after AllComb segments or triplets are not supported");
    ev->curAlgo = (stage == SDL_STAGE_TCP || stage == SDL_STAGE_MD_PSEUDO) ? 0 : algo;
    if (tlStage) {
        ev->tlAlgo = algo;
        ev->tlNm1 = algo == TLA_NM1 ? 1 : 0;
    }
    if (hipSetDevice(ev->g->device) != hipSuccess) return fail(SDL_ERR_HIP, "hipSetDevice failed");
    if (!ev->hitsReady) {
        if (ev->hostStaged && ev->hoff.size() == 1) ev->hoff.push_back(0);
        if (int r = beginBatch(ev)) return r;
        if (ev->E <= 0) return fail(SDL_ERR_STATE, "no event");
        ev->ranMask = 0;
        ev->acMask = 0;
    }
    ev->forceExact = ev->curAlgo != 0 || ev->acMask != 0 || ev->tlAlgo != TLA_DEFAULT;
    const int r = runStages(ev, stage, stage);
    ev->forceExact = false;
    if (r) return r;
    if (hipGetLastError() != hipSuccess) return fail(SDL_ERR_HIP, "kernel launch failed");
    ev->stage = stage;
    if (ev->curAlgo == (tlStage ? TLA_ALLCOMB : 1)) ev->acMask |= 1u << stage;
    return 0;
}

int sdl_event_run_chain(sdl_event* ev) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev) return fail(SDL_ERR_ARG, "null event");
    ev->tlNm1 = 0;  // the chain runs the default algorithms
    ev->tlAlgo = TLA_DEFAULT;
    ev->curAlgo = 0;
    ev->acMask = 0;
    if (hipSetDevice(ev->g->device) != hipSuccess) return fail(SDL_ERR_HIP, "hipSetDevice failed");
    ev->stage = -1;
    ev->hitsReady = false;
    ev->ranMask = 0;
    if (ev->hostStaged && ev->hoff.size() == 1) ev->hoff.push_back(0);
    if (int r = beginBatch(ev)) return r;
    if (ev->E <= 0) return fail(SDL_ERR_STATE, "no event");
    if (int r = runStages(ev, SDL_STAGE_MD, SDL_STAGE_TCL)) return r;
    if (hipGetLastError() != hipSuccess) return fail(SDL_ERR_HIP, "kernel launch failed");
    ev->stage = SDL_STAGE_TCL;
    return 0;
}

int sdl_event_sync(sdl_event* ev) {
    if (!ev) return fail(SDL_ERR_ARG, "null event");
    HIP_OK(hipStreamSynchronize(ev->st));
    return 0;
}

int sdl_event_layer_counts(sdl_event* ev, int e, unsigned long long out[110]) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || !out || e < 0 || e >= ev->E) return fail(SDL_ERR_ARG, "bad arguments");
    std::vector<unsigned long long> all;
    if (int r = readCounters(ev, all)) return r;
    std::memcpy(out, all.data() + size_t(e) * kNCounters * kNBins, sizeof(unsigned long long) * kNCounters * kNBins);
    return 0;
}

int sdl_event_counts(sdl_event* ev, int e, unsigned long long out[10]) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || !out || e < -1 || e >= ev->E) return fail(SDL_ERR_ARG, "bad arguments");
    std::vector<unsigned long long> all;
    if (int r = readCounters(ev, all)) return r;
    for (int k = 0; k < kNCounters; ++k) out[k] = 0;
    for (int ee = 0; ee < ev->E; ++ee) {
        if (e >= 0 && ee != e) continue;
        for (int k = 0; k < 5; ++k)
            for (int b = 0; b < kNBins; ++b) out[k] += all[(size_t(ee) * kNCounters + k) * kNBins + b];
    }
    // Object totals are the sizes of the object lists (Event.cc:1080-1092), not
    // sums of the per-layer counters, which some stages leave untouched.
    const unsigned m = ev->ranMask;
    const bool ran[5] = {bool(m & ((1u << SDL_STAGE_MD) | (1u << SDL_STAGE_MD_PSEUDO))),
                         bool(m & ((1u << SDL_STAGE_SG) | (1u << SDL_STAGE_SG_LAYERS))),
                         bool(m & (1u << SDL_STAGE_TP)),
                         bool(m & ((1u << SDL_STAGE_TL) | (1u << SDL_STAGE_TL_LAYERS) | (1u << SDL_STAGE_TL_NAV))),
                         bool(m & ((1u << SDL_STAGE_TCP) | (1u << SDL_STAGE_TCL) | (1u << SDL_STAGE_TC_LAYERS)))};
    const int stages[5] = {SDL_STAGE_MD, SDL_STAGE_SG, SDL_STAGE_TP, SDL_STAGE_TL, SDL_STAGE_TCP};
    for (int k = 0; k < 5; ++k) {
        if (!ran[k]) continue;
        if (e < 0) {
            const long long tcp = (m & (1u << SDL_STAGE_TCP)) ? ev->nTCP : 0;
            const long long tcl = (m & ((1u << SDL_STAGE_TCL) | (1u << SDL_STAGE_TC_LAYERS))) ? ev->nTCL : 0;
            const long long tot[5] = {ev->nMD, ev->nSG, ev->nTP, ev->nTL, tcp + tcl};
            out[5 + k] = (unsigned long long)tot[k];
            continue;
        }
        long long a, n;
        if (k < 4) {
            if (int r = stageRange(ev, e, stages[k], a, n)) return r;
            out[5 + k] = (unsigned long long)n;
            continue;
        }
        // TCs: from triplets and from tracklets, each only if its stage ran
        if (m & (1u << SDL_STAGE_TCP)) {
            if (int r = stageRange(ev, e, SDL_STAGE_TCP, a, n)) return r;
            out[9] += (unsigned long long)n;
        }
        if (m & ((1u << SDL_STAGE_TCL) | (1u << SDL_STAGE_TC_LAYERS))) {
            if (int r = stageRange(ev, e, SDL_STAGE_TCL, a, n)) return r;
            out[9] += (unsigned long long)n;
        }
    }
    return 0;
}

namespace {

// The body of sdl_event_export and sdl_event_export_noalloc: gathers event e's
// objects of `stage` as hit-index tuples into `staging` (device; null: the
// event's own export buffer, grown as needed) and copies them to `out`.
int exportObjects(sdl_event* ev, int e, int stage, int* out, long long cap, long long* n, int* staging,
                  long long stagingCap) {
    if (!ev || !n || e < 0 || e >= ev->E || stage < 0 || stage > SDL_STAGE_TCL) return fail(SDL_ERR_ARG, "bad arguments");
    // the layer-pair stages store SG, TL and TCL objects
    unsigned need = 1u << stage;
    if (stage == SDL_STAGE_MD) need |= 1u << SDL_STAGE_MD_PSEUDO;
    if (stage == SDL_STAGE_SG) need |= 1u << SDL_STAGE_SG_LAYERS;
    if (stage == SDL_STAGE_TL) need |= (1u << SDL_STAGE_TL_LAYERS) | (1u << SDL_STAGE_TL_NAV);
    if (stage == SDL_STAGE_TCL) need |= 1u << SDL_STAGE_TC_LAYERS;
    if (!(ev->ranMask & need)) return fail(SDL_ERR_STATE, "stage has not run");
    long long a = 0, cnt = 0;
    if (int r = stageRange(ev, e, stage, a, cnt)) return r;
    *n = cnt;
    const int width[6] = {2, 4, 8, 8, 16, 16};
    if (!out) return 0;
    if (cap < cnt * width[stage]) return fail(SDL_ERR_CAPACITY, "export buffer too small");
    if (cnt == 0) return 0;
    if (staging) {
        if (cnt * width[stage] > stagingCap)
            return fail(SDL_ERR_CAPACITY, "sdl_event_export_noalloc: the objects (" +
                                              std::to_string(cnt * width[stage]) + " ints) exceed its staging array (" +
                                              std::to_string(stagingCap) + " ints); use sdl_event_export");
    } else {
        if (int r = ev->exportBuf.ensure(cnt * width[stage])) return r;
        staging = ev->exportBuf.p;
    }
    const unsigned grid = gridFor(cnt, kBlock);
    if (stage == SDL_STAGE_MD)
        hipLaunchKernelGGL(k_export_md, dim3(grid), dim3(kBlock), 0, ev->st, a, cnt, ev->mdLo.p, ev->mdUp.p, ev->pidx,
                           staging);
    else {
        const int kind = stage == SDL_STAGE_SG ? 1 : (stage == SDL_STAGE_TP || stage == SDL_STAGE_TL) ? 2
                                                                                                     : (stage == SDL_STAGE_TCP ? 3 : 4);
        const int* pA = stage == SDL_STAGE_TP ? ev->tpIn.p : ev->tlIn.p;
        const int* pB = stage == SDL_STAGE_TP ? ev->tpOut.p : ev->tlOut.p;
        const int2* pC = stage == SDL_STAGE_TCP ? ev->tcp.p : ev->tcl.p;
        hipLaunchKernelGGL(k_export, dim3(grid), dim3(kBlock), 0, ev->st, a, cnt, kind, pA, pB, pC, ev->tpIn.p,
                           ev->tpOut.p, ev->tlIn.p, ev->tlOut.p, ev->sgIn.p, ev->sgOut.p, ev->mdLo.p, ev->mdUp.p,
                           ev->pidx, staging);
    }
    HIP_OK(hipMemcpyAsync(out, staging, sizeof(int) * cnt * width[stage], hipMemcpyDeviceToHost, ev->st));
    HIP_OK(hipStreamSynchronize(ev->st));
    return 0;
}

// Staging of sdl_event_export_noalloc: a file-scope device array (each GPU the
// library runs on holds its own copy until the process ends).  Host code uses
// it only through the per-device address from hipGetSymbolAddress.
constexpr long long kExportScratchInts = 16LL << 20;  // 64 MiB
constexpr int kMaxDevices = 64;
__device__ __align__(256) int scratch[kExportScratchInts];
std::mutex g_scratchMutex;
int* g_scratchPtr[kMaxDevices] = {};

// The address of `scratch` on `device` (current), looked up once per device.
int scratchOn(int device, int** ptr) {
    if (device < 0 || device >= kMaxDevices) return fail(SDL_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess || !p)
            return fail(SDL_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(r));
        g_scratchPtr[device] = static_cast<int*>(p);
    }
    *ptr = g_scratchPtr[device];
    return 0;
}

}  // namespace

int sdl_event_export(sdl_event* ev, int e, int stage, int* out, long long cap, long long* n) {
    OnStream onStream(ev ? ev->st : nullptr);
    return exportObjects(ev, e, stage, out, cap, n, nullptr, 0);
}

int sdl_event_export_noalloc(sdl_event* ev, int e, int stage, int* out, long long cap, long long* n) {
    if (!ev) return fail(SDL_ERR_ARG, "bad arguments");
    if (hipSetDevice(ev->g->device) != hipSuccess) return fail(SDL_ERR_HIP, "hipSetDevice failed");
    int* staging = nullptr;
    if (int r = scratchOn(ev->g->device, &staging)) return r;
    return exportObjects(ev, e, stage, out, cap, n, staging, kExportScratchInts);
}

int sdl_object_var_count(int stage) {
    return stage == SDL_STAGE_MD ? kMDVars : stage == SDL_STAGE_SG ? kSGVars : stage == SDL_STAGE_TL ? kTLVars
           : stage == SDL_STAGE_TCL ? kTCLVars : 0;
}

const char* sdl_object_var_names(int stage) {
    switch (stage) {
        case SDL_STAGE_MD:
            return "dz,shiftedDz,deltaPhi,deltaPhiChange,deltaPhiNoShift,deltaPhiChangeNoShift,miniCut";
        case SDL_STAGE_SG:
            return "rtIn,rtOut,zIn,zOut,zLo,zHi,rtLo,rtHi,deltaPhi,deltaPhiMin,deltaPhiMax,deltaPhiChange,"
                   "deltaPhiMinChange,deltaPhiMaxChange,sdCut,sdSlope,sdMuls,sdPVoff,reco_deltaPhi,"
                   "dAlpha_innerMD_segment,dAlpha_outerMD_segment,dAlpha_innerMD_outerMD";
        case SDL_STAGE_TL:  // typed values, then the reco-variable map's keys (prefixed reco_)
            return "betaIn,betaInCut,betaOut,betaOutCut,deltaBeta,deltaBetaCut,"
                   "reco_algo,reco_alpha1GeV_OutLo,reco_betaAv,reco_betaAv_0th,reco_betaAv_1st,reco_betaAv_2nd,"
                   "reco_betaAv_3rd,reco_betaAv_4th,reco_betaIn,reco_betaInRHmax,reco_betaInRHmin,reco_betaIn_0th,"
                   "reco_betaIn_1st,reco_betaIn_1stCorr,reco_betaIn_2nd,reco_betaIn_2ndCorr,reco_betaIn_3rd,"
                   "reco_betaIn_3rdCorr,reco_betaIn_4th,reco_betaIn_cut,reco_betaOut,reco_betaOutRHmax,"
                   "reco_betaOutRHmin,reco_betaOut_0th,reco_betaOut_1st,reco_betaOut_1stCorr,reco_betaOut_2nd,"
                   "reco_betaOut_2ndCorr,reco_betaOut_3rd,reco_betaOut_3rdCorr,reco_betaOut_4th,reco_betaPt_0th,"
                   "reco_betaPt_1st,reco_betaPt_2nd,reco_betaPt_3rd,reco_betaPt_4th,reco_betacormode,reco_dBeta,"
                   "reco_dBetaCut2,reco_dBetaLum2,reco_dBetaMuls,reco_dBetaRIn2,reco_dBetaROut2,reco_dBetaRes,"
                   "reco_dBeta_0th,reco_dBeta_1st,reco_dBeta_2nd,reco_dBeta_3rd,reco_dBeta_4th,reco_dBeta_midPoint,"
                   "reco_dPhi,reco_deltaPhiPos,reco_deltaZLum,reco_dr,reco_drt_InSeg,reco_drt_tl_axis,"
                   "reco_dzDrtScale,reco_hit1_x,reco_hit1_y,reco_hit2_x,reco_hit2_y,reco_hit3_x,reco_hit3_y,"
                   "reco_hit4_x,reco_hit4_y,reco_innerSgInnerMdDetId,reco_innerSgOuterMdDetId,reco_k2Rinv1GeVf,"
                   "reco_kRinv1GeVf,reco_kZ,reco_outerSgInnerMdDetId,reco_outerSgOuterMdDetId,reco_pixelPSZpitch,"
                   "reco_ptCut,reco_pt_beta,reco_pt_betaIn,reco_pt_betaOut,reco_rawBetaIn,reco_rawBetaInCorrection,"
                   "reco_rawBetaOut,reco_rawBetaOutCorrection,reco_rtHi_point,reco_rtIn,reco_rtLo_point,reco_rtOut,"
                   "reco_rtOut_o_rtIn,reco_rtRatio_OutLoInLo,reco_rt_InLo,reco_rt_OutLo,reco_sdIn_alpha,reco_sdIn_d,"
                   "reco_sdOut_alphaOut,reco_sdOut_d,reco_sdlCut,reco_sdlSlope,reco_sinAlphaMax,reco_strip2SZpitch,"
                   "reco_zGeom,reco_zHi,reco_zHiPointed,reco_zIn,reco_zLo,reco_zLoPointed,reco_zOut,reco_z_InLo,"
                   "reco_z_OutLo,reco_zpitch_InLo,reco_zpitch_OutLo";
        case SDL_STAGE_TCL: return "dR,innerR";
        default: return "";
    }
}

int sdl_event_export_vars(sdl_event* ev, int e, int stage, float* out, long long cap, long long* n) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || !n || e < 0 || e >= ev->E || sdl_object_var_count(stage) == 0) return fail(SDL_ERR_ARG, "bad arguments");
    unsigned need = 1u << stage;
    if (stage == SDL_STAGE_MD) need |= 1u << SDL_STAGE_MD_PSEUDO;
    if (stage == SDL_STAGE_SG) need |= 1u << SDL_STAGE_SG_LAYERS;
    if (stage == SDL_STAGE_TL) need |= (1u << SDL_STAGE_TL_LAYERS) | (1u << SDL_STAGE_TL_NAV);
    if (stage == SDL_STAGE_TCL) need |= 1u << SDL_STAGE_TC_LAYERS;
    if (!(ev->ranMask & need)) return fail(SDL_ERR_STATE, "stage has not run");
    long long a = 0, cnt = 0;
    if (int r = stageRange(ev, e, stage, a, cnt)) return r;
    *n = cnt;
    const int nv = sdl_object_var_count(stage);
    if (!out) return 0;
    if (cap < cnt * nv) return fail(SDL_ERR_CAPACITY, "export buffer too small");
    if (cnt == 0) return 0;
    // objects an AllComb algorithm accepted keep their constructors' values (0)
    // and set no reco variable (NaN): MiniDoublet.cc:41-51, Segment.cc:35-56,
    // Tracklet.cc:23-33, TrackCandidate.cc:121-141
    const unsigned acBits = stage == SDL_STAGE_SG ? (1u << SDL_STAGE_SG) | (1u << SDL_STAGE_SG_LAYERS)
                            : stage == SDL_STAGE_TL ? (1u << SDL_STAGE_TL) | (1u << SDL_STAGE_TL_LAYERS) | (1u << SDL_STAGE_TL_NAV)
                            : stage == SDL_STAGE_TCL ? (1u << SDL_STAGE_TCL) | (1u << SDL_STAGE_TC_LAYERS) : 0u;
    if (ev->acMask & acBits) {
        const int typed = stage == SDL_STAGE_SG ? 14 : stage == SDL_STAGE_TL ? kTLTyped : 0;
        for (long long i = 0; i < cnt; ++i)
            for (int k = 0; k < nv; ++k) out[i * nv + k] = k < typed ? 0.f : __builtin_nanf("");
        return 0;
    }
    if (int r = ev->exportVars.ensure(cnt * nv)) return r;
    sdl_geom* G = ev->g;
    const unsigned grid = gridFor(cnt, kBlock);
    if (stage == SDL_STAGE_MD) {
        const bool pseudo = !(ev->ranMask & (1u << SDL_STAGE_MD)) || (ev->acMask & (1u << SDL_STAGE_MD));
        hipLaunchKernelGGL(k_vars_md, dim3(grid), dim3(kBlock), 0, ev->st, a, cnt, G->NM, G->dMods, hitsView(ev),
                           mdView(ev), pseudo, ev->exportVars.p);
    } else if (stage == SDL_STAGE_SG) {
        hipLaunchKernelGGL(k_vars_sg, dim3(grid), dim3(kBlock), 0, ev->st, a, cnt, G->NM, G->dMods, hitsView(ev),
                           mdView(ev), sgView(ev), ev->exportVars.p);
    } else if (stage == SDL_STAGE_TL) {
        hipLaunchKernelGGL(k_vars_tl, dim3(grid), dim3(kBlock), 0, ev->st, a, cnt, G->NM, G->dMods, hitsView(ev),
                           mdView(ev), sgView(ev), ev->tlIn.p, ev->tlOut.p, ev->exportVars.p);
    } else {
        hipLaunchKernelGGL(k_vars_tcl, dim3(grid), dim3(kBlock), 0, ev->st, a, cnt, mdView(ev), sgView(ev),
                           ev->tlIn.p, ev->tlOut.p, ev->tcl.p, ev->exportVars.p);
    }
    HIP_OK(hipMemcpyAsync(out, ev->exportVars.p, sizeof(float) * cnt * nv, hipMemcpyDeviceToHost, ev->st));
    HIP_OK(hipStreamSynchronize(ev->st));
    return 0;
}

int sdl_event_num_events(const sdl_event* ev) { return ev ? ev->E : 0; }

int sdl_event_set_timing(sdl_event* ev, int on) {
    if (!ev) return fail(SDL_ERR_ARG, "null event");
    ev->timing = on != 0;
    return 0;
}

int sdl_event_kernel_ms(sdl_event* ev, double* out, int n) {
    if (!ev || !out || n < 0) return fail(SDL_ERR_ARG, "bad arguments");
    for (int k = 0; k < T_NTIMERS; ++k) timerCollect(ev, k);
    for (int k = 0; k < n && k < T_NTIMERS; ++k) out[k] = ev->kernelMs[k];
    return 0;
}

void* sdl_event_stream(sdl_event* ev) { return ev ? (void*)ev->st : nullptr; }

int sdl_event_set_graphs(sdl_event* ev, int on) {
    if (!ev) return fail(SDL_ERR_ARG, "null event");
    ev->useGraphs = on != 0;
    if (!ev->useGraphs) dropGraph(ev);
    return 0;
}

int sdl_event_get_tuning(sdl_event* ev, int* values, int n) {
    if (!ev || !values || n < TUNE_N) return fail(SDL_ERR_ARG, "bad arguments");
    for (int k = 0; k < TUNE_N; ++k) values[k] = ev->tune.v[k];
    return TUNE_N;
}

int sdl_event_set_tuning(sdl_event* ev, const int* values, int n) {
    if (!ev || !values || n != TUNE_N) return fail(SDL_ERR_ARG, "bad arguments");
    for (int k = 0; k < TUNE_N; ++k) {
        bool ok = false;
        for (int j = 0; kTuneValues[k][j]; ++j) ok = ok || values[k] == kTuneValues[k][j];
        if (!ok) return fail(SDL_ERR_ARG, std::string("unsupported value for ") + kTuneNames[k]);
    }
    for (int k = 0; k < TUNE_N; ++k) ev->tune.v[k] = values[k];
    return 0;
}

int sdl_event_autotune(sdl_event* ev, int reps, char* report, int cap) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || reps < 1) return fail(SDL_ERR_ARG, "bad arguments");
    using clk = std::chrono::steady_clock;
    // one timed chain: wall time of run_chain, which ends with its host round trip
    auto timeChain = [&](double& ms) -> int {
        const auto t0 = clk::now();
        if (int r = sdl_event_run_chain(ev)) return r;
        ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        return 0;
    };
    auto median = [&](double& out) -> int {
        double warm;
        if (int r = timeChain(warm)) return r;  // sizes, graph capture
        std::vector<double> t(reps);
        for (int k = 0; k < reps; ++k)
            if (int r = timeChain(t[k])) return r;
        std::sort(t.begin(), t.end());
        out = t[reps / 2];
        return 0;
    };
    std::string rep = "{";
    double best;
    if (int r = median(best)) return r;
    if (ev->NH <= 0) return fail(SDL_ERR_STATE, "autotune needs a batch with hits");
    char buf[160];
    std::snprintf(buf, sizeof(buf), "\"start_ms\": %.4f", best);
    rep += buf;
    // one knob at a time, in launch order, keeping the fastest value
    for (int k = 0; k < TUNE_N; ++k) {
        const int keep = ev->tune.v[k];
        int bestV = keep;
        for (int j = 0; kTuneValues[k][j]; ++j) {
            const int v = kTuneValues[k][j];
            if (v == keep) continue;
            ev->tune.v[k] = v;
            double ms;
            if (int r = median(ms)) {
                ev->tune.v[k] = keep;
                return r;
            }
            std::snprintf(buf, sizeof(buf), ", \"%s=%d\": %.4f", kTuneNames[k], v, ms);
            rep += buf;
            if (ms < best * 0.99) {  // a change must gain at least 1 %
                best = ms;
                bestV = v;
            }
        }
        ev->tune.v[k] = bestV;
    }
    std::snprintf(buf, sizeof(buf), ", \"best_ms\": %.4f", best);
    rep += buf;
    for (int k = 0; k < TUNE_N; ++k) {
        std::snprintf(buf, sizeof(buf), ", \"%s\": %d", kTuneNames[k], ev->tune.v[k]);
        rep += buf;
    }
    rep += "}";
    if (report && cap > 0) {
        const size_t m = std::min<size_t>(rep.size(), size_t(cap - 1));
        std::memcpy(report, rep.data(), m);
        report[m] = 0;
    }
    return 0;
}

int sdl_event_set_headroom(sdl_event* ev, int fromStage, double factor) {
    if (!ev || factor <= 0.0) return fail(SDL_ERR_ARG, "bad arguments");
    ev->headroomFrom = fromStage;
    ev->headroomFactor = factor;
    return 0;
}

// ---------------------------------------------------------------- learned proposal
int sdl_mlp_create(sdl_mlp** out) {
    if (!out) return fail(SDL_ERR_ARG, "null argument");
    auto* m = new sdl_mlp();
    (void)hipGetDevice(&m->device);
    const size_t pb = sizeof(float) * sdl::mlp::kNParams;
    if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess || hipMalloc(&m->dParams, pb) != hipSuccess ||
        hipMalloc(&m->dGrad, pb) != hipSuccess || hipMalloc(&m->dM, pb) != hipSuccess ||
        hipMalloc(&m->dV, pb) != hipSuccess || hipMalloc(&m->dLoss, sizeof(float)) != hipSuccess ||
        hipMalloc(&m->dPacked, sizeof(sdl::mlp::Packed)) != hipSuccess) {
        sdl_mlp_free(m);
        return fail(SDL_ERR_HIP, "MLP allocation failed");
    }
    for (float* p : {m->dParams, m->dGrad, m->dM, m->dV})
        hipLaunchKernelGGL(k_zero<float>, dim3(gridFor(sdl::mlp::kNParams, kBlock)), dim3(kBlock), 0, m->st, p,
                           (long long)sdl::mlp::kNParams);
    if (int r = repackMlp(m)) {
        sdl_mlp_free(m);
        return r;
    }
    *out = m;
    return 0;
}

void sdl_mlp_free(sdl_mlp* m) {
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    for (void* p : {(void*)m->dParams, (void*)m->dGrad, (void*)m->dM, (void*)m->dV, (void*)m->dLoss, (void*)m->dPacked})
        if (p) (void)hipFree(p);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

int sdl_mlp_num_params(void) { return sdl::mlp::kNParams; }

int sdl_mlp_set_params(sdl_mlp* m, const float* params, int n) {
    if (!m || !params || n != sdl::mlp::kNParams) return fail(SDL_ERR_ARG, "bad arguments");
    HIP_OK(hipMemcpyAsync(m->dParams, params, sizeof(float) * n, hipMemcpyHostToDevice, m->st));
    for (float* p : {m->dGrad, m->dM, m->dV})
        hipLaunchKernelGGL(k_zero<float>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, m->st, p, (long long)n);
    m->adamSteps = 0;
    return repackMlp(m);
}

int sdl_mlp_get_params(sdl_mlp* m, float* params, int n) {
    if (!m || !params || n != sdl::mlp::kNParams) return fail(SDL_ERR_ARG, "bad arguments");
    HIP_OK(hipMemcpyAsync(params, m->dParams, sizeof(float) * n, hipMemcpyDeviceToHost, m->st));
    HIP_OK(hipStreamSynchronize(m->st));
    return 0;
}

int sdl_mlp_gradient(sdl_mlp* m, const float* dX, const float* dY, long long n, float posWeight, float* grad,
                     float* loss) {
    using namespace sdl::mlp;
    if (!m || !dX || !dY || n <= 0 || !grad || !loss) return fail(SDL_ERR_ARG, "bad arguments");
    hipLaunchKernelGGL(k_zero<float>, dim3(gridFor(kNParams, kBlock)), dim3(kBlock), 0, m->st, m->dGrad,
                       (long long)kNParams);
    hipLaunchKernelGGL(k_zero<float>, dim3(1), dim3(kBlock), 0, m->st, m->dLoss, 1LL);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(gridFor(n, kTrainSamples)), dim3(kTrainThreads), 0, m->st, dX, dY,
                       (const int*)nullptr, 0LL, n, n, m->dParams, posWeight, m->dGrad, m->dLoss);
    HIP_OK(hipMemcpyAsync(grad, m->dGrad, sizeof(float) * kNParams, hipMemcpyDeviceToHost, m->st));
    HIP_OK(hipMemcpyAsync(loss, m->dLoss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    hipLaunchKernelGGL(k_zero<float>, dim3(gridFor(kNParams, kBlock)), dim3(kBlock), 0, m->st, m->dGrad,
                       (long long)kNParams);
    HIP_OK(hipStreamSynchronize(m->st));
    return 0;
}

int sdl_mlp_train(sdl_mlp* m, const float* dX, const float* dY, const int* dIdx, long long start, int batch, int steps,
                  float lr, float posWeight, float* lastLoss) {
    using namespace sdl::mlp;
    if (!m || !dX || !dY || batch <= 0 || steps < 0 || start < 0) return fail(SDL_ERR_ARG, "bad arguments");
    const float b1 = 0.9f, b2 = 0.999f;
    for (int k = 0; k < steps; ++k) {
        hipLaunchKernelGGL(k_zero<float>, dim3(1), dim3(kBlock), 0, m->st, m->dLoss, 1LL);
        hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(gridFor(batch, kTrainSamples)), dim3(kTrainThreads), 0, m->st, dX, dY,
                           dIdx, start + (long long)k * batch, (long long)batch, (long long)batch, m->dParams,
                           posWeight, m->dGrad, m->dLoss);
        ++m->adamSteps;
        const float bc1 = 1.f - std::pow(b1, float(m->adamSteps)), bc2 = 1.f - std::pow(b2, float(m->adamSteps));
        hipLaunchKernelGGL(k_mlp_adam, dim3(gridFor(kNParams, kBlock)), dim3(kBlock), 0, m->st, m->dParams, m->dGrad,
                           m->dM, m->dV, lr, b1, b2, 1e-8f, bc1, bc2);
    }
    if (lastLoss) HIP_OK(hipMemcpyAsync(lastLoss, m->dLoss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    HIP_OK(hipStreamSynchronize(m->st));
    return repackMlp(m);
}

int sdl_mlp_forward(sdl_mlp* m, const float* dX, long long n, float* dLogit) {
    if (!m || !dX || !dLogit || n < 0) return fail(SDL_ERR_ARG, "bad arguments");
    if (n == 0) return 0;
    const unsigned grid = std::min<unsigned>(gridFor((n + 63) / 64, kWavesPerBlock), 8192u);
    hipLaunchKernelGGL(sdl::mlp::k_mlp_forward, dim3(grid), dim3(kBlock), 0, m->st, dX, n, m->dPacked, dLogit);
    HIP_OK(hipStreamSynchronize(m->st));
    return 0;
}

int sdl_event_set_proposal(sdl_event* ev, sdl_mlp* m, float threshold) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || (m && !(threshold > 0.f && threshold < 1.f))) return fail(SDL_ERR_ARG, "bad arguments");
    ev->proposal = m;
    ev->proposalLogit = m ? std::log(threshold / (1.f - threshold)) : 0.f;
    return 0;
}

int sdl_event_run_stats(sdl_event* ev, long long out[4]) {
    if (!ev || !out) return fail(SDL_ERR_ARG, "bad arguments");
    for (int k = 0; k < 4; ++k) out[k] = ev->stats[k];
    return 0;
}

int sdl_event_tl_work(sdl_event* ev, long long out[4]) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || !out) return fail(SDL_ERR_ARG, "bad arguments");
    long long t[TOT_N];
    HIP_OK(hipMemcpyAsync(t, ev->tot.p, sizeof(t), hipMemcpyDeviceToHost, ev->st));
    HIP_OK(hipStreamSynchronize(ev->st));
    out[0] = t[TOT_TLPAIRS];
    out[1] = t[TOT_TLCANDS];
    out[2] = t[TOT_TL];
    out[3] = 0;
    if (ev->proposal && ev->tlScoredStride > 0) {
        std::vector<int> w(size_t(kNVariantsTL * ev->tlScoredStride));
        HIP_OK(hipMemcpyAsync(w.data(), ev->tlScored.p, sizeof(int) * w.size(), hipMemcpyDeviceToHost, ev->st));
        HIP_OK(hipStreamSynchronize(ev->st));
        for (int v = 0; v < kNVariantsTL; ++v) {
            const long long nw = (t[TOT_TLP0 + v] + 63) / 64;
            for (long long k = 0; k < nw && k < ev->tlScoredStride; ++k) out[3] += w[size_t(v * ev->tlScoredStride + k)];
        }
    }
    return 0;
}

int sdl_event_tl_pair_rows(sdl_event* ev, float* dX, float* dY, long long cap, long long* n) {
    OnStream onStream(ev ? ev->st : nullptr);
    if (!ev || !n) return fail(SDL_ERR_ARG, "bad arguments");
    if (ev->stage < SDL_STAGE_TL) return fail(SDL_ERR_STATE, "the tracklet stage has not run");
    if (ev->proposal) return fail(SDL_ERR_STATE, "training rows come from a plain chain (no proposal set)");
    long long t[TOT_N];
    HIP_OK(hipMemcpyAsync(t, ev->tot.p, sizeof(t), hipMemcpyDeviceToHost, ev->st));
    HIP_OK(hipStreamSynchronize(ev->st));
    const long long total = t[TOT_TLPAIRS];
    *n = total;
    if (!dX) return 0;
    if (!dY || cap < total) return fail(SDL_ERR_CAPACITY, "row buffer too small");
    const MDDev D = mdView(ev);
    const SGDev S = sgView(ev);
    const long long ns = std::max(ev->nSG, 1LL);
    for (int v = 0; v < kNVariantsTL; ++v) {
        const long long np = t[TOT_TLP0 + v];
        if (np == 0) continue;
        const VarRange R{ev->tlPOff.p, nullptr, ns, v, np, ev->tot.p};
        const dim3 grid(gridFor(np, kBlock));
        if (v == 0)
            hipLaunchKernelGGL(k_tl_pair_rows<TL_BBBB>, grid, dim3(kBlock), 0, ev->st, R, D, S, ev->sgStart.p,
                               ev->tlPair.p, ev->tlCOff.p, ev->tlFlag.p, dX, dY);
        else if (v == 1)
            hipLaunchKernelGGL(k_tl_pair_rows<TL_BBEE>, grid, dim3(kBlock), 0, ev->st, R, D, S, ev->sgStart.p,
                               ev->tlPair.p, ev->tlCOff.p, ev->tlFlag.p, dX, dY);
        else
            hipLaunchKernelGGL(k_tl_pair_rows<TL_EEEE>, grid, dim3(kBlock), 0, ev->st, R, D, S, ev->sgStart.p,
                               ev->tlPair.p, ev->tlCOff.p, ev->tlFlag.p, dX, dY);
    }
    HIP_OK(hipStreamSynchronize(ev->st));
    return 0;
}

}  // extern "C"
