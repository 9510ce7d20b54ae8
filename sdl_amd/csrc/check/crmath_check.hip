// Checks of sdl_crmath.h (test support, not part of the product library).
//
// Each fast function must return float(f(double(v))) of the libm for every
// argument.  The host scans compare against the host libm, the device scans
// against the device libm, over ranges of float bit patterns (all 2^32 of them
// for the one-argument functions) or over seeded argument pairs for atan2 and
// a / sin(b).  Every scan also counts how often the rounding test falls back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../sdl_crmath.h"

namespace {

enum { F_ASIN = 0, F_SIN = 1, F_COS = 2, F_ATAN = 3, F_ATAN2 = 4, F_DIVSIN = 5, F_SQRT = 6 };

__host__ __device__ inline float bitsToFloat(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__host__ __device__ inline uint32_t floatToBits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

// The value the libm path gives (float(f(double(v)))), the fast-path result,
// and whether the fast path decided.
struct Eval {
    float ref, got;
    bool fast;
};

__host__ __device__ inline Eval eval1(int fn, float v) {
    using namespace sdl::crm;
    Eval r;
    float f = 0.f;
    switch (fn) {
        case F_ASIN: r.ref = float(asin(double(v))); r.fast = asinFast(v, f); r.got = asinCR(v); break;
        case F_SIN: r.ref = float(sin(double(v))); r.fast = sinFast(v, f); r.got = sinCR(v); break;
        case F_COS: r.ref = float(cos(double(v))); r.fast = cosFast(v, f); r.got = cosCR(v); break;
        case F_ATAN: r.ref = float(atan(double(v))); r.fast = atanFast(v, f); r.got = atanCR(v); break;
        default: r.ref = float(sqrt(double(v))); r.fast = true; r.got = sqrtCR(v); break;
    }
    return r;
}

__host__ __device__ inline Eval eval2(int fn, float a, float b) {
    using namespace sdl::crm;
    Eval r;
    float f = 0.f;
    if (fn == F_ATAN2) {
        r.ref = float(atan2(double(a), double(b)));
        r.fast = atan2Fast(a, b, f);
        r.got = atan2CR(a, b);
    } else {
        r.ref = float(double(a) / sin(double(b)));
        r.fast = divSinFast(a, b, f);
        r.got = divSinCR(a, b);
    }
    return r;
}

// NaNs of either sign and payload compare equal; everything else bit for bit.
__host__ __device__ inline bool same(float a, float b) {
    if (a != a && b != b) return true;
    return floatToBits(a) == floatToBits(b);
}

// Seeded argument pairs: a counter-based hash (SplitMix64) of (seed, i).
__host__ __device__ inline uint64_t mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// mode 0: uniform bit patterns; 1: the cut algebra's ranges (coordinates up to
// 300 cm for atan2, |b| < 3 pi / 4 and a up to 10 for a / sin(b)).
__host__ __device__ inline void pairOf(int fn, int mode, uint64_t seed, uint64_t i, float& a, float& b) {
    const uint64_t h = mix(seed * 0x100000001b3ull + i), g = mix(h);
    if (mode == 0) {
        a = bitsToFloat(uint32_t(h));
        b = bitsToFloat(uint32_t(g));
        return;
    }
    const double ua = double(h >> 11) * 0x1p-53, ub = double(g >> 11) * 0x1p-53;
    if (fn == F_ATAN2) {
        a = float((ua * 2.0 - 1.0) * 300.0);
        b = float((ub * 2.0 - 1.0) * 300.0);
    } else {
        a = float(ua * 10.0);
        b = float((ub * 2.0 - 1.0) * 2.35);
    }
}

__global__ void k_scan1(int fn, uint64_t lo, uint64_t hi, unsigned long long* cnt, uint32_t* firstBad) {
    unsigned long long bad = 0, fb = 0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t u = lo + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; u < hi; u += stride) {
        const Eval r = eval1(fn, bitsToFloat(uint32_t(u)));
        if (!r.fast) ++fb;
        if (!same(r.ref, r.got)) {
            ++bad;
            atomicMin(firstBad, uint32_t(u));
        }
    }
    atomicAdd(&cnt[0], bad);
    atomicAdd(&cnt[1], fb);
}

__global__ void k_scan2(int fn, int mode, uint64_t seed, uint64_t n, unsigned long long* cnt, uint32_t* firstBad) {
    unsigned long long bad = 0, fb = 0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        float a, b;
        pairOf(fn, mode, seed, i, a, b);
        if (fn == F_ATAN2 && b == 0.f) continue;  // the callers handle x == 0 first
        const Eval r = eval2(fn, a, b);
        if (!r.fast) ++fb;
        if (!same(r.ref, r.got)) {
            ++bad;
            atomicMin(firstBad, uint32_t(i));
        }
    }
    atomicAdd(&cnt[0], bad);
    atomicAdd(&cnt[1], fb);
}

template <class F>
void hostParallel(uint64_t n, int threads, F body) {
    threads = std::max(1, threads);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([=] {
            const uint64_t a = n * t / threads, b = n * (t + 1) / threads;
            body(t, a, b);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

// One-argument functions (0 asin, 1 sin, 2 cos, 3 atan, 6 sqrt) over the bit patterns
// lo, lo + stride, ... < hi on the host: out[0] mismatches, out[1] fallbacks,
// out[2] the first mismatching pattern (or 2^32), out[3] patterns checked.
int crm_host_scan1(int fn, unsigned long long lo, unsigned long long hi, unsigned long long stride, int threads,
                   unsigned long long out[4]) {
    if (fn < 0 || fn > 6 || fn == F_ATAN2 || fn == F_DIVSIN || stride == 0 || hi < lo || hi > (1ull << 32)) return -1;
    const uint64_t n = (hi - lo + stride - 1) / stride;
    std::atomic<unsigned long long> bad{0}, fb{0}, first{1ull << 32};
    hostParallel(n, threads, [&](int, uint64_t a, uint64_t b) {
        unsigned long long lb = 0, lf = 0, lfirst = 1ull << 32;
        for (uint64_t k = a; k < b; ++k) {
            const uint64_t u = lo + k * stride;
            const Eval r = eval1(fn, bitsToFloat(uint32_t(u)));
            if (!r.fast) ++lf;
            if (!same(r.ref, r.got)) {
                ++lb;
                lfirst = std::min<unsigned long long>(lfirst, u);
            }
        }
        bad += lb;
        fb += lf;
        unsigned long long cur = first.load();
        while (lfirst < cur && !first.compare_exchange_weak(cur, lfirst)) {
        }
    });
    out[0] = bad;
    out[1] = fb;
    out[2] = first;
    out[3] = n;
    return 0;
}

// Two-argument functions (4 atan2(a, b), 5 a / sin(b)) on n seeded pairs
// (mode 0 uniform bit patterns, 1 the cut algebra's ranges) on the host.
int crm_host_scan2(int fn, int mode, unsigned long long seed, unsigned long long n, int threads,
                   unsigned long long out[4]) {
    if (fn < 4 || fn > 5 || mode < 0 || mode > 1) return -1;
    std::atomic<unsigned long long> bad{0}, fb{0}, first{~0ull}, checked{0};
    hostParallel(n, threads, [&](int, uint64_t a, uint64_t b) {
        unsigned long long lb = 0, lf = 0, lc = 0, lfirst = ~0ull;
        for (uint64_t i = a; i < b; ++i) {
            float x, y;
            pairOf(fn, mode, seed, i, x, y);
            if (fn == F_ATAN2 && y == 0.f) continue;
            ++lc;
            const Eval r = eval2(fn, x, y);
            if (!r.fast) ++lf;
            if (!same(r.ref, r.got)) {
                ++lb;
                lfirst = std::min<unsigned long long>(lfirst, i);
            }
        }
        bad += lb;
        fb += lf;
        checked += lc;
        unsigned long long cur = first.load();
        while (lfirst < cur && !first.compare_exchange_weak(cur, lfirst)) {
        }
    });
    out[0] = bad;
    out[1] = fb;
    out[2] = first;
    out[3] = checked;
    return 0;
}

// The pair the host and device scans draw at index i.
int crm_pair(int fn, int mode, unsigned long long seed, unsigned long long i, float* a, float* b) {
    pairOf(fn, mode, seed, i, *a, *b);
    return 0;
}

// Fast and reference values of one call on the host: out = {libm path, function, fast decided}.
int crm_host_eval(int fn, float a, float b, float out[3]) {
    const Eval r = (fn <= 3 || fn == F_SQRT) ? eval1(fn, a)
                                              : eval2(fn, a, b);
    out[0] = r.ref;
    out[1] = r.got;
    out[2] = r.fast ? 1.f : 0.f;
    return 0;
}

// The same scans on the current HIP device, against the device libm
// (out[2] of scan2: first mismatching pair index, or 2^32 - 1 when none).
int crm_device_scan1(int fn, unsigned long long lo, unsigned long long hi, unsigned long long out[4]) {
    if (fn < 0 || fn > 6 || fn == F_ATAN2 || fn == F_DIVSIN || hi < lo || hi > (1ull << 32)) return -1;
    unsigned long long* dc = nullptr;
    uint32_t* df = nullptr;
    if (hipMalloc(&dc, 2 * sizeof(unsigned long long)) != hipSuccess) return -3;
    if (hipMalloc(&df, sizeof(uint32_t)) != hipSuccess) return -3;
    const uint32_t init = 0xffffffffu;
    hipMemset(dc, 0, 2 * sizeof(unsigned long long));
    hipMemcpy(df, &init, 4, hipMemcpyHostToDevice);
    k_scan1<<<8192, 256>>>(fn, lo, hi, dc, df);
    unsigned long long c[2];
    uint32_t f;
    const hipError_t e = hipDeviceSynchronize();
    hipMemcpy(c, dc, sizeof c, hipMemcpyDeviceToHost);
    hipMemcpy(&f, df, 4, hipMemcpyDeviceToHost);
    hipFree(dc);
    hipFree(df);
    if (e != hipSuccess) return -3;
    out[0] = c[0];
    out[1] = c[1];
    out[2] = c[0] ? f : (1ull << 32);
    out[3] = hi - lo;
    return 0;
}

int crm_device_scan2(int fn, int mode, unsigned long long seed, unsigned long long n, unsigned long long out[4]) {
    if (fn < 4 || fn > 5 || mode < 0 || mode > 1) return -1;
    unsigned long long* dc = nullptr;
    uint32_t* df = nullptr;
    if (hipMalloc(&dc, 2 * sizeof(unsigned long long)) != hipSuccess) return -3;
    if (hipMalloc(&df, sizeof(uint32_t)) != hipSuccess) return -3;
    const uint32_t init = 0xffffffffu;
    hipMemset(dc, 0, 2 * sizeof(unsigned long long));
    hipMemcpy(df, &init, 4, hipMemcpyHostToDevice);
    k_scan2<<<8192, 256>>>(fn, mode, seed, n, dc, df);
    unsigned long long c[2];
    uint32_t f;
    const hipError_t e = hipDeviceSynchronize();
    hipMemcpy(c, dc, sizeof c, hipMemcpyDeviceToHost);
    hipMemcpy(&f, df, 4, hipMemcpyDeviceToHost);
    hipFree(dc);
    hipFree(df);
    if (e != hipSuccess) return -3;
    out[0] = c[0];
    out[1] = c[1];
    out[2] = f;
    out[3] = n;
    return 0;
}

}  // extern "C"
