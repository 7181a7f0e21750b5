// Shared device helpers for the dtfft_amd kernels.
//
// This file (and the other *.hip files in this directory) is compiled at plan
// creation by hipRTC for the GPU's gfx950 target; the host library embeds the
// text at build time.  Only HIP builtins are used (no standard headers).
#ifndef DTFFT_AMD_DEVICE_COMMON
#define DTFFT_AMD_DEVICE_COMMON

typedef long long i64;
typedef unsigned int u32;
typedef unsigned long long u64;

// Complex value; aligned to its size so a line of them is read with one
// 8-byte (float) or 16-byte (double) access per element.
template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cx {
    T x, y;
};

template <typename T> __device__ __forceinline__ cx<T> mk(T x, T y) { cx<T> r; r.x = x; r.y = y; return r; }
template <typename T> __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cx<T> conj(cx<T> a) { return mk(a.x, -a.y); }
// a * (-i)
template <typename T> __device__ __forceinline__ cx<T> mul_mi(cx<T> a) { return mk(a.y, -a.x); }
// a * (c + i s) with real constants
template <typename T> __device__ __forceinline__ cx<T> cmulc(cx<T> a, T c, T s) {
    return mk(a.x * c - a.y * s, a.x * s + a.y * c);
}
template <typename T> __device__ __forceinline__ cx<T> scale(cx<T> a, T s) { return mk(a.x * s, a.y * s); }

// Twiddle factors.  FP32 plans store each twiddle as an FP32 hi+lo pair of
// the FP64 value (hi = fl(w), lo = fl(w - hi)) and multiply with a
// compensated product; FP64 plans store FP64 twiddles.
struct __attribute__((aligned(16))) twf { float rh, ih, rl, il; };

template <typename T> struct tw_of;
template <> struct tw_of<float> { typedef twf type; };
template <> struct tw_of<double> { typedef cx<double> type; };

__device__ __forceinline__ cx<float> twmul(cx<float> a, const twf w) {
    // (a.x + i a.y) * ((rh + rl) + i (ih + il)); lo terms first, hi terms fused last.
    float re = __builtin_fmaf(a.x, w.rl, -a.y * w.il);
    float im = __builtin_fmaf(a.x, w.il, a.y * w.rl);
    re = __builtin_fmaf(-a.y, w.ih, re);
    im = __builtin_fmaf(a.y, w.rh, im);
    re = __builtin_fmaf(a.x, w.rh, re);
    im = __builtin_fmaf(a.x, w.ih, im);
    return mk(re, im);
}
__device__ __forceinline__ twf conj_tw(twf w) { w.ih = -w.ih; w.il = -w.il; return w; }
__device__ __forceinline__ cx<double> conj_tw(cx<double> w) { w.y = -w.y; return w; }

__device__ __forceinline__ cx<double> twmul(cx<double> a, const cx<double> w) {
    double re = __builtin_fma(a.x, w.x, -a.y * w.y);
    double im = __builtin_fma(a.x, w.y, a.y * w.x);
    return mk(re, im);
}

// Non-temporal store of one complex value (streaming output that no later instruction of the
// kernel reads back).  Measured on MI355X at 512^3 FP32: the LDS-DMA column-in / row-out pass
// went from 430 to 422 us with its row stores non-temporal, the row passes from 377-395 to
// 365-376 us; column-in passes of the plain kernel and column-out stores did not gain.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(cx<float>* p, cx<float> v) {
    __builtin_nontemporal_store(*(const u64*)&v, (u64*)p);
}
__device__ __forceinline__ void store_nt(cx<double>* p, cx<double> v) {
    __builtin_nontemporal_store(*(const u64x2*)&v, (u64x2*)p);
}

#endif
