// R2R transforms (DCT / DST types 1-4) with FFTW's unnormalised REDFT / RODFT
// definitions, the kinds dtFFT passes to FFTW (include/dtfft_config.h.in:66-73):
// a pre-process kernel maps each real line to a complex line of length M, the Stockham
// C2C stage of device/fft.hip transforms it (forward, or backward for types 3), and a
// post-process kernel extracts the n real outputs.  Per kind (x = input, n points):
//   DCT-2 (5):  M = n,  c_m = x_2m (m < ceil(n/2)), x_{2(n-1-m)+1} otherwise;
//               y_k = 2 Re(e^{-i pi k / 2n} C_k)                       (Makhoul)
//   DST-2 (9):  DCT-2 of (-1)^j x_j, read back reversed (k -> n-1-k)
//   DCT-3 (4):  M = n,  w_m = (x_m - i x_{n-m}) e^{i pi m / 2n} (x_n = 0), backward FFT;
//               y_2m = Re W_m, y_2m+1 = Re W_{n-1-m}
//   DST-3 (8):  (-1)^k DCT-3 of the reversed input
//   DCT-4 (6):  M = 2n, z_j = x_j e^{-i pi j / 2n} (zero above n);
//               y_k = 2 Re(e^{-i pi (2k+1) / 4n} Z_k)
//   DST-4 (10): (-1)^k DCT-4 of the reversed input
//   DCT-1 (3):  M = 2(n-1), even extension; y_k = Re E_k
//   DST-1 (7):  M = 2(n+1), odd extension [0, x, 0, -reverse(x)]; y_k = -Im O_{k+1}
// Twiddles e^{i pi a} are evaluated in FP64 (sincospi) for both precisions.  Every one of
// them is t4(j) = e^{-i pi j / 4n} or its conjugate, j < 2n; the fused kernel reads them from
// a table of that function, the three-launch path evaluates it in place.
//
// Lines of unit element stride take the fused kernel fft_r2r_kernel: the real lines go to
// LDS, the pre-process fills the complex work line in LDS, the Stockham passes of fft_tile
// run there, and the post-process writes the real outputs: one read and one write of each
// real element.  Other strides take the three launches through a complex work buffer.
#include "fft.hip"  // fft_tile: the Stockham passes, here run on work lines held in LDS

struct r2r_desc {
    i64 nl, nb;                // lines (l, b): line start l * ls + b * bs
    i64 in_ls, in_bs, in_es;   // input real line / batch / element strides
    i64 out_ls, out_bs, out_es;
    i64 n, m;                  // real length, complex work length
    i64 wdist;                 // complex work line stride (>= m)
    // Chunk-blocked forms of the fused kernel (0: natural layout).  in_ts: columns in, tile t
    // of a batch starts at t * in_ts (its LPB lines one contiguous block, element j of line l
    // at j * in_es + l).  out_seg: rows out with OSEG > 0, element k of a line at
    // (k / OSEG) * out_seg + k % OSEG from the line's start.
    i64 in_ts;
    i64 out_seg;
};

template <typename T>
__device__ __forceinline__ cx<T> cis_pi(double a) {  // e^{i pi a}
    double s, c;
    sincospi(a, &s, &c);
    return mk((T)c, (T)s);
}

// Element m (< M) of the complex work line.  x(j) returns real input j of the line;
// t4(j) returns e^{-i pi j / 4n}.
template <typename T, int KIND, class XF, class TF>
__device__ __forceinline__ cx<T> r2r_pre_elem(const XF& x, i64 n, i64 M, i64 m, const TF& t4) {
    cx<T> z = mk((T)0, (T)0);
    if constexpr (KIND == 5 || KIND == 9) {
        const T sg = (KIND == 9) ? (T)-1 : (T)1;
        z.x = m < (n + 1) / 2 ? x(2 * m) : sg * x(2 * (n - 1 - m) + 1);
    } else if constexpr (KIND == 4 || KIND == 8) {
        T a, bb;
        if (KIND == 4) { a = x(m); bb = m > 0 ? x(n - m) : (T)0; }
        else { a = x(n - 1 - m); bb = m > 0 ? x(m - 1) : (T)0; }
        const cx<T> w = conj(t4(2 * m));                   // e^{i pi m / 2n}
        z = mk(a * w.x + bb * w.y, a * w.y - bb * w.x);  // (a - i bb) w
    } else if constexpr (KIND == 6 || KIND == 10) {
        if (m < n) {
            const T a = KIND == 6 ? x(m) : x(n - 1 - m);
            const cx<T> w = t4(2 * m);  // e^{-i pi m / 2n}
            z = mk(a * w.x, a * w.y);
        }
    } else if constexpr (KIND == 3) {
        z.x = m < n ? x(m) : x(M - m);
    } else if constexpr (KIND == 7) {
        if (m >= 1 && m <= n) z.x = x(m - 1);
        else if (m >= n + 2) z.x = -x(M - 1 - m);
    }
    return z;
}

// Real output k (< n) from the transformed work line W(k).
template <typename T, int KIND, class WF, class TF>
__device__ __forceinline__ T r2r_post_elem(const WF& W, i64 n, i64 k, const TF& t4) {
    if constexpr (KIND == 5 || KIND == 9) {
        const i64 kk = KIND == 9 ? n - 1 - k : k;
        const cx<T> w = t4(2 * kk);  // e^{-i pi kk / 2n}
        const cx<T> c = W(kk);
        return (T)2 * (w.x * c.x - w.y * c.y);
    } else if constexpr (KIND == 4 || KIND == 8) {
        const i64 idx = (k % 2 == 0) ? k / 2 : n - 1 - (k - 1) / 2;
        T y = W(idx).x;
        if (KIND == 8 && (k & 1)) y = -y;
        return y;
    } else if constexpr (KIND == 6 || KIND == 10) {
        const cx<T> w = t4(2 * k + 1);  // e^{-i pi (2k + 1) / 4n}
        const cx<T> c = W(k);
        T y = (T)2 * (w.x * c.x - w.y * c.y);
        if (KIND == 10 && (k & 1)) y = -y;
        return y;
    } else if constexpr (KIND == 3) {
        return W(k).x;
    } else {
        return -W(k + 1).y;
    }
}

template <typename T, int KIND>
__global__ void __launch_bounds__(256) r2r_pre(const T* __restrict__ in, cx<T>* __restrict__ work, r2r_desc d) {
    const i64 n = d.n, M = d.m, total = d.nl * d.nb * M;
    auto t4 = [&](i64 j) { return cis_pi<T>(-(double)j / (double)(4 * n)); };
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (i64)gridDim.x * blockDim.x) {
        const i64 li = e / M, m = e % M;
        const i64 l = li % d.nl, b = li / d.nl;
        const T* x = in + l * d.in_ls + b * d.in_bs;
        auto xf = [&](i64 j) { return x[j * d.in_es]; };
        work[li * d.wdist + m] = r2r_pre_elem<T, KIND>(xf, n, M, m, t4);
    }
}

template <typename T, int KIND>
__global__ void __launch_bounds__(256) r2r_post(const cx<T>* __restrict__ work, T* __restrict__ out, r2r_desc d) {
    const i64 n = d.n, total = d.nl * d.nb * n;
    auto t4 = [&](i64 j) { return cis_pi<T>(-(double)j / (double)(4 * n)); };
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (i64)gridDim.x * blockDim.x) {
        const i64 li = e / n, k = e % n;
        const i64 l = li % d.nl, b = li / d.nl;
        const cx<T>* W = work + li * d.wdist;
        auto wf = [&](i64 j) { return W[j]; };
        out[l * d.out_ls + b * d.out_bs + k * d.out_es] = r2r_post_elem<T, KIND>(wf, n, k, t4);
    }
}

// t[j] = e^{-i pi j / 4n}, j < 2n: every twiddle of the pre- and post-processing
template <typename T>
__global__ void r2r_table(cx<T>* t, i64 n) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < 2 * n) t[j] = cis_pi<T>(-(double)j / (double)(4 * n));
}

// Real length of a KIND transform whose work line has M complex points.
template <int KIND, int M>
struct r2r_len {
    static constexpr int n = KIND == 3 ? M / 2 + 1 : KIND == 7 ? M / 2 - 1 : (KIND == 6 || KIND == 10) ? M / 2 : M;
};

// Fused R2R stage over lines of unit element stride, evenly spaced: line li starts at
// li * ls (the host folds batches into the line range, so d.nb == 1).  Persistent: block b takes tiles of LPB lines b,
// b + gridDim.x, ...  Both twiddle tables are copied to LDS once per block, and the next
// tile's real lines are loaded into registers while the current tile is transformed, so no
// phase of a tile waits for global memory.  The prefetch is issued on every path (line and
// tile indices clamped; past the end it re-reads the last line), which keeps the compiler's
// wait counts exact.  FULL: every tile has LPB lines, so the stores need no guard (a guard
// becomes a branch, and the compiler then waits for the stores before the next prefetch).
//
// HALF (types 2 and 3, n = 2M even): the n-point complex transform of the real sequence v of
// Makhoul's map (types 2) or of its Hermitian spectrum V (types 3) runs as an M-point
// complex transform, halving the butterflies.  Types 2: u_j = v_2j + i v_2j+1 is transformed
// and untangled, V_k = ((U_k + conj U_{M-k}) - i W^k (U_k - conj U_{M-k})) / 2 for k <= M and
// V_{n-k} = conj V_k, W = e^{-2 pi i / n} (the R2C untangling of fft_r2c_kernel).  Types 3:
// Z_k = (V_k + conj V_{M-k}) + i (V_k - conj V_{M-k}) conj(W^k), k < M, whose backward
// M-point transform is v_2j + i v_2j+1 (the C2R folding of fft_c2r_kernel).  rtwtab holds
// W^k, k <= M (gen_rtwiddles); it is unused otherwise.
template <typename T, class S, int TPL, int LPB, int PAD, int KIND, bool FULL, bool HALF = false,
          bool IN_COLS = false, bool OUT_COLS = false, int OSEG = 0>
__global__ void __launch_bounds__(TPL * LPB)
fft_r2r_kernel(const T* __restrict__ in, T* __restrict__ out, const typename tw_of<T>::type* __restrict__ tw,
               const cx<T>* __restrict__ t4tab, const typename tw_of<T>::type* __restrict__ rtwtab, r2r_desc d) {
    typedef dtfft::fft_tile<T, S, TPL, LPB, false, false, PAD> F;
    typedef typename tw_of<T>::type tw_t;
    constexpr int M = S::N;
    constexpr int NX = HALF ? 2 * M : r2r_len<KIND, M>::n;
    constexpr bool INV = (KIND == 4 || KIND == 8);  // types 3 run the backward transform
    static_assert(!HALF || KIND == 4 || KIND == 5 || KIND == 8 || KIND == 9, "half length: types 2 and 3");
    static_assert(!(IN_COLS || OUT_COLS) || FULL, "column layouts run whole tiles");
    static_assert(OSEG == 0 || (!IN_COLS && !OUT_COLS), "chunked output: the row form");
    constexpr int XIT = (LPB * NX + F::NT - 1) / F::NT;  // real elements per thread
    constexpr int MIT = (LPB * M + F::NT - 1) / F::NT;   // work elements per thread
    constexpr int TWN = S::TW_SIZE > 0 ? S::TW_SIZE : 1;
    // The column forms run 1024-thread blocks at the 128-VGPR limit: their pre- and
    // post-process loops are unrolled twice, which keeps the FP32 form's spills to a few
    // registers (61-72 fully unrolled).  Measured on MI355X at 512^3 FP32: the column stages
    // 313 / 309 / 328 us unrolled twice, against 322 / 312 / 330 us rolled and 313 / 307 /
    // 344 us unrolled four times.
    constexpr int MU = (IN_COLS || OUT_COLS) ? 2 : MIT, XU = (IN_COLS || OUT_COLS) ? 2 : XIT;
    // Types 3, HALF, M even, columns in: the fold works on pairs (m, M - m).  Measured on
    // MI355X at 512^3 FP32 (DCT-3): the column-in stages 311-312 / 333-334 against 334 /
    // 353 us; the row stage and the same pairing of the types-2 untangling were no faster
    // (the column form of the latter 363 against 324 us, at the 128-VGPR limit).
    constexpr bool PAIRED = HALF && INV && IN_COLS && M % 2 == 0;
    constexpr int H = M / 2;
    constexpr int PIT = (LPB * H + F::NT - 1) / F::NT > 0 ? (LPB * H + F::NT - 1) / F::NT : 1;
    __shared__ cx<T> lds[F::LDS_ELEMS];
    __shared__ T xs[LPB * NX];  // real lines: [line][j], or [j][line] when read across lines
    __shared__ tw_t twl[TWN];
    __shared__ cx<T> t4l[2 * NX];
    const i64 lines = d.nl * d.nb, ntiles = (lines + LPB - 1) / LPB;
    dtfft::lines_desc ld{};  // only nl is read by the LDS-to-LDS passes
    ld.nl = lines;
    for (int i = threadIdx.x; i < S::TW_SIZE; i += F::NT) twl[i] = tw[i];
    for (int i = threadIdx.x; i < 2 * NX; i += F::NT) t4l[i] = t4tab[i];
    auto t4 = [&](i64 j) { return t4l[j]; };
    __shared__ tw_t rtwl[HALF ? M + 1 : 1];
    if constexpr (HALF)
        for (int i = threadIdx.x; i <= M; i += F::NT) rtwl[i] = rtwtab[i];
    // Line li of the stage is line li % nl of batch li / nl.  Rows: element j of a line at
    // j, lines in_ls apart (the host folds evenly spaced batches into one range).  Columns
    // (IN_COLS / OUT_COLS): LPB adjacent lines (ls 1), element j at j * es; nl is a multiple
    // of LPB, so a tile never straddles two batches.
    auto in_base = [&](i64 l0) {
        const i64 b = l0 / d.nl;
        return b * d.in_bs + (d.in_ts ? (l0 - b * d.nl) / LPB * d.in_ts : (l0 - b * d.nl) * d.in_ls);
    };
    auto out_base = [&](i64 l0) { const i64 b = l0 / d.nl; return b * d.out_bs + (l0 - b * d.nl) * d.out_ls; };
    // element q of this thread in a tile: rows (line, j) = ((tid + q NT) / NX, % NX); columns
    // (line, j) = ((tid + q NT) % LPB, / LPB), so a wave reads whole LPB-line pieces of rows
    T pf[XIT];
    auto fetch = [&](i64 t) {
        const i64 l0 = (t < ntiles ? t : ntiles - 1) * LPB;
        if constexpr (IN_COLS) {
            const T* src = in + in_base(l0);
#pragma unroll
            for (int q = 0; q < XIT; q++) {
                const int e = (int)threadIdx.x + q * F::NT;
                const int ee = XIT * F::NT == LPB * NX || e < LPB * NX ? e : 0;
                pf[q] = src[(i64)(ee % LPB) * d.in_ls + (i64)(ee / LPB) * d.in_es];
            }
        } else {
#pragma unroll
            for (int q = 0; q < XIT; q++) {
                const int e = (int)threadIdx.x + q * F::NT;
                const int line = e / NX, j = e % NX;
                const i64 li = l0 + line < lines ? l0 + line : lines - 1;
                pf[q] = in[li * d.in_ls + (XIT * F::NT == LPB * NX || e < LPB * NX ? j : 0)];
            }
        }
    };
    auto stage_lines = [&]() {  // the prefetched lines into LDS
#pragma unroll
        for (int q = 0; q < XIT; q++) {
            const int e = (int)threadIdx.x + q * F::NT;
            if (XIT * F::NT == LPB * NX || e < LPB * NX) xs[e] = pf[q];
        }
    };
    i64 t = blockIdx.x;
    fetch(t);
    stage_lines();
    while (t < ntiles) {
        const i64 l0 = t * LPB;
        fetch(t + gridDim.x);
        __syncthreads();  // the tile's lines (and, the first time, the twiddles) are in LDS
        if constexpr (PAIRED) {
            // Pairs (m, M - m), m < M / 2, share their inputs: with e2 = V_m + conj V_{M-m} and
            // o2 = (V_m - conj V_{M-m}) conj(W^m), Z_m = e2 + i o2 and, as W^{M-m} = -conj W^m,
            // Z_{M-m} = conj(e2) + i conj(o2).  Item 0 takes m = 0 (V_0 and V_M real) and the
            // self-paired M / 2.  The work line holds conj Z (backward transform).
#pragma unroll 1
            for (int q = 0; q < PIT; q++) {
                const int e = (int)threadIdx.x + q * F::NT;
                if (PIT * F::NT == LPB * H || e < LPB * H) {
                    const int line = IN_COLS ? e % LPB : e / H, m = IN_COLS ? e / LPB : e % H;
                    auto xf = [&](i64 j) { return IN_COLS ? xs[j * LPB + line] : xs[line * NX + j]; };
                    auto V = [&](int j) { return r2r_pre_elem<T, KIND>(xf, (i64)NX, (i64)NX, (i64)j, t4); };
                    cx<T>* L = lds + line * F::LD;
                    cx<T> e2, o2;
                    auto fold = [&](int mm, const cx<T>& a, const cx<T>& b) {
                        const cx<T> bc = conj(b);
                        e2 = a + bc;
                        o2 = twmul(a - bc, conj_tw(rtwl[mm]));
                    };
                    if (m == 0) {
                        cx<T> a = V(0), b = V(M);
                        a.y = (T)0;
                        b.y = (T)0;
                        fold(0, a, b);
                        L[F::lidx(0)] = mk(e2.x - o2.y, -(e2.y + o2.x));
                        const cx<T> h = V(H);
                        fold(H, h, h);
                        L[F::lidx(H)] = mk(e2.x - o2.y, -(e2.y + o2.x));
                    } else {
                        fold(m, V(m), V(M - m));
                        L[F::lidx(m)] = mk(e2.x - o2.y, -(e2.y + o2.x));
                        L[F::lidx(M - m)] = mk(e2.x + o2.y, e2.y - o2.x);
                    }
                }
            }
        } else {
#pragma unroll MU
        for (int q = 0; q < MIT; q++) {
            const int e = (int)threadIdx.x + q * F::NT;
            if (MIT * F::NT == LPB * M || e < LPB * M) {
                // columns: lines fastest, so the reads of a wave go across xs's line axis
                const int line = IN_COLS ? e % LPB : e / M, m = IN_COLS ? e / LPB : e % M;
                auto xf = [&](i64 j) { return IN_COLS ? xs[j * LPB + line] : xs[line * NX + j]; };
                cx<T> z;
                if constexpr (!HALF) {
                    z = r2r_pre_elem<T, KIND>(xf, (i64)NX, (i64)M, (i64)m, t4);
                } else if constexpr (INV) {
                    // V of the full-length map at m and M - m (V_M is real), folded
                    cx<T> a = r2r_pre_elem<T, KIND>(xf, (i64)NX, (i64)NX, (i64)m, t4);
                    cx<T> b = r2r_pre_elem<T, KIND>(xf, (i64)NX, (i64)NX, (i64)(M - m), t4);
                    if (m == 0) { a.y = (T)0; b.y = (T)0; }
                    b = conj(b);
                    const cx<T> e2 = a + b;
                    const cx<T> o2 = twmul(a - b, conj_tw(rtwl[m]));
                    z = mk(e2.x - o2.y, e2.y + o2.x);
                } else {
                    z = mk(r2r_pre_elem<T, KIND>(xf, (i64)NX, (i64)NX, (i64)(2 * m), t4).x,
                           r2r_pre_elem<T, KIND>(xf, (i64)NX, (i64)NX, (i64)(2 * m + 1), t4).x);
                }
                // backward transform = conj(forward(conj z))
                lds[line * F::LD + F::lidx(m)] = INV ? conj(z) : z;
            }
        }
        }
        __syncthreads();
        F::template passes<0, false, false>(nullptr, nullptr, twl, lds, 0, 0, ld, l0, false);
        __syncthreads();
        // chunked rows (OSEG): the host keeps nl a multiple of LPB, so a tile lies in one batch
        // and its lines' bases follow from the tile's (no division per store)
        const i64 ob = (IN_COLS || OUT_COLS || OSEG > 0) ? out_base(l0) : 0;
#pragma unroll XU
        for (int q = 0; q < XIT; q++) {
            const int e = (int)threadIdx.x + q * F::NT;
            if (XIT * F::NT == LPB * NX || e < LPB * NX) {
                // rows out: a wave writes consecutive k of a line; columns out: lines fastest
                const int line = OUT_COLS ? e % LPB : e / NX, k = OUT_COLS ? e / LPB : e % NX;
                const i64 li = l0 + line;
                if (FULL || li < lines) {
                    const cx<T>* W = lds + line * F::LD;
                    T y;
                    if constexpr (!HALF) {
                        auto wf = [&](i64 j) {
                            const cx<T> c = W[F::lidx((int)j)];
                            return INV ? conj(c) : c;
                        };
                        y = r2r_post_elem<T, KIND>(wf, (i64)NX, (i64)k, t4);
                    } else if constexpr (INV) {
                        // v_j is the real (j even) or imaginary (j odd) part of u_{j / 2}
                        auto wf = [&](i64 j) {
                            const cx<T> c = conj(W[F::lidx((int)(j >> 1))]);
                            return mk((j & 1) ? c.y : c.x, (T)0);
                        };
                        y = r2r_post_elem<T, KIND>(wf, (i64)NX, (i64)k, t4);
                    } else {
                        // V_k untangled from U, V_{n-k} = conj V_k
                        auto vf = [&](i64 kf) {
                            const int kk = kf <= M ? (int)kf : NX - (int)kf;
                            const cx<T> a = W[F::lidx(kk == M ? 0 : kk)];
                            const cx<T> b = conj(W[F::lidx(kk == 0 ? 0 : M - kk)]);
                            const cx<T> s = a + b, df = a - b;
                            const cx<T> w = twmul(df, rtwl[kk]);
                            const T half = (T)0.5;
                            const cx<T> v = mk(half * (s.x + w.y), half * (s.y - w.x));
                            return kf <= M ? v : conj(v);
                        };
                        y = r2r_post_elem<T, KIND>(vf, (i64)NX, (i64)k, t4);
                    }
                    // rows out: non-temporal stores (measured on MI355X at 512^3 FP32: the
                    // column-in / row-out stages 324-326 against 332 us, the row stages
                    // unchanged within the noise); columns out: plain stores
                    if constexpr (OUT_COLS) out[ob + (i64)line * d.out_ls + (i64)k * d.out_es] = y;
                    else if constexpr (IN_COLS)
                        __builtin_nontemporal_store(y, &out[ob + (i64)line * d.out_ls + (i64)k * d.out_es]);
                    else if constexpr (OSEG > 0)
                        __builtin_nontemporal_store(y, &out[ob + (i64)line * d.out_ls + (i64)(k / OSEG) * d.out_seg + k % OSEG]);
                    else __builtin_nontemporal_store(y, &out[li * d.out_ls + k]);
                }
            }
        }
        t += gridDim.x;
        if (t >= ntiles) break;
        // Only this path reaches here, with the tile's stores issued after the prefetch, so
        // the wait for the prefetch lets those stores stay in flight.  The real lines were
        // last read by the pre-process, two barriers ago; the work lines that the
        // post-process reads are next written after the coming barrier.
        stage_lines();
    }
}
