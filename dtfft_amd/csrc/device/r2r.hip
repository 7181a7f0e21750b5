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
// Twiddles e^{i pi a} are evaluated in FP64 (sincospi) for both precisions.
#include "common.hip"

struct r2r_desc {
    i64 nl, nb;                // lines (l, b): line start l * ls + b * bs
    i64 in_ls, in_bs, in_es;   // input real line / batch / element strides
    i64 out_ls, out_bs, out_es;
    i64 n, m;                  // real length, complex work length
    i64 wdist;                 // complex work line stride (>= m)
};

template <typename T>
__device__ __forceinline__ cx<T> cis_pi(double a) {  // e^{i pi a}
    double s, c;
    sincospi(a, &s, &c);
    return mk((T)c, (T)s);
}

template <typename T, int KIND>
__global__ void __launch_bounds__(256) r2r_pre(const T* __restrict__ in, cx<T>* __restrict__ work, r2r_desc d) {
    const i64 n = d.n, M = d.m, total = d.nl * d.nb * M;
    const i64 h = (n + 1) / 2;
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (i64)gridDim.x * blockDim.x) {
        const i64 li = e / M, m = e % M;
        const i64 l = li % d.nl, b = li / d.nl;
        const T* x = in + l * d.in_ls + b * d.in_bs;
        cx<T> z = mk((T)0, (T)0);
        if constexpr (KIND == 5 || KIND == 9) {
            const T sg = (KIND == 9) ? (T)-1 : (T)1;
            z.x = m < h ? x[(2 * m) * d.in_es] : sg * x[(2 * (n - 1 - m) + 1) * d.in_es];
        } else if constexpr (KIND == 4 || KIND == 8) {
            T a, bb;
            if (KIND == 4) { a = x[m * d.in_es]; bb = m > 0 ? x[(n - m) * d.in_es] : (T)0; }
            else { a = x[(n - 1 - m) * d.in_es]; bb = m > 0 ? x[(m - 1) * d.in_es] : (T)0; }
            const cx<T> w = cis_pi<T>((double)m / (double)(2 * n));
            z = mk(a * w.x + bb * w.y, a * w.y - bb * w.x);  // (a - i bb) e^{i pi m / 2n}
        } else if constexpr (KIND == 6 || KIND == 10) {
            if (m < n) {
                const T a = KIND == 6 ? x[m * d.in_es] : x[(n - 1 - m) * d.in_es];
                const cx<T> w = cis_pi<T>(-(double)m / (double)(2 * n));
                z = mk(a * w.x, a * w.y);
            }
        } else if constexpr (KIND == 3) {
            z.x = m < n ? x[m * d.in_es] : x[(M - m) * d.in_es];
        } else if constexpr (KIND == 7) {
            if (m >= 1 && m <= n) z.x = x[(m - 1) * d.in_es];
            else if (m >= n + 2) z.x = -x[(M - 1 - m) * d.in_es];
        }
        work[li * d.wdist + m] = z;
    }
}

template <typename T, int KIND>
__global__ void __launch_bounds__(256) r2r_post(const cx<T>* __restrict__ work, T* __restrict__ out, r2r_desc d) {
    const i64 n = d.n, total = d.nl * d.nb * n;
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (i64)gridDim.x * blockDim.x) {
        const i64 li = e / n, k = e % n;
        const i64 l = li % d.nl, b = li / d.nl;
        const cx<T>* W = work + li * d.wdist;
        T y;
        if constexpr (KIND == 5 || KIND == 9) {
            const i64 kk = KIND == 9 ? n - 1 - k : k;
            const cx<T> w = cis_pi<T>(-(double)kk / (double)(2 * n));
            const cx<T> c = W[kk];
            y = (T)2 * (w.x * c.x - w.y * c.y);
        } else if constexpr (KIND == 4 || KIND == 8) {
            const i64 idx = (k % 2 == 0) ? k / 2 : n - 1 - (k - 1) / 2;
            y = W[idx].x;
            if (KIND == 8 && (k & 1)) y = -y;
        } else if constexpr (KIND == 6 || KIND == 10) {
            const cx<T> w = cis_pi<T>(-(double)(2 * k + 1) / (double)(4 * n));
            const cx<T> c = W[k];
            y = (T)2 * (w.x * c.x - w.y * c.y);
            if (KIND == 10 && (k & 1)) y = -y;
        } else if constexpr (KIND == 3) {
            y = W[k].x;
        } else {
            y = -W[k + 1].y;
        }
        out[l * d.out_ls + b * d.out_bs + k * d.out_es] = y;
    }
}
