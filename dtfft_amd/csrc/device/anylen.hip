// Transforms of any length: C2C, R2C and C2R lines whose length has a prime factor above 13
// (or needs more radix passes than a schedule holds), and real transforms of odd length.
// cuFFT and FFTW, the executors this library stands in for, accept every length
// (src/interfaces/fft/cufft/dtfft_executor_cufft_m.F90:77-93), and so must this one.
//
// A line of logical length n is carried through an M-point Stockham transform held in LDS
// (the passes of fft_tile in device/fft.hip):
//   CHIRP = false, M = n: the plain transform; real transforms of odd length take this
//     form, since the half-length complex trick of fft_r2c_kernel needs n even.
//   CHIRP = true, M >= 2n - 1 (Bluestein's chirp z-transform).  With jk = (j^2 + k^2 -
//     (k - j)^2) / 2 and the chirp c_j = exp(-i pi j^2 / n),
//       X_k = c_k sum_j (x_j c_j) conj(c_{k-j}),
//     a linear convolution that a cyclic one of length M >= 2n - 1 holds exactly:
//       X_k = c_k IDFT_M(DFT_M(a) DFT_M(b))_k,  a_j = x_j c_j (j < n, else 0),
//       b_j = conj(c_j) for j < n, b_{M-j} = conj(c_j) for 0 < j < n, else 0.
//     bh = DFT_M(b) / M is a table made once per (n, M); the inverse transform is
//     conj(DFT_M(conj(.))), so both transforms run the forward passes.
// KIND 0: C2C, complex in and out; the backward sign conjugates on load and store.
// KIND 1: R2C, n reals in, n/2 + 1 complex out (forward).
// KIND 2: C2R, n/2 + 1 complex in, n reals out (backward, FFTW c2r semantics: the input is
//         extended to the Hermitian spectrum X_{n-k} = conj X_k, Im X_0 and, for even n,
//         Im X_{n/2} do not contribute).
// Lines are addressed as in fft_tile, (l, b) with line start l * ls + b * bs, element
// stride ns, in units of the element type of each side (reals for the real side).
// Each block reads its lines whole before it writes, so in place is safe.
#include "fft.hip"

template <typename T, class S, int TPL, int LPB, int PAD, int KIND, bool CHIRP>
__global__ void __launch_bounds__(TPL * LPB)
fft_any_kernel(const void* __restrict__ in_, void* __restrict__ out_,
               const typename tw_of<T>::type* __restrict__ tw,
               const typename tw_of<T>::type* __restrict__ chirp,
               const typename tw_of<T>::type* __restrict__ bh, dtfft::lines_desc d, int inv, int n) {
    typedef dtfft::fft_tile<T, S, TPL, LPB, false, false, PAD> F;
    constexpr int M = S::N;
    __shared__ cx<T> lds[F::LDS_ELEMS];
    const bool iv = KIND == 2 ? true : (KIND == 1 ? false : inv != 0);
    const int nout = KIND == 1 ? n / 2 + 1 : n;
    // thread mapping along the lines where lines are adjacent in memory, else along them
    const bool lfast_in = d.in_ls == 1 && d.in_ns != 1;
    const bool lfast_out = d.out_ls == 1 && d.out_ns != 1;
    for (i64 t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        i64 l0, ibase, obase;
        F::tile_bases(d, t, l0, ibase, obase);
        for (int e = threadIdx.x; e < LPB * M; e += F::NT) {
            int line, m;
            if (lfast_in) { line = e % LPB; m = e / LPB; }
            else { line = e / M; m = e % M; }
            cx<T> v = mk((T)0, (T)0);
            if (m < n && l0 + line < d.nl) {
                const i64 lb = ibase + (i64)line * d.in_ls;
                if constexpr (KIND == 0) {
                    v = ((const cx<T>*)in_)[lb + (i64)m * d.in_ns];
                    if (iv) v = conj(v);
                } else if constexpr (KIND == 1) {
                    v = mk(((const T*)in_)[lb + (i64)m * d.in_ns], (T)0);
                } else {
                    const cx<T>* in = (const cx<T>*)in_;
                    const int h = n / 2;
                    if (m <= h) v = in[lb + (i64)m * d.in_ns];
                    else v = conj(in[lb + (i64)(n - m) * d.in_ns]);
                    if (m == 0 || 2 * m == n) v.y = (T)0;
                    v = conj(v);  // backward = conj(forward(conj))
                }
                if constexpr (CHIRP) v = twmul(v, chirp[m]);
            }
            lds[line * F::LD + F::lidx(m)] = v;
        }
        __syncthreads();
        F::template passes<0, false, false>(nullptr, nullptr, tw, lds, 0, 0, d, l0, false);
        __syncthreads();
        if constexpr (CHIRP) {
            // pointwise product with DFT_M(b) / M, conjugated for the inverse transform
            for (int e = threadIdx.x; e < LPB * M; e += F::NT) {
                const int line = e / M, k = e % M;
                cx<T>* r = &lds[line * F::LD + F::lidx(k)];
                *r = conj(twmul(*r, bh[k]));
            }
            __syncthreads();
            F::template passes<0, false, false>(nullptr, nullptr, tw, lds, 0, 0, d, l0, false);
            __syncthreads();
        }
        for (int e = threadIdx.x; e < LPB * nout; e += F::NT) {
            int line, k;
            if (lfast_out) { line = e % LPB; k = e / LPB; }
            else { line = e / nout; k = e % nout; }
            if (l0 + line >= d.nl) continue;
            cx<T> y = lds[line * F::LD + F::lidx(k)];
            if constexpr (CHIRP) y = twmul(conj(y), chirp[k]);
            const i64 o = obase + (i64)line * d.out_ls + (i64)k * d.out_ns;
            if constexpr (KIND == 0) ((cx<T>*)out_)[o] = iv ? conj(y) : y;
            else if constexpr (KIND == 1) ((cx<T>*)out_)[o] = y;
            else ((T*)out_)[o] = y.x;  // real part; the conjugation leaves it unchanged
        }
        __syncthreads();
    }
}

// Chirp c_j = exp(-i pi j^2 / n), j < n, with j^2 reduced mod 2n in integers so that the
// FP64 angle stays exact for every j.
__device__ __forceinline__ void store_tw(twf* t, i64 i, double c, double s) {
    twf w;
    w.rh = (float)c; w.rl = (float)(c - (double)w.rh);
    w.ih = (float)s; w.il = (float)(s - (double)w.ih);
    t[i] = w;
}
__device__ __forceinline__ void store_tw(cx<double>* t, i64 i, double c, double s) { t[i] = mk(c, s); }

template <typename T>
__global__ void gen_chirp(typename tw_of<T>::type* out, int n) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const i64 r = (j * j) % (2 * (i64)n);
    double s, c;
    sincospi(-(double)r / (double)n, &s, &c);
    store_tw(out, j, c, s);
}

// bh_k = DFT_M(b)_k / M = (1 + 2 sum_{j=1}^{n-1} cos(2 pi j k / M) exp(i pi j^2 / n)) / M:
// the two halves of b pair up as exp(-2 pi i j k / M) + exp(+2 pi i j k / M).  One block per
// k, FP64 sums over j reduced through LDS; j k and j^2 are reduced in integers.
template <typename T>
__global__ void __launch_bounds__(256) gen_bluestein_bh(typename tw_of<T>::type* out, int n, int M) {
    const i64 k = blockIdx.x;
    double sr = 0.0, si = 0.0;
    for (i64 j = 1 + threadIdx.x; j < n; j += blockDim.x) {
        const i64 a = (j * k) % (i64)M;
        const i64 r = (j * j) % (2 * (i64)n);
        const double cw = cospi(2.0 * (double)a / (double)M);
        double s, c;
        sincospi((double)r / (double)n, &s, &c);
        sr += cw * c;
        si += cw * s;
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) store_tw(out, k, (1.0 + 2.0 * red[0][0]) / (double)M, 2.0 * red[1][0] / (double)M);
}
