// Batched 1D C2C Stockham FFT kernels (register butterflies, LDS exchange).
//
// One thread block transforms LPB lines of length N.  The N-point transform
// is a sequence of radix-R passes (R in 2,3,4,5,7,8,11,13,16); pass p with span
// L = R_0*...*R_{p-1} reads x[j + r*N/R], multiplies by w_{L*R}^{r*(j%L)},
// runs an R-point DFT in registers and writes y[(j - j%L)*R + j%L + s*L].
// The first pass reads global memory directly and the last pass writes it
// directly; the passes in between exchange through LDS.
//
// Lines are addressed as (l, b): element n of line (l, b) sits at
//     base + l*ls + b*bs + n*ns
// so the same kernel covers the contiguous executor layout (ns = 1) and the
// "transpose fused into the next FFT" layout where the FFT axis is strided and
// consecutive lines are contiguous (ls = 1).  IN_COLS / OUT_COLS select which
// thread mapping makes the global access coalesced: along n (rows) or along l
// (columns).
//
// Forward transforms use exp(-2*pi*i/N); backward ones conjugate on load and
// store (IDFT(x) = conj(DFT(conj(x)))), both unnormalised as in FFTW.
#include "common.hip"

namespace dtfft {

// ---------------------------------------------------------------------------
// Small DFT butterflies, forward sign, in place on v[0..R-1].
// ---------------------------------------------------------------------------
template <typename T, int R> struct bfly;

template <typename T> struct bfly<T, 2> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        cx<T> a = v[0] + v[1], b = v[0] - v[1];
        v[0] = a; v[1] = b;
    }
};

template <typename T> struct bfly<T, 4> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        cx<T> s0 = v[0] + v[2], d0 = v[0] - v[2];
        cx<T> s1 = v[1] + v[3], d1 = mul_mi(v[1] - v[3]);
        v[0] = s0 + s1; v[2] = s0 - s1;
        v[1] = d0 + d1; v[3] = d0 - d1;
    }
};

template <typename T> struct bfly<T, 8> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        const T h = (T)0.7071067811865476;
        cx<T> a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) { a[i] = v[i] + v[i + 4]; b[i] = v[i] - v[i + 4]; }
        b[1] = cmulc(b[1], h, -h);
        b[2] = mul_mi(b[2]);
        b[3] = cmulc(b[3], -h, -h);
        bfly<T, 4>::run(a);
        bfly<T, 4>::run(b);
#pragma unroll
        for (int i = 0; i < 4; i++) { v[2 * i] = a[i]; v[2 * i + 1] = b[i]; }
    }
};

template <typename T> struct bfly<T, 16> {
    static __device__ __forceinline__ void run(cx<T>* v) {
        // 16 = 4 x 4: DFT over n2 for each n1, twiddle w16^(n1*k0), DFT over n1.
        const T c1 = (T)0.9238795325112867, s1 = (T)0.3826834323650898, h = (T)0.7071067811865476;
        cx<T> u[4][4];
#pragma unroll
        for (int n1 = 0; n1 < 4; n1++) {
            cx<T> t[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
            bfly<T, 4>::run(t);
#pragma unroll
            for (int k0 = 0; k0 < 4; k0++) u[k0][n1] = t[k0];
        }
        // twiddles w16^(n1*k0) = exp(-2*pi*i*n1*k0/16)
        u[1][1] = cmulc(u[1][1], c1, -s1);
        u[1][2] = cmulc(u[1][2], h, -h);
        u[1][3] = cmulc(u[1][3], s1, -c1);
        u[2][1] = cmulc(u[2][1], h, -h);
        u[2][2] = mul_mi(u[2][2]);
        u[2][3] = cmulc(u[2][3], -h, -h);
        u[3][1] = cmulc(u[3][1], s1, -c1);
        u[3][2] = cmulc(u[3][2], -h, -h);
        u[3][3] = cmulc(u[3][3], -c1, s1);
#pragma unroll
        for (int k0 = 0; k0 < 4; k0++) {
            bfly<T, 4>::run(u[k0]);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) v[k0 + 4 * k1] = u[k0][k1];
        }
    }
};

// Odd prime radices: y_0 = sum x_r; for s = 1..(p-1)/2
//   a_s = x_0 + sum_r (x_r + x_{p-r}) cos(2 pi r s / p)
//   b_s = sum_r (x_r - x_{p-r}) sin(2 pi r s / p)
//   y_s = a_s - i b_s, y_{p-s} = a_s + i b_s.
template <int P> struct prime_consts;
template <> struct prime_consts<3> {
    static __device__ __forceinline__ double c(int k) { return -0.5; }
    static __device__ __forceinline__ double s(int k) { return 0.8660254037844386; }
};
template <> struct prime_consts<5> {
    static __device__ __forceinline__ double c(int k) {
        return k == 1 ? 0.30901699437494745 : -0.8090169943749475;
    }
    static __device__ __forceinline__ double s(int k) {
        return k == 1 ? 0.9510565162951535 : 0.5877852522924731;
    }
};
template <> struct prime_consts<7> {
    static __device__ __forceinline__ double c(int k) {
        return k == 1 ? 0.6234898018587335 : (k == 2 ? -0.2225209339563144 : -0.9009688679024191);
    }
    static __device__ __forceinline__ double s(int k) {
        return k == 1 ? 0.7818314824680298 : (k == 2 ? 0.9749279121818236 : 0.4338837391175581);
    }
};
template <> struct prime_consts<11> {
    static __device__ __forceinline__ double c(int k) {
        const double t[5] = {0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                             -0.6548607339452851, -0.9594929736144974};
        return t[k - 1];
    }
    static __device__ __forceinline__ double s(int k) {
        const double t[5] = {0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                             0.7557495743542583, 0.28173255684142967};
        return t[k - 1];
    }
};
template <> struct prime_consts<13> {
    static __device__ __forceinline__ double c(int k) {
        const double t[6] = {0.8854560256532099, 0.5680647467311558, 0.12053668025532305,
                             -0.3546048870425356, -0.7485107481711011, -0.970941817426052};
        return t[k - 1];
    }
    static __device__ __forceinline__ double s(int k) {
        const double t[6] = {0.46472317204376856, 0.8229838658936564, 0.992708874098054,
                             0.9350162426854148, 0.6631226582407952, 0.23931566428755777};
        return t[k - 1];
    }
};

template <typename T, int P> struct bfly_prime {
    static __device__ __forceinline__ void run(cx<T>* v) {
        constexpr int H = (P - 1) / 2;
        cx<T> sp[H], sm[H];
        cx<T> y0 = v[0];
#pragma unroll
        for (int r = 1; r <= H; r++) {
            sp[r - 1] = v[r] + v[P - r];
            sm[r - 1] = v[r] - v[P - r];
            y0 = y0 + sp[r - 1];
        }
        cx<T> out[P];
        out[0] = y0;
#pragma unroll
        for (int s = 1; s <= H; s++) {
            cx<T> a = v[0], b = mk((T)0, (T)0);
#pragma unroll
            for (int r = 1; r <= H; r++) {
                const int q = (r * s) % P;  // cos/sin(2 pi q / P), folded into 1..H
                const int qq = q <= H ? q : P - q;
                const T c = (T)prime_consts<P>::c(qq);
                const T sn = (T)(q <= H ? prime_consts<P>::s(qq) : -prime_consts<P>::s(qq));
                a = a + scale(sp[r - 1], c);
                b = b + scale(sm[r - 1], sn);
            }
            // y_s = a - i b ; y_{P-s} = a + i b
            out[s] = mk(a.x + b.y, a.y - b.x);
            out[P - s] = mk(a.x - b.y, a.y + b.x);
        }
#pragma unroll
        for (int i = 0; i < P; i++) v[i] = out[i];
    }
};
template <typename T> struct bfly<T, 3> : bfly_prime<T, 3> {};
template <typename T> struct bfly<T, 5> : bfly_prime<T, 5> {};
template <typename T> struct bfly<T, 7> : bfly_prime<T, 7> {};
template <typename T> struct bfly<T, 11> : bfly_prime<T, 11> {};
template <typename T> struct bfly<T, 13> : bfly_prime<T, 13> {};

// ---------------------------------------------------------------------------
// Radix schedule (compile time).
// ---------------------------------------------------------------------------
template <int N_, int R0, int R1 = 1, int R2 = 1, int R3 = 1, int R4 = 1, int R5 = 1, int R6 = 1, int R7 = 1,
          int R8 = 1, int R9 = 1>
struct sched {
    static constexpr int N = N_;
    static constexpr int NP = (R0 > 1) + (R1 > 1) + (R2 > 1) + (R3 > 1) + (R4 > 1) + (R5 > 1) + (R6 > 1) + (R7 > 1) +
                              (R8 > 1) + (R9 > 1);
    static constexpr __device__ __host__ int radix(int p) {
        return p == 0 ? R0 : p == 1 ? R1 : p == 2 ? R2 : p == 3 ? R3 : p == 4 ? R4 : p == 5 ? R5 : p == 6 ? R6
               : p == 7 ? R7 : p == 8 ? R8 : R9;
    }
    static constexpr __device__ __host__ int span(int p) {  // L_p = product of radices before p
        return p == 0 ? 1 : span(p - 1) * radix(p - 1);
    }
    // offset of pass p's twiddles in the table: sum over q < p of L_q * (R_q - 1)
    static constexpr __device__ __host__ int tw_off(int p) {
        return p == 0 ? 0 : tw_off(p - 1) + span(p - 1) * (radix(p - 1) - 1);
    }
    static constexpr int TW_SIZE = tw_off(NP);
};

// Launch geometry of one kernel: (l, b) line space and strides in elements.
struct lines_desc {
    i64 nl, nb;          // lines per row, rows
    i64 in_ls, in_bs, in_ns;
    i64 out_ls, out_bs, out_ns;
    i64 ntl;             // ceil(nl / LPB)
    i64 ntiles;          // ntl * nb
    i64 in_seg_stride;   // SEG > 0: input element n sits at (n / SEG) * in_seg_stride + (n % SEG) * in_ns
    // Column-in tiles: input tile tl starts at tl * in_ts (0: tl * LPB * in_ls, lines side by
    // side).  A chunk-blocked input, where the LPB lines of a tile are stored as one contiguous
    // block of N rows of LPB elements (in_ns = LPB), has in_ts = N * LPB.
    i64 in_ts;
    i64 out_seg_stride;  // OSEG > 0: output element n sits at (n / OSEG) * out_seg_stride + (n % OSEG) * out_ns
};

// SEG > 0 reads the FFT axis in segments of SEG elements that lie in_seg_stride apart:
// the per-peer chunks of a receive buffer, so that unpack and FFT are one pass.
// OSEG > 0 writes the FFT axis in segments of OSEG elements that lie out_seg_stride apart:
// the chunk-blocked layout that a following column-in pass reads as contiguous tiles.
template <typename T, class S, int TPL, int LPB, bool IN_COLS, bool OUT_COLS, int PAD, int SEG = 0, int OSEG = 0>
struct fft_tile {
    static __device__ __forceinline__ i64 in_elem(int n, const lines_desc& d) {
        if constexpr (SEG > 0) return (i64)(n / SEG) * d.in_seg_stride + (i64)(n % SEG) * d.in_ns;
        else return (i64)n * d.in_ns;
    }
    static __device__ __forceinline__ i64 out_elem(int n, const lines_desc& d) {
        if constexpr (OSEG > 0) return (i64)(n / OSEG) * d.out_seg_stride + (i64)(n % OSEG) * d.out_ns;
        else return (i64)n * d.out_ns;
    }
    typedef typename tw_of<T>::type tw_t;
    static constexpr int N = S::N;
    static constexpr int NP = S::NP;
    // LDS layout of a line.  PAD 0 / 1: element n at n, line stride N + PAD.  PAD 2 / 3
    // ("spread"): one spare slot after every 16 elements, n at n + n / 16, line stride
    // N + N / 16 + (PAD & 1) made odd.  A Stockham pass writes its outputs R elements apart
    // across lanes (pass 0: n = R j + s); unspread, 8-byte elements R = 8 apart land on every
    // 16th bank and serialise up to 16 ways; spread, consecutive lanes shift by one bank pair.
    static constexpr bool SPREAD = (PAD & 2) != 0;
    static __device__ __host__ constexpr int lidx(int n) { return SPREAD ? n + (n >> 4) : n; }
    // LDS distance of m elements (m a multiple of 16): lidx(n + m) == lidx(n) + lspan(m)
    static __device__ __host__ constexpr int lspan(int m) { return SPREAD ? m + m / 16 : m; }
    static constexpr int LD = SPREAD ? ((N + (N >> 4) + (PAD & 1)) | 1) : N + PAD;  // line stride
    static constexpr int LDS_ELEMS = LPB * LD;
    static constexpr int TW_SIZE = S::TW_SIZE;  // twiddles of the schedule
    static constexpr int NT = TPL * LPB;

    // thread -> (line, j) mapping: rows (j fastest) or columns (line fastest)
    template <bool COLS>
    static __device__ __forceinline__ void map(int tid, int& line, int& j) {
        if (COLS) { line = tid % LPB; j = tid / LPB; }
        else { line = tid / TPL; j = tid % TPL; }
    }

    // One radix pass.  SRC_G: read global (first pass); DST_G: write global (last pass);
    // otherwise the pass reads / writes the line in LDS (in place, with a barrier between).
    template <int p, bool SRC_G, bool DST_G, bool COLS>
    static __device__ __forceinline__ void pass(const cx<T>* __restrict__ gin, cx<T>* __restrict__ gout,
                                                const tw_t* __restrict__ tw, cx<T>* lds,
                                                i64 ibase, i64 obase, const lines_desc& d, i64 l0,
                                                bool inv) {
        constexpr int R = S::radix(p);
        constexpr int L = S::span(p);
        constexpr int NB = N / R;
        constexpr int IT = (NB + TPL - 1) / TPL;
        constexpr bool EXACT = (IT * TPL == NB);
        const int tid = threadIdx.x;
        // threads beyond NT exist when several tile shapes share one block (batched
        // launch); they only take part in the barriers
        const bool act = (int)blockDim.x == NT || tid < NT;
        int line, j;
        map<COLS>(tid, line, j);
        const bool lv = act && (l0 + line) < d.nl;
        cx<T> v[IT][R];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int b = j + it * TPL;
            if (act && (EXACT || b < NB)) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int n = b + r * NB;
                    if (SRC_G) {
                        cx<T> x = mk((T)0, (T)0);
                        if (lv) x = gin[ibase + (i64)line * d.in_ls + in_elem(n, d)];
                        v[it][r] = inv ? conj(x) : x;
                    } else {
                        // NB a multiple of 16: the spare slots advance by NB / 16 per r
                        const int q = (NB % 16 == 0) ? lidx(b) + r * lspan(NB) : lidx(n);
                        v[it][r] = lds[line * LD + q];
                    }
                }
            }
        }
        if (!SRC_G && !DST_G) __syncthreads();  // everyone has read before the in-place write
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int b = j + it * TPL;
            if (act && (EXACT || b < NB)) {
                const int k = b % L;
                if (L > 1) {
#pragma unroll
                    for (int r = 1; r < R; r++)
                        v[it][r] = twmul(v[it][r], tw[S::tw_off(p) + k * (R - 1) + (r - 1)]);
                }
                bfly<T, R>::run(v[it]);
                const int base = (b - k) * R + k;
#pragma unroll
                for (int s = 0; s < R; s++) {
                    const int n = base + s * L;
                    if (DST_G) {
                        if (lv) {
                            cx<T> y = v[it][s];
                            cx<T>* o = &gout[obase + (i64)line * d.out_ls + out_elem(n, d)];
                            // rows in and rows out: non-temporal stores (measured faster on MI355X
                            // for the row passes, slower for column-in and column-out passes)
                            if constexpr (!COLS && !IN_COLS) store_nt(o, inv ? conj(y) : y);
                            else *o = inv ? conj(y) : y;
                        }
                    } else {
                        // L a multiple of 16: the spare slots advance by L / 16 per s; L = 1
                        // with R dividing 16: base is a multiple of R, so base .. base + R - 1
                        // share one 16-element group
                        const int q = (L % 16 == 0) ? lidx(base) + s * lspan(L)
                                      : (L == 1 && 16 % R == 0) ? lidx(base) + s : lidx(n);
                        lds[line * LD + q] = v[it][s];
                    }
                }
            }
        }
    }

    // Runs passes p..NP-1.  FIRST_G: pass 0 reads global (else LDS was filled and synced by
    // the caller).  LAST_G: the last pass writes global (else it leaves the line in LDS;
    // the caller syncs before reading it).
    template <int p, bool FIRST_G, bool LAST_G>
    static __device__ __forceinline__ void passes(const cx<T>* __restrict__ gin, cx<T>* __restrict__ gout,
                                                  const tw_t* __restrict__ tw, cx<T>* lds, i64 ibase,
                                                  i64 obase, const lines_desc& d, i64 l0, bool inv) {
        if constexpr (NP == 0 && p == 0 && FIRST_G != LAST_G) {
            // length-1 transform: only move the line between global memory and LDS
            for (int line = threadIdx.x; line < LPB && (int)threadIdx.x < NT; line += NT) {
                if (l0 + line >= d.nl) continue;
                if (FIRST_G) {
                    cx<T> x = gin[ibase + (i64)line * d.in_ls];
                    lds[line * LD + lidx(0)] = inv ? conj(x) : x;
                } else {
                    cx<T> y = lds[line * LD + lidx(0)];
                    gout[obase + (i64)line * d.out_ls] = inv ? conj(y) : y;
                }
            }
        }
        if constexpr (p < NP) {
            constexpr bool first = (p == 0);
            constexpr bool last = (p == NP - 1);
            if constexpr (first && FIRST_G && last) {
                // single pass from global: stage the input through LDS so that in-place use
                // is safe and the read and write mappings may differ
                for (int e = threadIdx.x; e < LPB * N && (int)threadIdx.x < NT; e += NT) {
                    int line, n;
                    if (IN_COLS) { line = e % LPB; n = e / LPB; } else { line = e / N; n = e % N; }
                    cx<T> x = mk((T)0, (T)0);
                    if (l0 + line < d.nl) x = gin[ibase + (i64)line * d.in_ls + in_elem(n, d)];
                    lds[line * LD + lidx(n)] = inv ? conj(x) : x;
                }
                __syncthreads();
                pass<p, false, LAST_G, OUT_COLS>(gin, gout, tw, lds, ibase, obase, d, l0, inv);
            } else if constexpr (first && FIRST_G) {
                pass<p, true, false, IN_COLS>(gin, gout, tw, lds, ibase, obase, d, l0, inv);
            } else if constexpr (first) {
                pass<p, false, last && LAST_G, (last && LAST_G) ? OUT_COLS : false>(gin, gout, tw, lds, ibase,
                                                                                   obase, d, l0, inv);
            } else if constexpr (last) {
                __syncthreads();
                pass<p, false, LAST_G, LAST_G ? OUT_COLS : false>(gin, gout, tw, lds, ibase, obase, d, l0, inv);
            } else {
                __syncthreads();
                pass<p, false, false, false>(gin, gout, tw, lds, ibase, obase, d, l0, inv);
            }
            passes<p + 1, FIRST_G, LAST_G>(gin, gout, tw, lds, ibase, obase, d, l0, inv);
        }
    }

    static __device__ __forceinline__ void tile_bases(const lines_desc& d, i64 tile, i64& l0, i64& ibase,
                                                      i64& obase) {
        const i64 tl = tile % d.ntl;
        const i64 b = tile / d.ntl;
        l0 = tl * LPB;
        ibase = (d.in_ts ? tl * d.in_ts : l0 * d.in_ls) + b * d.in_bs;
        obase = l0 * d.out_ls + b * d.out_bs;
    }

    static __device__ __forceinline__ void run(const cx<T>* __restrict__ gin, cx<T>* __restrict__ gout,
                                               const tw_t* __restrict__ tw, cx<T>* lds,
                                               const lines_desc& d, i64 tile, bool inv) {
        i64 l0, ibase, obase;
        tile_bases(d, tile, l0, ibase, obase);
        if constexpr (NP == 0) {
            return;
        } else {
            passes<0, true, true>(gin, gout, tw, lds, ibase, obase, d, l0, inv);
        }
    }
};

}  // namespace dtfft

// Plain batched C2C launch, grid-stride over tiles.  TWL = false: one tile per block, the
// passes read the twiddles from global memory.  TWL = true: persistent blocks (as many as are
// resident) copy the twiddle table into LDS once; used for long lines, whose table would
// otherwise share the L1 cache with the tile data.
//
// Column-in launches map tiles to blocks XCD by XCD.  Workgroups are dealt to the 8 XCDs of
// the chip in turn (block b on XCD b % 8), and each XCD has its own L2.  A column tile reads
// one LPB-line piece of every row; when rows are not a multiple of 128 bytes long (513 lines
// of an R2C pencil) every piece straddles two 128-byte lines, whose other halves belong to
// the neighbouring tiles.  Dealt in tile order, neighbours sit on different XCDs and each
// fetches both lines (about twice the bytes).  So within each round of G = gridDim.x slots,
// XCD x takes the G / 8 consecutive tiles [x G / 8, (x + 1) G / 8), and neighbours share an
// L2.  The slots run over whole rounds so that the map is a permutation of the tiles.
namespace dtfft {
__device__ __forceinline__ i64 xcd_tile(i64 slot, i64 G) {
    if (G % 8 != 0) return slot;
    const i64 r = slot / G, b = slot % G;
    return r * G + (b % 8) * (G / 8) + b / 8;
}
}  // namespace dtfft

template <typename T, class S, int TPL, int LPB, bool IN_COLS, bool OUT_COLS, int PAD, int SEG = 0, bool TWL = false,
          int OSEG = 0>
__global__ void __launch_bounds__(TPL * LPB)
fft_c2c_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
               const typename tw_of<T>::type* __restrict__ tw, dtfft::lines_desc d, int inv) {
    typedef dtfft::fft_tile<T, S, TPL, LPB, IN_COLS, OUT_COLS, PAD, SEG, OSEG> F;
    typedef typename tw_of<T>::type tw_t;
    __shared__ cx<T> lds[F::LDS_ELEMS];
    const tw_t* twp = tw;
    if constexpr (TWL) {
        __shared__ tw_t twl[F::TW_SIZE > 0 ? F::TW_SIZE : 1];
        for (int i = threadIdx.x; i < F::TW_SIZE; i += blockDim.x) twl[i] = tw[i];
        __syncthreads();
        twp = twl;
    }
    const i64 G = gridDim.x;
    if constexpr (IN_COLS) {
        const i64 slots = (d.ntiles + G - 1) / G * G;
        for (i64 s = blockIdx.x; s < slots; s += G) {
            const i64 t = dtfft::xcd_tile(s, G);
            if (t < d.ntiles) F::run(in, out, twp, lds, d, t, inv != 0);
            __syncthreads();
        }
    } else {
        for (i64 t = blockIdx.x; t < d.ntiles; t += G) {
            F::run(in, out, twp, lds, d, t, inv != 0);
            __syncthreads();
        }
    }
}

// Real-to-complex along contiguous lines of even length n = 2M (rows layout).
// The line is read as M complex values z_k = x_2k + i x_2k+1, transformed with the
// M-point schedule S, and untangled:
//   X_k = ((Z_k + conj Z_{M-k}) - i W^k (Z_k - conj Z_{M-k})) / 2,  W = exp(-2 pi i / n),
// for k = 0..M (Z_M = Z_0).  Strides in `d` are in complex units of the input view
// (in_ls = n/2) and of the output (out_ls = odist).  rtw holds W^k, k = 0..M.
// OSEG > 0: output k at (k / OSEG) * out_seg_stride + k % OSEG (the chunk-blocked input of a
// following column-in pass, as fft_c2c_kernel's OSEG).
template <typename T, class S, int TPL, int LPB, int PAD, int OSEG = 0>
__global__ void __launch_bounds__(TPL * LPB)
fft_r2c_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
               const typename tw_of<T>::type* __restrict__ tw,
               const typename tw_of<T>::type* __restrict__ rtw, dtfft::lines_desc d) {
    typedef dtfft::fft_tile<T, S, TPL, LPB, false, false, PAD, 0, OSEG> F;
    constexpr int M = S::N;
    __shared__ cx<T> lds[F::LDS_ELEMS];
    for (i64 t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        i64 l0, ibase, obase;
        F::tile_bases(d, t, l0, ibase, obase);
        F::template passes<0, true, false>(in, out, tw, lds, ibase, obase, d, l0, false);
        __syncthreads();
        const T half = (T)0.5;
        auto untangle = [&](int line, int k) {
            const cx<T> a = lds[line * F::LD + F::lidx(k == M ? 0 : k)];
            const cx<T> b = conj(lds[line * F::LD + F::lidx(k == 0 ? 0 : M - k)]);
            const cx<T> s = a + b, df = a - b;
            const cx<T> w = twmul(df, rtw[k]);                            // W^k (Z_k - conj Z_{M-k})
            const cx<T> x = mk(half * (s.x + w.y), half * (s.y - w.x));  // (s - i w) / 2
            // non-temporal (measured on MI355X, 1024 x 512 x 256 FP64 R2C rows: 394 against
            // 399-403 us; FP32 unchanged within the noise)
            store_nt(&out[obase + (i64)line * d.out_ls + F::out_elem(k, d)], x);
        };
        // LPB * M items, fully unrolled; the thread of k = 0 also writes X_M, so the loop has
        // no mostly idle last round (measured on MI355X, 1024 x 512 x 256: FP32 rows 239
        // against 254 us, FP64 380 against 397 us for a loop over LPB * (M + 1) items)
        constexpr int UIT = (LPB * M + F::NT - 1) / F::NT;
#pragma unroll
        for (int q = 0; q < UIT; q++) {
            const int e = (int)threadIdx.x + q * F::NT;
            if (UIT * F::NT == LPB * M || e < LPB * M) {
                const int line = e / M, k = e % M;
                if (l0 + line < d.nl) {
                    untangle(line, k);
                    if (k == 0) untangle(line, M);
                }
            }
        }
        __syncthreads();
    }
}

// Complex-to-real inverse of fft_r2c_kernel (FFTW c2r semantics: Im X_0 and Im X_M are
// ignored):  2E_k = X_k + conj X_{M-k},  2O_k = (X_k - conj X_{M-k}) conj(W^k),
// Z_k = 2E_k + i 2O_k, k < M; the backward M-point transform of Z gives
// y_2j + i y_2j+1.  Input strides in complex units (in_ls = idist), output strides in
// complex units of the real output view (out_ls = n/2).
// IN_COLS: the complex input is read across lines (consecutive lines adjacent, element
// stride in_ns) -- the transpose that precedes the C2R pass is fused into its load.
// TWL: persistent blocks copy both twiddle tables into LDS once (as fft_c2c_kernel's TWL).
namespace dtfft {
template <typename T, class S, int TPL, int LPB, int PAD, bool IN_COLS>
struct c2r_tiles {
    typedef fft_tile<T, S, TPL, LPB, false, false, PAD> F;
    typedef typename tw_of<T>::type tw_t;
    static __device__ __forceinline__ void run(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                                               const tw_t* tw, const tw_t* rtw, cx<T>* lds, const lines_desc& d) {
    constexpr int M = S::N;
    for (i64 t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        i64 l0, ibase, obase;
        F::tile_bases(d, t, l0, ibase, obase);
        for (int e = threadIdx.x; e < LPB * M; e += F::NT) {
            int line, k;
            if constexpr (IN_COLS) { line = e % LPB; k = e / LPB; }
            else { line = e / M; k = e % M; }
            cx<T> z = mk((T)0, (T)0);
            if (l0 + line < d.nl) {
                const i64 lb = ibase + (i64)line * d.in_ls;
                cx<T> a, b;
                if constexpr (IN_COLS) {
                    a = in[lb + (i64)k * d.in_ns];
                    b = in[lb + (i64)(M - k) * d.in_ns];
                } else {
                    a = in[lb + k];
                    b = in[lb + (M - k)];
                }
                if (k == 0) { a.y = (T)0; b.y = (T)0; }  // Im X_0, Im X_M do not contribute
                b = conj(b);
                const cx<T> e2 = a + b;
                const cx<T> o2 = twmul(a - b, conj_tw(rtw[k]));
                z = mk(e2.x - o2.y, e2.y + o2.x);  // 2E + i 2O
            }
            // backward transform = conj(forward(conj z)): store conj(z); the last pass
            // conjugates on its global write (inv = true)
            lds[line * F::LD + F::lidx(k)] = conj(z);
        }
        __syncthreads();
        F::template passes<0, false, true>(in, out, tw, lds, ibase, obase, d, l0, true);
        __syncthreads();
    }
    }
};
}  // namespace dtfft

template <typename T, class S, int TPL, int LPB, int PAD, bool IN_COLS = false, bool TWL = false>
__global__ void __launch_bounds__(TPL * LPB)
fft_c2r_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
               const typename tw_of<T>::type* __restrict__ tw,
               const typename tw_of<T>::type* __restrict__ rtw, dtfft::lines_desc d) {
    typedef dtfft::c2r_tiles<T, S, TPL, LPB, PAD, IN_COLS> K;
    typedef typename tw_of<T>::type tw_t;
    __shared__ cx<T> lds[K::F::LDS_ELEMS];
    if constexpr (TWL) {
        constexpr int TWN = S::TW_SIZE > 0 ? S::TW_SIZE : 1;
        __shared__ tw_t twl[TWN];
        __shared__ tw_t rtwl[S::N + 1];
        for (int i = threadIdx.x; i < S::TW_SIZE; i += blockDim.x) twl[i] = tw[i];
        for (int i = threadIdx.x; i <= S::N; i += blockDim.x) rtwl[i] = rtw[i];
        __syncthreads();
        K::run(in, out, twl, rtwl, lds, d);
    } else {
        K::run(in, out, tw, rtw, lds, d);
    }
}

// ---------------------------------------------------------------------------
// Batched persistent launch: one stage of many independent grids.
// Each grid has a descriptor (its lines in the packed buffers, its twiddles and the
// index of its tile shape in the template list F...).  A fixed set of resident blocks
// walks a host-sorted list of (grid, tile) work items; each item dispatches to the
// tile shape of its grid's length.  blockDim is the largest shape; smaller shapes
// leave the extra threads idle (they only join the barriers).
// ---------------------------------------------------------------------------
namespace dtfft {
struct grid_stage {
    lines_desc d;
    i64 in_off, out_off;  // element offsets of the grid in the packed buffers
    const void* tw;
    int cls;
    int pad_;
};

template <typename T, int I, class F0, class... Fs>
__device__ __forceinline__ void dispatch(int cls, const cx<T>* in, cx<T>* out, const void* tw, cx<T>* lds,
                                         const lines_desc& d, i64 tile, bool inv) {
    if (cls == I) {
        F0::run(in, out, (const typename tw_of<T>::type*)tw, lds, d, tile, inv);
    } else if constexpr (sizeof...(Fs) > 0) {
        dispatch<T, I + 1, Fs...>(cls, in, out, tw, lds, d, tile, inv);
    }
}

template <class F0, class... Fs> struct max_lds {
    static constexpr int v = F0::LDS_ELEMS > max_lds<Fs...>::v ? F0::LDS_ELEMS : max_lds<Fs...>::v;
};
template <class F0> struct max_lds<F0> { static constexpr int v = F0::LDS_ELEMS; };
template <class F0, class... Fs> struct first_of { typedef F0 type; };
}  // namespace dtfft

template <typename T, int NTMAX, class... F>
__global__ void __launch_bounds__(NTMAX)
batched_fft_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, const dtfft::grid_stage* __restrict__ gs,
                   const int2* __restrict__ items, int nitems, int inv) {
    __shared__ cx<T> lds[dtfft::max_lds<F...>::v];
    typedef typename dtfft::first_of<F...>::type F0;
    typedef typename tw_of<T>::type tw_t;
    // 160 KiB of LDS per workgroup on gfx950; 16 bytes of slack for the table's alignment
    constexpr bool TW_FITS = (long long)dtfft::max_lds<F...>::v * (long long)sizeof(cx<T>) +
                                 (long long)(F0::TW_SIZE > 0 ? F0::TW_SIZE : 1) * (long long)sizeof(tw_t) + 16 <=
                             160 * 1024;
    if constexpr (sizeof...(F) == 1 && TW_FITS) {
        // One tile shape, so one length: every grid of the launch uses the same twiddle table
        // (the host passes one table to all of them).  The persistent block keeps it in LDS,
        // so the passes of its tiles read no twiddle from global memory.  (A tile that leaves
        // no room for the table reads it from global memory, in the branch below.)
        __shared__ tw_t twl[F0::TW_SIZE > 0 ? F0::TW_SIZE : 1];
        if (nitems > 0) {
            const tw_t* tw = (const tw_t*)gs[0].tw;
            for (int i = threadIdx.x; i < F0::TW_SIZE; i += blockDim.x) twl[i] = tw[i];
        }
        __syncthreads();
        for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
            const int2 w = items[it];
            const dtfft::grid_stage& g = gs[w.x];
            F0::run(in + g.in_off, out + g.out_off, twl, lds, g.d, w.y, inv != 0);
            __syncthreads();
        }
    } else {
        for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
            const int2 w = items[it];
            const dtfft::grid_stage& g = gs[w.x];
            dtfft::dispatch<T, 0, F...>(g.cls, in + g.in_off, out + g.out_off, g.tw, lds, g.d, w.y, inv != 0);
            __syncthreads();
        }
    }
}

// Twiddle table of one pass: tw[off + k*(R-1) + (r-1)] = exp(-2*pi*i*r*k/(L*R)),
// generated in FP64 (sincospi) and stored as FP32 hi+lo pairs or FP64.
__global__ void gen_twiddles_f32(twf* tw, int off, int L, int R) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L * (R - 1)) return;
    const int k = e / (R - 1), r = e % (R - 1) + 1;
    double s, c;
    sincospi(-2.0 * (double)(r * k) / (double)(L * R), &s, &c);
    twf w;
    w.rh = (float)c; w.rl = (float)(c - (double)w.rh);
    w.ih = (float)s; w.il = (float)(s - (double)w.ih);
    tw[off + e] = w;
}

__global__ void gen_twiddles_f64(cx<double>* tw, int off, int L, int R) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L * (R - 1)) return;
    const int k = e / (R - 1), r = e % (R - 1) + 1;
    double s, c;
    sincospi(-2.0 * (double)(r * k) / (double)(L * R), &s, &c);
    tw[off + e] = mk(c, s);
}

// R2C / C2R untangling twiddles: tw[off + k] = exp(-2*pi*i*k/n), k < cnt.
__global__ void gen_rtwiddles_f32(twf* tw, int off, int n, int cnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    double s, c;
    sincospi(-2.0 * (double)k / (double)n, &s, &c);
    twf w;
    w.rh = (float)c; w.rl = (float)(c - (double)w.rh);
    w.ih = (float)s; w.il = (float)(s - (double)w.ih);
    tw[off + k] = w;
}

__global__ void gen_rtwiddles_f64(cx<double>* tw, int off, int n, int cnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    double s, c;
    sincospi(-2.0 * (double)k / (double)n, &s, &c);
    tw[off + k] = mk(c, s);
}

// ---------------------------------------------------------------------------
// Column-in / row-out FFT pass (the fused transpose+FFT) with a double-buffered LDS tile
// filled by direct global->LDS loads (global_load_lds_dwordx4), so that the loads of the
// next tile overlap the passes and stores of this one.  A tile is LPB adjacent lines of the
// whole N-point axis, LPB * sizeof(cx<T>) = 128 bytes per FFT row.
//   * load layout (per buffer): 128-byte rows, 8 rows per 1 KiB group (one wave-wide load,
//     lane l -> byte 16 l), groups 1152 bytes apart; element (n, line) at byte
//       (n / 8) * 1152 + (n % 8) * 128 + line * sizeof(cx<T>);
//   * the radix passes run in place in that layout with lanes across lines;
//   * the last pass writes a line-major staging layout (odd line stride LD) into the same
//     buffer, from which whole rows are stored with lanes along the row.
// OUT_COLS: the output is column-major too (the FFT along columns in place); the last pass
// then writes the load layout back and the store mirrors the load, 16 bytes per lane.
// Barriers inside a tile synchronise LDS only (lgkmcnt), so the next tile's loads stay in
// flight; they are waited for at the top of the next tile.  Twiddles live in LDS (no global
// loads other than the tile loads).  The host uses this kernel only when nl % LPB == 0,
// N % 8 == 0, in_ls == 1, out_ns == 1 (out_ls == 1 for OUT_COLS) and every row segment is
// 16-byte aligned.
// ---------------------------------------------------------------------------
#ifndef DTFFT_LDS_BARRIER
#define DTFFT_LDS_BARRIER() asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory")
#endif
#ifndef DTFFT_WAIT_ALL
#define DTFFT_WAIT_ALL() __builtin_amdgcn_s_waitcnt(0)
#endif
// The tile load is issued through inline assembly so that the compiler does not track it:
// its own LDS-DMA bookkeeping cannot always tell the two tile buffers apart and would wait
// for the prefetch before reading the other buffer.  The kernels wait for it explicitly
// (DTFFT_WAIT_ALL).  M0 holds the wave-uniform LDS destination; lane l lands at M0 + 16 l.
#ifndef DTFFT_LOAD_LDS16
#define DTFFT_LOAD_LDS16(gptr, lptr)                                                              \
    do {                                                                                          \
        const unsigned lds_ = __builtin_amdgcn_readfirstlane(                                     \
            (unsigned)(unsigned long long)((__attribute__((address_space(3))) void*)(lptr)));     \
        asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off"                       \
                     :                                                                            \
                     : "s"(lds_), "v"((const void*)(gptr))                                        \
                     : "memory", "m0");                                                           \
    } while (0)
#endif

namespace dtfft {
template <typename T, class S, int TPL, int LPB, int PAD, bool OUT_COLS>
struct cols_dma {
    typedef typename tw_of<T>::type tw_t;
    static constexpr int ES = (int)sizeof(cx<T>);
    static_assert(LPB * ES == 128, "a tile row is one 128-byte segment");
    static constexpr int N = S::N;
    static constexpr int NP = S::NP;
    static_assert(N % 8 == 0 && NP > 0, "rows come in groups of 8");
    static constexpr int NT = TPL * LPB;
    static_assert(NT % 64 == 0, "whole waves");
    static constexpr int NW = NT / 64;
    static constexpr int GROUPB = 1024 + 128;
    static constexpr int NG = N / 8;
    static constexpr int LD = (N + PAD) | 1;
    static constexpr int LOADB = NG * GROUPB;
    static constexpr int STAGEB = LPB * LD * ES;
    static constexpr int BUFE = (LOADB > STAGEB ? LOADB : STAGEB) / ES;
    static constexpr int TWN = S::TW_SIZE > 0 ? S::TW_SIZE : 1;

    // element index of (n, line) in the load layout
    static __device__ __forceinline__ int a_off(int n, int line) {
        return (n >> 3) * (GROUPB / ES) + (n & 7) * (128 / ES) + line;
    }

    static __device__ __forceinline__ void bases(const lines_desc& d, i64 t, i64& ibase, i64& obase) {
        const i64 tl = t % d.ntl, b = t / d.ntl;
        ibase = tl * (d.in_ts ? d.in_ts : (i64)LPB) + b * d.in_bs;
        obase = tl * LPB * d.out_ls + b * d.out_bs;
    }

    // this wave's share of the loads of the tile at ibase into buf
    static __device__ __forceinline__ void load(const cx<T>* in, cx<T>* buf, const lines_desc& d, i64 ibase) {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int row = lane >> 3, chunk = lane & 7;
        for (int q = wave; q < NG; q += NW) {
            const char* g = (const char*)(in + ibase + (i64)(q * 8 + row) * d.in_ns) + chunk * 16;
            DTFFT_LOAD_LDS16(g, (char*)buf + q * GROUPB);
        }
    }

    template <int p>
    static __device__ __forceinline__ void pass(cx<T>* buf, const tw_t* twl, bool inv) {
        constexpr int R = S::radix(p);
        constexpr int L = S::span(p);
        constexpr int NB = N / R;
        constexpr int IT = (NB + TPL - 1) / TPL;
        constexpr bool EXACT = (IT * TPL == NB);
        constexpr bool LAST = (p == NP - 1);
        const int line = threadIdx.x % LPB, j = threadIdx.x / LPB;
        cx<T> v[IT][R];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int b = j + it * TPL;
            if (EXACT || b < NB) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    cx<T> x = buf[a_off(b + r * NB, line)];
                    if (p == 0 && inv) x = conj(x);
                    v[it][r] = x;
                }
            }
        }
        DTFFT_LDS_BARRIER();  // every read of this pass precedes the in-place writes
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int b = j + it * TPL;
            if (EXACT || b < NB) {
                const int k = b % L;
                if (L > 1) {
#pragma unroll
                    for (int r = 1; r < R; r++)
                        v[it][r] = twmul(v[it][r], twl[S::tw_off(p) + k * (R - 1) + (r - 1)]);
                }
                bfly<T, R>::run(v[it]);
                const int base = (b - k) * R + k;
#pragma unroll
                for (int s = 0; s < R; s++) {
                    const int n = base + s * L;
                    if (LAST && !OUT_COLS) buf[line * LD + n] = v[it][s];
                    else buf[a_off(n, line)] = v[it][s];
                }
            }
        }
        DTFFT_LDS_BARRIER();
    }

    template <int p>
    static __device__ __forceinline__ void passes(cx<T>* buf, const tw_t* twl, bool inv) {
        if constexpr (p < NP) {
            pass<p>(buf, twl, inv);
            passes<p + 1>(buf, twl, inv);
        }
    }

    // passes on a loaded tile, then whole rows to global memory
    static __device__ __forceinline__ void tile(cx<T>* buf, cx<T>* out, const tw_t* twl, const lines_desc& d,
                                                i64 t, bool inv) {
        passes<0>(buf, twl, inv);
        store(buf, out, d, t, inv);
    }

    // the transformed tile (staging layout, or the load layout for OUT_COLS) to global memory
    static __device__ __forceinline__ void store(const cx<T>* buf, cx<T>* out, const lines_desc& d, i64 t,
                                                 bool inv) {
        i64 ibase, obase;
        bases(d, t, ibase, obase);
        if constexpr (OUT_COLS) {
            // mirror of load(): lane l of group q stores bytes 16 l .. 16 l + 15 of the group
            const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
            const int row = lane >> 3, chunk = lane & 7;
            constexpr int PER = 16 / ES;  // complex values per 16-byte chunk
            for (int q = wave; q < NG; q += NW) {
                const cx<T>* src = buf + q * (GROUPB / ES) + row * (128 / ES) + chunk * PER;
                cx<T>* dst = out + obase + (i64)(q * 8 + row) * d.out_ns + chunk * PER;
#pragma unroll
                for (int e = 0; e < PER; e++) {
                    cx<T> y = src[e];
                    if (inv) y = conj(y);
                    dst[e] = y;
                }
            }
        } else {
            const int line = threadIdx.x / TPL, j = threadIdx.x % TPL;
            cx<T>* o = out + obase + (i64)line * d.out_ls;
#pragma unroll 4
            for (int n = j; n < N; n += TPL) {
                cx<T> y = buf[line * LD + n];
                if (inv) y = conj(y);
                store_nt(&o[n], y);
            }
        }
    }
};
}  // namespace dtfft

// Persistent: block b handles tiles b, b + gridDim.x, ...; the two LDS buffers are separate
// arrays so that the compiler can tell the tile being loaded from the one being processed.
template <typename T, class S, int TPL, int LPB, int PAD, bool OUT_COLS = false>
__global__ void __launch_bounds__(TPL * LPB)
fft_cols_dma_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                    const typename tw_of<T>::type* __restrict__ tw, dtfft::lines_desc d, int inv) {
    typedef dtfft::cols_dma<T, S, TPL, LPB, PAD, OUT_COLS> K;
    __shared__ cx<T> buf_a[K::BUFE];
    __shared__ cx<T> buf_b[K::BUFE];
    __shared__ typename tw_of<T>::type twl[K::TWN];
    for (int i = threadIdx.x; i < S::TW_SIZE; i += K::NT) twl[i] = tw[i];
    const bool iv = inv != 0;
    const i64 G = gridDim.x;
    i64 t = blockIdx.x, ib, ob;
    if (t < d.ntiles) { K::bases(d, t, ib, ob); K::load(in, buf_a, d, ib); }
    DTFFT_WAIT_ALL();
    __syncthreads();  // first tile and twiddles in LDS
    i64 tn = t + G;
    if (tn < d.ntiles) { K::bases(d, tn, ib, ob); K::load(in, buf_b, d, ib); }
    // Three stages in flight: the passes of tile t run while the next tile's loads and the
    // previous tile's stores are outstanding.  The one wait per tile (for everything issued,
    // since loads and stores share one counter) comes after the passes, when both were issued
    // a whole pass sequence earlier; the freed buffer then takes the tile after next.
    for (;;) {
        if (t >= d.ntiles) break;
        K::template passes<0>(buf_a, twl, iv);
        DTFFT_WAIT_ALL();
        __syncthreads();  // tile tn is in buf_b
        K::store(buf_a, out, d, t, iv);
        DTFFT_LDS_BARRIER();  // every wave has read its part of buf_a
        i64 tnn = tn + G;
        if (tnn < d.ntiles) { K::bases(d, tnn, ib, ob); K::load(in, buf_a, d, ib); }
        t = tn;
        tn = tnn;
        if (t >= d.ntiles) break;
        K::template passes<0>(buf_b, twl, iv);
        DTFFT_WAIT_ALL();
        __syncthreads();  // tile tn is in buf_a
        K::store(buf_b, out, d, t, iv);
        DTFFT_LDS_BARRIER();
        tnn = tn + G;
        if (tnn < d.ntiles) { K::bases(d, tnn, ib, ob); K::load(in, buf_b, d, ib); }
        t = tn;
        tn = tnn;
    }
}

// Single-buffer variant: one tile buffer per block (about half the LDS), so that two blocks
// share a CU and overlap each other's loads with their passes and stores.
template <typename T, class S, int TPL, int LPB, int PAD, bool OUT_COLS = false>
__global__ void __launch_bounds__(TPL * LPB)
fft_cols_dma1_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                     const typename tw_of<T>::type* __restrict__ tw, dtfft::lines_desc d, int inv) {
    typedef dtfft::cols_dma<T, S, TPL, LPB, PAD, OUT_COLS> K;
    __shared__ cx<T> buf[K::BUFE];
    __shared__ typename tw_of<T>::type twl[K::TWN];
    for (int i = threadIdx.x; i < S::TW_SIZE; i += K::NT) twl[i] = tw[i];
    const bool iv = inv != 0;
    for (i64 t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        i64 ib, ob;
        K::bases(d, t, ib, ob);
        K::load(in, buf, d, ib);
        DTFFT_WAIT_ALL();
        __syncthreads();
        K::tile(buf, out, twl, d, t, iv);
        __syncthreads();  // the staged rows are read before the next tile's loads land
    }
}

// ---------------------------------------------------------------------------
// Fused Z->X transpose + C2R (the backward of a fused R2C plan) through one LDS tile filled
// by direct global->LDS loads, one tile per block (two blocks per CU).  A tile is LPB
// adjacent lines (one 128-byte piece per input row) of the M + 1 input rows X_0 .. X_M.
//   1. load: rows 0 .. M in the load layout of cols_dma (groups of 8 rows, 1152 bytes apart);
//   2. untangle in place: one thread owns the pair of rows (k, M - k) of a line and turns
//      X_k, X_{M-k} into conj(Z_k), conj(Z_{M-k}) (Z_k = 2E_k + i 2O_k as fft_c2r_kernel),
//      so every input element is read from HBM once;
//   3. the M-point passes of cols_dma in place, the last one writing the line-major staging
//      layout, and the rows of M complex (2M reals) stored whole, conjugated.
// TWLDS: the FFT twiddle table in LDS (one block per CU at M = 512) or read from global
// memory (the tile alone fits two blocks per CU); the untangling twiddles are read from
// global memory.  The host uses this kernel only when nl % LPB == 0, in_ls == 1, out_ns == 1,
// every input row piece is 16-byte aligned and the LDS fits.
// ---------------------------------------------------------------------------
namespace dtfft {
template <typename T, class S, int TPL, int LPB, int PAD>
struct c2r_dma {
    typedef cols_dma<T, S, TPL, LPB, PAD, false> C;
    typedef typename tw_of<T>::type tw_t;
    static constexpr int ES = C::ES;
    static constexpr int M = S::N;
    static constexpr int ROWS = M + 1;
    static constexpr int NGL = (ROWS + 7) / 8;
    // the last group holds ROWS - 8 (NGL - 1) rows and needs no padding after them
    static constexpr int LOADB = (NGL - 1) * C::GROUPB + (ROWS - 8 * (NGL - 1)) * 128;
    static constexpr int BUFE = (LOADB > C::STAGEB ? LOADB : C::STAGEB) / ES;
    static constexpr int RTWN = M / 2 + 1;  // untangling twiddles kept: W^k, k <= M/2

    // this wave's share of the M + 1 rows of the tile at ibase
    static __device__ __forceinline__ void load(const cx<T>* in, cx<T>* buf, const lines_desc& d, i64 ibase) {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int row = lane >> 3, chunk = lane & 7;
        for (int q = wave; q < NGL; q += C::NW) {
            if (q * 8 + row < ROWS) {
                const char* g = (const char*)(in + ibase + (i64)(q * 8 + row) * d.in_ns) + chunk * 16;
                DTFFT_LOAD_LDS16(g, (char*)buf + q * C::GROUPB);
            }
        }
    }

    // conj(Z_k) from X_k (a) and X_{M-k} (b): 2E_k = a + conj b, 2O_k = (a - conj b) conj(W^k)
    static __device__ __forceinline__ cx<T> untangle(cx<T> a, cx<T> b, const tw_t w) {
        b = conj(b);
        const cx<T> e2 = a + b;
        const cx<T> o2 = twmul(a - b, conj_tw(w));
        return conj(mk(e2.x - o2.y, e2.y + o2.x));
    }

    // -conj(w): W^{M-k} = exp(-i pi (M - k) / M) = -conj(W^k)
    static __device__ __forceinline__ twf mirror(twf w) { w.rh = -w.rh; w.rl = -w.rl; return w; }
    static __device__ __forceinline__ cx<double> mirror(cx<double> w) { w.x = -w.x; return w; }

    // rtw holds W^k for k <= M / 2 (the other half follows by the mirror)
    static __device__ __forceinline__ void untangle_tile(cx<T>* buf, const tw_t* rtw) {
        const int line = threadIdx.x % LPB;
        for (int k = threadIdx.x / LPB; k <= M / 2; k += C::NT / LPB) {
            cx<T> a = buf[C::a_off(k, line)];
            cx<T> b = buf[C::a_off(M - k, line)];
            if (k == 0) { a.y = (T)0; b.y = (T)0; }  // Im X_0, Im X_M do not contribute
            const tw_t w = rtw[k];
            const cx<T> zk = untangle(a, b, w);
            if (k > 0 && 2 * k < M) buf[C::a_off(M - k, line)] = untangle(b, a, mirror(w));
            buf[C::a_off(k, line)] = zk;
        }
    }

    // The untangling fused into the first radix pass.  Butterfly b of pass 0 reads rows
    // b + r NB0; the mirror of row b + r NB0 is row (NB0 - b) + (R0 - 1 - r) NB0, a row of
    // butterfly NB0 - b (butterflies 0 and NB0 / 2 mirror onto themselves, row 0 onto the
    // extra row M).  With two butterflies per thread, thread j takes b = j and b = NB0 - j
    // (thread 0: b = 0 and NB0 / 2), so every mirror pair is in one thread's registers and
    // the untangled values go straight into the butterflies: one LDS sweep less per tile.
    static constexpr int R0 = S::radix(0);
    static constexpr int NB0 = M / R0;
    static constexpr bool FUSE0 = NB0 % 2 == 0 && NB0 == 2 * TPL;

    static __device__ __forceinline__ tw_t wk(const tw_t* rtw, int n) {
        return 2 * n <= M ? rtw[n] : mirror(rtw[M - n]);
    }

    static __device__ __forceinline__ void pass0_untangle(cx<T>* buf, const tw_t* rtw) {
        const int line = threadIdx.x % LPB, j = threadIdx.x / LPB;
        const int b0 = j, b1 = j == 0 ? NB0 / 2 : NB0 - j;
        cx<T> x0[R0], x1[R0];
#pragma unroll
        for (int r = 0; r < R0; r++) {
            x0[r] = buf[C::a_off(b0 + r * NB0, line)];
            x1[r] = buf[C::a_off(b1 + r * NB0, line)];
        }
        const cx<T> xm = buf[C::a_off(M, line)];  // X_M, the mirror of row 0 (thread 0)
        DTFFT_LDS_BARRIER();  // every read of the tile precedes the in-place writes
        cx<T> z0[R0], z1[R0];
#pragma unroll
        for (int r = 0; r < R0; r++) {
            cx<T> a = x0[r];
            cx<T> m = j == 0 ? (r == 0 ? xm : x0[(R0 - r) % R0]) : x1[R0 - 1 - r];
            if (j == 0 && r == 0) { a.y = (T)0; m.y = (T)0; }  // Im X_0, Im X_M do not contribute
            z0[r] = untangle(a, m, wk(rtw, b0 + r * NB0));
            const cx<T> m1 = j == 0 ? x1[R0 - 1 - r] : x0[R0 - 1 - r];
            z1[r] = untangle(x1[r], m1, wk(rtw, b1 + r * NB0));
        }
        bfly<T, R0>::run(z0);
        bfly<T, R0>::run(z1);
#pragma unroll
        for (int s = 0; s < R0; s++) {
            if constexpr (S::NP == 1) {  // pass 0 is the last: the line-major staging layout
                buf[line * C::LD + b0 * R0 + s] = z0[s];
                buf[line * C::LD + b1 * R0 + s] = z1[s];
            } else {
                buf[C::a_off(b0 * R0 + s, line)] = z0[s];
                buf[C::a_off(b1 * R0 + s, line)] = z1[s];
            }
        }
        DTFFT_LDS_BARRIER();
    }

    // untangling and the M-point passes on a loaded tile; the transformed lines are left in
    // the staging layout
    static __device__ __forceinline__ void transform(cx<T>* buf, const tw_t* twl, const tw_t* rtw) {
        if constexpr (FUSE0) {
            pass0_untangle(buf, rtw);
            C::template passes<1>(buf, twl, false);
        } else {
            untangle_tile(buf, rtw);
            DTFFT_LDS_BARRIER();
            C::template passes<0>(buf, twl, false);
        }
    }
};
}  // namespace dtfft

template <typename T, class S, int TPL, int LPB, int PAD, bool TWLDS>
__global__ void __launch_bounds__(TPL * LPB)
fft_c2r_dma_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                   const typename tw_of<T>::type* __restrict__ tw,
                   const typename tw_of<T>::type* __restrict__ rtw, dtfft::lines_desc d) {
    typedef dtfft::c2r_dma<T, S, TPL, LPB, PAD> K;
    typedef typename K::C C;
    typedef typename tw_of<T>::type tw_t;
    __shared__ cx<T> buf[K::BUFE];
    const tw_t* twp = tw;
    if constexpr (TWLDS) {
        __shared__ tw_t twl[C::TWN];
        for (int i = threadIdx.x; i < S::TW_SIZE; i += C::NT) twl[i] = tw[i];
        twp = twl;
    }
    for (i64 t = blockIdx.x; t < d.ntiles; t += gridDim.x) {
        i64 ib, ob;
        C::bases(d, t, ib, ob);
        K::load(in, buf, d, ib);
        DTFFT_WAIT_ALL();
        __syncthreads();  // the tile (and the twiddles) in LDS
        K::transform(buf, twp, rtw);
        C::store(buf, out, d, t, true);  // backward = conj(forward(conj z))
        __syncthreads();  // the staged rows are read before the next tile's loads land
    }
}

// Persistent double-buffered variant: one block per CU holds two tile buffers and fills the
// next tile with direct loads while the current one is untangled, transformed and stored (the
// order of fft_cols_dma_kernel).  No global load may be issued while a prefetch is in flight
// (its wait would wait for the prefetch too), so both twiddle tables live in LDS: the FFT
// table and the untangling half W^k, k <= M/2.
template <typename T, class S, int TPL, int LPB, int PAD>
__global__ void __launch_bounds__(TPL * LPB)
fft_c2r_dma2_kernel(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                    const typename tw_of<T>::type* __restrict__ tw,
                    const typename tw_of<T>::type* __restrict__ rtw, dtfft::lines_desc d) {
    typedef dtfft::c2r_dma<T, S, TPL, LPB, PAD> K;
    typedef typename K::C C;
    typedef typename tw_of<T>::type tw_t;
    __shared__ cx<T> buf_a[K::BUFE];
    __shared__ cx<T> buf_b[K::BUFE];
    __shared__ tw_t twl[C::TWN];
    __shared__ tw_t rtwl[K::RTWN];
    for (int i = threadIdx.x; i < S::TW_SIZE; i += C::NT) twl[i] = tw[i];
    for (int i = threadIdx.x; i < K::RTWN; i += C::NT) rtwl[i] = rtw[i];
    const i64 G = gridDim.x;
    i64 t = blockIdx.x, ib, ob;
    if (t < d.ntiles) { C::bases(d, t, ib, ob); K::load(in, buf_a, d, ib); }
    DTFFT_WAIT_ALL();
    __syncthreads();  // first tile and the tables in LDS
    i64 tn = t + G;
    if (tn < d.ntiles) { C::bases(d, tn, ib, ob); K::load(in, buf_b, d, ib); }
    for (;;) {
        if (t >= d.ntiles) break;
        K::transform(buf_a, twl, rtwl);
        DTFFT_WAIT_ALL();
        __syncthreads();  // tile tn is in buf_b
        C::store(buf_a, out, d, t, true);
        DTFFT_LDS_BARRIER();  // every wave has read its part of buf_a
        i64 tnn = tn + G;
        if (tnn < d.ntiles) { C::bases(d, tnn, ib, ob); K::load(in, buf_a, d, ib); }
        t = tn;
        tn = tnn;
        if (t >= d.ntiles) break;
        K::transform(buf_b, twl, rtwl);
        DTFFT_WAIT_ALL();
        __syncthreads();  // tile tn is in buf_a
        C::store(buf_b, out, d, t, true);
        DTFFT_LDS_BARRIER();
        tnn = tn + G;
        if (tnn < d.ntiles) { C::bases(d, tnn, ib, ob); K::load(in, buf_b, d, ib); }
        t = tn;
        tn = tnn;
    }
}
