// Exact data-movement kernels behind the kernel plugin: every permute, pack
// and unpack of dtFFT is a move of a 3D box of elements
//     out[dout + i0*os0 + i1*os1 + i2*os2] = in[din + i0 + i1*is1 + i2*is2]
// with axis 0 contiguous in the input.  Two kernels cover all of them:
//   * box_copy       -- os0 == 1: rows stay contiguous (pack, unpack, the
//                       second half of the two-step backward permute);
//   * box_transpose  -- os1 == 1: the output-contiguous axis is axis 1, so each
//                       (i0, i1) plane is transposed through an LDS tile; axis 2
//                       is the batch axis.
// Elements are moved as raw 4/8/16-byte words, so results are bit-exact.
// All offsets are 64-bit: a 512^3 complex128 buffer is 2 GiB.
#include "common.hip"

template <int ES> struct word;
template <> struct word<4> { typedef u32 type; };
template <> struct word<8> { typedef u64 type; };
struct __attribute__((aligned(16))) u128 { u64 a, b; };
template <> struct word<16> { typedef u128 type; };

struct box_desc {
    i64 e0, e1, e2;          // extents
    i64 is1, is2;            // input strides of axes 1, 2 (axis 0 has stride 1)
    i64 os0, os1, os2;       // output strides
    i64 din, dout;           // element offsets
    i64 t0, t1;              // tiles along axes 0 and 1 (box_transpose) / row blocks (box_copy)
    i64 nblocks;             // total work items
};

// Rows of e0 elements, VEC elements per thread access (VEC*ES <= 16 bytes).
// Block = 256 threads; each block moves ROWS_PER_BLOCK rows.
template <int ES, int VEC>
__global__ void __launch_bounds__(256) box_copy(void* __restrict__ vout, const void* __restrict__ vin, box_desc d) {
    typedef typename word<ES>::type W;
    const W* __restrict__ in = (const W*)vin;
    W* __restrict__ out = (W*)vout;
    const i64 nrow_chunks = (d.e0 + VEC * 64 - 1) / (VEC * 64);  // 64-thread chunks per row
    // each wave handles one (row, chunk); 4 waves per block
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const i64 nrows = d.e1 * d.e2;
    const i64 nitems = nrows * nrow_chunks;
    for (i64 it = (i64)blockIdx.x * 4 + wave; it < nitems; it += (i64)gridDim.x * 4) {
        const i64 row = it / nrow_chunks, ch = it % nrow_chunks;
        const i64 i1 = row % d.e1, i2 = row / d.e1;
        const i64 ib = d.din + i1 * d.is1 + i2 * d.is2;
        const i64 ob = d.dout + i1 * d.os1 + i2 * d.os2;
        const i64 x0 = (ch * 64 + lane) * VEC;
        if constexpr (VEC > 1) {
            if (x0 + VEC <= d.e0) {
                // aligned vector move (host guarantees offsets/strides are multiples of VEC)
                typedef W vw __attribute__((ext_vector_type(VEC)));
                *(vw*)(out + ob + x0) = *(const vw*)(in + ib + x0);
                continue;
            }
        }
#pragma unroll
        for (int v = 0; v < VEC; v++)
            if (x0 + v < d.e0) out[ob + x0 + v] = in[ib + x0 + v];
    }
}

// Tiled transpose of (i0, i1) planes: TILE x TILE elements, TILE * ROWS threads.
// Loads read along i0 (contiguous input), stores write along i1 (contiguous output).
template <int ES, int TILE, int ROWS, int PAD>
__global__ void __launch_bounds__(TILE * ROWS)
box_transpose(void* __restrict__ vout, const void* __restrict__ vin, box_desc d) {
    typedef typename word<ES>::type W;
    const W* __restrict__ in = (const W*)vin;
    W* __restrict__ out = (W*)vout;
    __shared__ W tile[TILE][TILE + PAD];
    const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
    for (i64 blk = blockIdx.x; blk < d.nblocks; blk += gridDim.x) {
        const i64 b0 = blk % d.t0;
        const i64 r = blk / d.t0;
        const i64 b1 = r % d.t1;
        const i64 i2 = r / d.t1;
        const i64 x0 = b0 * TILE, y0 = b1 * TILE;
        const i64 ib = d.din + i2 * d.is2;
        const i64 ob = d.dout + i2 * d.os2;
        // load: tile[i1 - y0][i0 - x0]
#pragma unroll
        for (int k = 0; k < TILE; k += ROWS) {
            const i64 i0 = x0 + tx, i1 = y0 + ty + k;
            if (i0 < d.e0 && i1 < d.e1) tile[ty + k][tx] = in[ib + i0 + i1 * d.is1];
        }
        __syncthreads();
        // store: lanes along i1
#pragma unroll
        for (int k = 0; k < TILE; k += ROWS) {
            const i64 i1 = y0 + tx, i0 = x0 + ty + k;
            if (i0 < d.e0 && i1 < d.e1) out[ob + i1 + i0 * d.os0] = tile[tx][ty + k];
        }
        __syncthreads();
    }
}

// box_transpose with 16-byte lanes for 4- and 8-byte elements (V = 16 / ES elements per
// vector): a TILE x TILE tile, (TILE / V) x ROWS threads.  Loads: each lane reads V
// consecutive elements along i0 (the contiguous input axis) as one 16-byte access and writes
// them into the LDS tile row i1; stores: each lane gathers V consecutive i1 of one i0 from
// the tile and writes them as one 16-byte access along i1 (the contiguous output axis).
// With PAD = 1 the gather of a wave lands on distinct LDS banks (lane stride V (TILE + 1)
// words).  The host launches it only when every row start on both sides is a multiple of V
// elements and both pointers are 16-byte aligned; vectors that cross the box edge fall back
// to per-element moves.
template <int ES, int V, int TILE, int ROWS, int PAD>
__global__ void __launch_bounds__((TILE / V) * ROWS)
box_transpose_v(void* __restrict__ vout, const void* __restrict__ vin, box_desc d) {
    typedef typename word<ES>::type W;
    typedef W vec __attribute__((ext_vector_type(V)));
    static_assert(V * ES == 16 && TILE % V == 0 && TILE % ROWS == 0, "16-byte lanes over whole tiles");
    constexpr int TX = TILE / V;
    const W* __restrict__ in = (const W*)vin;
    W* __restrict__ out = (W*)vout;
    __shared__ W tile[TILE][TILE + PAD];
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (i64 blk = blockIdx.x; blk < d.nblocks; blk += gridDim.x) {
        const i64 b0 = blk % d.t0;
        const i64 r = blk / d.t0;
        const i64 b1 = r % d.t1;
        const i64 i2 = r / d.t1;
        const i64 x0 = b0 * TILE, y0 = b1 * TILE;
        const i64 ib = d.din + i2 * d.is2;
        const i64 ob = d.dout + i2 * d.os2;
        // load: tile[i1 - y0][i0 - x0], V elements of a row per lane
#pragma unroll
        for (int k = 0; k < TILE; k += ROWS) {
            const i64 i0 = x0 + tx * V, i1 = y0 + ty + k;
            if (i1 < d.e1) {
                const W* src = in + ib + i0 + i1 * d.is1;
                if (i0 + V <= d.e0) {
                    const vec w = *(const vec*)src;
#pragma unroll
                    for (int v = 0; v < V; v++) tile[ty + k][tx * V + v] = w[v];
                } else {
#pragma unroll
                    for (int v = 0; v < V; v++)
                        if (i0 + v < d.e0) tile[ty + k][tx * V + v] = src[v];
                }
            }
        }
        __syncthreads();
        // store: V consecutive i1 of one i0 per lane
#pragma unroll
        for (int k = 0; k < TILE; k += ROWS) {
            const i64 i0 = x0 + ty + k, i1 = y0 + tx * V;
            if (i0 < d.e0) {
                W* dst = out + ob + i1 + i0 * d.os0;
                if (i1 + V <= d.e1) {
                    vec w;
#pragma unroll
                    for (int v = 0; v < V; v++) w[v] = tile[tx * V + v][ty + k];
                    *(vec*)dst = w;
                } else {
#pragma unroll
                    for (int v = 0; v < V; v++)
                        if (i1 + v < d.e1) dst[v] = tile[tx * V + v][ty + k];
                }
            }
        }
        __syncthreads();
    }
}

// Scratch data for the autotuner: finite pseudo-random values in +-[0.5, 1) (FP32 words, or
// FP64 values as pairs of words), so that candidates are timed on data that toggles as many
// bits as real input does (an all-zero buffer draws less power, and the GPU's clocks follow it).
__global__ void __launch_bounds__(256) fill_tuning_data(u32* __restrict__ p, i64 nwords, int dbl) {
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < nwords; i += (i64)gridDim.x * 256) {
        u32 h = (u32)i * 2654435761u;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        u32 v;
        if (dbl) v = (i & 1) ? ((h & 0x800FFFFFu) | 0x3FE00000u) : h;  // high word: sign, exponent of [0.5, 1)
        else v = (h & 0x807FFFFFu) | 0x3F000000u;
        p[i] = v;
    }
}
