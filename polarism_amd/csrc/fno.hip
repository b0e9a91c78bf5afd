// BF16 MFMA kernels of the Fourier-neural-operator surrogate.
//
// pgpe_gemm: batched strided C = alpha op(A) op(B) + beta C with bf16 A and
// B, fp32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16), and a fused epilogue (bias along rows or
// columns, GELU, residual add) storing fp32 or bf16.  Every FNO layer --
// the truncated DFTs, the per-mode channel mixing, the pointwise channel
// linear maps -- and every backward product is one call of it.
//
// Tiling: a 256-thread workgroup (4 wavefronts) owns a 128 x 128 tile of C;
// each wavefront a 64 x 64 quarter made of 2 x 2 MFMA 32 x 32 blocks.  K
// advances 32 at a time through LDS tiles stored k-contiguous (As[m][k],
// Bs[n][k], padded) so each lane fetches its 8 consecutive k values as one
// 16-byte LDS read -- the MFMA operand layout: lane l holds row (l % 32),
// k = 8 (l / 32) .. +7.  The 16 accumulators of a lane are rows
// 8 (r / 4) + 4 (l / 32) + (r % 4), column l % 32.
#include <hip/hip_bf16.h>

#include "common.hpp"

namespace pgpe {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
// Staging type of the large-tile loads: a native vector, so that the prefetched tile stays in
// registers (a copy of the uint4 struct goes through memcpy and keeps its array in scratch).
typedef uint32_t vu32x4 __attribute__((ext_vector_type(4)));

constexpr int GT_M = 128, GT_N = 128, GT_K = 32, GT_PAD = 8;

struct GemmParams {
    const __hip_bfloat16* A;
    const __hip_bfloat16* B;
    void* C;
    const float* bias;
    const float* residual;   // fp32, same strides as C (may alias C only if beta == 0)
    int M, N, K;
    int64_t sam, sak, sab;   // A strides (elements): row, k, batch
    int64_t sbk, sbn, sbb;
    int64_t scm, scn, scb;
    int64_t srb;             // residual batch stride
    float alpha, beta;
    int bias_mode;           // 0 none, 1 per row (m), 2 per column (n)
    int gelu;                // 1: apply tanh-free GELU (erf form)
    int out_bf16;
    int vec_a, vec_b;        // 16-byte vector loads are aligned for A / B
    int batch;
};

__device__ __forceinline__ float gelu_erf(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752440f)); }

// Stage a 128 (rows r) x 32 (k) bf16 tile of a strided operand into LDS as
// S[r][k].  A k-contiguous operand (sk == 1) is read as 16-byte vectors of 8
// consecutive k; an r-contiguous one (sr == 1) as vectors of 8 consecutive r,
// scattered into the k-contiguous LDS rows.  ``vec`` says the base pointer and
// strides keep every vector 16-byte aligned; edge vectors fall back to
// element loads with bounds checks.
__device__ __forceinline__ void load_tile(__hip_bfloat16 (*S)[GT_K + GT_PAD], const __hip_bfloat16* __restrict__ G,
                                          int r0, int R, int64_t sr, int k0, int K, int64_t sk, int vec, int tid) {
    const __hip_bfloat16 zero = __float2bfloat16(0.0f);
    if (sk == 1) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int v = tid + 256 * i;        // 512 vectors: 128 rows x 4 vectors of 8 k
            const int rr = v >> 2, kq = (v & 3) * 8;
            const int gr = r0 + rr, gk = k0 + kq;
            if (vec && gr < R && gk + 8 <= K) {
                *reinterpret_cast<uint4*>(&S[rr][kq]) = *reinterpret_cast<const uint4*>(G + gr * sr + gk);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    S[rr][kq + j] = (gr < R && gk + j < K) ? G[gr * sr + gk + j] : zero;
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int v = tid + 256 * i;        // 512 vectors: 32 k x 16 vectors of 8 rows
            const int kk = v >> 4, rq = (v & 15) * 8;
            const int gr = r0 + rq, gk = k0 + kk;
            if (vec && gk < K && gr + 8 <= R) {
                union { uint4 u; __hip_bfloat16 h[8]; } t;
                t.u = *reinterpret_cast<const uint4*>(G + gk * sk + gr);
#pragma unroll
                for (int j = 0; j < 8; ++j) S[rq + j][kk] = t.h[j];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    S[rq + j][kk] = (gk < K && gr + j < R) ? G[gk * sk + (gr + j) * sr] : zero;
            }
        }
    }
}

// Epilogue of a wavefront's BI x BJ blocks of 32 x 32 (block (i, j) at rows m_base + 32 i, columns
// n_base + 32 j), one block at a time: the block's old C and residual values are all loaded
// before its first store (16 loads in flight per lane instead of one round trip per value).
// Every output element belongs to one lane, so this stays exact when the residual is C itself.
// Per element: alpha acc (+ beta C) (+ bias) (+ residual), then GELU, in that order.
template <int BI, int BJ>
__device__ __forceinline__ void gemm_epilogue(const GemmParams& p, const f32x16 (&acc)[BI][BJ], int bz, int m_base,
                                              int n_base, int lane) {
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
            const int gn = n_base + 32 * j + (lane & 31);
            const int gm0 = m_base + 32 * i + 4 * (lane >> 5);
            float old[16], res[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gm = gm0 + 8 * (r >> 2) + (r & 3);
                old[r] = 0.0f;
                res[r] = 0.0f;
                if (gm >= p.M || gn >= p.N) continue;
                const int64_t off = bz * p.scb + gm * p.scm + gn * p.scn;
                if (p.beta != 0.0f)
                    old[r] = p.out_bf16 ? __bfloat162float(((const __hip_bfloat16*)p.C)[off]) : ((const float*)p.C)[off];
                if (p.residual) res[r] = p.residual[bz * p.srb + gm * p.scm + gn * p.scn];
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gm = gm0 + 8 * (r >> 2) + (r & 3);
                if (gm >= p.M || gn >= p.N) continue;
                const int64_t off = bz * p.scb + gm * p.scm + gn * p.scn;
                float v = p.alpha * acc[i][j][r];
                if (p.beta != 0.0f) v += p.beta * old[r];
                if (p.bias_mode == 1) v += p.bias[gm];
                else if (p.bias_mode == 2) v += p.bias[gn];
                if (p.residual) v += res[r];
                if (p.gelu) v = gelu_erf(v);
                if (p.out_bf16) ((__hip_bfloat16*)p.C)[off] = __float2bfloat16(v);
                else ((float*)p.C)[off] = v;
            }
        }
}

__global__ void __launch_bounds__(256) gemm_bf16_kernel(const GemmParams p) {
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 As[GT_M][GT_K + GT_PAD];
    __shared__ __attribute__((aligned(16))) __hip_bfloat16 Bs[GT_N][GT_K + GT_PAD];
    const int bz = blockIdx.z;
    const int m0 = blockIdx.y * GT_M;
    const int n0 = blockIdx.x * GT_N;
    const __hip_bfloat16* A = p.A + bz * p.sab;
    const __hip_bfloat16* B = p.B + bz * p.sbb;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    for (int k0 = 0; k0 < p.K; k0 += GT_K) {
        load_tile(As, A, m0, p.M, p.sam, k0, p.K, p.sak, p.vec_a, tid);
        load_tile(Bs, B, n0, p.N, p.sbn, k0, p.K, p.sbk, p.vec_b, tid);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GT_K; ks += 16) {
            bf16x8 a[2], b[2];
            const int kk = ks + 8 * (lane >> 5);
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 32 * i + (lane & 31)][kk]);
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 32 * j + (lane & 31)][kk]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    gemm_epilogue<2, 2>(p, acc, bz, m0 + wm, n0 + wn, lane);
}

// ---------------------------------------------------------------- large tiles
// Large products (M a multiple of 256, N of 128, K of 64; A k-contiguous; B k- or n-contiguous):
// a 256-thread workgroup owns a 256 x 128 tile of C, each wavefront 128 x 64 of it as 4 x 2
// MFMA blocks of 32 x 32 (8 accumulators of 16 per lane).  K advances 64 at a time through two
// LDS buffers: the next K-tile's operands are loaded into registers before the current tile's
// 32 MFMAs per wavefront and written to the other buffer after them, one barrier per K-tile.
// LDS rows hold 64 k plus 8 of padding (As[m][k], Bs[n][k]: the MFMA operand of a lane is one
// 16-byte read).  A row-major (n-contiguous) B is transposed on the way in: each thread loads a
// 4 (k) x 8 (n) block as four 16-byte rows and recombines them with byte permutes into eight
// 4-element k columns, written as 8-byte LDS stores.  Workgroups take tiles in an XCD-aware
// order (consecutive tiles on one XCD) grouped by four row panels, so the workgroups an XCD runs
// together share their A and B panels in its L2.
constexpr int BG_M = 256, BG_N = 128, BG_K = 64, BG_LD = BG_K + 8;
constexpr size_t kBigLdsBytes = (size_t)2 * (BG_M + BG_N) * BG_LD * sizeof(__hip_bfloat16);

// 16-bit halves of two dwords: (lo(a), lo(b)) and (hi(a), hi(b)).
__device__ __forceinline__ uint32_t pack_lo(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x05040100u); }
__device__ __forceinline__ uint32_t pack_hi(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x07060302u); }

// Row-major B is transposed in registers: thread tid holds the 4 k x 8 n block at (kb, nb) and
// writes it as eight 8-byte pieces to rows nb .. nb + 7 of the k-contiguous LDS tile (row stride
// 144 bytes).  Lanes of one wavefront on 16 different n blocks would put their rows 8 x 144 bytes
// apart, all on two bank offsets (an 8-way conflict); on the 16 k blocks of 4 n blocks they
// cover every bank twice, the minimum for 512 bytes.
__device__ __forceinline__ int b_block_n(int tid) { return ((tid >> 6) * 4 + ((tid >> 4) & 3)) * 8; }

// NT = 256: four waves of 128 x 64 per tile (one wave per SIMD); NT = 512: eight waves of 64 x 64
// (two per SIMD), the form launched.
template <bool B_KCONTIG, int NT>
__global__ void __launch_bounds__(NT) gemm_bf16_big_kernel(const GemmParams p) {
    constexpr int WROWS = NT / 128;           // waves along m (two along n)
    constexpr int WM = BG_M / WROWS;          // rows of a wave's tile
    constexpr int BI = WM / 32;               // its 32 x 32 accumulators along m
    constexpr int NA = BG_M * BG_K / 8 / NT;  // A vectors of 8 k per thread
    constexpr int NB = BG_N * BG_K / 8 / NT;  // k-contiguous B vectors per thread
    extern __shared__ __attribute__((aligned(16))) unsigned char big_smem[];
    typedef __hip_bfloat16 Row[BG_LD];
    Row* As = reinterpret_cast<Row*>(big_smem);   // [2][BG_M]
    Row* Bs = As + 2 * BG_M;                      // [2][BG_N]
    const int tiles_m = p.M / BG_M, tiles_n = p.N / BG_N;
    const int per_batch = tiles_m * tiles_n;
    const int t = xcd_tile(blockIdx.x, per_batch * p.batch);
    if (t < 0) return;
    const int bz = t / per_batch;
    const int tt = t - bz * per_batch;
    // grouped raster: four row panels per group, columns inside
    const int group = tt / (4 * tiles_n);
    const int first_m = group * 4;
    const int group_m = min(4, tiles_m - first_m);
    const int in_group = tt - group * 4 * tiles_n;
    const int m0 = (first_m + in_group % group_m) * BG_M;
    const int n0 = (in_group / group_m) * BG_N;
    const __hip_bfloat16* __restrict__ A = p.A + bz * p.sab;
    const __hip_bfloat16* __restrict__ B = p.B + bz * p.sbb;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * 64;
    // row-major B: 256 blocks of 4 k x 8 n, one per thread of the first 256
    const bool b_owner = B_KCONTIG || tid < 256;

    vu32x4 ra[NA], rb[B_KCONTIG ? NB : 4];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int v = tid + NT * i;  // 2048 vectors: 256 rows x 8 vectors of 8 k
            ra[i] = *reinterpret_cast<const vu32x4*>(A + (int64_t)(m0 + (v >> 3)) * p.sam + k0 + (v & 7) * 8);
        }
        if constexpr (B_KCONTIG) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int v = tid + NT * i;  // 1024 vectors: 128 rows (n) x 8 vectors of 8 k
                rb[i] = *reinterpret_cast<const vu32x4*>(B + (int64_t)(n0 + (v >> 3)) * p.sbn + k0 + (v & 7) * 8);
            }
        } else if (b_owner) {
            // block of 4 k x 8 n: 16 blocks along n, 16 along k; a wavefront's lanes take all 16
            // k blocks of 4 n blocks (b_block_n)
            const int nb = b_block_n(tid), kb = (tid & 15) * 4;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                rb[i] = *reinterpret_cast<const vu32x4*>(B + (int64_t)(k0 + kb + i) * p.sbk + n0 + nb);
        }
    };
    auto store = [&](int buf) {
        Row* as = As + buf * BG_M;
        Row* bs = Bs + buf * BG_N;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int v = tid + NT * i;
            *reinterpret_cast<vu32x4*>(&as[v >> 3][(v & 7) * 8]) = ra[i];
        }
        if constexpr (B_KCONTIG) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int v = tid + NT * i;
                *reinterpret_cast<vu32x4*>(&bs[v >> 3][(v & 7) * 8]) = rb[i];
            }
        } else if (b_owner) {
            const int nb = b_block_n(tid), kb = (tid & 15) * 4;
            const uint32_t w[4][4] = {{rb[0].x, rb[0].y, rb[0].z, rb[0].w},
                                      {rb[1].x, rb[1].y, rb[1].z, rb[1].w},
                                      {rb[2].x, rb[2].y, rb[2].z, rb[2].w},
                                      {rb[3].x, rb[3].y, rb[3].z, rb[3].w}};
#pragma unroll
            for (int d = 0; d < 4; ++d) {  // dword d of a row holds columns nb + 2d, nb + 2d + 1
                uint2 even, odd;          // k = kb .. kb + 3 of column nb + 2d (even), nb + 2d + 1 (odd)
                even.x = pack_lo(w[0][d], w[1][d]);
                even.y = pack_lo(w[2][d], w[3][d]);
                odd.x = pack_hi(w[0][d], w[1][d]);
                odd.y = pack_hi(w[2][d], w[3][d]);
                *reinterpret_cast<uint2*>(&bs[nb + 2 * d][kb]) = even;
                *reinterpret_cast<uint2*>(&bs[nb + 2 * d + 1][kb]) = odd;
            }
        }
    };

    f32x16 acc[BI][2];
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int ktiles = p.K / BG_K;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < ktiles; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < ktiles) load((kt + 1) * BG_K);
        const Row* as = As + buf * BG_M;
        const Row* bs = Bs + buf * BG_N;
#pragma unroll
        for (int ks = 0; ks < BG_K; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            bf16x8 a[BI], b[2];
#pragma unroll
            for (int i = 0; i < BI; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&as[wm + 32 * i + (lane & 31)][kk]);
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&bs[wn + 32 * j + (lane & 31)][kk]);
#pragma unroll
            for (int i = 0; i < BI; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < ktiles) store(buf ^ 1);
        __syncthreads();
    }
    gemm_epilogue<BI, 2>(p, acc, bz, m0 + wm, n0 + wn, lane);
}

// ---------------------------------------------------------------- elementwise
__global__ void f32_to_bf16_kernel(const float* __restrict__ x, __hip_bfloat16* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = __float2bfloat16(x[i]);
}

// y = gelu(x) (fp32 in, fp32 and optional bf16 copy out)
__global__ void gelu_fwd_kernel(const float* __restrict__ x, float* __restrict__ y, __hip_bfloat16* __restrict__ yb,
                                int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = gelu_erf(x[i]);
    if (y) y[i] = v;
    if (yb) yb[i] = __float2bfloat16(v);
}

// dx = dy * gelu'(x), gelu'(x) = Phi(x) + x phi(x)
__global__ void gelu_bwd_kernel(const float* __restrict__ x, const float* __restrict__ dy, float* __restrict__ dx,
                                __hip_bfloat16* __restrict__ dxb, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = x[i];
    const float cdf = 0.5f * (1.0f + erff(v * 0.70710678118654752440f));
    const float pdf = 0.39894228040143267794f * expf(-0.5f * v * v);
    const float g = dy[i] * (cdf + v * pdf);
    if (dx) dx[i] = g;
    if (dxb) dxb[i] = __float2bfloat16(g);
}

// Per-mode complex channel scale on planar (re, im) spectra laid out
// [ky][c][b][kx]: u = d[c, ky, kx] v (conj(d) when conj), fp32 and bf16 copies.
__global__ void mode_scale_kernel(const float* __restrict__ vr, const float* __restrict__ vi,
                                  const float* __restrict__ dr, const float* __restrict__ di, float* __restrict__ ur,
                                  float* __restrict__ ui, __hip_bfloat16* __restrict__ urb,
                                  __hip_bfloat16* __restrict__ uib, int my, int C, int B, int mx, int conj) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)my * C * B * mx;
    if (i >= n) return;
    const int kx = (int)(i % mx);
    const int c = (int)((i / ((int64_t)B * mx)) % C);
    const int ky = (int)(i / ((int64_t)C * B * mx));
    const int64_t q = ((int64_t)c * my + ky) * mx + kx;
    const float wr = dr[q], wi = conj ? -di[q] : di[q];
    const float a = vr[i], b = vi[i];
    const float re = wr * a - wi * b, im = wr * b + wi * a;
    if (ur) ur[i] = re;
    if (ui) ui[i] = im;
    if (urb) urb[i] = __float2bfloat16(re);
    if (uib) uib[i] = __float2bfloat16(im);
}

// gd[c, ky, kx] = sum_b conj(v) du  (complex), v and du planar [ky][c][b][kx].
__global__ void mode_scale_grad_kernel(const float* __restrict__ vr, const float* __restrict__ vi,
                                       const float* __restrict__ dur, const float* __restrict__ dui,
                                       float* __restrict__ gdr, float* __restrict__ gdi, int my, int C, int B, int mx) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)C * my * mx) return;
    const int kx = (int)(q % mx);
    const int ky = (int)((q / mx) % my);
    const int c = (int)(q / ((int64_t)my * mx));
    float sr = 0.0f, si = 0.0f;
    for (int b = 0; b < B; ++b) {
        const int64_t i = (((int64_t)ky * C + c) * B + b) * mx + kx;
        sr += vr[i] * dur[i] + vi[i] * dui[i];
        si += vr[i] * dui[i] - vi[i] * dur[i];
    }
    gdr[q] = sr;
    gdi[q] = si;
}

// out[n] = sum over rows of x[row, n] (bias gradients), rows x cols fp32 row-major.
__global__ void colsum_kernel(const float* __restrict__ x, float* __restrict__ out, int64_t rows, int64_t cols,
                              int64_t row_stride, int64_t col_stride) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int64_t r = 0; r < rows; ++r) s += x[r * row_stride + c * col_stride];
    out[c] = s;
}

// out[c] = sum_r x[c * col_stride + r] for a column stored contiguously
// (row_stride 1): one workgroup per column, strided partial sums, then a
// wavefront-shuffle and LDS reduction in a fixed order (deterministic).
__global__ void __launch_bounds__(256) colsum_contig_kernel(const float* __restrict__ x, float* __restrict__ out,
                                                             int64_t rows, int64_t col_stride) {
    __shared__ float part[4];
    const int64_t c = blockIdx.x;
    const float* __restrict__ col = x + c * col_stride;
    float s = 0.0f;
    for (int64_t r = threadIdx.x; r < rows; r += 256) s += col[r];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[c] = (part[0] + part[1]) + (part[2] + part[3]);
}

static dim3 blocks1d(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace pgpe

using namespace pgpe;

extern "C" {

int pgpe_gemm_bf16(const pgpe_gemm_args* a, void* stream) {
    if (!a || !a->A || !a->B || !a->C) return fail(PGPE_ERR_ARG, "pgpe_gemm_bf16: null pointer");
    if (a->M <= 0 || a->N <= 0 || a->K <= 0 || a->batch <= 0) return fail(PGPE_ERR_ARG, "pgpe_gemm_bf16: empty shape");
    if ((a->sak != 1 && a->sam != 1) || (a->sbk != 1 && a->sbn != 1))
        return fail(PGPE_ERR_ARG, "pgpe_gemm_bf16: A and B need a unit stride along k or along m / n");
    if (a->bias_mode && !a->bias) return fail(PGPE_ERR_ARG, "pgpe_gemm_bf16: bias_mode without bias");
    GemmParams p;
    p.A = (const __hip_bfloat16*)a->A;
    p.B = (const __hip_bfloat16*)a->B;
    p.C = a->C;
    p.bias = a->bias;
    p.residual = a->residual;
    p.M = a->M;
    p.N = a->N;
    p.K = a->K;
    p.sam = a->sam; p.sak = a->sak; p.sab = a->sab;
    p.sbk = a->sbk; p.sbn = a->sbn; p.sbb = a->sbb;
    p.scm = a->scm; p.scn = a->scn; p.scb = a->scb;
    p.srb = a->srb;
    p.alpha = a->alpha;
    p.beta = a->beta;
    p.bias_mode = a->bias_mode;
    p.gelu = a->gelu;
    p.out_bf16 = a->out_bf16;
    p.batch = a->batch;
    // 16-byte vectors along the unit-stride axis need a 16-byte aligned base and strides
    // that are multiples of 8 elements on the other axes.
    auto aligned = [](const void* ptr, int64_t s1, int64_t s2) {
        return ((uintptr_t)ptr % 16 == 0) && (s1 % 8 == 0) && (s2 % 8 == 0);
    };
    p.vec_a = a->sak == 1 ? aligned(a->A, a->sam, a->sab) : aligned(a->A, a->sak, a->sab);
    p.vec_b = a->sbk == 1 ? aligned(a->B, a->sbn, a->sbb) : aligned(a->B, a->sbk, a->sbb);
    // large tiles: A k-contiguous, B k- or n-contiguous, all on 16-byte vectors, whole tiles
    const bool big = a->sak == 1 && p.vec_a && p.vec_b && a->M % BG_M == 0 && a->N % BG_N == 0 && a->K % BG_K == 0 &&
                     (int64_t)(a->M / BG_M) * (a->N / BG_N) * a->batch <= 0x7fffffff;
    if (big) {
        // eight waves of 64 x 64 per tile: two per SIMD hide the LDS reads between MFMAs that one
        // wave of 128 x 64 left exposed (8192^3: 690 against 606 TFLOP/s row-major B, 812 against
        // 695 K-contiguous)
        constexpr int NT = 512;
        auto kern = a->sbk == 1 ? gemm_bf16_big_kernel<true, NT> : gemm_bf16_big_kernel<false, NT>;
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kBigLdsBytes);
        if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(gemm_bf16_big_kernel)");
        const int tiles = (a->M / BG_M) * (a->N / BG_N);
        hipLaunchKernelGGL(kern, dim3(xcd_grid(tiles * a->batch)), dim3(NT), kBigLdsBytes, (hipStream_t)stream, p);
        PGPE_LAUNCHED("gemm_bf16_big_kernel");
        return PGPE_OK;
    }
    const dim3 grid((a->N + GT_N - 1) / GT_N, (a->M + GT_M - 1) / GT_M, a->batch);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, (hipStream_t)stream, p);
    PGPE_LAUNCHED("gemm_bf16_kernel");
    return PGPE_OK;
}

int pgpe_f32_to_bf16(const float* x, void* y, int64_t n, void* stream) {
    if (n <= 0) return PGPE_OK;
    if (!x || !y) return fail(PGPE_ERR_ARG, "pgpe_f32_to_bf16: null pointer");
    hipLaunchKernelGGL(f32_to_bf16_kernel, blocks1d(n), dim3(256), 0, (hipStream_t)stream, x, (__hip_bfloat16*)y, n);
    PGPE_LAUNCHED("f32_to_bf16_kernel");
    return PGPE_OK;
}

int pgpe_gelu_fwd(const float* x, float* y, void* yb, int64_t n, void* stream) {
    if (n <= 0) return PGPE_OK;
    if (!x || (!y && !yb)) return fail(PGPE_ERR_ARG, "pgpe_gelu_fwd: null pointer");
    hipLaunchKernelGGL(gelu_fwd_kernel, blocks1d(n), dim3(256), 0, (hipStream_t)stream, x, y, (__hip_bfloat16*)yb, n);
    PGPE_LAUNCHED("gelu_fwd_kernel");
    return PGPE_OK;
}

int pgpe_gelu_bwd(const float* x, const float* dy, float* dx, void* dxb, int64_t n, void* stream) {
    if (n <= 0) return PGPE_OK;
    if (!x || !dy || (!dx && !dxb)) return fail(PGPE_ERR_ARG, "pgpe_gelu_bwd: null pointer");
    hipLaunchKernelGGL(gelu_bwd_kernel, blocks1d(n), dim3(256), 0, (hipStream_t)stream, x, dy, dx,
                       (__hip_bfloat16*)dxb, n);
    PGPE_LAUNCHED("gelu_bwd_kernel");
    return PGPE_OK;
}

int pgpe_mode_scale(const float* vr, const float* vi, const float* dr, const float* di, float* ur, float* ui,
                    void* urb, void* uib, int my, int C, int B, int mx, int conj, void* stream) {
    if (!vr || !vi || !dr || !di || (!ur && !urb) || my <= 0 || C <= 0 || B <= 0 || mx <= 0)
        return fail(PGPE_ERR_ARG, "pgpe_mode_scale: bad argument");
    const int64_t n = (int64_t)my * C * B * mx;
    hipLaunchKernelGGL(mode_scale_kernel, blocks1d(n), dim3(256), 0, (hipStream_t)stream, vr, vi, dr, di, ur, ui,
                       (__hip_bfloat16*)urb, (__hip_bfloat16*)uib, my, C, B, mx, conj);
    PGPE_LAUNCHED("mode_scale_kernel");
    return PGPE_OK;
}

int pgpe_mode_scale_grad(const float* vr, const float* vi, const float* dur, const float* dui, float* gdr,
                         float* gdi, int my, int C, int B, int mx, void* stream) {
    if (!vr || !vi || !dur || !dui || !gdr || !gdi || my <= 0 || C <= 0 || B <= 0 || mx <= 0)
        return fail(PGPE_ERR_ARG, "pgpe_mode_scale_grad: bad argument");
    hipLaunchKernelGGL(mode_scale_grad_kernel, blocks1d((int64_t)C * my * mx), dim3(256), 0, (hipStream_t)stream,
                       vr, vi, dur, dui, gdr, gdi, my, C, B, mx);
    PGPE_LAUNCHED("mode_scale_grad_kernel");
    return PGPE_OK;
}

int pgpe_colsum(const float* x, float* out, int64_t rows, int64_t cols, int64_t row_stride, int64_t col_stride,
                void* stream) {
    if (rows <= 0 || cols <= 0) return fail(PGPE_ERR_ARG, "pgpe_colsum: empty");
    if (!x || !out) return fail(PGPE_ERR_ARG, "pgpe_colsum: null pointer");
    if (row_stride == 1 && cols <= 0x7fffffff) {
        hipLaunchKernelGGL(colsum_contig_kernel, dim3((unsigned)cols), dim3(256), 0, (hipStream_t)stream, x, out, rows,
                           col_stride);
        PGPE_LAUNCHED("colsum_contig_kernel");
        return PGPE_OK;
    }
    hipLaunchKernelGGL(colsum_kernel, blocks1d(cols), dim3(256), 0, (hipStream_t)stream, x, out, rows, cols,
                       row_stride, col_stride);
    PGPE_LAUNCHED("colsum_kernel");
    return PGPE_OK;
}

}  // extern "C"
