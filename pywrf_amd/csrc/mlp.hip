// Dense layers of the WSM6 column surrogate on the matrix cores (gfx950):
// BF16 operands, FP32 accumulation, v_mfma_f32_32x32x16_bf16.
//
// One GEMM form serves the forward and backward passes:
//   C[M x N] = A[M x K] . B[N x K]^T
// with both operands K-contiguous (row-major A, row-major B = the layer's
// weight [out x in]); the backward products use transposed copies (the
// optimiser keeps W^T next to W, activations and output gradients are
// transposed by k_transpose), so every operand is read along K.
//
// Tiling: a 256-thread workgroup (4 waves, 2 x 2) owns a 128 x 128 tile of C;
// each wave owns 64 x 64 = 2 x 2 MFMA tiles of 32 x 32.  K advances 128 at a
// time: the slab's A and B tiles (128 x 128 bf16 each, rows padded to 136
// bf16 so that the 16-byte fragment reads of neighbouring rows fall in
// different banks) are loaded from global into registers one slab ahead, so
// the loads of slab t + 1 are in flight while the 32 MFMAs per wave of slab t
// run from LDS; a slab this deep keeps the matrix cores busy across the
// global-load latency with two workgroups per CU.  Columns of A and B at or
// beyond K enter LDS as zeros; rows beyond M and N are clamped loads whose
// products are never stored.
// Blocks are numbered so that the N-tiles of one M-tile go to one XCD (they
// share A, the large operand).
//
// Epilogues: FP32 C (+ bias), FP32 C^T (+ bias; the surrogate's output, read
// per feature across columns), BF16 relu(C + bias), BF16 C * [H > 0] (the
// ReLU derivative, H the layer's forward output) and FP32 atomic accumulation
// for split-K weight gradients (grid.z splits K; C must start at zero).
#include "common.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define MLP_BM 128
#define MLP_BN 128
#ifndef MLP_BK
#define MLP_BK 128
#endif
#ifndef MLP_NBUF
#define MLP_NBUF (MLP_BK >= 128 ? 1 : 2)  // LDS buffers: one for deep slabs, two (ping-pong) for shallow ones
#endif
#define MLP_LDS_ROW (MLP_BK + 8)  // bf16 per LDS row (8 padding)
#define MLP_KQ 32                 // K must be a multiple of this
#define MLP_CH (MLP_BM * MLP_BK / 8 / 256)  // 16-byte pieces of A (and of B) per thread per slab

enum { EPI_F32 = 0, EPI_BF16_RELU = 1, EPI_BF16_DRELU = 2, EPI_F32_ATOMIC = 3, EPI_F32_T = 4 };

struct GemmArgs {
  int M, N, K, lda, ldb, ldc, ldh, kchunk;
  const bf16* A;
  const bf16* B;
  void* C;
  const float* bias;
  const bf16* H;
};

// 16 bytes of a row-major bf16 matrix at (row, col), both clamped into the
// matrix ([0, rows) x [0, cols - 8]).  The load is unconditional: a load the
// compiler may predicate is waited for before the predicated region closes,
// which would serialise a slab's sixteen loads.  Rows clamped beyond M or N
// only feed outputs that are not stored; columns beyond K are replaced by
// zeros when the slab is written to LDS.
__device__ inline uint4 load16(const bf16* base, int ld, int row, int rows, int col, int cols) {
  const int r = row < rows ? row : rows - 1, c = col < cols - 8 ? col : cols - 8;
  return *(const uint4*)(base + (long)r * ld + c);
}

template <int EPI>
__global__ void __launch_bounds__(256) k_mlp_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[MLP_NBUF][MLP_BM * MLP_LDS_ROW];
  __shared__ __attribute__((aligned(16))) bf16 Bs[MLP_NBUF][MLP_BN * MLP_LDS_ROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // tile order: blockIdx.x counts the N-tiles, blockIdx.y the M-tiles, and
  // each XCD takes a contiguous run of that order, so all N-tiles of one
  // M-tile run on one XCD and read their A tile (the activations, the large
  // operand) once from memory into that XCD's L2; B (a weight matrix) fits
  // in every L2
  int bx = (int)blockIdx.x, by = (int)blockIdx.y;
  xcd_block_2d(bx, by);
  const int m0 = by * MLP_BM, n0 = bx * MLP_BN;
  const int kbeg = (int)blockIdx.z * g.kchunk;
  const int kend = imin(g.K, kbeg + g.kchunk);
  const int nk = kend > kbeg ? (kend - kbeg + MLP_BK - 1) / MLP_BK : 0;

  // piece p = tid + 256 r of a slab covers row p / (BK / 8), columns 8 (p % (BK / 8)) .. + 8
  uint4 ra[MLP_CH], rb[MLP_CH];
#define MLP_GLOAD(k0)                                                                   \
  do {                                                                                  \
    _Pragma("unroll") for (int r = 0; r < MLP_CH; ++r) {                                \
      const int p = tid + 256 * r, row = p / (MLP_BK / 8), col = (k0) + (p % (MLP_BK / 8)) * 8; \
      ra[r] = load16(g.A, g.lda, m0 + row, g.M, col, kend);                             \
      rb[r] = load16(g.B, g.ldb, n0 + row, g.N, col, kend);                             \
    }                                                                                   \
  } while (0)
  // the slab at k0 into LDS, columns at or beyond K as zeros
#define MLP_LSTORE(k0, buf)                                                             \
  do {                                                                                  \
    const int lim = kend - (k0);                                                        \
    _Pragma("unroll") for (int r = 0; r < MLP_CH; ++r) {                                \
      const int p = tid + 256 * r, row = p / (MLP_BK / 8), col = (p % (MLP_BK / 8)) * 8; \
      const bool in = col < lim;                                                        \
      uint4 a = ra[r], b = rb[r];                                                       \
      a.x = in ? a.x : 0u; a.y = in ? a.y : 0u; a.z = in ? a.z : 0u; a.w = in ? a.w : 0u; \
      b.x = in ? b.x : 0u; b.y = in ? b.y : 0u; b.z = in ? b.z : 0u; b.w = in ? b.w : 0u; \
      *(uint4*)&As[buf][row * MLP_LDS_ROW + col] = a;                                   \
      *(uint4*)&Bs[buf][row * MLP_LDS_ROW + col] = b;                                   \
    }                                                                                   \
  } while (0)

  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.0f;

  const int lr = lane & 31, lk = (lane >> 5) * 8;
#define MLP_COMPUTE(buf)                                                                \
  _Pragma("unroll") for (int ks = 0; ks < MLP_BK; ks += 16) {                           \
    bf16x8 af[2], bfr[2];                                                               \
    _Pragma("unroll") for (int t = 0; t < 2; ++t) {                                     \
      af[t] = *(const bf16x8*)&As[buf][(wm * 64 + t * 32 + lr) * MLP_LDS_ROW + ks + lk]; \
      bfr[t] = *(const bf16x8*)&Bs[buf][(wn * 64 + t * 32 + lr) * MLP_LDS_ROW + ks + lk]; \
    }                                                                                   \
    _Pragma("unroll") for (int a = 0; a < 2; ++a)                                       \
      _Pragma("unroll") for (int b = 0; b < 2; ++b)                                     \
        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0); \
  }
  if (MLP_NBUF == 1) {
    // one buffer: store the prefetched slab, then fetch the next one while
    // the MFMAs of this one run (two barriers per slab)
    if (nk > 0) MLP_GLOAD(kbeg);
    for (int kt = 0; kt < nk; ++kt) {
      __syncthreads();               // every wave is done reading the previous slab
      MLP_LSTORE(kbeg + kt * MLP_BK, 0);
      __syncthreads();
      if (kt + 1 < nk) MLP_GLOAD(kbeg + (kt + 1) * MLP_BK);
      MLP_COMPUTE(0)
    }
  } else {
    // two buffers: fetch slab t + 1, run slab t from one buffer, store slab
    // t + 1 into the other (one barrier per slab)
    if (nk > 0) {
      MLP_GLOAD(kbeg);
      MLP_LSTORE(kbeg, 0);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < nk) MLP_GLOAD(kbeg + (kt + 1) * MLP_BK);
      MLP_COMPUTE(cur)
      if (kt + 1 < nk) MLP_LSTORE(kbeg + (kt + 1) * MLP_BK, cur ^ 1);
      __syncthreads();
    }
  }
#undef MLP_COMPUTE
#undef MLP_GLOAD
#undef MLP_LSTORE

  // accumulator element e of lane l holds C[8 (e / 4) + 4 (l / 32) + e % 4][l % 32] of the 32 x 32 tile.
  // Every operand of the epilogue (bias, ReLU mask) is loaded, at clamped
  // indices, before the first store: stores and loads share one counter, and
  // a load needed after predicated stores would make each store wait for all
  // earlier ones.
  float bv[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int n = imin(n0 + wn * 64 + b * 32 + lr, g.N - 1);
    bv[b] = (EPI == EPI_F32 || EPI == EPI_BF16_RELU || EPI == EPI_F32_T) && g.bias ? g.bias[n] : 0.0f;
  }
  if (EPI == EPI_BF16_DRELU) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int n = imin(n0 + wn * 64 + b * 32 + lr, g.N - 1);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int m = imin(m0 + wm * 64 + a * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3), g.M - 1);
          if (!((float)g.H[(long)m * g.ldh + n] > 0.0f)) acc[a][b][e] = 0.0f;
        }
      }
  }
  // the values, formed unconditionally (so the operand loads are waited for
  // once, here, and none is pending at the stores)
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        float v = acc[a][b][e] + bv[b];
        if (EPI == EPI_BF16_RELU) v = v > 0.0f ? v : 0.0f;
        acc[a][b][e] = v;
      }
  // wait here for every load (s_waitcnt vmcnt(0)): the compiler sinks the
  // arithmetic above into the predicated stores below and would otherwise
  // wait for the operand loads -- and so for every earlier store -- in each
  // of them
  __builtin_amdgcn_s_waitcnt(0x0F70);
  // a tile wholly inside C stores without per-element predicates
  const bool full = m0 + MLP_BM <= g.M && n0 + MLP_BN <= g.N;
  // (the A buffer holds the 128 x 128 tile only when a slab row is at least 128 wide)
  if ((EPI == EPI_BF16_RELU || EPI == EPI_BF16_DRELU) && MLP_BK >= MLP_BN && full && (g.ldc & 7) == 0 &&
      ((uintptr_t)g.C & 15) == 0) {
    // the 128 x 128 BF16 tile through LDS (the A slab buffer, free once every
    // wave is past its last MFMA), then out as 16-byte pieces of whole rows:
    // 8 stores a thread instead of 64 two-byte ones
    bf16* const T = As[0];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int ml = wm * 64 + a * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
          T[ml * MLP_LDS_ROW + wn * 64 + b * 32 + lr] = (bf16)acc[a][b][e];
        }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < MLP_BM * MLP_BN / 8 / 256; ++r) {
      const int p = tid + 256 * r, row = p / (MLP_BN / 8), col = (p % (MLP_BN / 8)) * 8;
      *(uint4*)((bf16*)g.C + (long)(m0 + row) * g.ldc + n0 + col) = *(const uint4*)&T[row * MLP_LDS_ROW + col];
    }
    return;
  }
  if (EPI == EPI_F32_T && full && (g.ldc & 3) == 0 && ((uintptr_t)g.C & 15) == 0) {
    // C^T: a lane's elements 4r .. 4r + 3 are four consecutive rows m of one
    // column n, i.e. four consecutive floats of C^T's row n: one 16-byte store
    // (m is a multiple of 4, ldc too)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int n = n0 + wn * 64 + b * 32 + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wm * 64 + a * 32 + 8 * r + 4 * (lane >> 5);
          float4 v;
          v.x = acc[a][b][4 * r];
          v.y = acc[a][b][4 * r + 1];
          v.z = acc[a][b][4 * r + 2];
          v.w = acc[a][b][4 * r + 3];
          *(float4*)((float*)g.C + (long)n * g.ldc + m) = v;
        }
      }
    return;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = n0 + wn * 64 + b * 32 + lr;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int m = m0 + wm * 64 + a * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        if (!full && (m >= g.M || n >= g.N)) continue;
        const float v = acc[a][b][e];
        const long q = (long)m * g.ldc + n;
        if (EPI == EPI_F32) {
          ((float*)g.C)[q] = v;
        } else if (EPI == EPI_F32_T) {
          ((float*)g.C)[(long)n * g.ldc + m] = v;
        } else if (EPI == EPI_BF16_RELU || EPI == EPI_BF16_DRELU) {
          ((bf16*)g.C)[q] = (bf16)v;
        } else {
          atomicAdd(&((float*)g.C)[q], v);
        }
      }
    }
}

extern "C" int pywrf_mlp_gemm(int epi, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C,
                              int ldc, const float* bias, const void* H, int ldh, int ksplit, pywrf_stream_t s) {
  if (M < 0 || N < 0 || K < 0 || K % MLP_KQ || lda < K || ldb < K || !A || !B || !C) return PYWRF_E_RANGE;
  if (epi == EPI_F32_T ? ldc < M : ldc < N) return PYWRF_E_RANGE;
  if (lda % 8 || ldb % 8 || ((uintptr_t)A & 15) || ((uintptr_t)B & 15)) return PYWRF_E_RANGE;  // 16-byte rows
  if (epi == EPI_BF16_DRELU && (!H || ldh < N)) return PYWRF_E_RANGE;
  if (epi < 0 || epi > 4 || ksplit < 1 || (ksplit > 1 && epi != EPI_F32_ATOMIC)) return PYWRF_E_ARG;
  if ((long)(epi == EPI_F32_T ? N : M) * ldc >= (1L << 31) || (long)M * lda >= (1L << 31) ||
      (long)N * ldb >= (1L << 31))
    return PYWRF_E_RANGE;
  if (M == 0 || N == 0) return 0;
  GemmArgs g;
  g.M = M; g.N = N; g.K = K; g.lda = lda; g.ldb = ldb; g.ldc = ldc; g.ldh = ldh;
  g.A = (const bf16*)A; g.B = (const bf16*)B; g.C = C; g.bias = bias; g.H = (const bf16*)H;
  const int slabs = (K + MLP_BK - 1) / MLP_BK;
  const int per = (slabs + ksplit - 1) / ksplit;
  g.kchunk = per * MLP_BK;
  const int splits = K == 0 ? 1 : (slabs + per - 1) / per;
  dim3 grid((unsigned)((N + MLP_BN - 1) / MLP_BN), (unsigned)((M + MLP_BM - 1) / MLP_BM), (unsigned)splits);
  switch (epi) {
    case EPI_F32: hipLaunchKernelGGL(k_mlp_gemm<EPI_F32>, grid, dim3(256), 0, (hipStream_t)s, g); break;
    case EPI_BF16_RELU: hipLaunchKernelGGL(k_mlp_gemm<EPI_BF16_RELU>, grid, dim3(256), 0, (hipStream_t)s, g); break;
    case EPI_BF16_DRELU: hipLaunchKernelGGL(k_mlp_gemm<EPI_BF16_DRELU>, grid, dim3(256), 0, (hipStream_t)s, g); break;
    case EPI_F32_T: hipLaunchKernelGGL(k_mlp_gemm<EPI_F32_T>, grid, dim3(256), 0, (hipStream_t)s, g); break;
    default: hipLaunchKernelGGL(k_mlp_gemm<EPI_F32_ATOMIC>, grid, dim3(256), 0, (hipStream_t)s, g); break;
  }
  PW_RET();
}

// ------------------------------------------------------------------ fragment order
// W [N x K] (row-major, ldw) rewritten in the order the B operand of
// v_mfma_f32_32x32x16_bf16 is read: for row tile nt (32 rows) and k-step ks
// (16 columns), lane l's 8 values W[32 nt + l % 32][16 ks + 8 (l / 32) + ..]
// at out[((nt (K / 16) + ks) 64 + l) 8], so that a wave's fragment is one
// contiguous 1 KB read.  Rows at or beyond N are zero.
__global__ void k_mlp_frag_pack(int N, int K, const bf16* W, int ldw, bf16* out, long total) {
  const int nks = K / 16;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
    const int l = (int)(q % 64);
    const long fr = q / 64;
    const int ks = (int)(fr % nks), nt = (int)(fr / nks);
    const int row = 32 * nt + (l & 31), col = 16 * ks + 8 * (l >> 5);
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (row < N) v = *(const uint4*)(W + (long)row * ldw + col);
    *(uint4*)(out + q * 8) = v;
  }
}

extern "C" int pywrf_mlp_frag_pack(int N, int K, const void* W, int ldw, void* out, pywrf_stream_t s) {
  if (N < 0 || K < 0 || K % 16 || ldw < K || ldw % 8 || !W || !out) return PYWRF_E_RANGE;
  if (((uintptr_t)W & 15) || ((uintptr_t)out & 15) || (long)N * ldw >= (1L << 31)) return PYWRF_E_RANGE;
  const long total = (long)((N + 31) / 32) * (K / 16) * 64;
  if (total == 0) return 0;
  const unsigned blocks = (unsigned)imin((int)((total + 255) / 256), 4096);
  hipLaunchKernelGGL(k_mlp_frag_pack, dim3(blocks), dim3(256), 0, (hipStream_t)s, N, K, (const bf16*)W, ldw,
                     (bf16*)out, total);
  PW_RET();
}

// ------------------------------------------------------------------ transpose
// out[c][r] = in[r][c] for a bf16 matrix, through 64 x 64 LDS tiles (reads and
// writes along rows of 64 contiguous elements).  Entries of out's rows beyond
// R (up to ldo) are left untouched.
__global__ void __launch_bounds__(256) k_mlp_transpose(int R, int C, const bf16* in, int ldi, bf16* out, int ldo) {
  __shared__ bf16 t[64][65];
  const int r0 = (int)blockIdx.y * 64, c0 = (int)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int y = ty; y < 64; y += 4) {
    const int r = r0 + y, c = c0 + tx;
    t[y][tx] = (r < R && c < C) ? in[(long)r * ldi + c] : (bf16)0.0f;
  }
  __syncthreads();
  for (int y = ty; y < 64; y += 4) {
    const int c = c0 + y, r = r0 + tx;
    if (c < C && r < R) out[(long)c * ldo + r] = t[tx][y];
  }
}

extern "C" int pywrf_mlp_transpose(int R, int C, const void* in, int ldi, void* out, int ldo, pywrf_stream_t s) {
  if (R < 0 || C < 0 || ldi < C || ldo < R || !in || !out) return PYWRF_E_RANGE;
  if ((long)R * ldi >= (1L << 31) || (long)C * ldo >= (1L << 31)) return PYWRF_E_RANGE;
  if (R == 0 || C == 0) return 0;
  hipLaunchKernelGGL(k_mlp_transpose, dim3((unsigned)((C + 63) / 64), (unsigned)((R + 63) / 64)), dim3(256), 0,
                     (hipStream_t)s, R, C, (const bf16*)in, ldi, (bf16*)out, ldo);
  PW_RET();
}

// ------------------------------------------------------------------ row sums (bias gradients)
// out[r] = sum over c < C of in[r][c] (bf16 in, FP32 sums): one workgroup per row.
__global__ void __launch_bounds__(256) k_mlp_rowsum(int C, const bf16* in, int ld, float* out) {
  __shared__ float part[256];
  const int r = (int)blockIdx.x;
  float acc = 0.0f;
  for (int c = threadIdx.x; c < C; c += 256) acc += (float)in[(long)r * ld + c];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[r] = part[0];
}

extern "C" int pywrf_mlp_rowsum(int R, int C, const void* in, int ld, float* out, pywrf_stream_t s) {
  if (R < 0 || C < 0 || ld < C || !in || !out || (long)R * ld >= (1L << 31)) return PYWRF_E_RANGE;
  if (R == 0) return 0;
  hipLaunchKernelGGL(k_mlp_rowsum, dim3((unsigned)R), dim3(256), 0, (hipStream_t)s, C, (const bf16*)in, ld, out);
  PW_RET();
}

// ------------------------------------------------------------------ loss
// Weighted mean squared error over B rows of N outputs and its gradient:
//   loss += sum_{b,n} w[n] (Y - T)^2 * scale,   dY = bf16(2 w[n] (Y - T) * scale)
// (scale = 1 / (B * sum w) gives the mean).  loss accumulates atomically
// (start it at zero); dY entries of columns n in [N, lddy) are set to zero.
__global__ void __launch_bounds__(256) k_mlp_mse(int B, int N, const float* Y, int ldy, const float* T, int ldt,
                                                 const float* w, float scale, bf16* dY, int lddy, float* loss) {
  __shared__ float part[256];
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)B * lddy;
  float l = 0.0f;
  if (idx < total) {
    const int b = (int)(idx / lddy), n = (int)(idx % lddy);
    float g = 0.0f;
    if (n < N) {
      const float d = Y[(long)b * ldy + n] - T[(long)b * ldt + n];
      const float wn = w[n];
      l = wn * d * d * scale;
      g = 2.0f * wn * d * scale;
    }
    dY[idx] = (bf16)g;
  }
  part[threadIdx.x] = l;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) part[threadIdx.x] += part[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) atomicAdd(loss, part[0]);
}

extern "C" int pywrf_mlp_mse(int B, int N, const float* Y, int ldy, const float* T, int ldt, const float* w,
                             double scale, void* dY, int lddy, float* loss, pywrf_stream_t s) {
  if (B < 0 || N < 0 || ldy < N || ldt < N || lddy < N || !Y || !T || !w || !dY) return PYWRF_E_RANGE;
  const long total = (long)B * lddy;
  if (total >= (1L << 31)) return PYWRF_E_RANGE;
  if (total == 0) return 0;
  hipLaunchKernelGGL(k_mlp_mse, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)s, B, N, Y, ldy, T,
                     ldt, w, (float)scale, (bf16*)dY, lddy, loss);
  PW_RET();
}

// ------------------------------------------------------------------ Adam
// One Adam step (Kingma & Ba, bias-corrected) on an FP32 master array of
// rows x cols, gradient g * gscale; writes the updated master, and when wb /
// wtb are given its BF16 copy (rows x cols, row-major) and transposed BF16
// copy (cols x rows).
__global__ void __launch_bounds__(256) k_mlp_adam(int n, int cols, float* w, const float* g, float* m, float* v,
                                                  float lr_t, float b1, float b2, float eps, float gscale, bf16* wb,
                                                  bf16* wtb, int rows) {
  const int idx = (int)(blockIdx.x * 256 + threadIdx.x);
  if (idx >= n) return;
  const float gr = g[idx] * gscale;
  const float mm = b1 * m[idx] + (1.0f - b1) * gr;
  const float vv = b2 * v[idx] + (1.0f - b2) * gr * gr;
  m[idx] = mm;
  v[idx] = vv;
  const float nw = w[idx] - lr_t * mm / (sqrtf(vv) + eps);
  w[idx] = nw;
  if (wb) wb[idx] = (bf16)nw;
  if (wtb) {
    const int r = idx / cols, c = idx % cols;
    wtb[(long)c * rows + r] = (bf16)nw;
  }
}

extern "C" int pywrf_mlp_adam(int rows, int cols, float* w, const float* g, float* m, float* v, double lr,
                              double beta1, double beta2, double eps, double gscale, int t, void* wb, void* wtb,
                              pywrf_stream_t s) {
  if (rows < 0 || cols < 0 || t < 1 || !w || !g || !m || !v || (long)rows * cols >= (1L << 31)) return PYWRF_E_RANGE;
  const int n = rows * cols;
  if (n == 0) return 0;
  // bias corrections folded into the step size (double on the host)
  const double c1 = 1.0 - pow(beta1, t), c2 = 1.0 - pow(beta2, t);
  const double lr_t = lr * sqrt(c2) / c1;
  const double eps_t = eps * sqrt(c2);
  hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, n, cols, w, g, m, v,
                     (float)lr_t, (float)beta1, (float)beta2, (float)eps_t, (float)gscale, (bf16*)wb, (bf16*)wtb, rows);
  PW_RET();
}
