// WSM6 column surrogate: a 369 -> 512 -> 512 -> 287 ReLU network over each
// microphysics column, run inline in place of wsm6 (pywrf/wrf_physics.py:
// 137-1214, called at solver.py:1714-1756), with a per-column fallback to the
// plain scheme.
//
// Per column (the scheme's tile, the same columns and levels as wsm6):
//   input  41 levels x 9 fields  T = th * pii, p, rho, qv, qc, qr, qi, qs, qg
//          (feature var * 41 + k; level 40 and features 369..383 are padding)
//   output 41 levels x 7 tendencies  T, qv, qc, qr, qi, qs, qg  (feature
//          var * 41 + k, per second; 287 used, 288 stored)
// Inputs are standardised (x - mean) * rstd and rounded to BF16; outputs are
// standardised tendencies, out * std + mean.  The three layers run on the
// matrix cores (mlp.hip, BF16 operands, FP32 accumulation).
//
// Checks (a column that fails any of them gets the plain scheme instead):
//   * no mixing ratio below -qtol after the update (values in [-qtol, 0) are
//     set to 0);
//   * no water-mass gain: the column's total water sum_k rho dz (qv + qc + qr
//     + qi + qs + qg) may not rise by more than wtol times its old value
//     (the scheme only removes water, by precipitation);
//   * every updated value finite;
//   * the column inside the training envelope: each input the update reads
//     (T, rho and the six mixing ratios at every level) within the range the
//     training set showed for that feature, widened by (envelope - 1) times
//     its width on each side, and each tendency within envelope times the
//     largest magnitude the training set showed for it;
//   * no level heated or cooled by more than dtmax in the call (the deep
//     convective columns, where errors of the network feed back fastest,
//     are left to the plain scheme).
// A passing column's water loss sum_k rho dz (old - new) [kg m-2 = mm] is its
// surface precipitation of the step (rainncv, added to rainnc; the surrogate
// does not split it into snow and graupel, so snowncv = graupelncv = sr = 0).
//
// The whole surrogate is one launch (k_sg_fused: inputs, hidden layers and
// outputs stay in LDS, the checks and the update run in the same workgroup);
// its only workspace is the fallback list (int [ncol + 1]), which the plain
// scheme then takes.
#include <mutex>
#include "common.h"

typedef __bf16 bf16;

#define SG_LEV 41
#define SG_NIN 9
#define SG_NOUT 7
#define SG_KIN 384   // 369 padded to a multiple of 32
#define SG_H 512
#define SG_NO 288    // 287 padded
#define SG_COLS 64   // columns per pack workgroup
#ifndef SG_PW
#define SG_PW 8      // waves per pack workgroup (each takes every SG_PW-th level)
#endif

extern "C" int pywrf_mlp_gemm(int epi, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C,
                              int ldc, const float* bias, const void* H, int ldh, int ksplit, pywrf_stream_t s);
extern "C" int pywrf_wsm6_columns(const pywrf_grid_t* g, float* th, float* q, float* qc, float* qr, float* qi,
                                  float* qs, float* qg, const float* den, const float* pii, const float* p,
                                  const float* delz, double delt, float* rain, float* rainncv, float* snow,
                                  float* snowncv, float* sr, float* graupel, float* graupelncv, const int* cols,
                                  const int* ncols, pywrf_stream_t s);

struct Tile {
  Dims D;
  int j0, i0, ni, ncol, km;
};

struct Fields {
  float *th, *qv, *qc, *qr, *qi, *qs, *qg;
  const float *rho, *pii, *p, *dz;
  float *rain, *rainncv, *snowncv, *sr, *graupelncv;
};

// X[c - c_base][feature] for the columns c_base + [0, m) of the chunk: one
// workgroup per 64 columns, lane = column (rows of the fields are read
// contiguously), its SG_PW waves taking every SG_PW-th level (8: five levels
// a wave; 4 left too few waves in flight, 16 gained nothing more); all ten
// fields of a level (th, pii and the eight others) are loaded before any is
// used.  The
// standardised values go to LDS as [feature][column] and leave as 16-byte
// pieces of whole rows of X.
__global__ void __launch_bounds__(64 * SG_PW) k_sg_pack(Tile T, Fields f, const float* mean, const float* rstd, int c_base,
                                                 int m, bf16* X) {
  __shared__ __attribute__((aligned(16))) bf16 lds[SG_KIN][SG_COLS + 2];
  const int c0 = (int)blockIdx.x * SG_COLS;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int cl = c0 + tx, c = c_base + cl;
  const bool live = cl < m;
  const int j = T.j0 + (live ? c / T.ni : 0), i = T.i0 + (live ? c % T.ni : 0);
  for (int k = ty; k < SG_LEV; k += SG_PW) {
    const bool on = live && k < T.km;
    const long q = T.D.ix(on ? k : 0, j, i);
    float v[SG_NIN];                                    // loads at a valid (clamped) point, no branch
    const float th = f.th[q], pii = f.pii[q];
    v[1] = f.p[q]; v[2] = f.rho[q]; v[3] = f.qv[q]; v[4] = f.qc[q]; v[5] = f.qr[q];
    v[6] = f.qi[q]; v[7] = f.qs[q]; v[8] = f.qg[q];
    v[0] = th * pii;
#pragma unroll
    for (int var = 0; var < SG_NIN; ++var) {
      const int feat = var * SG_LEV + k;
      lds[feat][tx] = (bf16)(on ? (v[var] - mean[feat]) * rstd[feat] : 0.0f);
    }
  }
  for (int feat = SG_NIN * SG_LEV + ty; feat < SG_KIN; feat += SG_PW) lds[feat][tx] = (bf16)0.0f;
  __syncthreads();
  // 64 rows of 48 pieces of 8 features; consecutive threads write consecutive pieces
  for (int id = threadIdx.x; id < SG_COLS * (SG_KIN / 8); id += 64 * SG_PW) {
    const int row = id / (SG_KIN / 8), f8 = (id % (SG_KIN / 8)) * 8;
    if (c0 + row >= m) continue;
    typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
    bf16x8v piece;
#pragma unroll
    for (int e = 0; e < 8; ++e) piece[e] = lds[f8 + e][row];
    *(bf16x8v*)(X + (long)(c0 + row) * SG_KIN + f8) = piece;
  }
}

// The network's tendencies applied with the checks.  A column's levels are
// taken in SA_LG groups of SA_LPG consecutive levels; for each group the
// updated values are formed in registers, the per-level checks run and the
// group's column water is summed (levels in order); the groups' flags and
// sums then meet in LDS and are added in group order, and a column that
// passes every check stores its values, one that fails any is appended to the
// fallback list.
struct Envelope {
  const float *lo, *hi, *amax;
  float scale, dtmax;
};

// input feature var * 41 + k inside the widened training range
__device__ inline bool in_range(const Envelope& e, int feat, float x) {
  const float w = (e.hi[feat] - e.lo[feat]) * (e.scale - 1.0f);
  return x >= e.lo[feat] - w && x <= e.hi[feat] + w;
}

struct SgApply {
  const float *omean, *ostd;      // de-standardisation of the outputs
  Envelope env;
  float dt, qtol, wtol;
  int *fall, *nfall;              // fallback list and its count
};

#define SA_LG 16                  // level groups of a column
#define SA_LPG 3                  // levels per group (16 x 3 >= 41)

// level group g of column (j, i): out(o) is the network's standardised output o = var * 41 + k
template <class Out>
__device__ inline void sg_group(const Tile& T, const Fields& f, const SgApply& a, int g, int j, int i, Out out,
                                bool& bad, float& wo, float& wn, float (&nth)[SA_LPG],
                                float (&nq)[SA_LPG][SG_NOUT - 1]) {
  const Envelope& env = a.env;
  auto tend = [&](int v, int k) {
    const int o = v * SG_LEV + k;
    return out(o) * a.ostd[o] + a.omean[o];
  };
  bad = false;
  wo = 0.0f;
  wn = 0.0f;
#pragma unroll
  for (int kk = 0; kk < SA_LPG; ++kk) {
    const int k = g * SA_LPG + kk;
    if (k >= T.km) break;
    const long q = T.D.ix(k, j, i);
    const float th = f.th[q], pii = f.pii[q], rho = f.rho[q], mass = rho * f.dz[q];
    const float old[SG_NOUT - 1] = {f.qv[q], f.qc[q], f.qr[q], f.qi[q], f.qs[q], f.qg[q]};
    const float t0 = th * pii, tt = tend(0, k);
    const float tn = t0 + a.dt * tt;
    bad |= !(tn - tn == 0.0f) || !(fabsf(tn - t0) <= env.dtmax);   // NaN, infinite or too large
    bad |= !in_range(env, k, t0) || !in_range(env, 2 * SG_LEV + k, rho) || !(fabsf(tt) <= env.scale * env.amax[k]);
    nth[kk] = tn / pii;
    float so = 0.0f, sn = 0.0f;
#pragma unroll
    for (int v = 1; v < SG_NOUT; ++v) {
      const float tv = tend(v, k);
      const float x = old[v - 1] + a.dt * tv;
      bad |= !(x - x == 0.0f) || x < -a.qtol;
      bad |= !in_range(env, (v + 2) * SG_LEV + k, old[v - 1]) || !(fabsf(tv) <= env.scale * env.amax[v * SG_LEV + k]);
      nq[kk][v - 1] = x > 0.0f ? x : 0.0f;
      so += old[v - 1];
      sn += nq[kk][v - 1];
    }
    wo += mass * so;
    wn += mass * sn;
  }
}

// the stores of a passing column's level group g
__device__ inline void sg_group_store(const Tile& T, const Fields& f, int g, int j, int i, const float (&nth)[SA_LPG],
                                      const float (&nq)[SA_LPG][SG_NOUT - 1]) {
  float* dst[SG_NOUT - 1] = {f.qv, f.qc, f.qr, f.qi, f.qs, f.qg};
#pragma unroll
  for (int kk = 0; kk < SA_LPG; ++kk) {
    const int k = g * SA_LPG + kk;
    if (k >= T.km) break;
    const long q = T.D.ix(k, j, i);
    f.th[q] = nth[kk];
#pragma unroll
    for (int v = 0; v < SG_NOUT - 1; ++v) dst[v][q] = nq[kk][v];
  }
}

// the column's surface precipitation of the step: its water loss
__device__ inline void sg_precip(const Tile& T, const Fields& f, int j, int i, float w_old, float w_new) {
  const long q2 = T.D.ix2(j, i);
  const float pr = w_old - w_new > 0.0f ? w_old - w_new : 0.0f;
  f.rainncv[q2] = pr;
  f.rain[q2] = f.rain[q2] + pr;
  f.snowncv[q2] = 0.0f;
  f.graupelncv[q2] = 0.0f;
  f.sr[q2] = 0.0f;
}

__global__ void k_sg_zero_int(int* p) { *p = 0; }

// The network in one launch: a 512-thread workgroup (8 waves) takes SF_M = 128
// columns, packs their standardised inputs into LDS as BF16 rows (the
// k_sg_pack values), and runs the three layers out of LDS: each layer's
// output overwrites its input there (after a barrier) as the next layer's
// BF16 operand, so the hidden activations never leave the CU; the outputs
// either go through LDS to the checks and the update (APPLY, below) or to
// memory, feature-major as the GEMM's transposed epilogue writes them.
// Layers 1 and 2: wave w owns output features [64w, 64w + 64)
// of all 128 columns (4 x 2 MFMA tiles of 32 x 32: the A fragments come from
// LDS, each shared by two MFMAs, the B fragments -- L2-resident weights
// rewritten in fragment order, so each is one contiguous 1 KB read -- straight
// from global memory, each shared by four, fetched SF_KD k-steps ahead).
// Layer 3 (288 = 9 x 32 features): wave w owns feature tile w of the four
// column tiles, and waves 0..3 also tile 8 of column tile w (fragments fetched
// one k-step ahead).  Every 32 x 32
// output tile is accumulated from zero over k ascending in 16-deep MFMA
// steps, with the bias, ReLU and BF16 rounding of the GEMM epilogues, so the
// outputs equal the three-GEMM path's bit for bit.
#define SF_M 128
#define SF_LP (SG_H + 8)      // LDS row pitch in bf16: rows 4 banks apart, 16-byte fragment reads conflict-free
typedef __bf16 sf_bf16x8 __attribute__((ext_vector_type(8)));
typedef float sf_f32x16 __attribute__((ext_vector_type(16)));

// B fragment of row tile nt, k-step ks from a weight in fragment order
// (pywrf_mlp_frag_pack): one contiguous 1 KB read per wave
__device__ inline sf_bf16x8 sf_frag(const bf16* Wf, int nks, int nt, int ks, int lane) {
  return *(const sf_bf16x8*)(Wf + ((long)(nt * nks + ks) * 64 + lane) * 8);
}

// acc[t][u] += L[32 t + .., k] . W[n0 + 32 u + .., k] for k in [0, K), K a multiple of 32 SF_KD
#define SF_KD 2               // k-steps of B fragments fetched ahead (1, 2 and 4 time alike)

template <int NU>
__device__ inline void sf_layer(const bf16* L, const bf16* Wf, int K, int n0, int lane, sf_f32x16 (&acc)[4][NU]) {
  const int lr = lane & 31, lk = (lane >> 5) * 8, nks = K / 16, nt0 = n0 / 32;
  sf_bf16x8 b0[SF_KD][NU], b1[SF_KD][NU];
  auto fetch = [&](sf_bf16x8 (&b)[SF_KD][NU], int kb) {
#pragma unroll
    for (int s = 0; s < SF_KD; ++s)
#pragma unroll
      for (int u = 0; u < NU; ++u) b[s][u] = sf_frag(Wf, nks, nt0 + u, kb / 16 + s, lane);
  };
  auto run = [&](const sf_bf16x8 (&b)[SF_KD][NU], int kb) {
#pragma unroll
    for (int s = 0; s < SF_KD; ++s) {
      sf_bf16x8 af[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) af[t] = *(const sf_bf16x8*)&L[(32 * t + lr) * SF_LP + kb + 16 * s + lk];
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int u = 0; u < NU; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[t], b[s][u], acc[t][u], 0, 0, 0);
    }
  };
  constexpr int D = 16 * SF_KD;
  fetch(b0, 0);
#pragma unroll 1
  for (int kb = 0; kb < K; kb += 2 * D) {
    fetch(b1, kb + D);
    run(b0, kb);
    if (kb + 2 * D < K) fetch(b0, kb + 2 * D);
    run(b1, kb + D);
  }
}

// acc -> bf16(relu(acc + bias)) into L[row][n0 + ..] (rows of the 4 column tiles)
template <int NU>
__device__ inline void sf_store_hidden(bf16* L, const float* bias, int n0, int lane, const sf_f32x16 (&acc)[4][NU]) {
  const int lr = lane & 31;
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const float bv = bias[n0 + 32 * u + lr];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = 32 * t + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        float v = acc[t][u][e] + bv;
        v = v > 0.0f ? v : 0.0f;
        L[row * SF_LP + n0 + 32 * u + lr] = (bf16)v;
      }
  }
}

// column tile t's 32 x 32 output tile of features [n0, n0 + 32) + bias into Yt (feature-major)
__device__ inline void sf_store_out(float* Yt, int ldy, int c0, int mv, const float* bias, int n0, int t, int lane,
                                    const sf_f32x16& acc) {
  const int n = n0 + (lane & 31);
  const float bv = bias[n];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = 32 * t + 8 * r + 4 * (lane >> 5);
    float* dst = Yt + (long)n * ldy + c0 + m;
    if (m + 4 <= mv) {
      float4 v;
      v.x = acc[4 * r] + bv;
      v.y = acc[4 * r + 1] + bv;
      v.z = acc[4 * r + 2] + bv;
      v.w = acc[4 * r + 3] + bv;
      *(float4*)dst = v;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (m + e < mv) dst[e] = acc[4 * r + e] + bv;
    }
  }
}

// APPLY: the outputs go through LDS, 64 columns at a time (FP32, pitch
// SF_FP), to the checks and the update (sg_group, the column's 16 level groups
// taken by the 8 waves two each), and nothing but the updated fields, the
// precipitation and the fallback list leave the CU.  Otherwise the outputs are
// stored to Yt (feature-major, [288][ldy]).
#define SF_FP (SG_NO + 1)     // odd pitch: a wave's 64 columns fall in distinct banks
template <bool APPLY>
__global__ void __launch_bounds__(512) k_sg_fused(Tile T, Fields f, const float* mean, const float* rstd,
                                                   const bf16* w1, const float* b1, const bf16* w2, const float* b2,
                                                   const bf16* w3, const float* b3, int c_base, int m, float* Yt,
                                                   int ldy, SgApply a) {
  __shared__ __attribute__((aligned(16))) bf16 L[SF_M * SF_LP];
  static_assert(64 * SF_FP * 4 + 3 * SA_LG * 64 * 4 <= SF_M * SF_LP * 2, "the apply buffers fit in L");
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c0 = (int)blockIdx.x * SF_M;                 // chunk-local first column of the tile
  const int mv = imin(SF_M, m - c0);                     // columns of the tile in the chunk
  // inputs: thread = (column tid % 128, level group tid / 128), as k_sg_pack
  {
    const int cl = tid & (SF_M - 1), grp = tid >> 7;
    const bool live = cl < mv;
    const int c = c_base + c0 + (live ? cl : 0);
    const int j = T.j0 + c / T.ni, i = T.i0 + c % T.ni;
    bf16* row = L + cl * SF_LP;
    for (int k = grp; k < SG_LEV; k += 4) {
      const bool on = live && k < T.km;
      const long q = T.D.ix(on ? k : 0, j, i);
      float v[SG_NIN];
      const float th = f.th[q], pii = f.pii[q];
      v[1] = f.p[q]; v[2] = f.rho[q]; v[3] = f.qv[q]; v[4] = f.qc[q]; v[5] = f.qr[q];
      v[6] = f.qi[q]; v[7] = f.qs[q]; v[8] = f.qg[q];
      v[0] = th * pii;
#pragma unroll
      for (int var = 0; var < SG_NIN; ++var) {
        const int feat = var * SG_LEV + k;
        row[feat] = (bf16)(on ? (v[var] - mean[feat]) * rstd[feat] : 0.0f);
      }
    }
    for (int feat = SG_NIN * SG_LEV + grp; feat < SG_KIN; feat += 4) row[feat] = (bf16)0.0f;
  }
  __syncthreads();
  sf_f32x16 acc[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[t][u][e] = 0.0f;
  sf_layer<2>(L, w1, SG_KIN, 64 * w, lane, acc);
  __syncthreads();                                       // every wave is done reading the inputs
  sf_store_hidden<2>(L, b1, 64 * w, lane, acc);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[t][u][e] = 0.0f;
  sf_layer<2>(L, w2, SG_H, 64 * w, lane, acc);
  __syncthreads();
  sf_store_hidden<2>(L, b2, 64 * w, lane, acc);
  __syncthreads();
  // layer 3: feature tile w of the four column tiles, and for waves 0..3 tile 8 of column tile w
  sf_f32x16 o[4][1], ox;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int e = 0; e < 16; ++e) o[t][0][e] = 0.0f;
#pragma unroll
  for (int e = 0; e < 16; ++e) ox[e] = 0.0f;
  {
    const int lr = lane & 31, lk = (lane >> 5) * 8;
    const bool extra = w < 4;                            // uniform across the wave
    const int xr = 32 * (extra ? w : 0) + lr;            // the extra tile's column rows
    // the k-step's two B fragments, fetched one k-step ahead
    sf_bf16x8 b = sf_frag(w3, SG_H / 16, w, 0, lane), bx = sf_frag(w3, SG_H / 16, 8, 0, lane);
#pragma unroll 2
    for (int ks = 0; ks < SG_H; ks += 16) {
      const int kn = ks + 16 < SG_H ? ks + 16 : ks;      // the last k-step fetches its own again (unused)
      const sf_bf16x8 nb = sf_frag(w3, SG_H / 16, w, kn / 16, lane), nbx = sf_frag(w3, SG_H / 16, 8, kn / 16, lane);
      sf_bf16x8 af[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) af[t] = *(const sf_bf16x8*)&L[(32 * t + lr) * SF_LP + ks + lk];
#pragma unroll
      for (int t = 0; t < 4; ++t) o[t][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[t], b, o[t][0], 0, 0, 0);
      if (extra) {
        const sf_bf16x8 ax = *(const sf_bf16x8*)&L[xr * SF_LP + ks + lk];
        ox = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ax, bx, ox, 0, 0, 0);
      }
      b = nb;
      bx = nbx;
    }
    if (!APPLY) {
      if (c0 >= m) return;
#pragma unroll
      for (int t = 0; t < 4; ++t) sf_store_out(Yt, ldy, c0, mv, b3, 32 * w, t, lane, o[t][0]);
      if (extra) sf_store_out(Yt, ldy, c0, mv, b3, 8 * 32, w, lane, ox);
      return;
    }
  }
  float* F = (float*)L;                                  // [64][SF_FP] outputs (+ bias) of half the columns
  float* s_wo = F + 64 * SF_FP;                          // [SA_LG][64] group water before / after, flags
  float* s_wn = s_wo + SA_LG * 64;
  int* s_bad = (int*)(s_wn + SA_LG * 64);
  const int lr = lane & 31;
  const int col = tid & 63, grp = tid >> 6;             // apply: column of the half, level groups grp, grp + 8
  const float bw = b3[32 * w + lr], bx = b3[8 * 32 + lr];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();                                     // layer 3 done with L / the previous half done with F
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = 32 * tt + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        F[r * SF_FP + 32 * w + lr] = o[2 * h + tt][0][e] + bw;
      }
    if (w < 4 && (w >> 1) == h)                          // the extra tile (column tile w, features 256..287)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = 32 * (w - 2 * h) + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        F[r * SF_FP + 8 * 32 + lr] = ox[e] + bx;
      }
    __syncthreads();
    const int cl = 64 * h + col;
    const bool live = cl < mv;
    const int c = c_base + c0 + (live ? cl : 0);
    const int j = T.j0 + c / T.ni, i = T.i0 + c % T.ni;
    auto out = [&](int oi) { return F[col * SF_FP + oi]; };
    bool bad0, bad1;
    float wo0, wn0, wo1, wn1;
    float nth0[SA_LPG], nq0[SA_LPG][SG_NOUT - 1], nth1[SA_LPG], nq1[SA_LPG][SG_NOUT - 1];
    sg_group(T, f, a, grp, j, i, out, bad0, wo0, wn0, nth0, nq0);
    sg_group(T, f, a, grp + 8, j, i, out, bad1, wo1, wn1, nth1, nq1);
    s_wo[grp * 64 + col] = wo0;
    s_wn[grp * 64 + col] = wn0;
    s_bad[grp * 64 + col] = bad0 ? 1 : 0;
    s_wo[(grp + 8) * 64 + col] = wo1;
    s_wn[(grp + 8) * 64 + col] = wn1;
    s_bad[(grp + 8) * 64 + col] = bad1 ? 1 : 0;
    __syncthreads();
    float w_old = 0.0f, w_new = 0.0f;
    int nbad = 0;
#pragma unroll
    for (int gq = 0; gq < SA_LG; ++gq) {
      w_old += s_wo[gq * 64 + col];
      w_new += s_wn[gq * 64 + col];
      nbad += s_bad[gq * 64 + col];
    }
    if (live) {
      if (nbad || w_new - w_old > a.wtol * w_old) {
        if (grp == 0) a.fall[atomicAdd(a.nfall, 1)] = c;
      } else {
        sg_group_store(T, f, grp, j, i, nth0, nq0);
        sg_group_store(T, f, grp + 8, j, i, nth1, nq1);
        if (grp == 0) sg_precip(T, f, j, i, w_old, w_new);
      }
    }
  }
}

// workspace bytes: the fallback list and its count (the fused kernel keeps
// inputs, hidden layers and outputs on chip)
static long sg_list_bytes(long ncol) { return ((ncol + 1) * 4 + 255) / 256 * 256; }

static int sg_tile(const pywrf_grid_t* g, Tile& T, long& ncol) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  const int km = (g->kte - 1) - g->kts;
  if (g->kts != 0 || km < 4 || km > SG_LEV || g->kte > g->nz) return PYWRF_E_RANGE;
  if (g->jts < 0 || g->its < 0 || g->jte - 1 > g->ny || g->ite - 1 > g->nx) return PYWRF_E_RANGE;
  const int ni = (g->ite - 1) - g->its, nj = (g->jte - 1) - g->jts;
  ncol = ni > 0 && nj > 0 ? (long)ni * nj : 0;
  T.D = dims_of(g); T.j0 = g->jts; T.i0 = g->its; T.ni = ni; T.ncol = (int)ncol; T.km = km;
  return 0;
}

// apply: run the checks and the update in the same launch (Y unused); else
// store the outputs to Y
static void sg_fused_launch(const Tile& T, const Fields& f, const pywrf_surrogate_t* net, int c_base, int m, float* Y,
                            int ldy, const SgApply* apply, pywrf_stream_t s) {
  const dim3 grid((unsigned)((m + SF_M - 1) / SF_M)), block(512);
  const bf16 *w1 = (const bf16*)net->w1f, *w2 = (const bf16*)net->w2f, *w3 = (const bf16*)net->w3f;
  if (apply)
    hipLaunchKernelGGL(k_sg_fused<true>, grid, block, 0, (hipStream_t)s, T, f, net->in_mean, net->in_rstd, w1, net->b1,
                       w2, net->b2, w3, net->b3, c_base, m, Y, ldy, *apply);
  else
    hipLaunchKernelGGL(k_sg_fused<false>, grid, block, 0, (hipStream_t)s, T, f, net->in_mean, net->in_rstd, w1,
                       net->b1, w2, net->b2, w3, net->b3, c_base, m, Y, ldy, SgApply{});
}

static int sg_run(const pywrf_grid_t* g, const pywrf_surrogate_t* net, float* th, const float* rho,
                  const float* pii, const float* p, const float* delz, double dt, float* qv, float* qc, float* qr,
                  float* qi, float* qs, float* qg, float* rain, float* rainncv, float* snow, float* snowncv, float* sr,
                  float* graupel, float* graupelncv, char* ws, long ws_bytes, int* nfall, pywrf_stream_t s) {
  if (!net || !ws || !nfall) return PYWRF_E_RANGE;
  if (!net->w1f || !net->w2f || !net->w3f) return PYWRF_E_ARG;
  Tile T;
  long ncol;
  if (int e = sg_tile(g, T, ncol)) return e;
  if (ncol == 0) return 0;
  if (ws_bytes < sg_list_bytes(ncol)) return PYWRF_E_RANGE;
  int* fall = (int*)ws;
  Fields f;
  f.th = th; f.qv = qv; f.qc = qc; f.qr = qr; f.qi = qi; f.qs = qs; f.qg = qg;
  f.rho = rho; f.pii = pii; f.p = p; f.dz = delz;
  f.rain = rain; f.rainncv = rainncv; f.snowncv = snowncv; f.sr = sr; f.graupelncv = graupelncv;
  (void)snow; (void)graupel;
  if (!net->in_lo || !net->in_hi || !net->out_absmax || !(net->envelope >= 1.0f)) return PYWRF_E_ARG;
  SgApply a;
  a.omean = net->out_mean; a.ostd = net->out_std;
  a.env.lo = net->in_lo; a.env.hi = net->in_hi; a.env.amax = net->out_absmax; a.env.scale = net->envelope;
  a.env.dtmax = net->dtmax;
  a.dt = (float)dt; a.qtol = net->qtol; a.wtol = net->wtol; a.fall = fall; a.nfall = nfall;
  hipLaunchKernelGGL(k_sg_zero_int, dim3(1), dim3(1), 0, (hipStream_t)s, nfall);
  sg_fused_launch(T, f, net, 0, (int)ncol, nullptr, 0, &a, s);
  // the plain scheme on the columns that failed a check
  return pywrf_wsm6_columns(g, th, qv, qc, qr, qi, qs, qg, rho, pii, p, delz, dt, rain, rainncv, snow, snowncv, sr,
                            graupel, graupelncv, fall, nfall, s);
}

// The network's outputs for every tile column, feature-major (Y[f * ldy + c]),
// by the fused kernel (fused = 1) or by k_sg_pack and three GEMM launches
// through `work` (fused = 0: packed inputs and both hidden layers, 2816 bytes
// a column, columns rounded up to 128); the two give the same bits.
extern "C" int pywrf_wsm6_surrogate_outputs(const pywrf_grid_t* g, const pywrf_surrogate_t* net, const float* th,
                                            const float* rho, const float* pii, const float* p, const float* qv,
                                            const float* qc, const float* qr, const float* qi, const float* qs,
                                            const float* qg, float* Y, int ldy, int fused, void* work,
                                            long long work_bytes, pywrf_stream_t s) {
  if (!net || !Y) return PYWRF_E_RANGE;
  Tile T;
  long ncol;
  if (int e = sg_tile(g, T, ncol)) return e;
  if (ncol == 0) return 0;
  if (ldy < ncol || (long)SG_NO * ldy >= (1L << 31)) return PYWRF_E_RANGE;
  if (fused && (ldy % 4 || ((uintptr_t)Y & 15))) return PYWRF_E_RANGE;   // 16-byte stores of four columns
  Fields f = {};
  f.th = (float*)th; f.qv = (float*)qv; f.qc = (float*)qc; f.qr = (float*)qr; f.qi = (float*)qi;
  f.qs = (float*)qs; f.qg = (float*)qg; f.rho = rho; f.pii = pii; f.p = p;
  if (fused) {
    if (!net->w1f || !net->w2f || !net->w3f) return PYWRF_E_ARG;
    sg_fused_launch(T, f, net, 0, (int)ncol, Y, ldy, nullptr, s);
    PW_RET();
  }
  const long rows = (ncol + 127) / 128 * 128;
  if (!work || work_bytes < rows * (SG_KIN * 2L + 2L * SG_H * 2L)) return PYWRF_E_RANGE;
  bf16* X = (bf16*)work;
  bf16* H1 = X + rows * SG_KIN;
  bf16* H2 = H1 + rows * SG_H;
  const int m = (int)ncol;
  hipLaunchKernelGGL(k_sg_pack, dim3((unsigned)((m + SG_COLS - 1) / SG_COLS)), dim3(64 * SG_PW), 0, (hipStream_t)s,
                     T, f, net->in_mean, net->in_rstd, 0, m, X);
  int e = pywrf_mlp_gemm(1, m, SG_H, SG_KIN, X, SG_KIN, net->w1, SG_KIN, H1, SG_H, net->b1, nullptr, 0, 1, s);
  if (!e) e = pywrf_mlp_gemm(1, m, SG_H, SG_H, H1, SG_H, net->w2, SG_H, H2, SG_H, net->b2, nullptr, 0, 1, s);
  if (!e) e = pywrf_mlp_gemm(4, m, SG_NO, SG_H, H2, SG_H, net->w3, SG_H, Y, ldy, net->b3, nullptr, 0, 1, s);
  return e;
}

extern "C" int pywrf_wsm6_surrogate_workspace(const pywrf_grid_t* g, int chunk, long long* bytes) {
  if (!g || chunk < 128 || !bytes) return PYWRF_E_ARG;
  const long ncol = (long)imax((g->ite - 1) - g->its, 0) * imax((g->jte - 1) - g->jts, 0);
  *bytes = sg_list_bytes(ncol);
  return 0;
}

extern "C" int pywrf_wsm6_surrogate(const pywrf_grid_t* g, const pywrf_surrogate_t* net, float* th,
                                    const float* rho, const float* pii, const float* p, const float* delz, double dt,
                                    float* qv, float* qc, float* qr, float* qi, float* qs, float* qg, float* rain,
                                    float* rainncv, float* snow, float* snowncv, float* sr, float* graupel,
                                    float* graupelncv, void* workspace, long long workspace_bytes, int* nfallback,
                                    pywrf_stream_t s) {
  return sg_run(g, net, th, rho, pii, p, delz, dt, qv, qc, qr, qi, qs, qg, rain, rainncv, snow, snowncv, sr, graupel,
                graupelncv, (char*)workspace, (long)workspace_bytes, nfallback, s);
}

// ------------------------------------------------------------------ allocation-free variant
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops): the same computation through a file-scope device
// array instead of a caller's workspace.  It holds the fallback list of up to
// SG_NA_MAX_COLS columns (0.26 MB on every GPU the library runs on, for the
// life of the process).
// Every call on a GPU uses that GPU's copy, so calls that overlap on one GPU
// race; the fallback counter (the caller's *nfallback) is zeroed by a kernel
// at the start of each call and the list is written before it is read.
#define SG_NA_MAX_COLS 65536
#define SG_NA_BYTES ((((long)SG_NA_MAX_COLS + 1) * 4 + 255) / 256 * 256)
static __device__ __align__(256) char scratch[SG_NA_BYTES];

#define SG_NA_MAX_DEVICES 64
static void* sg_scratch_ptr[SG_NA_MAX_DEVICES];
static std::mutex sg_scratch_lock;

// The device address of `scratch` on the current device (the array's name is
// a host-side stand-in, not a device address); looked up once per device.
static int sg_scratch(char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SG_NA_MAX_DEVICES) return PYWRF_E_RANGE;
  std::lock_guard<std::mutex> guard(sg_scratch_lock);
  if (!sg_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return e != hipSuccess ? (int)e : PYWRF_E_RANGE;
    sg_scratch_ptr[dev] = ptr;
  }
  *out = (char*)sg_scratch_ptr[dev];
  return 0;
}

extern "C" int pywrf_wsm6_surrogate_noalloc(const pywrf_grid_t* g, const pywrf_surrogate_t* net, float* th,
                                            const float* rho, const float* pii, const float* p, const float* delz,
                                            double dt, float* qv, float* qc, float* qr, float* qi, float* qs,
                                            float* qg, float* rain, float* rainncv, float* snow, float* snowncv,
                                            float* sr, float* graupel, float* graupelncv, int* nfallback,
                                            pywrf_stream_t s) {
  if (!g) return PYWRF_E_RANGE;
  const long ncol = (long)imax((g->ite - 1) - g->its, 0) * imax((g->jte - 1) - g->jts, 0);
  if (ncol > SG_NA_MAX_COLS) return PYWRF_E_RANGE;     // more columns than the static list holds
  char* ws = nullptr;
  int e = sg_scratch(&ws);
  if (e) return e;
  return sg_run(g, net, th, rho, pii, p, delz, dt, qv, qc, qr, qi, qs, qg, rain, rainncv, snow, snowncv, sr, graupel,
                graupelncv, ws, SG_NA_BYTES, nfallback, s);
}

extern "C" int pywrf_wsm6_surrogate_noalloc_limits(int* max_columns, long long* scratch_bytes) {
  if (!max_columns || !scratch_bytes) return PYWRF_E_ARG;
  *max_columns = SG_NA_MAX_COLS;
  *scratch_bytes = SG_NA_BYTES;
  return 0;
}
