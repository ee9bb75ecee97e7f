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
// Columns are processed in chunks through a workspace: the fallback list
// (int [ncol]), then packed inputs X (bf16 [chunk x 384]), hidden layers
// (bf16 [chunk x 512] x 2) and outputs (f32, feature-major [288 x chunk]).
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

// The network's tendencies applied with the checks, on a workgroup of 64
// columns x SA_LG level groups (16: three levels a wave; 8 groups ran 35 us
// longer at 240 x 240 for want of waves in flight): lane = column (every load and store of a
// level is a row segment of the field), wave = a group of SA_LPG consecutive
// levels.  Each thread forms the updated values of its levels in registers,
// runs the per-level checks and sums its levels' column water; the groups'
// flags and sums meet in LDS (added in group order), and a column that
// passes every check stores its values, one that fails any is appended to
// the fallback list.  Yt holds the network's outputs feature-major,
// [288][ldy].
struct Envelope {
  const float *lo, *hi, *amax;
  float scale, dtmax;
};

// input feature var * 41 + k inside the widened training range
__device__ inline bool in_range(const Envelope& e, int feat, float x) {
  const float w = (e.hi[feat] - e.lo[feat]) * (e.scale - 1.0f);
  return x >= e.lo[feat] - w && x <= e.hi[feat] + w;
}

#ifndef SA_LG
#define SA_LG 16                                   // level groups (waves) per workgroup
#endif
#define SA_LPG ((SG_LEV + SA_LG - 1) / SA_LG)      // levels per group
__global__ void __launch_bounds__(64 * SA_LG) k_sg_apply(Tile T, Fields f, const float* Yt, int ldy,
                                                          const float* omean, const float* ostd, Envelope env,
                                                          float dt, float qtol, float wtol, int c_base, int m,
                                                          int* fall, int* nfall) {
  __shared__ float s_wo[SA_LG][64], s_wn[SA_LG][64];
  __shared__ int s_bad[SA_LG][64];
  const int tx = threadIdx.x, lg = threadIdx.y;
  const int cl = (int)blockIdx.x * 64 + tx;
  const bool live = cl < m;
  const int c = c_base + (live ? cl : 0);
  const int j = T.j0 + c / T.ni, i = T.i0 + c % T.ni;
  const int k0 = lg * SA_LPG;
  auto tend = [&](int v, int k) {
    const int o = v * SG_LEV + k;
    return Yt[(long)o * ldy + (live ? cl : 0)] * ostd[o] + omean[o];
  };
  bool bad = false;
  float wo = 0.0f, wn = 0.0f;
  float nth[SA_LPG], nq[SA_LPG][SG_NOUT - 1];
#pragma unroll
  for (int kk = 0; kk < SA_LPG; ++kk) {
    const int k = k0 + kk;
    if (k >= T.km) break;                      // uniform across the wave
    const long q = T.D.ix(k, j, i);
    const float th = f.th[q], pii = f.pii[q], rho = f.rho[q], mass = rho * f.dz[q];
    const float old[SG_NOUT - 1] = {f.qv[q], f.qc[q], f.qr[q], f.qi[q], f.qs[q], f.qg[q]};
    const float t0 = th * pii, tt = tend(0, k);
    const float tn = t0 + dt * tt;
    bad |= !(tn - tn == 0.0f) || !(fabsf(tn - t0) <= env.dtmax);   // NaN, infinite or too large
    bad |= !in_range(env, k, t0) || !in_range(env, 2 * SG_LEV + k, rho) || !(fabsf(tt) <= env.scale * env.amax[k]);
    nth[kk] = tn / pii;
    float so = 0.0f, sn = 0.0f;
#pragma unroll
    for (int v = 1; v < SG_NOUT; ++v) {
      const float tv = tend(v, k);
      const float x = old[v - 1] + dt * tv;
      bad |= !(x - x == 0.0f) || x < -qtol;
      bad |= !in_range(env, (v + 2) * SG_LEV + k, old[v - 1]) || !(fabsf(tv) <= env.scale * env.amax[v * SG_LEV + k]);
      nq[kk][v - 1] = x > 0.0f ? x : 0.0f;
      so += old[v - 1];
      sn += nq[kk][v - 1];
    }
    wo += mass * so;
    wn += mass * sn;
  }
  s_wo[lg][tx] = wo;
  s_wn[lg][tx] = wn;
  s_bad[lg][tx] = bad ? 1 : 0;
  __syncthreads();
  float w_old = 0.0f, w_new = 0.0f;
  int nbad = 0;
#pragma unroll
  for (int gq = 0; gq < SA_LG; ++gq) {
    w_old += s_wo[gq][tx];
    w_new += s_wn[gq][tx];
    nbad += s_bad[gq][tx];
  }
  if (!live) return;
  if (nbad || w_new - w_old > wtol * w_old) {
    if (lg == 0) fall[atomicAdd(nfall, 1)] = c;
    return;
  }
  float* dst[SG_NOUT - 1] = {f.qv, f.qc, f.qr, f.qi, f.qs, f.qg};
#pragma unroll
  for (int kk = 0; kk < SA_LPG; ++kk) {
    const int k = k0 + kk;
    if (k >= T.km) break;
    const long q = T.D.ix(k, j, i);
    f.th[q] = nth[kk];
#pragma unroll
    for (int v = 0; v < SG_NOUT - 1; ++v) dst[v][q] = nq[kk][v];
  }
  if (lg == 0) {
    const long q2 = T.D.ix2(j, i);
    const float pr = w_old - w_new > 0.0f ? w_old - w_new : 0.0f;
    f.rainncv[q2] = pr;
    f.rain[q2] = f.rain[q2] + pr;
    f.snowncv[q2] = 0.0f;
    f.graupelncv[q2] = 0.0f;
    f.sr[q2] = 0.0f;
  }
}

__global__ void k_sg_zero_int(int* p) { *p = 0; }

static long sg_chunk_bytes(long chunk) {
  return chunk * (SG_KIN * 2L + 2L * SG_H * 2L + SG_NO * 4L);
}

static int sg_run(const pywrf_grid_t* g, const pywrf_surrogate_t* net, float* th, const float* rho,
                  const float* pii, const float* p, const float* delz, double dt, float* qv, float* qc, float* qr,
                  float* qi, float* qs, float* qg, float* rain, float* rainncv, float* snow, float* snowncv, float* sr,
                  float* graupel, float* graupelncv, char* ws, long ws_bytes, int* nfall, pywrf_stream_t s) {
  if (!pw_size_ok(g) || !net || !ws || !nfall) return PYWRF_E_RANGE;
  const int km = (g->kte - 1) - g->kts;
  if (g->kts != 0 || km < 4 || km > SG_LEV || g->kte > g->nz) return PYWRF_E_RANGE;
  if (g->jts < 0 || g->its < 0 || g->jte - 1 > g->ny || g->ite - 1 > g->nx) return PYWRF_E_RANGE;
  const int ni = (g->ite - 1) - g->its, nj = (g->jte - 1) - g->jts;
  if (ni <= 0 || nj <= 0) return 0;
  const long ncol = (long)ni * nj;
  // workspace: fallback list and counter first, then the chunk buffers
  const long list_bytes = ((ncol + 1) * 4 + 255) / 256 * 256;
  const long avail = ws_bytes - list_bytes;
  long chunk = avail > 0 ? avail / sg_chunk_bytes(1) : 0;
  chunk = chunk / 128 * 128;
  if (chunk < 128) return PYWRF_E_RANGE;
  if (chunk > ncol) chunk = (ncol + 127) / 128 * 128;
  int* fall = (int*)ws;
  char* b = ws + list_bytes;
  bf16* X = (bf16*)b;
  bf16* H1 = X + chunk * SG_KIN;
  bf16* H2 = H1 + chunk * SG_H;
  float* Y = (float*)(H2 + chunk * SG_H);
  Tile T;
  T.D = dims_of(g); T.j0 = g->jts; T.i0 = g->its; T.ni = ni; T.ncol = (int)ncol; T.km = km;
  Fields f;
  f.th = th; f.qv = qv; f.qc = qc; f.qr = qr; f.qi = qi; f.qs = qs; f.qg = qg;
  f.rho = rho; f.pii = pii; f.p = p; f.dz = delz;
  f.rain = rain; f.rainncv = rainncv; f.snowncv = snowncv; f.sr = sr; f.graupelncv = graupelncv;
  (void)snow; (void)graupel;
  if (!net->in_lo || !net->in_hi || !net->out_absmax || !(net->envelope >= 1.0f)) return PYWRF_E_ARG;
  Envelope env;
  env.lo = net->in_lo; env.hi = net->in_hi; env.amax = net->out_absmax; env.scale = net->envelope;
  env.dtmax = net->dtmax;
  hipLaunchKernelGGL(k_sg_zero_int, dim3(1), dim3(1), 0, (hipStream_t)s, nfall);
  for (long c0 = 0; c0 < ncol; c0 += chunk) {
    const int m = (int)(ncol - c0 < chunk ? ncol - c0 : chunk);
    hipLaunchKernelGGL(k_sg_pack, dim3((unsigned)((m + SG_COLS - 1) / SG_COLS)), dim3(64 * SG_PW), 0, (hipStream_t)s, T, f,
                       net->in_mean, net->in_rstd, (int)c0, m, X);
    int e = pywrf_mlp_gemm(1, m, SG_H, SG_KIN, X, SG_KIN, net->w1, SG_KIN, H1, SG_H, net->b1, nullptr, 0, 1, s);
    if (!e) e = pywrf_mlp_gemm(1, m, SG_H, SG_H, H1, SG_H, net->w2, SG_H, H2, SG_H, net->b2, nullptr, 0, 1, s);
    // the outputs feature-major (Y^T, [288][chunk]) for the column-per-thread update
    if (!e) e = pywrf_mlp_gemm(4, m, SG_NO, SG_H, H2, SG_H, net->w3, SG_H, Y, (int)chunk, net->b3, nullptr, 0, 1, s);
    if (e) return e;
    hipLaunchKernelGGL(k_sg_apply, dim3((unsigned)((m + 63) / 64)), dim3(64, SA_LG), 0, (hipStream_t)s, T, f, Y,
                       (int)chunk, net->out_mean, net->out_std, env, (float)dt, net->qtol, net->wtol, (int)c0, m,
                       fall, nfall);
  }
  // the plain scheme on the columns that failed a check
  return pywrf_wsm6_columns(g, th, qv, qc, qr, qi, qs, qg, rho, pii, p, delz, dt, rain, rainncv, snow, snowncv, sr,
                            graupel, graupelncv, fall, nfall, s);
}

extern "C" int pywrf_wsm6_surrogate_workspace(const pywrf_grid_t* g, int chunk, long long* bytes) {
  if (!g || chunk < 128 || !bytes) return PYWRF_E_ARG;
  const long ncol = (long)imax((g->ite - 1) - g->its, 0) * imax((g->jte - 1) - g->jts, 0);
  *bytes = ((ncol + 1) * 4 + 255) / 256 * 256 + sg_chunk_bytes(chunk / 128 * 128);
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
// SG_NA_MAX_COLS columns and the buffers of one chunk of SG_NA_CHUNK columns
// (65.3 MB on every GPU the library runs on, for the life of the process).
// Every call on a GPU uses that GPU's copy, so calls that overlap on one GPU
// race; the fallback counter (the caller's *nfallback) is zeroed by a kernel
// at the start of each call and the list is written before it is read.
#define SG_NA_MAX_COLS 65536
#define SG_NA_CHUNK 16384
#define SG_NA_LIST_BYTES ((((long)SG_NA_MAX_COLS + 1) * 4 + 255) / 256 * 256)
#define SG_NA_BYTES (SG_NA_LIST_BYTES + (long)SG_NA_CHUNK * (SG_KIN * 2 + 2 * SG_H * 2 + SG_NO * 4))
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
