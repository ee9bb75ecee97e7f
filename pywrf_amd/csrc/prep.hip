// RK-stage preparation, open lateral boundaries, first-stage physics prep and
// the Smagorinsky coefficients (pywrf/wrf_dynamics.py:30-1076).
//
// Each function is one or two launches.  A thread owns one output point and
// evaluates the reference's statements for that point in their original
// order, recomputing stencil intermediates (hat, hatavg, es, qvs ...) from the
// inputs instead of staging them in full-grid scratch arrays; only where a
// later statement reads an earlier statement's output at a neighbouring point
// (edge copies, column recurrences) is the function split into launches.
#include "common.h"
#include "openbc.h"
#include "consts.h"

#define AT(p, k, j, i) at_b((p), D.ob((k), (j), (i)))
#define AT2(p, j, i) at_b((p), D.ob2((j), (i)))

// Generic index-set check: every index in range and the stencil reach `r`
// (cells beyond the outermost index used) inside the allocation.
int pw_check_grid(const pywrf_grid_t* g, int r) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  int ilo = imin(g->ids, g->its), ihi = imax(g->ide, g->ite);
  int jlo = imin(g->jds, g->jts), jhi = imax(g->jde, g->jte);
  if (ilo - r < 0 || ihi + r > g->nx || jlo - r < 0 || jhi + r > g->ny) return PYWRF_E_RANGE;
  if (g->kts < 0 || g->kds < 0 || g->kte > g->nz || g->kde > g->nz) return PYWRF_E_RANGE;
  if (g->ide - g->ids < 8 || g->jde - g->jds < 8) return PYWRF_E_RANGE;
  return 0;
}
#define CHECK(g, r) do { int _e = pw_check_grid((g), (r)); if (_e) return _e; } while (0)

// ---------------------------------------------------------------- calculate_full, calc_mu_uv, couple_momentum
__global__ void k_calculate_full(Box b, Dims D, float* mut, const float* mub, const float* mu) {
  int j, i;
  if (!box_column(b, j, i)) return;
  AT2(mut, j, i) = AT2(mub, j, i) + AT2(mu, j, i);
}

extern "C" int pywrf_calculate_full(const pywrf_grid_t* g, float* mut, const float* mub, const float* mu,
                                    pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(0, 1, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  PW_LAUNCH_COLS(k_calculate_full, b, s, dims_of(g), mut, mub, mu);
  PW_RET();
}

// Box over [jts, jte) x [its, ite): each thread writes muu and/or muv at its point.
__global__ void k_calc_mu_uv(Box b, Dims D, const float* mu, const float* mub, float* muu, float* muv,
                             int jts, int jtf, int jte, int its, int itf, int ite) {
  int j, i;
  if (!box_column(b, j, i)) return;
  if (j < jtf) {
    if (i >= its + 1 && i < ite - 1)
      AT2(muu, j, i) = 0.5f * (AT2(mu, j, i) + AT2(mu, j, i - 1) + AT2(mub, j, i) + AT2(mub, j, i - 1));
    else if (i == its)
      AT2(muu, j, i) = AT2(mu, j, i) + AT2(mub, j, i);
    if (i == ite - 1) AT2(muu, j, i) = AT2(mu, j, ite - 2) + AT2(mub, j, ite - 2);
  }
  if (i < itf) {
    if (j >= jts + 1 && j < jte - 1)
      AT2(muv, j, i) = 0.5f * (AT2(mu, j, i) + AT2(mu, j - 1, i) + AT2(mub, j, i) + AT2(mub, j - 1, i));
    else if (j == jts)
      AT2(muv, j, i) = AT2(mu, j, i) + AT2(mub, j, i);
    if (j == jte - 1) AT2(muv, j, i) = AT2(mu, jte - 2, i) + AT2(mub, jte - 2, i);
  }
}

extern "C" int pywrf_calc_mu_uv(const pywrf_grid_t* g, const float* mu, const float* mub, float* muu, float* muv,
                                pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  int jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  Box b = mkbox(0, 1, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_COLS(k_calc_mu_uv, b, s, dims_of(g), mu, mub, muu, muv, g->jts, jtf, g->jte, g->its, itf, g->ite);
  PW_RET();
}

__global__ void k_couple_momentum(Box b, Dims D, Box bu, Box bv, Box bw, const float* muu, float* ru,
                                  const float* u, const float* msfu, const float* muv, float* rv, const float* v,
                                  const float* msfvx_inv, const float* mut, float* rw, const float* w,
                                  const float* msft) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long p = D.ix(k, j, i);
  const bool hu = bu.has(k, j, i), hv = bv.has(k, j, i), hw = bw.has(k, j, i);
  const float xu = hu ? u[p] * AT2(muu, j, i) / AT2(msfu, j, i) : 0.0f;
  const float xv = hv ? v[p] * AT2(muv, j, i) * AT2(msfvx_inv, j, i) : 0.0f;
  const float xw = hw ? w[p] * AT2(mut, j, i) / AT2(msft, j, i) : 0.0f;
  if (hu) ru[p] = xu;   // stores last: no load waits on them
  if (hv) rv[p] = xv;
  if (hw) rw[p] = xw;
}

extern "C" int pywrf_couple_momentum(const pywrf_grid_t* g, const float* muu, float* ru, const float* u,
                                     const float* msfu, const float* muv, float* rv, const float* v,
                                     const float* msfv, const float* msfvx_inv, const float* mut, float* rw,
                                     const float* w, const float* msft, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  (void)msfv;
  CHECK(g, 0);
  int ktf = imin(g->kte, g->kde - 1), jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  Box bu = mkbox(g->kts, ktf, g->jts, jtf, g->its, g->ite);
  Box bv = mkbox(g->kts, ktf, g->jts, g->jte, g->its, itf);
  Box bw = mkbox(g->kts, g->kte, g->jts, jtf, g->its, itf);
  Box b = mkbox(g->kts, g->kte, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_BOX(k_couple_momentum, b, s, dims_of(g), bu, bv, bw, muu, ru, u, msfu, muv, rv, v, msfvx_inv, mut,
                rw, w, msft);
  PW_RET();
}

// ---------------------------------------------------------------- calc_ww_cp
__device__ inline float ww_divv(const Dims& D, int k, int j, int i, const float* u, const float* v,
                                const float* muu, const float* muv, const float* msftx, const float* msfuy,
                                const float* msfvx_inv, const float* dnw, float rdx, float rdy) {
  float a = AT2(muu, j, i + 1) * AT(u, k, j, i + 1) / AT2(msfuy, j, i + 1);
  float b = AT2(muu, j, i) * AT(u, k, j, i) / AT2(msfuy, j, i);
  float c = AT2(muv, j + 1, i) * AT(v, k, j + 1, i) * AT2(msfvx_inv, j + 1, i);
  float d = AT2(muv, j, i) * AT(v, k, j, i) * AT2(msfvx_inv, j, i);
  return AT2(msftx, j, i) * dnw[k] * (rdx * (a - b) + rdy * (c - d));
}

// Column block (a 512-thread workgroup owns 64 i-columns of one row j): the
// divergence of every level is computed level-parallel into LDS; one wave then
// forms the column sum in torch's order and runs the eta-velocity recurrence.
#define WW_TI 64
#define WW_TB 512
#define WW_TG (WW_TB / WW_TI)
struct WwCp {
  const float *u, *v, *muu, *muv, *c1h, *msftx, *msfuy, *msfvx_inv, *dnw;
  float* ww;
  float rdx, rdy;
};

__global__ void __launch_bounds__(WW_TB) k_calc_ww_cp(Box b, Dims D, WwCp a) {
  extern __shared__ float s_dv[];   // [nz][WW_TI]
  __shared__ float s_dc[256];       // dnw[k] * c1h[k] (nz <= 256: the LDS check bounds nz by 256)
  for (int t = threadIdx.x; t < D.nz; t += WW_TB) s_dc[t] = a.dnw[t] * a.c1h[t];
  const int lane = threadIdx.x % WW_TI, grp = threadIdx.x / WW_TI;
  int bx, by;
  xcd_block_2d(bx, by);   // rows of one XCD contiguous (neighbouring rows share an L2)
  const int j = b.j0 + by, i = b.i0 + bx * WW_TI + lane;
  const bool ok = i < b.i1;
  if (ok)
    for (int k = grp; k < D.nz; k += WW_TG)
      s_dv[k * WW_TI + lane] = (k >= b.k0 && k < b.k1) ? ww_divv(D, k, j, i, a.u, a.v, a.muu, a.muv, a.msftx,
                                                                 a.msfuy, a.msfvx_inv, a.dnw, a.rdx, a.rdy)
                                                       : 0.0f;
  __syncthreads();
  if (!ok || grp != 0) return;
  // column sum over all nz levels of the divergence (zero outside [k0, k1))
  const float dmdt = torch_outer_sum(D.nz, [&](int k) { return s_dv[k * WW_TI + lane]; });
  float wk = AT(a.ww, 0, j, i);
  for (int k = 0; k < b.k1 - 1; ++k) {
    wk = wk - s_dc[k] * dmdt - s_dv[k * WW_TI + lane];
    AT(a.ww, k + 1, j, i) = wk;
  }
}

extern "C" int pywrf_calc_ww_cp(const pywrf_grid_t* g, const float* u, const float* v, const float* muu,
                                const float* muv, const float* c1h, float* ww, double rdx, double rdy,
                                const float* msftx, const float* msfuy, const float* msfvx_inv, const float* dnw,
                                pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  int ktf = imin(g->kte, g->kde - 1), jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  Box b = mkbox(g->kts, ktf, g->jts, jtf, g->its, itf);
  if (g->kts != 0) return PYWRF_E_RANGE;
  if ((size_t)g->nz * WW_TI * sizeof(float) > 64 * 1024 || g->nz > 256) return PYWRF_E_RANGE;
  if (b.nj() <= 0 || b.ni() <= 0) PW_RET();
  WwCp a = {u, v, muu, muv, c1h, msftx, msfuy, msfvx_inv, dnw, ww, F(rdx), F(rdy)};
  hipLaunchKernelGGL(k_calc_ww_cp, dim3((unsigned)((b.ni() + WW_TI - 1) / WW_TI), (unsigned)b.nj()), dim3(WW_TB),
                     (size_t)g->nz * WW_TI * sizeof(float), (hipStream_t)s, b, dims_of(g), a);
  PW_RET();
}

// ---------------------------------------------------------------- calc_cq, calc_alt, calc_php
// In-order float sum of the n species at (k, j, i); up to 8 species the loads
// are issued together.
__device__ inline float qsum(const Dims& D, const float* moist, int n, int k, int j, int i) {
  const long p = D.ix(k, j, i), st = (long)D.nz * D.plane();
  if (n <= 8) {
    float v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = m < n ? moist[p + m * st] : 0.0f;
    float s = v[0];
#pragma unroll
    for (int m = 1; m < 8; ++m)
      if (m < n) s += v[m];
    return s;
  }
  float s = moist[p];
  for (int m = 1; m < n; ++m) s += moist[p + m * st];
  return s;
}

// Tiles of 63 i x 4 j points marching CQ_KC levels: lane l sums the species
// at i0 - 1 + l, so the point at i takes the sum at i - 1 from the lane to
// its left, and the sum at k - 1 is carried from the level below; only the
// sum at j - 1 is evaluated again.  Each sum is the same in-order float sum.
#define CQ_OX 63
#define CQ_KC 4
__global__ void k_calc_cq(Box b, Dims D, Box bu, Box bv, Box bw, const float* moist, int n, float* cqu,
                          float* cqv, float* cqw) {
  const int i = b.i0 - 1 + (int)(blockIdx.x * CQ_OX + threadIdx.x);
  const int j = b.j0 + (int)(blockIdx.y * PW_TY + threadIdx.y);
  if (j >= b.j1) return;                      // a whole wave (one row) leaves together
  const int k0 = b.k0 + (int)blockIdx.z * CQ_KC, k1 = imin(k0 + CQ_KC, b.k1);
  const bool live = i < b.i1;
  const bool own = threadIdx.x >= 1 && live;
  // every sum of the tile first (no store in between, so all loads can be in flight)
  float sdown = (live && k0 >= 1) ? qsum(D, moist, n, k0 - 1, j, i) : 0.0f;
  float s[CQ_KC], sn[CQ_KC];
#pragma unroll
  for (int kk = 0; kk < CQ_KC; ++kk) {
    const int k = k0 + kk;
    s[kk] = (live && k < k1) ? qsum(D, moist, n, k, j, i) : 0.0f;
    sn[kk] = (own && k < k1 && bv.has(k, j, i)) ? qsum(D, moist, n, k, j - 1, i) : 0.0f;
  }
#pragma unroll
  for (int kk = 0; kk < CQ_KC; ++kk) {
    const int k = k0 + kk;
    if (k >= k1) break;                       // uniform across the wave
    const float s0 = s[kk];
    const float sw = __shfl_up(s0, 1, PW_TX);
    if (own) {
      const int p = D.ix(k, j, i);
      if (bu.has(k, j, i)) cqu[p] = 1.0f / (1.0f + 0.5f * (s0 + sw));
      if (bv.has(k, j, i)) cqv[p] = 1.0f / (1.0f + 0.5f * (s0 + sn[kk]));
      if (bw.has(k, j, i)) cqw[p] = 0.5f * (s0 + sdown);
    }
    sdown = s0;
  }
}

extern "C" int pywrf_calc_cq(const pywrf_grid_t* g, const float* moist, int n_moist, float* cqu, float* cqv,
                             float* cqw, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (n_moist < 1) return PYWRF_E_ARG;
  int ktf = imin(g->kte, g->kde - 1), jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  Box bu = mkbox(g->kts, ktf, g->jts, jtf, g->its, g->ite);
  Box bv = mkbox(g->kts, ktf, g->jts, g->jte, g->its, itf);
  Box bw = mkbox(g->kts + 1, ktf, g->jts, jtf, g->its, itf);
  Box b = mkbox(g->kts, ktf, g->jts, g->jte, g->its, g->ite);
  if (b.size() > 0) {
    dim3 grid((unsigned)((b.ni() + CQ_OX - 1) / CQ_OX), (unsigned)((b.nj() + PW_TY - 1) / PW_TY),
              (unsigned)((b.nk() + CQ_KC - 1) / CQ_KC));
    hipLaunchKernelGGL(k_calc_cq, grid, dim3(PW_TX, PW_TY), 0, (hipStream_t)s, b, dims_of(g), bu, bv, bw, moist,
                       n_moist, cqu, cqv, cqw);
  }
  PW_RET();
}

__global__ void k_calc_alt(Box b, Dims D, float* alt, const float* al, const float* alb) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long p = D.ix(k, j, i);
  alt[p] = al[p] + alb[p];
}

extern "C" int pywrf_calc_alt(const pywrf_grid_t* g, float* alt, const float* al, const float* alb,
                              pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  PW_LAUNCH_BOX(k_calc_alt, b, s, dims_of(g), alt, al, alb);
  PW_RET();
}

__global__ void k_calc_php(Box b, Dims D, float* php, const float* ph, const float* phb) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long p = D.ix(k, j, i), q = D.ix(k + 1, j, i);
  php[p] = 0.5f * (phb[p] + phb[q] + ph[p] + ph[q]);
}

extern "C" int pywrf_calc_php(const pywrf_grid_t* g, float* php, const float* ph, const float* phb,
                              pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  if (b.k1 + 1 > g->nz) return PYWRF_E_RANGE;
  PW_LAUNCH_BOX(k_calc_php, b, s, dims_of(g), php, ph, phb);
  PW_RET();
}

// ---------------------------------------------------------------- set_physical_bc3d / bc2d
__global__ void k_open_bc(Dims D, float* dat, int k0, BcSpec c) {
  open_bc_plane(dat + (long)(k0 + (int)blockIdx.x) * D.plane(), D.nx, c);
}

static int bc_launch(const pywrf_grid_t* g, float* dat, int k0, int k1, const BcSpec& c, pywrf_stream_t s) {
  // every row / column touched must lie inside the allocation
  int rlo = imin(c.xr0, c.ustag ? c.ur0 : c.xr0), rhi = imax(c.xr1, c.ustag ? c.ur1 : c.xr1);
  if (rlo < 0 || rhi > g->ny || c.ids - 3 < 0 || c.ide + 3 > g->nx) return PYWRF_E_RANGE;
  if (c.yc0 < 0 || c.yc1 > g->nx || c.jds - 3 < 0 || c.jde + 3 > g->ny) return PYWRF_E_RANGE;
  if (k0 < 0 || k1 > g->nz) return PYWRF_E_RANGE;
  if (k1 <= k0) return 0;
  hipLaunchKernelGGL(k_open_bc, dim3(k1 - k0), dim3(BC_THREADS), 0, (hipStream_t)s, dims_of(g), dat, k0, c);
  PW_RET();
}

extern "C" int pywrf_set_physical_bc2d(const pywrf_grid_t* g, float* dat, char variable, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  char v = lower(variable);
  if (!in(v, "uvwtxyrp")) return PYWRF_E_ARG;
  int istag = in(v, "uxde") ? 0 : -1, jstag = in(v, "vydf") ? 0 : -1;
  BcSpec c;
  c.ids = g->ids; c.ide = g->ide; c.jds = g->jds; c.jde = g->jde;
  c.xr0 = c.ur0 = g->jds;
  c.xr1 = c.ur1 = g->jde + jstag;
  c.ustag = in(v, "ux");
  c.vstag = in(v, "vy");
  c.yc0 = imax(g->ids, g->its - 1);
  c.yc1 = imin(g->ite + 1, g->ide + istag);
  Dims D2; (void)D2;
  pywrf_grid_t g2 = *g;
  g2.nz = 1;
  return bc_launch(&g2, dat, 0, 1, c, s);
}

extern "C" int pywrf_set_physical_bc3d(const pywrf_grid_t* g, float* dat, char variable, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  char v = variable;
  if (v == 'U') v = 'u';
  else if (v == 'V') v = 'v';
  if (!in(v, "uvwtdexyfrp")) return PYWRF_E_ARG;
  int istag = in(v, "uxde") ? 0 : -1, jstag = in(v, "vydf") ? 0 : -1;
  int k1 = in(v, "efw") ? imin(g->kde, g->kte) : imin(g->kde - 1, g->kte);
  BcSpec c;
  c.ids = g->ids; c.ide = g->ide; c.jds = g->jds; c.jde = g->jde;
  c.xr0 = g->jts - 4;
  c.xr1 = g->jde + jstag + 4;
  c.ur0 = g->jts - 5;
  c.ur1 = g->jde + jstag + 4;
  c.ustag = in(v, "ux");
  c.vstag = in(v, "vy");
  c.yc0 = g->ids;
  c.yc1 = g->ide + istag;
  return bc_launch(g, dat, g->kts, k1, c, s);
}

// Several open-boundary fills in one launch: the reference applies
// set_physical_bc3d / set_physical_bc2d to a group of distinct fields in a
// row; each (field, level) pair here is one workgroup running the same
// x-then-y copy as k_open_bc.
#define PW_BC_GROUP_MAX 32
struct BcJobs {
  int n;
  BcJob job[PW_BC_GROUP_MAX];
};

__global__ void k_open_bc_group(long plane, int nx, BcJobs J) {
  if ((int)blockIdx.y >= J.n) return;
  const BcJob& jb = J.job[blockIdx.y];
  if ((int)blockIdx.x >= jb.nk) return;
  open_bc_plane(jb.p + (long)(jb.k0 + (int)blockIdx.x) * plane, nx, jb.c);
}

extern "C" int pywrf_set_physical_bc_group(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                                           const char* variables, const int* is2d, pywrf_stream_t s) {
  if (!pw_size_ok(g3) || !pw_size_ok(g2)) return PYWRF_E_RANGE;
  if (n < 0 || n > PW_BC_GROUP_MAX || g3->ny != g2->ny || g3->nx != g2->nx) return PYWRF_E_ARG;
  BcJobs J;
  J.n = n;
  int maxk = 1;
  for (int f = 0; f < n; ++f) {
    BcJob& jb = J.job[f];
    jb.p = (float*)fields[f];
    int e, k1 = 1;
    if (is2d[f]) {
      e = bc_spec2(g2, variables[f], jb.c);
      jb.k0 = 0;
      k1 = 1;
      if (!e && !bc_spec_ok(g2->ny, g2->nx, 1, 0, 1, jb.c)) e = PYWRF_E_RANGE;
    } else {
      e = bc_spec3(g3, variables[f], jb.c, k1);
      jb.k0 = g3->kts;
      if (!e && !bc_spec_ok(g3->ny, g3->nx, g3->nz, g3->kts, k1, jb.c)) e = PYWRF_E_RANGE;
    }
    if (e) return e;
    jb.nk = k1 - jb.k0 > 0 ? k1 - jb.k0 : 0;
    maxk = imax(maxk, jb.nk);
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_open_bc_group, dim3(maxk, n), dim3(BC_THREADS), 0, (hipStream_t)s, (long)g3->ny * g3->nx,
                     g3->nx, J);
  PW_RET();
}

// ---------------------------------------------------------------- phy_prep
// Launch 1 (whole array): th_phy, p_phy on the box; pi_phy, t_phy everywhere
// (fresh arrays); rho, u_phy, v_phy, z_at_w on the box.
__global__ void k_phy_prep_pt(Box all, Dims D, Box b, const float* u, const float* v, const float* p,
                              const float* pb, const float* alt, const float* ph, const float* phb, const float* t,
                              const float* qv, float* rho, float* th_phy, float* p_phy, float* u_phy, float* v_phy,
                              float* z_at_w, float* pi_phy, float* t_phy) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  // every value is formed first and stored last, so that no load waits on a store
  float th, pp, ro = 0.0f, up = 0.0f, vp = 0.0f, zw = 0.0f;
  bool inb = b.has(k, j, i);
  if (inb) {
    th = t[q] + F(pw::t0);
    pp = p[q] + pb[q];
    ro = 1.0f / alt[q] * (1.0f + qv[q]);
    up = 0.5f * (u[q] + AT(u, k, j, i + 1));
    vp = 0.5f * (v[q] + AT(v, k, j + 1, i));
    zw = (phb[q] + ph[q]) / F(pw::g);
  } else {
    th = th_phy[q];
    pp = p_phy[q];
  }
  float pi = powf(pp / F(pw::p1000mb), F(pw::rcp));
  if (inb) {
    th_phy[q] = th;
    p_phy[q] = pp;
  }
  pi_phy[q] = pi;
  t_phy[q] = th * pi;
  if (inb) {
    rho[q] = ro;
    u_phy[q] = up;
    v_phy[q] = vp;
    z_at_w[q] = zw;
  }
}

// Launch 2 (column block: a 512-thread workgroup owns 64 i-columns of one
// row j): dz8w, z, p8w, t8w and the per-level moisture sums level-parallel
// (sums staged in LDS), then one wave runs the downward hydrostatic
// recurrence p_hyd_w per column from LDS, then all waves write p_hyd.
#define PP_TI 64
#define PP_TB 512
#define PP_TG (PP_TB / PP_TI)
struct PhyCol {
  const float *mut, *moist, *p_phy, *t_phy, *z_at_w, *dnw, *fzm, *fzp;
  float *dz8w, *z, *p8w, *t8w, *p_hyd, *p_hyd_w;
  float p_top;
  int n, kte, kde;
};

__global__ void __launch_bounds__(PP_TB) k_phy_prep_col(Box b, Dims D, PhyCol a) {
  extern __shared__ float sm[];
  float* s_qt = sm;                  // [nz][PP_TI]
  float* s_pw = sm + D.nz * PP_TI;   // [nz][PP_TI]
  __shared__ float s_dnw[128];       // nz <= 128 (two LDS arrays of nz x 64 floats within 64 KB)
  for (int t = threadIdx.x; t < D.nz; t += PP_TB) s_dnw[t] = a.dnw[t];
  const int lane = threadIdx.x % PP_TI, grp = threadIdx.x / PP_TI;
  int bx, by;
  xcd_block_2d(bx, by);   // rows of one XCD contiguous (neighbouring rows share an L2)
  const int j = b.j0 + by, i = b.i0 + bx * PP_TI + lane;
  const bool ok = i < b.i1;
  const int k0 = b.k0, k1 = b.k1, kte = a.kte, kde = a.kde;
  auto ZW = [&](int kk) { return AT(a.z_at_w, kk, j, i); };
  auto ZM = [&](int kk) { return 0.5f * (ZW(kk) + ZW(kk + 1)); };
  auto PP = [&](int kk) { return AT(a.p_phy, kk, j, i); };
  auto TP = [&](int kk) { return AT(a.t_phy, kk, j, i); };
  if (ok) {
    for (int k = grp; k < D.nz; k += PP_TG) {
      const long q = D.ix(k, j, i);
      if (k >= k0 && k < kte - 1) a.dz8w[q] = ZW(k + 1) - ZW(k);
      if (k == kte - 1) a.dz8w[q] = 0.0f;
      if (k >= k0 && k < k1) a.z[q] = ZM(k);
      if (k >= k0 + 1 && k < k1) {
        a.p8w[q] = a.fzm[k] * PP(k) + a.fzp[k] * PP(k - 1);
        a.t8w[q] = a.fzm[k] * TP(k) + a.fzp[k] * TP(k - 1);
      }
      if (k >= k0 && k < kte - 1) s_qt[k * PP_TI + lane] = (k < k1 - 1) ? qsum(D, a.moist, a.n, k, j, i) : 0.0f;
    }
    if (grp == 1) {
      float z0 = ZW(0), z1 = ZM(0), z2 = ZM(1);
      float w1 = (z0 - z2) / (z1 - z2);
      float w2 = 1.0f - w1;
      AT(a.p8w, 0, j, i) = w1 * PP(0) + w2 * PP(1);
      AT(a.t8w, 0, j, i) = w1 * TP(0) + w2 * TP(1);
    }
    if (grp == 2) {
      float z0 = ZW(kte - 1), z1 = ZM(k1 - 1), z2 = ZM(k1 - 2);
      float w1 = (z0 - z2) / (z1 - z2);
      float w2 = 1.0f - w1;
      AT(a.p8w, kde - 1, j, i) = expf(w1 * logf(PP(kde - 2)) + w2 * logf(PP(kde - 3)));
      AT(a.t8w, kde - 1, j, i) = w1 * TP(kde - 2) + w2 * TP(kde - 3);
    }
  }
  __syncthreads();
  if (ok && grp == 0) {
    const float m = AT2(a.mut, j, i);
    float pw_ = a.p_top;
    s_pw[(kte - 1) * PP_TI + lane] = pw_;
    AT(a.p_hyd_w, kte - 1, j, i) = pw_;
    for (int k = kte - 2; k >= k0; --k) {
      pw_ = pw_ - (1.0f + s_qt[k * PP_TI + lane]) * m * s_dnw[k];
      s_pw[k * PP_TI + lane] = pw_;
      AT(a.p_hyd_w, k, j, i) = pw_;
    }
  }
  __syncthreads();
  if (!ok) return;
  for (int k = k0 + grp; k < k1; k += PP_TG)
    AT(a.p_hyd, k, j, i) = 0.5f * (s_pw[k * PP_TI + lane] + s_pw[(k + 1) * PP_TI + lane]);
}

extern "C" int pywrf_phy_prep(const pywrf_grid_t* g, const float* mut, const float* u, const float* v,
                              const float* p, const float* pb, const float* alt, const float* ph, const float* phb,
                              const float* t, const float* moist, int n_moist, float* rho, float* th_phy,
                              float* p_phy, float* u_phy, float* v_phy, float* p8w, float* t8w, float* z,
                              float* z_at_w, float* dz8w, float* p_hyd, float* p_hyd_w, const float* dnw,
                              const float* fzm, const float* fzp, double p_top, float* pi_phy_out,
                              float* t_phy_out, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (n_moist < 2 || g->kte != g->nz || g->kde != g->nz || g->kts != 0) return PYWRF_E_RANGE;
  const size_t lds = 2 * (size_t)g->nz * PP_TI * sizeof(float);
  if (lds > 64 * 1024) return PYWRF_E_RANGE;
  Dims D = dims_of(g);
  int k1 = imin(g->kte, g->kde - 1);
  Box b = mkbox(g->kts, k1, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  Box all = mkbox(0, g->nz, 0, g->ny, 0, g->nx);
  const float* qv = moist + (long)pw::P_QV * D.nz * D.plane();
  PW_LAUNCH_BOX(k_phy_prep_pt, all, s, D, b, u, v, p, pb, alt, ph, phb, t, qv, rho, th_phy, p_phy, u_phy, v_phy,
                z_at_w, pi_phy_out, t_phy_out);
  PhyCol c = {mut, moist, p_phy, t_phy_out, z_at_w, dnw, fzm, fzp, dz8w, z, p8w, t8w, p_hyd, p_hyd_w, F(p_top),
              n_moist, g->kte, g->kde};
  if (b.nj() <= 0 || b.ni() <= 0) PW_RET();
  hipLaunchKernelGGL(k_phy_prep_col, dim3((unsigned)((b.ni() + PP_TI - 1) / PP_TI), (unsigned)b.nj()), dim3(PP_TB),
                     lds, (hipStream_t)s, b, D, c);
  PW_RET();
}

// ---------------------------------------------------------------- calculate_phy_tend
__global__ void k_scale_by_mut(Box b, Dims D, const float* mut, float* a0, float* a1, float* a2, float* a3,
                               float* a4, float* a5, float* sc, int nsc, int divide) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i), st = (long)D.nz * D.plane();
  float m = AT2(mut, j, i);
  float* arr[6] = {a0, a1, a2, a3, a4, a5};
  float v[6 + 8];   // loads before stores (up to 8 scalar species together)
#pragma unroll
  for (int n = 0; n < 6; ++n) v[n] = arr[n][q];
#pragma unroll
  for (int n = 0; n < 8; ++n) v[6 + n] = n < nsc ? sc[q + n * st] : 0.0f;
#pragma unroll
  for (int n = 0; n < 6; ++n) arr[n][q] = divide ? v[n] / m : m * v[n];
#pragma unroll
  for (int n = 0; n < 8; ++n)
    if (n < nsc) sc[q + n * st] = divide ? v[6 + n] / m : m * v[6 + n];
  for (int n = 8; n < nsc; ++n) sc[q + n * st] = divide ? sc[q + n * st] / m : m * sc[q + n * st];
}

extern "C" int pywrf_calculate_phy_tend(const pywrf_grid_t* g, const float* mut, float* rublten, float* rvblten,
                                        float* rthblten, float* rqvblten, float* rqcblten, float* rqiblten,
                                        float* scalar_tend, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  PW_LAUNCH_BOX(k_scale_by_mut, b, s, dims_of(g), mut, rublten, rvblten, rthblten, rqvblten, rqcblten, rqiblten,
                scalar_tend, 3, 0);
  PW_RET();
}

// ---------------------------------------------------------------- compute_diff_metrics
__device__ inline float zw_at(const Dims& D, const float* ph, const float* phb, int k, int j, int i) {
  return (AT(ph, k, j, i) + AT(phb, k, j, i)) / F(pw::g);
}

__global__ void k_diff_metrics(Box all, Dims D, const float* ph, const float* phb, float* z, float* rdz,
                               float* rdzw, float* zx, float* zy, float rdx, float rdy, int kts, int kte, int ktf,
                               int jts, int jte, int jtf, int its, int itf, int ids, int ide, int jds, int jde) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  // zw lives on [kts, kte) x [jts-1, jte) x [its, ide-1)
  bool zcol = (j >= jts - 1 && j < jte && i >= its && i < ide - 1);
  if (zcol && k >= kts && k < ktf) rdzw[q] = 1.0f / (zw_at(D, ph, phb, k + 1, j, i) - zw_at(D, ph, phb, k, j, i));
  if (zcol && k >= kts + 1 && k < ktf)
    rdz[q] = 2.0f / (zw_at(D, ph, phb, k + 1, j, i) - zw_at(D, ph, phb, k - 1, j, i));
  if (zcol && k == 0) rdz[q] = 2.0f / (zw_at(D, ph, phb, 1, j, i) - zw_at(D, ph, phb, 0, j, i));
  const float G = F(pw::g);
  if (k >= kts && k < kte && j >= jts && j < jtf && i >= ids + 1 && i < itf)
    zx[q] = rdx * (phb[q] - AT(phb, k, j, i - 1)) / G + rdx * (ph[q] - AT(ph, k, j, i - 1)) / G;
  if (k >= kts && k < kte && j >= jds + 1 && j < jtf && i >= its && i < itf)
    zy[q] = rdy * (phb[q] - AT(phb, k, j - 1, i)) / G + rdy * (ph[q] - AT(ph, k, j - 1, i)) / G;
  if (k < ktf && j >= jts && j < jtf && (i == ide - 1 || i == ids - 1)) zx[q] = 0.0f;
  if (k < ktf && i >= its && i < itf && (j == jde - 1 || j == jds - 1)) zy[q] = 0.0f;
  if (k < ktf && j >= jts && j < jtf && i >= its && i < itf)
    z[q] = 0.5f * (ph[q] + phb[q] + AT(ph, k + 1, j, i) + AT(phb, k + 1, j, i));
}

extern "C" int pywrf_compute_diff_metrics(const pywrf_grid_t* g, const float* ph, const float* phb, float* z,
                                          float* rdz, float* rdzw, float* zx, float* zy, double rdx, double rdy,
                                          pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 2);
  int ktf = imin(g->kte, g->kde - 1), jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  if (ktf + 1 > g->nz) return PYWRF_E_RANGE;
  Box all = mkbox(0, g->kte, g->jts - 1, imax(g->jte, g->jde), imin(g->its, g->ids - 1), imax(g->ide, itf));
  PW_LAUNCH_BOX(k_diff_metrics, all, s, dims_of(g), ph, phb, z, rdz, rdzw, zx, zy, F(rdx), F(rdy), g->kts, g->kte,
                ktf, g->jts, g->jte, jtf, g->its, itf, g->ids, g->ide, g->jds, g->jde);
  PW_RET();
}

// ---------------------------------------------------------------- cal_deform_and_div
struct DefCoef {
  const float *cf1, *cf2, *cf3, *dn, *dnw;
  int ktes1, ktes2, kte, ktf;
};
struct DefW {
  float cf1, cf2, cf3, cft1, cft2;
};

__device__ inline DefW def_weights(const DefCoef& c) {
  DefW w;
  w.cf1 = *c.cf1;
  w.cf2 = *c.cf2;
  w.cf3 = *c.cf3;
  w.cft2 = -0.5f * c.dnw[c.ktes1] / c.dn[c.ktes1];
  w.cft1 = 1.0f - w.cft2;
  return w;
}

// w-level average of hat(k, ja, ia) and hat(k, jb, ib), hat = X / M (the
// three hatavg statements that cal_deform_and_div repeats for D11, D22, D12).
__device__ inline float hat_pair_avg(const Dims& D, const float* X, const float* M, const float* fnm,
                                     const float* fnp, const DefCoef& c, const DefW& w, int k, int ja, int ia,
                                     int jb, int ib) {
  auto h = [&](int kk, int jj, int ii) { return AT(X, kk, jj, ii) / AT2(M, jj, ii); };
  if (k == 0)
    return 0.5f * (w.cf1 * h(0, ja, ia) + w.cf2 * h(1, ja, ia) + w.cf3 * h(2, ja, ia) + w.cf1 * h(0, jb, ib) +
                   w.cf2 * h(1, jb, ib) + w.cf3 * h(2, jb, ib));
  if (k == c.kte - 1)
    return 0.5f * (w.cft1 * (h(c.ktes1, ja, ia) + h(c.ktes1, jb, ib)) +
                   w.cft2 * (h(c.ktes2, ja, ia) + h(c.ktes2, jb, ib)));
  return 0.5f * (fnm[k] * (h(k, ja, ia) + h(k, jb, ib)) + fnp[k] * (h(k - 1, ja, ia) + h(k - 1, jb, ib)));
}

__global__ void k_deform(Box all, Dims D, const float* u, const float* v, const float* w, float* div, float* d11,
                         float* d22, float* d33, float* d12, float* d13, float* d23, const float* msfux,
                         const float* msfuy, const float* msfvx, const float* msfvy, const float* msftx,
                         const float* msfty, float rdx, float rdy, const float* rdz, const float* rdzw,
                         const float* fnm, const float* fnp, const float* zx, const float* zy, DefCoef c, int jts,
                         int jte, int jtf, int its, int ite, int itf, int ide, int jde) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  DefW cw = def_weights(c);
  long q = D.ix(k, j, i);
  const int kts = 0, ktf = c.ktf;
  // every value is formed first and stored last, so that no load waits on a store
  float o11 = 0.0f, o22 = 0.0f, o33 = 0.0f, odv = 0.0f, o12 = 0.0f, o13 = 0.0f, o23 = 0.0f;
  // D11, D22, D33 and div on [kts, ktf) x [jts, jtf) x [its, itf)
  const bool s1 = k < ktf && j >= jts && j < jtf && i >= its && i < itf;
  const bool s2 = k < ktf && j >= jts + 1 && j < jte - 1 && i >= its + 1 && i < ite - 1;
  const bool s3 = k <= ktf && j >= jts && j < jde - 1 && i >= its && i < ide - 1;
  if (s1) {
    float mm = AT2(msftx, j, i) * AT2(msfty, j, i);
    float ha1 = hat_pair_avg(D, u, msfuy, fnm, fnp, c, cw, k + 1, j, i, j, i + 1);
    float ha0 = hat_pair_avg(D, u, msfuy, fnm, fnp, c, cw, k, j, i, j, i + 1);
    float t1 = (ha1 - ha0) * (0.25f * (AT(zx, k, j, i) + AT(zx, k, j, i + 1) + AT(zx, k + 1, j, i) +
                                       AT(zx, k + 1, j, i + 1)) * rdzw[q]);
    float hp = AT(u, k, j, i + 1) / AT2(msfuy, j, i + 1), h0 = u[q] / AT2(msfuy, j, i);
    t1 = mm * (rdx * (hp - h0) - t1);
    o11 = 2.0f * t1;
    float dv = t1 + 0.0f;
    ha1 = hat_pair_avg(D, v, msfvx, fnm, fnp, c, cw, k + 1, j, i, j + 1, i);
    ha0 = hat_pair_avg(D, v, msfvx, fnm, fnp, c, cw, k, j, i, j + 1, i);
    hp = AT(v, k, j + 1, i) / AT2(msfvx, j + 1, i);
    h0 = v[q] / AT2(msfvx, j, i);
    t1 = mm * (rdy * (hp - h0) - (ha1 - ha0) * (0.25f * (AT(zy, k, j, i) + AT(zy, k, j + 1, i) +
                                                         AT(zy, k + 1, j, i) + AT(zy, k + 1, j + 1, i)) *
                                                rdzw[q]));
    o22 = 2.0f * t1;
    dv = t1 + dv;
    t1 = (AT(w, k + 1, j, i) - w[q]) * rdzw[q];
    o33 = 2.0f * t1;
    odv = t1 + dv;
  }
  // D12 on [kts, ktf) x [jts+1, jte-1) x [its+1, ite-1)
  if (s2) {
    float mm = 0.25f * (AT2(msfux, j - 1, i) + AT2(msfux, j, i)) * (AT2(msfvy, j, i - 1) + AT2(msfvy, j, i));
    float ha1 = hat_pair_avg(D, u, msfux, fnm, fnp, c, cw, k + 1, j - 1, i, j, i);
    float ha0 = hat_pair_avg(D, u, msfux, fnm, fnp, c, cw, k, j - 1, i, j, i);
    float t1 = (ha1 - ha0) * (0.25f * (AT(zy, k, j, i - 1) + AT(zy, k, j, i) + AT(zy, k + 1, j, i - 1) +
                                       AT(zy, k + 1, j, i)) * 0.25f *
                              (rdzw[q] + AT(rdzw, k, j, i - 1) + AT(rdzw, k, j - 1, i - 1) + AT(rdzw, k, j - 1, i)));
    float h0 = u[q] / AT2(msfux, j, i), hm = AT(u, k, j - 1, i) / AT2(msfux, j - 1, i);
    float r = mm * (rdy * (h0 - hm) - t1);
    ha1 = hat_pair_avg(D, v, msfvy, fnm, fnp, c, cw, k + 1, j, i - 1, j, i);
    ha0 = hat_pair_avg(D, v, msfvy, fnm, fnp, c, cw, k, j, i - 1, j, i);
    t1 = (ha1 - ha0) * (0.25f * (AT(zx, k, j - 1, i) + AT(zx, k, j, i) + AT(zx, k + 1, j - 1, i) +
                                 AT(zx, k + 1, j, i)) * 0.25f *
                        (rdzw[q] + AT(rdzw, k, j - 1, i) + AT(rdzw, k, j - 1, i - 1) + AT(rdzw, k, j, i - 1)));
    h0 = v[q] / AT2(msfvy, j, i);
    hm = AT(v, k, j, i - 1) / AT2(msfvy, j, i - 1);
    o12 = r + mm * (rdx * (h0 - hm) - t1);
  }
  // D13, D23 on [kts, ktf] x [jts, jde-1) x [its, ide-1)
  if (s3) {
    if (k == kts || k == ktf) {
      o13 = 0.0f;
      o23 = 0.0f;
    } else {
      float mm = AT2(msfux, j, i) * AT2(msfuy, j, i);
      auto ht = [&](int kk, int jj, int ii) { return AT(w, kk, jj, ii) / AT2(msfty, jj, ii); };
      float hv0 = 0.25f * (ht(k, j, i) + ht(k + 1, j, i) + ht(k, j, i - 1) + ht(k + 1, j, i - 1));
      float hv1 = 0.25f * (ht(k - 1, j, i) + ht(k, j, i) + ht(k - 1, j, i - 1) + ht(k, j, i - 1));
      float t1 = (hv0 - hv1) * AT(zx, k, j, i) * 0.5f * (rdz[q] + AT(rdz, k, j, i - 1));
      float r = mm * (rdx * (ht(k, j, i) - ht(k, j, i - 1)) - t1);
      t1 = (u[q] - AT(u, k - 1, j, i)) * 0.5f * (rdz[q] + AT(rdz, k, j, i - 1));
      o13 = r + t1;
      mm = AT2(msfvx, j, i) * AT2(msfvy, j, i);
      auto hx = [&](int kk, int jj, int ii) { return AT(w, kk, jj, ii) / AT2(msftx, jj, ii); };
      hv0 = 0.25f * (hx(k, j, i) + hx(k + 1, j, i) + hx(k, j - 1, i) + hx(k + 1, j - 1, i));
      hv1 = 0.25f * (hx(k - 1, j, i) + hx(k, j, i) + hx(k - 1, j - 1, i) + hx(k, j - 1, i));
      t1 = (hv0 - hv1) * AT(zy, k, j, i) * 0.5f * (rdz[q] + AT(rdz, k, j - 1, i));
      r = mm * (rdy * (hx(k, j, i) - hx(k, j - 1, i)) - t1);
      t1 = (v[q] - AT(v, k - 1, j, i)) * 0.5f * (rdz[q] + AT(rdz, k, j - 1, i));
      o23 = r + t1;
    }
  }
  if (s1) {
    d11[q] = o11;
    d22[q] = o22;
    d33[q] = o33;
    div[q] = odv;
  }
  if (s2) d12[q] = o12;
  if (s3) {
    d13[q] = o13;
    d23[q] = o23;
  }
}

// Edge copies of D12, D13, D23: columns ids and ide-1 over rows [jts, jte),
// then rows jds and jde-1 over columns [its, ite) (which read the corner
// cells the column copies wrote), composed into one gather per cell.
__global__ void k_deform_edges(Box all, Dims D, float* d12, float* d13, float* d23, int jts, int jte, int its,
                               int ite, int ids, int ide, int jds, int jde) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  int sj = j, si = i;
  bool row = (i >= its && i < ite) && (j == jds || j == jde - 1);
  bool col = (j >= jts && j < jte) && (i == ids || i == ide - 1);
  if (!row && !col) return;
  if (row) {
    sj = (j == jds) ? jds + 1 : jde - 2;
    if (sj >= jts && sj < jte) si = (i == ids) ? ids + 1 : (i == ide - 1 ? ide - 2 : i);
  } else {
    si = (i == ids) ? ids + 1 : ide - 2;
  }
  long q = D.ix(k, j, i), r = D.ix(k, sj, si);
  d12[q] = d12[r];
  d13[q] = d13[r];
  d23[q] = d23[r];
}

extern "C" int pywrf_cal_deform_and_div(const pywrf_grid_t* g, const float* u, const float* v, const float* w,
                                        float* div, float* defor11, float* defor22, float* defor33, float* defor12,
                                        float* defor13, float* defor23, const float* msfux, const float* msfuy,
                                        const float* msfvx, const float* msfvy, const float* msftx,
                                        const float* msfty, double rdx, double rdy, const float* dn,
                                        const float* dnw, const float* rdz, const float* rdzw, const float* fnm,
                                        const float* fnp, const float* cf, const float* zx, const float* zy,
                                        pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 2);
  if (g->kts != 0 || g->kte != g->nz || g->kde != g->nz) return PYWRF_E_RANGE;
  DefCoef c;
  c.cf1 = cf;
  c.cf2 = cf + 1;
  c.cf3 = cf + 2;
  c.dn = dn;
  c.dnw = dnw;
  c.ktes1 = g->kte - 2;
  c.ktes2 = g->kte - 3;
  c.kte = g->kte;
  c.ktf = g->kde - 1;
  int jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  Box all = mkbox(0, c.ktf + 1, g->jts, imax(g->jde - 1, g->jte), g->its, imax(g->ide - 1, g->ite));
  Dims D = dims_of(g);
  PW_LAUNCH_BOX(k_deform, all, s, D, u, v, w, div, defor11, defor22, defor33, defor12, defor13, defor23, msfux,
                msfuy, msfvx, msfvy, msftx, msfty, F(rdx), F(rdy), rdz, rdzw, fnm, fnp, zx, zy, c, g->jts, g->jte,
                jtf, g->its, g->ite, itf, g->ide, g->jde);
  Box e = mkbox(g->kts, g->kte, imin(g->jts, g->jds), imax(g->jte, g->jde), imin(g->its, g->ids),
                imax(g->ite, g->ide));
  PW_LAUNCH_BOX(k_deform_edges, e, s, D, defor12, defor13, defor23, g->jts, g->jte, g->its, g->ite, g->ids, g->ide,
                g->jds, g->jde);
  PW_RET();
}

// ---------------------------------------------------------------- calculate_km_kh
__device__ inline float sat_qvs(float t, float p) {
  float es = F(1000.0 * pw::SVP1) * expf(F(pw::SVP2) * (t - F(pw::SVPT0)) / (t - F(pw::SVP3)));
  return F(pw::EP_2) * es / (p - es);
}

__device__ inline float coefa_of(float qv, float t, float theta) {
  float xlvqv = F(pw::XLV) * qv;
  return (1.0f + xlvqv / F(pw::r_d) / t) / (1.0f + F(pw::XLV) * xlvqv / F(pw::Cp) / F(pw::r_v) / t / t) / theta;
}

__device__ inline float thetae(float theta, float qvs, float t) {
  return theta * (1.0f + F(pw::XLV) * qvs) / F(pw::Cp) / t;
}

struct KmArgs {
  const float *d11, *d22, *d12, *p8w, *t8w, *theta, *t, *p, *qv, *qc, *qi, *rdz, *rdzw, *cf, *msftx, *msfty;
  float *xkmh, *xkmv, *xkhh, *xkhv, *bn2, *xkmh_out;
  float dx, dy;
  int ktf, j0, j1, i0, i1, sj0, sj1, si0, si1;
};

// BN2 at (k, j, i) inside the column box (the reference's where-select of the
// saturated / unsaturated forms; zero outside [kts, ktf-1) and the box).
__device__ float bn2_at(const Dims& D, const KmArgs& a, int k, int j, int i) {
  const float G = F(pw::g), qc_cr = 0.00001f;
  auto L = [&](const float* f, int kk) { return AT(f, kk, j, i); };
  auto tmp1 = [&](int kk) { return L(a.qv, kk) + L(a.qc, kk) + L(a.qi, kk); };
  auto qvs = [&](int kk) { return sat_qvs(L(a.t, kk), L(a.p, kk)); };
  float cf1 = a.cf[0], cf2 = a.cf[1], cf3 = a.cf[2];
  bool cond = (L(a.qv, k) >= qvs(k)) || (L(a.qc, k) + 0.0f >= qc_cr);
  if (k >= 1 && k < a.ktf - 1) {
    float tmpdz = 1.0f / L(a.rdz, k) + 1.0f / L(a.rdz, k + 1);
    if (cond) {
      float coefa = coefa_of(L(a.qv, k), L(a.t, k), L(a.theta, k));
      float tp1 = thetae(L(a.theta, k + 1), qvs(k + 1), L(a.t, k + 1));
      float tm1 = thetae(L(a.theta, k - 1), qvs(k - 1), L(a.t, k - 1));
      return G * (coefa * (tp1 - tm1) / tmpdz - (tmp1(k + 1) - tmp1(k - 1)) / tmpdz);
    }
    return G * ((L(a.theta, k + 1) - L(a.theta, k - 1)) / L(a.theta, k) / tmpdz +
                1.61f * (L(a.qv, k + 1) - L(a.qv, k - 1)) / tmpdz - (tmp1(k + 1) - tmp1(k - 1)) / tmpdz);
  }
  // k == kts
  float tmp1sfc = cf1 * tmp1(0) + cf2 * tmp1(1) + cf3 * tmp1(2);
  if (cond) {
    float tmpdz = 1.0f / L(a.rdz, 1) + 0.5f / L(a.rdzw, 0);
    float thetasfc = L(a.t8w, 0) / powf(L(a.p8w, 0) / F(pw::p1000mb), F(pw::r_d / pw::Cp));
    float qvsfc = cf1 * qvs(0) + cf2 * qvs(1) + cf3 * qvs(2);
    float coefa = coefa_of(L(a.qv, 0), L(a.t, 0), L(a.theta, 0));
    float tp1 = thetae(L(a.theta, 1), qvs(1), L(a.t, 1));
    float thetaesfc = thetasfc * (1.0f + F(pw::XLV) * qvsfc / F(pw::Cp) / L(a.t8w, 0));
    return G * (coefa * (tp1 - thetaesfc) / tmpdz - (tmp1(1) - tmp1sfc) / tmpdz);
  }
  float qvsfc = cf1 * L(a.qv, 0) + cf2 * L(a.qv, 1) + cf3 * L(a.qv, 2);
  float tmpdz = 1.0f / L(a.rdzw, 0);
  return G * ((L(a.theta, 1) - L(a.theta, 0)) / L(a.theta, 0) / tmpdz + 1.61f * (L(a.qv, 1) - qvsfc) / tmpdz -
              (tmp1(1) - tmp1sfc) / tmpdz);
}

__global__ void k_km_kh(Box all, Dims D, KmArgs a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  // BN2: level ktf-1 repeats level ktf-2 over the whole plane
  int kb = (k == a.ktf - 1) ? a.ktf - 2 : k;
  bool incol = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  // every value is formed first and stored last, so that no load waits on a store
  const float bn = (incol && kb < a.ktf - 1) ? bn2_at(D, a, kb, j, i) : 0.0f;
  // 2-D Smagorinsky on the box, then the cap against 10 * mlen_h everywhere
  bool inb = k < a.ktf && j >= a.sj0 && j < a.sj1 && i >= a.si0 && i < a.si1;
  float cap = 0.0f, xk = a.xkmh[q];
  if (j >= a.sj0 && j < a.sj1 && i >= a.si0 && i < a.si1) {
    float mlen = sqrtf(a.dx / AT2(a.msftx, j, i) * a.dy / AT2(a.msfty, j, i));
    cap = 10.0f * mlen;
    if (inb) {
      float dd = AT(a.d11, k, j, i) - AT(a.d22, k, j, i);
      float def2 = 0.25f * (dd * dd);
      float tmp = 0.25f * (AT(a.d12, k, j, i) + AT(a.d12, k, j + 1, i) + AT(a.d12, k, j, i + 1) +
                           AT(a.d12, k, j + 1, i + 1));
      def2 = def2 + tmp * tmp;
      tmp = sqrtf(def2);
      xk = F(pw::c_s * pw::c_s) * mlen * mlen * tmp;
    }
  }
  float xn = xk < cap ? xk : cap;
  a.bn2[q] = bn;
  if (inb) a.xkmh[q] = xk;
  a.xkmh_out[q] = xn;
  a.xkmv[q] = 0.0f;
  a.xkhv[q] = 0.0f;
  if (inb) a.xkhh[q] = xn / F(pw::prandtl);
}

extern "C" int pywrf_calculate_km_kh(const pywrf_grid_t* g, float* xkmh, float* xkmv, float* xkhh, float* xkhv,
                                     const float* defor11, const float* defor22, const float* defor12,
                                     const float* p8w, const float* t8w, const float* theta, const float* t,
                                     const float* p, const float* moist, double dx, double dy, const float* rdz,
                                     const float* rdzw, const float* cf, const float* msftx, const float* msfty,
                                     float* BN2_out, float* xkmh_out, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->kts != 0 || g->kte > g->nz) return PYWRF_E_RANGE;
  Dims D = dims_of(g);
  long st = (long)D.nz * D.plane();
  KmArgs a;
  a.d11 = defor11; a.d22 = defor22; a.d12 = defor12; a.p8w = p8w; a.t8w = t8w; a.theta = theta; a.t = t; a.p = p;
  a.qv = moist + pw::P_QV * st; a.qc = moist + pw::P_QC * st; a.qi = moist + pw::P_QI * st;
  a.rdz = rdz; a.rdzw = rdzw; a.cf = cf; a.msftx = msftx; a.msfty = msfty;
  a.xkmh = xkmh; a.xkmv = xkmv; a.xkhh = xkhh; a.xkhv = xkhv; a.bn2 = BN2_out; a.xkmh_out = xkmh_out;
  a.dx = F(dx); a.dy = F(dy);
  a.ktf = imin(g->kte, g->kde - 1);
  a.j0 = g->jts; a.j1 = imin(g->jte, g->jde - 1); a.i0 = g->its; a.i1 = imin(g->ite, g->ide - 1);
  a.sj0 = g->jts + 1; a.sj1 = imin(g->jte, g->jde - 2); a.si0 = g->its + 1; a.si1 = imin(g->ite, g->ide - 2);
  Box all = mkbox(0, g->nz, 0, g->ny, 0, g->nx);
  PW_LAUNCH_BOX(k_km_kh, all, s, D, a);
  PW_RET();
}

// ---------------------------------------------------------------- update_phy_ten
__global__ void k_update_phy_ten(Box b, Dims D, Box ba, Box bu, Box bv, float* rt_tendf, float* ru_tendf,
                                 float* rv_tendf, float* mqv, float* mqc, float* mqi, const float* rth,
                                 const float* ru, const float* rv, const float* rqv, const float* rqc,
                                 const float* rqi) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  // every value is formed first and stored last, so that no load waits on a store
  const bool ha = ba.has(k, j, i), hu = bu.has(k, j, i), hv = bv.has(k, j, i);
  float nt = 0.0f, nqv = 0.0f, nqc = 0.0f, nqi = 0.0f, nu = 0.0f, nv = 0.0f;
  if (ha) {
    nt = rt_tendf[q] + rth[q];
    nqv = mqv[q] + rqv[q];
    nqc = mqc[q] + rqc[q];
    nqi = mqi[q] + rqi[q];
  }
  if (hu) nu = ru_tendf[q] + 0.5f * ru[q] + 0.5f * AT(ru, k, j, i - 1);
  if (hv) nv = rv_tendf[q] + 0.5f * rv[q] + 0.5f * AT(rv, k, j - 1, i);
  if (ha) {
    rt_tendf[q] = nt;
    mqv[q] = nqv;
    mqc[q] = nqc;
    mqi[q] = nqi;
  }
  if (hu) ru_tendf[q] = nu;
  if (hv) rv_tendf[q] = nv;
}

extern "C" int pywrf_update_phy_ten(const pywrf_grid_t* g, float* rt_tendf, float* ru_tendf, float* rv_tendf,
                                    float* moist_tendf, const float* rthblten, const float* rublten,
                                    const float* rvblten, const float* rqvblten, const float* rqcblten,
                                    const float* rqiblten, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  Dims D = dims_of(g);
  long st = (long)D.nz * D.plane();
  int ktf = imin(g->kte, g->kde - 1), jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  Box ba = mkbox(g->kts, ktf, g->jts, jtf, g->its, itf);
  Box bu = mkbox(g->kts, ktf, g->jts, jtf, g->its, g->ite);
  Box bv = mkbox(g->kts, ktf, g->jts, g->jte, g->its, itf);
  Box b = mkbox(g->kts, ktf, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_BOX(k_update_phy_ten, b, s, D, ba, bu, bv, rt_tendf, ru_tendf, rv_tendf, moist_tendf + pw::P_QV * st,
                moist_tendf + pw::P_QC * st, moist_tendf + pw::P_QI * st, rthblten, rublten, rvblten, rqvblten,
                rqcblten, rqiblten);
  PW_RET();
}
