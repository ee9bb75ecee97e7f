// Adjoints (transposes) of linear parts of the step, for gradients through it.
//
// * advect_scalar (wrf_dynamics.py:1841-2028): with the mass fluxes fixed the
//   tendency is linear in the field, T = A q, the face value of each face
//   being vel * sum_o c_o(face, sign) q(face + o).  k_advect_scalar_adj forms
//   A^T lam by gathering, for every point, the faces whose stencils hold it
//   (six in each direction), each face's sensitivity being formed from the
//   adjoint tendencies of the two points it separates.  No atomics; every
//   output point is written by one thread.
// * rk_update_scalar (wrf_dynamics.py:4055-4119): pointwise.
// * The column tridiagonal solve of advance_w (wrf_dynamics.py:3758-3765):
//   forward elimination w[k] = (w[k] - a[k] w[k-1]) alpha[k], k = 1..K, then
//   back substitution w[k] = w[k] - gamma[k] w[k+1], k = K-1..1.  The
//   adjoint runs the transposed sweeps in the opposite order (the back
//   substitution's transpose upwards, the elimination's downwards) and also
//   gives the sensitivities to a, alpha and gamma.
#include "common.h"
#include "consts.h"
#include "diff.h"
#include "openbc.h"

int pw_check_grid(const pywrf_grid_t* g, int r);
#define CHECK(g, r) do { int _e = pw_check_grid((g), (r)); if (_e) return _e; } while (0)

// ------------------------------------------------------------------ advect_scalar^T
struct AdjS {
  const float *lam, *ru, *rv, *rom, *msftx, *fzm, *fzp, *rdzw;
  float* out;
  float rdx, rdy;
  int ids, ide, jds, jde, ktf, accumulate;
  Box badv, by, bx;   // the forward's boxes: all terms, the y-term, the x-term
};

// d(face value) / d(q at face + o), o = -3..2, without the velocity factor:
// the 5th-order face (flux5) with upwind sign s = +-1.
__device__ inline float c5(int o, float s) {
  switch (o) {
    case -3: return (1.0f + s) / 60.0f;
    case -2: return (-8.0f - 5.0f * s) / 60.0f;
    case -1: return (37.0f + 10.0f * s) / 60.0f;
    case 0: return (37.0f - 10.0f * s) / 60.0f;
    case 1: return (-8.0f + 5.0f * s) / 60.0f;
    case 2: return (1.0f - s) / 60.0f;
  }
  return 0.0f;
}
// the 3rd-order face (flux3), o = -2..1
__device__ inline float c3(int o, float s) {
  switch (o) {
    case -2: return (-1.0f - s) / 12.0f;
    case -1: return (7.0f + 3.0f * s) / 12.0f;
    case 0: return (7.0f - 3.0f * s) / 12.0f;
    case 1: return (-1.0f + s) / 12.0f;
  }
  return 0.0f;
}

// Coefficient of q(face + o) in the value of the horizontal face f of a row
// or column whose domain is [d0, d1) (y: jds, jde; x: ids, ide): the
// reference's face rules (5th order inside, 2nd order next to the edge, 3rd
// order one further in; zero flux elsewhere).  s_lo / s_hi: the upwind signs
// the 3rd-order faces at d0 + 2 and d1 - 3 use.
__device__ inline float hface_coef(int f, int o, int d0, int d1, float s, float s_lo, float s_hi) {
  if (f >= d0 + 3 && f < d1 - 3) return c5(o, s);
  if (f == d0 + 1 || f == d1 - 2) return (o == 0 || o == -1) ? 0.5f : 0.0f;
  if (f == d0 + 2) return c3(o, s_lo);
  if (f == d1 - 3) return c3(o, s_hi);
  return 0.0f;
}

__global__ void k_advect_scalar_adj(Box all, Dims D, AdjS a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  float acc = 0.0f;
  // y-faces jf = j - o hold q(k, j, i) at offset o
  if (k >= a.badv.k0 && k < a.badv.k1 && i >= a.badv.i0 && i < a.badv.i1) {
#pragma unroll
    for (int o = -3; o <= 2; ++o) {
      const int jf = j - o;
      if (jf < 1 || jf >= D.ny) continue;
      // sensitivity of the face value: the points jf (face on its low side)
      // and jf - 1 (face on its high side) that take the y-term
      float gF = 0.0f;
      if (a.by.has(k, jf, i)) gF += (AT2(a.msftx, jf, i) * a.rdy) * AT(a.lam, k, jf, i);
      if (a.by.has(k, jf - 1, i)) gF -= (AT2(a.msftx, jf - 1, i) * a.rdy) * AT(a.lam, k, jf - 1, i);
      if (gF == 0.0f) continue;
      const float vel = AT(a.rv, k, jf, i);
      const float s = sgn1(vel);
      acc += gF * vel * hface_coef(jf, o, a.jds, a.jde, s, s, s);
    }
  }
  // x-faces f = i - o
  if (k >= a.badv.k0 && k < a.badv.k1 && j >= a.badv.j0 && j < a.badv.j1) {
#pragma unroll
    for (int o = -3; o <= 2; ++o) {
      const int f = i - o;
      if (f < 1 || f >= D.nx) continue;
      float gF = 0.0f;
      if (a.bx.has(k, j, f)) gF += (AT2(a.msftx, j, f) * a.rdx) * AT(a.lam, k, j, f);
      if (a.bx.has(k, j, f - 1)) gF -= (AT2(a.msftx, j, f - 1) * a.rdx) * AT(a.lam, k, j, f - 1);
      if (gF == 0.0f) continue;
      const float vel = AT(a.ru, k, j, f);
      const float s = sgn1(vel);
      // the face at ide - 3 takes its upwind sign from ru at ide - 2
      const float s_hi = (f == a.ide - 3 && f + 1 < D.nx) ? sgn1(AT(a.ru, k, j, f + 1)) : s;
      acc += gF * vel * hface_coef(f, o, a.ids, a.ide, s, s, s_hi);
    }
  }
  // z-faces kf = k - o (upwind sign from -rom)
  if (j >= a.badv.j0 && j < a.badv.j1 && i >= a.badv.i0 && i < a.badv.i1) {
#pragma unroll
    for (int o = -3; o <= 2; ++o) {
      const int kf = k - o;
      if (kf < 1 || kf >= a.ktf || kf >= D.nz) continue;
      float gF = 0.0f;
      if (kf < a.badv.k1) gF += a.rdzw[kf] * AT(a.lam, kf, j, i);
      if (kf - 1 >= a.badv.k0) gF -= a.rdzw[kf - 1] * AT(a.lam, kf - 1, j, i);
      if (gF == 0.0f) continue;
      const float vel = AT(a.rom, kf, j, i);
      const float s = sgn1(-vel);
      float c;
      if (kf >= 3 && kf < a.ktf - 2) c = c5(o, s);
      else if (kf == 1 || kf == a.ktf - 1) c = o == 0 ? a.fzm[kf] : (o == -1 ? a.fzp[kf] : 0.0f);
      else c = c3(o, s);   // kf == 2 or ktf - 2
      acc += gF * vel * c;
    }
  }
  // (each face sensitivity above already carries the sign with which the
  // forward subtracts the face difference from the tendency)
  float* out = &AT(a.out, k, j, i);
  *out = a.accumulate ? *out + acc : acc;
}

extern "C" int pywrf_advect_scalar_adjoint(const pywrf_grid_t* g, const float* lam_tend, const float* ru,
                                           const float* rv, const float* rom, const float* msftx, const float* fzm,
                                           const float* fzp, double rdx, double rdy, const float* rdzw,
                                           float* lam_field, int accumulate, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  AdjS a;
  a.lam = lam_tend; a.ru = ru; a.rv = rv; a.rom = rom; a.msftx = msftx; a.fzm = fzm; a.fzp = fzp; a.rdzw = rdzw;
  a.out = lam_field; a.rdx = F(rdx); a.rdy = F(rdy); a.accumulate = accumulate;
  a.ids = g->ids; a.ide = g->ide; a.jds = g->jds; a.jde = g->jde; a.ktf = imin(g->kte, g->kde - 1);
  if (g->kts != 0 || a.ktf + 1 > g->nz || g->kte > g->nz) return PYWRF_E_RANGE;
  const int i0 = g->its, i1 = imin(g->ite, g->ide - 1), j0 = g->jts, j1 = imin(g->jte, g->jde - 1);
  a.badv = mkbox(g->kts, a.ktf, j0, j1, i0, i1);
  a.by = mkbox(g->kts, a.ktf, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), i0, i1);
  a.bx = mkbox(g->kts, a.ktf, j0, j1, imax(g->ids + 1, g->its), imin(g->ide - 2, g->ite));
  Box all = whole(g);
  PW_LAUNCH_BOX(k_advect_scalar_adj, all, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ rk_update_scalar^T
struct AdjU {
  const float *lam, *msfty, *mu_old, *mu_new, *mu_base;
  float *lam_n, *lam_adv, *lam_sc, *lam_prev;
  float dt;
  Box bf, bi, bs;   // the updated box, the advection box (spec zone excluded), the sc_tend box
};

__global__ void k_rk_update_scalar_adj(Box all, Dims D, AdjU a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  const float l = AT(a.lam, k, j, i);
  if (a.bf.has(k, j, i)) {
    const float muold = AT2(a.mu_old, j, i) + AT2(a.mu_base, j, i);
    const float munew = AT2(a.mu_new, j, i) + AT2(a.mu_base, j, i);
    const float lt = (a.dt * l) / munew;        // d q_new / d tendency
    AT(a.lam_n, k, j, i) += (muold * l) / munew;
    AT(a.lam_adv, k, j, i) = a.bi.has(k, j, i) ? lt * AT2(a.msfty, j, i) : 0.0f;
    if (a.lam_sc) AT(a.lam_sc, k, j, i) = a.bs.has(k, j, i) ? lt : 0.0f;
    AT(a.lam_prev, k, j, i) = 0.0f;
  } else {
    AT(a.lam_adv, k, j, i) = 0.0f;
    if (a.lam_sc) AT(a.lam_sc, k, j, i) = 0.0f;
    AT(a.lam_prev, k, j, i) = l;                // points the update leaves alone
  }
}

extern "C" int pywrf_rk_update_scalar_adjoint(const pywrf_grid_t* g, const float* lam_new, const float* msfty,
                                              const float* mu_old, const float* mu_new, const float* mu_base,
                                              double dt, int spec_zone, float* lam_n, float* lam_adv, float* lam_sc,
                                              float* lam_prev, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kde - 1 > g->nz || g->jde - 1 > g->ny || g->ide - 1 > g->nx) return PYWRF_E_RANGE;
  AdjU a;
  a.lam = lam_new; a.msfty = msfty; a.mu_old = mu_old; a.mu_new = mu_new; a.mu_base = mu_base;
  a.lam_n = lam_n; a.lam_adv = lam_adv; a.lam_sc = lam_sc; a.lam_prev = lam_prev;
  a.dt = F(dt);
  a.bf = mkbox(g->kts, g->kde - 1, g->jts, g->jde - 1, g->its, g->ide - 1);
  a.bi = mkbox(g->kts, imin(g->kte, g->kde - 1), imax(g->jts, g->jds + spec_zone), imin(g->jte, g->jde - spec_zone - 1),
               imax(g->its, g->ids + spec_zone), imin(g->ite, g->ide - spec_zone - 1));
  a.bs = mkbox(g->kts, g->kte - 1, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  Box all = whole(g);
  PW_LAUNCH_BOX(k_rk_update_scalar_adj, all, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ tridiagonal solve and its adjoint
// One thread per column of advance_w's box; consecutive threads take
// consecutive i, so every level's access is a contiguous row segment.
struct Tri {
  const float *a, *alpha, *gamma;
  float* w;
  int K;   // levels 1..K are solved; level 0 holds the lower boundary value
};

__global__ void k_tridiag_solve(Box b, Dims D, Tri t) {
  int j, i;
  if (!box_column(b, j, i)) return;
  float prev = AT(t.w, 0, j, i);
  for (int k = 1; k <= t.K; ++k) {
    const float y = (AT(t.w, k, j, i) - AT(t.a, k, j, i) * prev) * AT(t.alpha, k, j, i);
    AT(t.w, k, j, i) = y;
    prev = y;
  }
  float next = prev;   // w[K]
  for (int k = t.K - 1; k >= 1; --k) {
    const float x = AT(t.w, k, j, i) - AT(t.gamma, k, j, i) * next;
    AT(t.w, k, j, i) = x;
    next = x;
  }
}

static bool tri_box(const pywrf_grid_t* g, Box& b, int& K) {
  K = g->kte - 1;
  b = mkbox(0, K + 1, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), imax(g->its, g->ids + 1),
            imin(g->ite, g->ide - 2));
  return g->kts == 0 && K >= 1 && K < g->nz;
}

extern "C" int pywrf_tridiag_solve(const pywrf_grid_t* g, const float* a, const float* alpha, const float* gamma,
                                   float* w, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  Box b;
  Tri t;
  if (!tri_box(g, b, t.K)) return PYWRF_E_RANGE;
  t.a = a; t.alpha = alpha; t.gamma = gamma; t.w = w;
  PW_LAUNCH_COLS(k_tridiag_solve, b, s, dims_of(g), t);
  PW_RET();
}

// Adjoint of the solve: given the right-hand side r (the input of the
// solve), the solution x and the sensitivity lam of x, writes the
// sensitivities of r (over lam, in place), a, alpha and gamma.  The forward
// elimination's intermediate y is recomputed from r.
struct TriAdj {
  const float *a, *alpha, *gamma, *r, *x;
  float *lam, *ga, *galpha, *ggamma, *y;
  int K;
};

__global__ void k_tridiag_adj(Box b, Dims D, TriAdj t) {
  int j, i;
  if (!box_column(b, j, i)) return;
  // recompute y (the elimination's output) into t.y
  float prev = AT(t.r, 0, j, i);
  AT(t.y, 0, j, i) = prev;
  for (int k = 1; k <= t.K; ++k) {
    const float y = (AT(t.r, k, j, i) - AT(t.a, k, j, i) * prev) * AT(t.alpha, k, j, i);
    AT(t.y, k, j, i) = y;
    prev = y;
  }
  // back substitution x[k] = y[k] - gamma[k] x[k+1], run for k = K-1..1:
  // transposed upwards; lam becomes the sensitivity of y
  AT(t.ggamma, 0, j, i) = 0.0f;
  AT(t.ggamma, t.K, j, i) = 0.0f;
  for (int k = 1; k <= t.K - 1; ++k) {
    const float lk = AT(t.lam, k, j, i);
    AT(t.ggamma, k, j, i) = -lk * AT(t.x, k + 1, j, i);
    AT(t.lam, k + 1, j, i) -= AT(t.gamma, k, j, i) * lk;
  }
  // elimination y[k] = (r[k] - a[k] y[k-1]) alpha[k], k = 1..K: transposed
  // downwards; lam becomes the sensitivity of r
  AT(t.ga, 0, j, i) = 0.0f;
  AT(t.galpha, 0, j, i) = 0.0f;
  for (int k = t.K; k >= 1; --k) {
    const float ly = AT(t.lam, k, j, i);
    const float al = AT(t.alpha, k, j, i);
    const float ym = AT(t.y, k - 1, j, i);
    const float lr = ly * al;
    AT(t.galpha, k, j, i) = ly * (AT(t.r, k, j, i) - AT(t.a, k, j, i) * ym);
    AT(t.ga, k, j, i) = -lr * ym;
    AT(t.lam, k, j, i) = lr;
    AT(t.lam, k - 1, j, i) -= AT(t.a, k, j, i) * lr;
  }
}

extern "C" int pywrf_tridiag_solve_adjoint(const pywrf_grid_t* g, const float* a, const float* alpha,
                                           const float* gamma, const float* r, const float* x, float* lam,
                                           float* grad_a, float* grad_alpha, float* grad_gamma, float* work,
                                           pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  TriAdj t;
  Box b;
  if (!tri_box(g, b, t.K)) return PYWRF_E_RANGE;
  t.a = a; t.alpha = alpha; t.gamma = gamma; t.r = r; t.x = x; t.lam = lam;
  t.ga = grad_a; t.galpha = grad_alpha; t.ggamma = grad_gamma; t.y = work;
  PW_LAUNCH_COLS(k_tridiag_adj, b, s, dims_of(g), t);
  PW_RET();
}

// adds v to *p (atomically: neighbouring columns share points); zero adds are skipped
__device__ inline void atomic_add(float* p, float v) { if (v != 0.0f) atomicAdd(p, v); }

__global__ void k_zero_n(float* p, long n) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) p[q] = 0.0f;
}

// ------------------------------------------------------------------ advance_w^T
// advance_w (wrf_dynamics.py:3641-3796) with respect to w, ph, t_2, t_2ave,
// u, v, ww, muts and muave (the coefficients -- mut, c2a, cqw, alt, the
// tridiagonal factors, the damping weights from ph_1 -- fixed).  It is linear
// in all but muts, whose derivatives take the forward's outputs t_2ave and w:
//   t2a[k]  = (0.5 ((1+e) t_2[k] + (1-e) t2a_in[k]) + muave t0) / (muts (t0 + t_1[k])),  k < K
//   rhs[k]  = ph[k] + msfty dts (ph_tend[k] + 0.5 g (1-e) w[k] - ...) / mut,            1 <= k <= K
//   wp[k]   = w[k] + ... (rhs, ph at k-1, k, k+1; t2a at k-1, k),                      1 <= k <= K
//   ws      = the column solve of wp (w[0] from the terrain slopes and u, v),
//   w_out   = (ws - dampwt mut w_save) / (1 + dampwt),
//   ph_out  = rhs + msfty 0.5 dts g (1+e) w_out / muts,
// with K = kte - 1, on the columns [jds+1, jde-2) x [ids+1, ide-2); every
// other point passes w, ph and t_2ave through.  w[0] is formed from the
// terrain slopes and u, v at levels 0..2 (cf1..cf3), and muave enters
// t2a's numerator and wp[k] as - dts g muave / msfty.  One thread per column
// walks these in reverse.  work, work2: 3-D scratch (the sensitivities of w
// and of rhs along the column).
struct AdjW {
  const float *lT, *lW, *lP;                      // sensitivities of the outputs t_2ave, w, ph
  const float *mut, *muts, *t_1, *ph_1, *phb, *c2a, *cqw, *alt, *a, *alpha, *gamma, *rdnw, *rdn, *msfty;
  const float *t2a_out, *w_out, *ht, *msftx, *cf;  // the forward's outputs t_2ave, w; terrain, map factor, cf1..3
  const float *fnm, *fnp;
  float *lw_in, *lph_in, *lt2, *lt2a_in, *work, *work2;
  float *lu, *lv, *lmuts, *lmuave, *lww;           // zeroed by the caller; u, v by atomic adds
  float *lrw, *lpht, *lws;                         // rw_tend, ph_tend, w_save (zeroed by the caller)
  float* lt1;                                      // t_1 (zeroed by the caller)
  const float *ww, *wsave;                         // the call's ww and w_save
  float* lph1;                                     // ph_1: vertical advection and damping (zeroed)
  float dts, t0, ep1, em1, dampmag, hpi, hdepth, rdx, rdy;
  int K, j0, j1, i0, i1;
};

__global__ void k_advance_w_adj(Box all, Dims D, AdjW a) {
  int j, i;
  if (!box_column(all, j, i)) return;
  const int nz = D.nz, K = a.K;
  const bool in = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  if (!in) {
    for (int k = 0; k < nz; ++k) {
      AT(a.lw_in, k, j, i) = AT(a.lW, k, j, i);
      AT(a.lph_in, k, j, i) = AT(a.lP, k, j, i);
      AT(a.lt2a_in, k, j, i) = AT(a.lT, k, j, i);
      AT(a.lt2, k, j, i) = 0.0f;
    }
    return;
  }
  const float G = F(pw::g);
  const float my = AT2(a.msfty, j, i), mi = 1.0f / my, mut = AT2(a.mut, j, i), muts = AT2(a.muts, j, i);
  // damping weights (as the forward forms them) and the sensitivity of the solved w
  const float htop = (AT(a.ph_1, K, j, i) + AT(a.phb, K, j, i)) / G, hbot = htop - a.hdepth;
  for (int k = 1; k <= K; ++k) {
    const float hk = (AT(a.ph_1, k, j, i) + AT(a.phb, k, j, i)) / G;
    float dw = 0.0f;
    if (hk >= hbot) {
      const float sn = sinf(a.hpi * (hk - hbot) / a.hdepth);
      dw = a.dampmag * (sn * sn);
    }
    const float lwd = AT(a.lW, k, j, i) + my * 0.5f * a.dts * G * a.ep1 / muts * AT(a.lP, k, j, i);
    AT(a.work, k, j, i) = lwd / (1.0f + dw);
    AT(a.lws, k, j, i) = -dw * mut / (1.0f + dw) * lwd;   // the damping's pull towards w_save
    // the damping weight's dependence on ph_1: through hk at this level and
    // hbot (from level K); at level K the two cancel
    if (hk >= hbot && k < K) {
      const float th = a.hpi * (hk - hbot) / a.hdepth;
      const float ddw = a.dampmag * 2.0f * sinf(th) * cosf(th) * a.hpi / a.hdepth;
      const float ldw = lwd * (-mut * AT(a.wsave, k, j, i) - AT(a.w_out, k, j, i)) / (1.0f + dw);
      AT(a.lph1, k, j, i) += ldw * ddw / G;
      AT(a.lph1, K, j, i) -= ldw * ddw / G;
    }
  }
  // the column solve, transposed: back substitution upwards, elimination downwards
  for (int k = 1; k <= K - 1; ++k) AT(a.work, k + 1, j, i) -= AT(a.gamma, k, j, i) * AT(a.work, k, j, i);
  for (int k = K; k >= 1; --k) {
    const float v = AT(a.work, k, j, i) * AT(a.alpha, k, j, i);
    AT(a.work, k, j, i) = v;
    if (k - 1 >= 1) AT(a.work, k - 1, j, i) -= AT(a.a, k, j, i) * v;
  }
  // the surface w (from the terrain and u, v at levels 0..2) enters the solve
  // at level 1 and is an output itself
  {
    const float lw0 = AT(a.lW, 0, j, i) - AT(a.a, 1, j, i) * AT(a.work, 1, j, i);
    const float htc = AT2(a.ht, j, i), mx = AT2(a.msftx, j, i);
    const float cy = my * 0.5f * a.rdy * lw0, cx = mx * 0.5f * a.rdx * lw0;
    const float vn = cy * (AT2(a.ht, j + 1, i) - htc), vs = cy * (htc - AT2(a.ht, j - 1, i));
    const float ue = cx * (AT2(a.ht, j, i + 1) - htc), uw = cx * (htc - AT2(a.ht, j, i - 1));
    for (int c = 0; c < 3; ++c) {
      atomic_add(&AT(a.lv, c, j + 1, i), a.cf[c] * vn);
      atomic_add(&AT(a.lv, c, j, i), a.cf[c] * vs);
      atomic_add(&AT(a.lu, c, j, i + 1), a.cf[c] * ue);
      atomic_add(&AT(a.lu, c, j, i), a.cf[c] * uw);
    }
  }
  // rw_tend enters the predictor as dts rw_tend
  for (int k = 1; k <= K; ++k) AT(a.lrw, k, j, i) = a.dts * AT(a.work, k, j, i);
  // muave in the predictor (- dts g muave / msfty at every level 1..K) and muts in ph_out
  float lmuave = 0.0f, lmuts = 0.0f;
  {
    const float cp = my * 0.5f * a.dts * G * a.ep1;
    for (int k = 1; k <= K; ++k) {
      lmuave -= a.dts * G * mi * AT(a.work, k, j, i);
      lmuts -= cp * AT(a.w_out, k, j, i) / (muts * muts) * AT(a.lP, k, j, i);
    }
  }
  // accumulators: rhs (work2), ph_in, t2a (in lt2a_in), w_in
  for (int k = 0; k < nz; ++k) {
    const bool upd = k >= 1 && k <= K;
    AT(a.work2, k, j, i) = upd ? AT(a.lP, k, j, i) : 0.0f;
    AT(a.lph_in, k, j, i) = upd ? 0.0f : AT(a.lP, k, j, i);
    AT(a.lt2a_in, k, j, i) = AT(a.lT, k, j, i);
    AT(a.lw_in, k, j, i) = upd ? AT(a.work, k, j, i) : 0.0f;    // w[0] is overwritten from the terrain
  }
  // the explicit predictor wp[k]
  for (int k = 1; k <= K; ++k) {
    const float L = AT(a.work, k, j, i);
    if (k < K) {
      const float C = mi * AT(a.cqw, k, j, i) * (0.5f * a.dts * G * a.rdn[k]);
      const float P = AT(a.c2a, k, j, i) * a.rdnw[k] / mut, M = AT(a.c2a, k - 1, j, i) * a.rdnw[k - 1] / mut;
      AT(a.work2, k + 1, j, i) += C * P * a.ep1 * L;
      AT(a.work2, k, j, i) -= C * (P + M) * a.ep1 * L;
      if (k - 1 >= 1) AT(a.work2, k - 1, j, i) += C * M * a.ep1 * L;
      AT(a.lph_in, k + 1, j, i) += C * P * a.em1 * L;
      AT(a.lph_in, k, j, i) -= C * (P + M) * a.em1 * L;
      AT(a.lph_in, k - 1, j, i) += C * M * a.em1 * L;
      const float tg = a.dts * G * mi * a.rdn[k];
      AT(a.lt2a_in, k, j, i) += tg * AT(a.c2a, k, j, i) * AT(a.alt, k, j, i) * L;
      AT(a.lt2a_in, k - 1, j, i) -= tg * AT(a.c2a, k - 1, j, i) * AT(a.alt, k - 1, j, i) * L;
    } else {
      const float rk = a.rdnw[K - 1];
      const float Dk = mi * (-0.5f * a.dts * G / mut * rk * rk * 2.0f * AT(a.c2a, K - 1, j, i));
      AT(a.work2, K, j, i) += Dk * a.ep1 * L;
      if (K - 1 >= 1) AT(a.work2, K - 1, j, i) -= Dk * a.ep1 * L;
      AT(a.lph_in, K, j, i) += Dk * a.em1 * L;
      AT(a.lph_in, K - 1, j, i) -= Dk * a.em1 * L;
      AT(a.lt2a_in, K - 1, j, i) -= mi * a.dts * G * 2.0f * rk * AT(a.c2a, K - 1, j, i) * AT(a.alt, K - 1, j, i) * L;
    }
  }
  // rhs[k] = ph[k] + msfty dts 0.5 g (1-e) w[k] / mut - msfty dts (fnm[k] wup[k] + fnp[k] wdn[k]) / mut
  // + (terms without w, ph, ww), wup / wdn = 0.5 (ww[k+1] + ww[k]) / (ww[k] + ww[k-1]) times rdnw and the
  // ph_1 + phb differences above / below level k (k < K)
  const float gam = my * a.dts * 0.5f * G * a.em1 / mut;
  for (int k = 1; k <= K; ++k) {
    const float lr = AT(a.work2, k, j, i);
    AT(a.lph_in, k, j, i) += lr;
    AT(a.lw_in, k, j, i) += gam * lr;
    AT(a.lpht, k, j, i) = my * a.dts / mut * lr;       // rhs holds msfty dts ph_tend / mut
    if (k < K) {
      const float li = -a.dts * my / mut * lr;
      const float d1 = AT(a.ph_1, k + 1, j, i) - AT(a.ph_1, k, j, i) + AT(a.phb, k + 1, j, i) - AT(a.phb, k, j, i);
      const float d0 = AT(a.ph_1, k, j, i) - AT(a.ph_1, k - 1, j, i) + AT(a.phb, k, j, i) - AT(a.phb, k - 1, j, i);
      const float cu = 0.5f * a.rdnw[k] * d1 * a.fnm[k] * li, cd = 0.5f * a.rdnw[k - 1] * d0 * a.fnp[k] * li;
      AT(a.lww, k + 1, j, i) += cu;
      AT(a.lww, k, j, i) += cu + cd;
      AT(a.lww, k - 1, j, i) += cd;
      // the same wdwn terms through ph_1
      const float pu = a.fnm[k] * li * 0.5f * (AT(a.ww, k + 1, j, i) + AT(a.ww, k, j, i)) * a.rdnw[k];
      const float pd = a.fnp[k] * li * 0.5f * (AT(a.ww, k, j, i) + AT(a.ww, k - 1, j, i)) * a.rdnw[k - 1];
      AT(a.lph1, k + 1, j, i) += pu;
      AT(a.lph1, k, j, i) += pd - pu;
      AT(a.lph1, k - 1, j, i) -= pd;
    }
  }
  // t2a[k] for k < K from t_2, t2a_in, muave and muts; levels from K on pass through
  for (int k = 0; k < nz; ++k) {
    const float l = AT(a.lt2a_in, k, j, i);
    if (k < K) {
      const float den = muts * (a.t0 + AT(a.t_1, k, j, i));
      AT(a.lt2, k, j, i) = 0.5f * a.ep1 / den * l;
      AT(a.lt2a_in, k, j, i) = 0.5f * a.em1 / den * l;
      lmuave += a.t0 / den * l;
      lmuts -= AT(a.t2a_out, k, j, i) / muts * l;
      AT(a.lt1, k, j, i) = -AT(a.t2a_out, k, j, i) / (a.t0 + AT(a.t_1, k, j, i)) * l;
    } else {
      AT(a.lt2, k, j, i) = 0.0f;
    }
  }
  AT2(a.lmuave, j, i) = lmuave;
  AT2(a.lmuts, j, i) = lmuts;
}

extern "C" int pywrf_advance_w_adjoint(const pywrf_grid_t* g, const float* lam_t2ave_out, const float* lam_w_out,
                                       const float* lam_ph_out, const float* mut, const float* muts,
                                       const float* t_1, const float* ph_1, const float* phb, const float* c2a,
                                       const float* cqw, const float* alt, const float* a, const float* alpha,
                                       const float* gamma, double dts, double t0, double epssm, const float* rdnw,
                                       const float* rdn, const float* msfty, const float* t2ave_out,
                                       const float* w_out, const float* ht, const float* msftx, const float* cf,
                                       double rdx, double rdy, const float* fnm, const float* fnp, float* lam_w,
                                       float* lam_ph, float* lam_t_2, float* lam_t2ave, float* lam_u,
                                       float* lam_v, float* lam_muts, float* lam_muave, float* lam_ww,
                                       float* lam_rw_tend, float* lam_ph_tend, float* lam_w_save, float* lam_t_1,
                                       const float* ww, const float* w_save, float* lam_ph_1, float* work,
                                       float* work2, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->kts != 0 || g->kte > g->nz || g->kte < 3) return PYWRF_E_RANGE;
  AdjW p;
  p.lT = lam_t2ave_out; p.lW = lam_w_out; p.lP = lam_ph_out;
  p.mut = mut; p.muts = muts; p.t_1 = t_1; p.ph_1 = ph_1; p.phb = phb; p.c2a = c2a; p.cqw = cqw; p.alt = alt;
  p.a = a; p.alpha = alpha; p.gamma = gamma; p.rdnw = rdnw; p.rdn = rdn; p.msfty = msfty;
  p.lw_in = lam_w; p.lph_in = lam_ph; p.lt2 = lam_t_2; p.lt2a_in = lam_t2ave; p.work = work; p.work2 = work2;
  p.t2a_out = t2ave_out; p.w_out = w_out; p.ht = ht; p.msftx = msftx; p.cf = cf;
  p.lu = lam_u; p.lv = lam_v; p.lmuts = lam_muts; p.lmuave = lam_muave; p.lww = lam_ww;
  p.lrw = lam_rw_tend; p.lpht = lam_ph_tend; p.lws = lam_w_save; p.lt1 = lam_t_1;
  p.ww = ww; p.wsave = w_save; p.lph1 = lam_ph_1;
  p.fnm = fnm; p.fnp = fnp;
  p.rdx = F(rdx); p.rdy = F(rdy);
  p.dts = F(dts); p.t0 = F(t0); p.ep1 = F(1. + epssm); p.em1 = F(1. - epssm);
  p.dampmag = F(dts * 0.2); p.hpi = F(0.5 * 3.14159265359); p.hdepth = 5000.0f;
  p.K = g->kte - 1;
  p.j0 = imax(g->jts, g->jds + 1); p.j1 = imin(g->jte, g->jde - 2);
  p.i0 = imax(g->its, g->ids + 1); p.i1 = imin(g->ite, g->ide - 2);
  if (p.j0 < 1 || p.i0 < 1 || p.j1 + 1 > g->ny || p.i1 + 1 > g->nx || g->nz < 3) return PYWRF_E_RANGE;
  const long n3 = (long)g->nz * g->ny * g->nx, n2 = (long)g->ny * g->nx;
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_u, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_v, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_ww, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_rw_tend, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_ph_tend, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_w_save, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_t_1, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_ph_1, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(256), dim3(256), 0, (hipStream_t)s, lam_muts, n2);
  hipLaunchKernelGGL(k_zero_n, dim3(256), dim3(256), 0, (hipStream_t)s, lam_muave, n2);
  Box all = whole(g);
  PW_LAUNCH_COLS(k_advance_w_adj, all, s, dims_of(g), p);
  PW_RET();
}

// ------------------------------------------------------------------ advance_uv^T
// advance_uv (wrf_dynamics.py:3391-3525) is linear in u, v, ru_tend, rv_tend,
// p, ph, al, mu and mudf once alt, pb, php, cqu / cqv, the column masses
// muu / muv and the map factors are fixed.  For a u point (k, j, i) of its box
//   u += dts ru_tend - dts cqu dpx + mudf_x,
//   dpx = A (ph[k+1,i] - ph[k+1,i-1] + ph[k,i] - ph[k,i-1]
//            + (alt[k,i] + alt[k,i-1]) (p[k,i] - p[k,i-1]) + (al[k,i] + al[k,i-1]) (pb[k,i] - pb[k,i-1]))
//       + B[k] (rdnw[k] (dpn[k+1] - dpn[k]) - 0.5 (mu[i-1] + mu[i])),
//   A = msfux / msfuy 0.5 rdx muu,  B[k] = msfux / msfuy rdx (php[k,i] - php[k,i-1]),
//   dpn[0] from p[0..2] with cf1..cf3, dpn[k] = 0.5 (fnm[k] (p[k,i] + p[k,i-1]) + fnp[k] (p[k-1,i] + p[k-1,i-1])),
//   dpn[kde-1] = 0,  mudf_x = -emdiv dx (mudf[i] - mudf[i-1]) / msfuy;
// v the same along j.  One thread per u (or v) column of the box walks its
// levels and adds its sensitivities into the two columns it reads (atomic
// adds: neighbouring columns share points); u, v pass through.
struct AdjUV {
  const float *lu, *lv, *alt, *pb, *php, *cqu, *cqv, *muu, *muv, *msfux, *msfuy, *msfvx, *msfvy, *msfvx_inv;
  const float *fnm, *fnp, *rdnw, *cf;
  float *gp, *gph, *gal, *gmu, *gmudf, *grut, *grvt;
  float dts, rdx, rdy, emdiv, dx, dy;
  int k0, k1, kde1;              // levels [k0, k1) of the box; dpn is zero at kde1 = kde - 1
  Box bu, bv;                    // the u and v boxes (j, i ranges)
};


template <bool U>
__global__ void k_advance_uv_adj(Box b, Dims D, AdjUV a) {
  int j, i;
  if (!box_column(b, j, i)) return;
  // the column this point's differences reach back to: i - 1 for u, j - 1 for v
  const int jm = U ? j : j - 1, im = U ? i - 1 : i;
  const float* lam = U ? a.lu : a.lv;
  const float mf = U ? AT2(a.msfux, j, i) / AT2(a.msfuy, j, i) : AT2(a.msfvy, j, i) / AT2(a.msfvx, j, i);
  const float rd = U ? a.rdx : a.rdy;
  const float Acoef = mf * 0.5f * rd * (U ? AT2(a.muu, j, i) : AT2(a.muv, j, i));
  const float* cq = U ? a.cqu : a.cqv;
  float* grt = U ? a.grut : a.grvt;
  float smu = 0.0f, sl = 0.0f, c_prev = 0.0f;
  for (int k = a.k0; k < a.k1; ++k) {
    const float l = AT(lam, k, j, i);
    sl += l;
    AT(grt, k, j, i) = a.dts * l;
    const float G = -a.dts * AT(cq, k, j, i) * l;             // sensitivity of dp at this point
    // the A terms: ph at k and k + 1, p with the alt sum, al with the pb difference
    const float AG = Acoef * G;
    atomic_add(&AT(a.gph, k + 1, j, i), AG);
    atomic_add(&AT(a.gph, k + 1, jm, im), -AG);
    atomic_add(&AT(a.gph, k, j, i), AG);
    atomic_add(&AT(a.gph, k, jm, im), -AG);
    const float salt = AT(a.alt, k, j, i) + AT(a.alt, k, jm, im);
    atomic_add(&AT(a.gp, k, j, i), AG * salt);
    atomic_add(&AT(a.gp, k, jm, im), -AG * salt);
    const float dpb = AT(a.pb, k, j, i) - AT(a.pb, k, jm, im);
    atomic_add(&AT(a.gal, k, j, i), AG * dpb);
    atomic_add(&AT(a.gal, k, jm, im), AG * dpb);
    // the B terms: rdnw[k] (dpn[k + 1] - dpn[k]) and the column mass
    const float BG = mf * rd * (AT(a.php, k, j, i) - AT(a.php, k, jm, im)) * G;
    smu += -0.5f * BG;
    const float c_k = a.rdnw[k] * BG;
    const float ldpn = c_prev - c_k;                          // sensitivity of dpn[k]
    if (k == 0) {
      for (int c = 0; c < 3; ++c) {                           // dpn[0] from p[0..2] with cf1..cf3
        const float h = 0.5f * a.cf[c] * ldpn;
        atomic_add(&AT(a.gp, c, j, i), h);
        atomic_add(&AT(a.gp, c, jm, im), h);
      }
    } else if (k < a.kde1) {
      const float hf = 0.5f * a.fnm[k] * ldpn, hb = 0.5f * a.fnp[k] * ldpn;
      atomic_add(&AT(a.gp, k, j, i), hf);
      atomic_add(&AT(a.gp, k, jm, im), hf);
      atomic_add(&AT(a.gp, k - 1, j, i), hb);
      atomic_add(&AT(a.gp, k - 1, jm, im), hb);
    }
    c_prev = c_k;                                             // dpn[k + 1] gets c_k (zero at the top)
  }
  atomic_add(&AT2(a.gmu, j, i), smu);
  atomic_add(&AT2(a.gmu, jm, im), smu);
  const float xy = -a.emdiv * (U ? a.dx / AT2(a.msfuy, j, i) : a.dy * AT2(a.msfvx_inv, j, i)) * sl;
  atomic_add(&AT2(a.gmudf, j, i), xy);
  atomic_add(&AT2(a.gmudf, jm, im), -xy);
}

// Outputs: lam_u / lam_v (= the sensitivities of the outputs: u and v pass
// through), lam_p, lam_ph, lam_al, lam_mu, lam_mudf, lam_ru_tend,
// lam_rv_tend; the library zeroes them first.

extern "C" int pywrf_advance_uv_adjoint(const pywrf_grid_t* g, const float* lam_u_out, const float* lam_v_out,
                                        const float* pb, const float* php, const float* alt, const float* muu,
                                        const float* cqu, const float* muv, const float* cqv, const float* msfux,
                                        const float* msfuy, const float* msfvx, const float* msfvx_inv,
                                        const float* msfvy, double rdx, double rdy, double dts, const float* cf,
                                        const float* fnm, const float* fnp, double emdiv, const float* rdnw,
                                        int spec_zone, float* lam_u, float* lam_v, float* lam_p, float* lam_ph,
                                        float* lam_al, float* lam_mu, float* lam_mudf, float* lam_ru_tend,
                                        float* lam_rv_tend, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->kts != 0 || g->nz < 3 || g->kde > g->nz || g->kte > g->nz) return PYWRF_E_RANGE;
  const long n3 = (long)g->nz * g->ny * g->nx, n2 = (long)g->ny * g->nx;
  float* z3[] = {lam_p, lam_ph, lam_al, lam_ru_tend, lam_rv_tend};
  for (float* z : z3) hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, z, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(256), dim3(256), 0, (hipStream_t)s, lam_mu, n2);
  hipLaunchKernelGGL(k_zero_n, dim3(256), dim3(256), 0, (hipStream_t)s, lam_mudf, n2);
  if (hipError_t e = hipMemcpyAsync(lam_u, lam_u_out, n3 * sizeof(float), hipMemcpyDeviceToDevice, (hipStream_t)s))
    return (int)e;
  if (hipError_t e = hipMemcpyAsync(lam_v, lam_v_out, n3 * sizeof(float), hipMemcpyDeviceToDevice, (hipStream_t)s))
    return (int)e;
  AdjUV a;
  a.lu = lam_u_out; a.lv = lam_v_out; a.alt = alt; a.pb = pb; a.php = php; a.cqu = cqu; a.cqv = cqv;
  a.muu = muu; a.muv = muv; a.msfux = msfux; a.msfuy = msfuy; a.msfvx = msfvx; a.msfvy = msfvy;
  a.msfvx_inv = msfvx_inv; a.fnm = fnm; a.fnp = fnp; a.rdnw = rdnw; a.cf = cf;
  a.gp = lam_p; a.gph = lam_ph; a.gal = lam_al; a.gmu = lam_mu; a.gmudf = lam_mudf;
  a.grut = lam_ru_tend; a.grvt = lam_rv_tend;
  a.dts = F(dts); a.rdx = F(rdx); a.rdy = F(rdy); a.emdiv = F(emdiv); a.dx = F(1.0 / rdx); a.dy = F(1.0 / rdy);
  a.k0 = g->kts; a.k1 = imin(g->kte, g->kde - 1); a.kde1 = g->kde - 1;
  const int i0 = imax(g->its, g->ids + spec_zone), i1 = imin(g->ite, g->ide - spec_zone - 1);
  const int j0 = imax(g->jts, g->jds + spec_zone), j1 = imin(g->jte, g->jde - spec_zone - 1);
  const int iu1 = imin(g->ite, g->ide - spec_zone), jv1 = imin(g->jte, g->jde - spec_zone);
  if (i0 < 1 || j0 < 1 || iu1 > g->nx || jv1 > g->ny || a.k1 + 1 > g->nz) return PYWRF_E_RANGE;
  Box bu = mkbox(0, 1, j0, j1, i0, iu1), bv = mkbox(0, 1, j0, jv1, i0, i1);
  PW_LAUNCH_COLS(k_advance_uv_adj<true>, bu, s, dims_of(g), a);
  PW_LAUNCH_COLS(k_advance_uv_adj<false>, bv, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------ advance_mu_t adjoint
// The transpose of advance_mu_t (wrf_dynamics.py:3528-3638) with respect to u,
// v, ww, mu, muave, muts, mudf, t, t_ave, ft and mu_tend, with t_1 and the
// map factors held fixed (u_1, v_1 and ww_1 enter additively).  On a column
// (j, i) of the box, levels k < k1 = kte - 1:
//   dvdxi[k] = msftx msfty (rdy (v[k,j+1] - v[k,j]) + rdx (u[k,i+1] - u[k,i]) + terms in u_1, v_1),
//   S = sum_k dnw[k] dvdxi[k] + mu_tend,  mu' = mu + dts S,  mudf' = S,  muts' = mut + mu',
//   muave' = 0.5 ((1 + epssm) mu' + (1 - epssm) mu),
//   ww'[0] = ww[0] - ww_1[0],  ww'[k] = ww[0] - sum_{m<k} dnw[m] (S + dvdxi[m]) / msfty - ww_1[k],
//   wdtn[k] = ww'[k] (fnm[k] t_1[k] + fnp[k] t_1[k-1]) (0 < k < k1; 0 at 0 and k1),
//   t_ave' = t,  t' = t + msfty dts ft - dts msfty (msftx (0.5 rdy (v[k,j+1] (t_1[j+1] + t_1) - v (t_1 + t_1[j-1]))
//                    + 0.5 rdx (u[i+1] (t_1[i+1] + t_1) - u (t_1 + t_1[i-1]))) + rdnw[k] (wdtn[k+1] - wdtn[k])).
// Outside the box every output is its input.  One thread per column: a first
// pass down the column forms the suffix sums of the ww sensitivities (through
// wdtn) and the total sensitivity of S, a second pass the sensitivities of
// dvdxi and t, added into u and v with atomics (neighbouring columns share
// faces).
struct AdjMuT {
  const float *lww, *lt, *ltave, *lmuave, *lmuts, *lmudf, *lmu;
  const float *t1, *msftx, *msfty, *dnw, *fnm, *fnp, *rdnw;
  float *gww, *gu, *gv, *gmu, *gmuave, *gmuts, *gmudf, *gt, *gtave, *gft, *gmut;
  // the large-step fields: u_1, v_1 in the divergence (times muu / msfuy, muv msfvx_inv), ww_1
  // subtracted from ww, t_1 in the theta fluxes (with u, v and the forward's ww)
  const float *u, *v, *wwo, *muu, *muv, *msfuy, *msfvx_inv;
  float *gu1, *gv1, *gww1, *gt1;                   // zeroed by the caller
  float rdx, rdy, dts, epssm;
  int k1;
};

__device__ inline float mut_lw(const AdjMuT& a, const Dims& D, int k, int j, int i, float c) {
  // the sensitivity of ww'[k]: its own and, through wdtn[k] (0 < k < k1), t's
  float lw = AT(a.lww, k, j, i);
  if (k >= 1) {
    const float lwdtn = c * (a.rdnw[k] * AT(a.lt, k, j, i) - a.rdnw[k - 1] * AT(a.lt, k - 1, j, i));
    lw += lwdtn * (a.fnm[k] * AT(a.t1, k, j, i) + a.fnp[k] * AT(a.t1, k - 1, j, i));
  }
  return lw;
}

__global__ void k_advance_mu_t_adj(Box b, Dims D, AdjMuT a) {
  int j, i;
  if (!box_column(b, j, i)) return;
  const float mfx = AT2(a.msftx, j, i), mfy = AT2(a.msfty, j, i);
  const float e = a.epssm;
  const float lmn = AT2(a.lmu, j, i) + AT2(a.lmuts, j, i) + 0.5f * (1.0f + e) * AT2(a.lmuave, j, i);
  AT2(a.gmu, j, i) = lmn + 0.5f * (1.0f - e) * AT2(a.lmuave, j, i);
  AT2(a.gmuave, j, i) = 0.0f;
  AT2(a.gmuts, j, i) = 0.0f;
  AT2(a.gmudf, j, i) = 0.0f;
  const float c = a.dts * mfy;
  float R = 0.0f, A = 0.0f;
  for (int k = a.k1 - 1; k >= 0; --k) {
    A += a.dnw[k] * R;                           // R = sum of the ww' sensitivities above k
    R += mut_lw(a, D, k, j, i, c);
  }
  AT(a.gww, 0, j, i) = R;
  for (int k = 1; k < a.k1; ++k) AT(a.gww, k, j, i) = 0.0f;
  const float lS = a.dts * lmn + AT2(a.lmudf, j, i) - A / mfy;
  AT2(a.gmut, j, i) = lS;
  R = 0.0f;
  for (int k = a.k1 - 1; k >= 0; --k) {
    const float ldv = a.dnw[k] * lS - a.dnw[k] / mfy * R;
    const float lwk = mut_lw(a, D, k, j, i, c);
    R += lwk;
    AT(a.gww1, k, j, i) = -lwk;                  // ww' = ww - ww_1
    const float lt = AT(a.lt, k, j, i);
    AT(a.gt, k, j, i) = lt + AT(a.ltave, k, j, i);
    AT(a.gtave, k, j, i) = 0.0f;
    AT(a.gft, k, j, i) = mfy * a.dts * lt;
    const float q = -a.dts * mfy * mfx * lt;
    const float du = mfx * mfy * a.rdx * ldv, dv = mfx * mfy * a.rdy * ldv;
    const float t0 = AT(a.t1, k, j, i);
    atomic_add(&AT(a.gu, k, j, i + 1), du + q * 0.5f * a.rdx * (AT(a.t1, k, j, i + 1) + t0));
    atomic_add(&AT(a.gu, k, j, i), -du - q * 0.5f * a.rdx * (t0 + AT(a.t1, k, j, i - 1)));
    atomic_add(&AT(a.gv, k, j + 1, i), dv + q * 0.5f * a.rdy * (AT(a.t1, k, j + 1, i) + t0));
    atomic_add(&AT(a.gv, k, j, i), -dv - q * 0.5f * a.rdy * (t0 + AT(a.t1, k, j - 1, i)));
    atomic_add(&AT(a.gu1, k, j, i + 1), du * (AT2(a.muu, j, i + 1) / AT2(a.msfuy, j, i + 1)));
    atomic_add(&AT(a.gu1, k, j, i), -du * (AT2(a.muu, j, i) / AT2(a.msfuy, j, i)));
    atomic_add(&AT(a.gv1, k, j + 1, i), dv * (AT2(a.muv, j + 1, i) * AT2(a.msfvx_inv, j + 1, i)));
    atomic_add(&AT(a.gv1, k, j, i), -dv * (AT2(a.muv, j, i) * AT2(a.msfvx_inv, j, i)));
    // t_1 in the horizontal theta fluxes
    const float ue = AT(a.u, k, j, i + 1), uw = AT(a.u, k, j, i), vn = AT(a.v, k, j + 1, i), vs = AT(a.v, k, j, i);
    const float hx = q * 0.5f * a.rdx, hy = q * 0.5f * a.rdy;
    atomic_add(&AT(a.gt1, k, j, i + 1), hx * ue);
    atomic_add(&AT(a.gt1, k, j, i - 1), -hx * uw);
    atomic_add(&AT(a.gt1, k, j + 1, i), hy * vn);
    atomic_add(&AT(a.gt1, k, j - 1, i), -hy * vs);
    atomic_add(&AT(a.gt1, k, j, i), hx * (ue - uw) + hy * (vn - vs));
    // and in the vertical flux wdtn[k] = ww'[k] (fnm[k] t_1[k] + fnp[k] t_1[k-1]), 0 < k < k1
    if (k >= 1) {
      const float lwdtn = c * (a.rdnw[k] * AT(a.lt, k, j, i) - a.rdnw[k - 1] * AT(a.lt, k - 1, j, i));
      const float wk = AT(a.wwo, k, j, i);
      atomic_add(&AT(a.gt1, k, j, i), lwdtn * wk * a.fnm[k]);
      atomic_add(&AT(a.gt1, k - 1, j, i), lwdtn * wk * a.fnp[k]);
    }
  }
}

extern "C" int pywrf_advance_mu_t_adjoint(const pywrf_grid_t* g, const float* lam_ww, const float* lam_t,
                                          const float* lam_t_ave, const float* lam_muave, const float* lam_muts,
                                          const float* lam_mudf, const float* lam_mu, const float* t_1,
                                          const float* msftx, const float* msfty, const float* dnw,
                                          const float* fnm, const float* fnp, const float* rdnw, double rdx,
                                          double rdy, double dts, double epssm, float* lam_ww_ori, float* lam_u,
                                          float* lam_v, float* lam_mu_ori, float* lam_muave_ori,
                                          float* lam_muts_ori, float* lam_mudf_ori, float* lam_t_ori,
                                          float* lam_t_ave_ori, float* lam_ft, float* lam_mu_tend,
                                          const float* ww_out, const float* u, const float* v, const float* muu,
                                          const float* muv, const float* msfuy, const float* msfvx_inv,
                                          float* lam_u_1, float* lam_v_1, float* lam_ww_1, float* lam_t_1,
                                          pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->kts != 0 || g->kte > g->nz || g->kte < 2) return PYWRF_E_RANGE;
  const long n3 = (long)g->nz * g->ny * g->nx, n2 = (long)g->ny * g->nx;
  const hipStream_t st = (hipStream_t)s;
  // outside the box every output passes its input's sensitivity through
  const struct { float* dst; const float* src; long n; } pass[] = {
      {lam_ww_ori, lam_ww, n3}, {lam_t_ori, lam_t, n3}, {lam_t_ave_ori, lam_t_ave, n3},
      {lam_mu_ori, lam_mu, n2}, {lam_muave_ori, lam_muave, n2}, {lam_muts_ori, lam_muts, n2},
      {lam_mudf_ori, lam_mudf, n2}};
  for (const auto& p : pass)
    if (hipError_t e = hipMemcpyAsync(p.dst, p.src, p.n * sizeof(float), hipMemcpyDeviceToDevice, st)) return (int)e;
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, st, lam_u, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, st, lam_v, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, st, lam_ft, n3);
  hipLaunchKernelGGL(k_zero_n, dim3(256), dim3(256), 0, st, lam_mu_tend, n2);
  for (float* z : {lam_u_1, lam_v_1, lam_ww_1, lam_t_1})
    hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, st, z, n3);
  AdjMuT a;
  a.lww = lam_ww; a.lt = lam_t; a.ltave = lam_t_ave; a.lmuave = lam_muave; a.lmuts = lam_muts;
  a.lmudf = lam_mudf; a.lmu = lam_mu; a.t1 = t_1; a.msftx = msftx; a.msfty = msfty; a.dnw = dnw;
  a.fnm = fnm; a.fnp = fnp; a.rdnw = rdnw;
  a.gww = lam_ww_ori; a.gu = lam_u; a.gv = lam_v; a.gmu = lam_mu_ori; a.gmuave = lam_muave_ori;
  a.gmuts = lam_muts_ori; a.gmudf = lam_mudf_ori; a.gt = lam_t_ori; a.gtave = lam_t_ave_ori; a.gft = lam_ft;
  a.gmut = lam_mu_tend;
  a.u = u; a.v = v; a.wwo = ww_out; a.muu = muu; a.muv = muv; a.msfuy = msfuy; a.msfvx_inv = msfvx_inv;
  a.gu1 = lam_u_1; a.gv1 = lam_v_1; a.gww1 = lam_ww_1; a.gt1 = lam_t_1;
  a.rdx = F(rdx); a.rdy = F(rdy); a.dts = F(dts); a.epssm = F(epssm);
  a.k1 = g->kte - 1;
  const int j0 = imax(g->jts, g->jds + 1), j1 = imin(g->jte, g->jde - 2);
  const int i0 = imax(g->its, g->ids + 1), i1 = imin(g->ite, g->ide - 2);
  if (j0 < 1 || i0 < 1 || j1 + 1 > g->ny || i1 + 1 > g->nx) return PYWRF_E_RANGE;
  Box b = mkbox(0, 1, j0, j1, i0, i1);
  PW_LAUNCH_COLS(k_advance_mu_t_adj, b, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ calc_p_rho^T
// calc_p_rho (wrf_dynamics.py:3269-3313) in its box (levels kts..min(kte,
// kde-1), the interior columns):
//   al = -(alt mu + rdnw[k] (ph[k+1] - ph[k])) / mut,
//   p  = c2a (alt (t_2 - mu t_1) / (mut (t0 + t_1)) - al),
// then, on a call with step != 0, p' = p + smdiv (p - pm1) and pm1' = p (in
// the box); at step 0 pm1' is a copy of the whole of p.  Outside the box al and
// p keep their incoming values.  alt, t_1 and c2a are held fixed.  One thread
// per column walks the levels: ph[k]'s sensitivity gathers al[k]'s and
// al[k-1]'s, those of mu and mut are summed down the column.
struct AdjPRho {
  const float *lal, *lp, *lpm1;                          // sensitivities of the results
  const float *ph, *alt, *t_2, *t_1, *c2a, *mu, *mut, *rdnw;
  float *gal, *gp, *gpm1, *gph, *gt2, *gmu, *gmut;     // sensitivities of the inputs
  float* gt1;                                           // t_1 (null: not wanted)
  float t0, smdiv;
  int step;
};

__global__ void k_calc_p_rho_adj(Box all, Dims D, Box b, AdjPRho a) {
  int j, i;
  if (!box_column(all, j, i)) return;
  const bool col = b.has2(j, i);
  const float mut = col ? AT2(a.mut, j, i) : 1.0f, mu = col ? AT2(a.mu, j, i) : 0.0f;
  float gmu = 0.0f, gmut = 0.0f, prev = 0.0f;   // prev: rdnw[k-1] (al[k-1]'s sensitivity) / mut
  for (int k = 0; k < D.nz; ++k) {
    const long q = D.ix(k, j, i);
    const float lal = a.lal[q], lp = a.lp[q], lpm = a.lpm1[q];
    float r = 0.0f;
    if (col && k >= b.k0 && k < b.k1) {
      const float ph0 = a.ph[q], ph1 = AT(a.ph, k + 1, j, i);
      const float altq = a.alt[q], t1 = a.t_1[q], t2 = a.t_2[q], c2 = a.c2a[q], rd = a.rdnw[k];
      const float al = -1.0f / mut * (altq * mu + rd * (ph1 - ph0));
      const float den = mut * (a.t0 + t1);
      const float pa = altq * (t2 - mu * t1) / den;
      float lpr;                                 // sensitivity of p before the damping
      if (a.step != 0) {
        lpr = lp * (1.0f + a.smdiv) + lpm;
        a.gpm1[q] = -a.smdiv * lp;
      } else {
        lpr = lp + lpm;
        a.gpm1[q] = 0.0f;
      }
      const float cl = c2 * lpr;
      const float L = lal - cl;                  // sensitivity of al, through p as well
      a.gt2[q] = cl * altq / den;
      if (a.gt1) a.gt1[q] = cl * (-altq * mu / den - pa / (a.t0 + t1));
      gmu += cl * (-(altq * t1) / den) + L * (-altq / mut);
      gmut += cl * (-pa / mut) + L * (-al / mut);
      r = rd * L / mut;
      a.gal[q] = 0.0f;
      a.gp[q] = 0.0f;
    } else {
      a.gal[q] = lal;
      a.gp[q] = a.step != 0 ? lp : lp + lpm;
      a.gpm1[q] = a.step != 0 ? lpm : 0.0f;
      a.gt2[q] = 0.0f;
      if (a.gt1) a.gt1[q] = 0.0f;
    }
    a.gph[q] = r - prev;
    prev = r;
  }
  AT2(a.gmu, j, i) = gmu;
  AT2(a.gmut, j, i) = gmut;
}

extern "C" int pywrf_calc_p_rho_adjoint(const pywrf_grid_t* g, const float* lam_al, const float* lam_p,
                                        const float* lam_pm1, const float* ph, const float* alt, const float* t_2,
                                        const float* t_1, const float* c2a, const float* mu, const float* mut,
                                        double t0, const float* rdnw, double smdiv, int step, float* lam_al_ori,
                                        float* lam_p_ori, float* lam_pm1_in, float* lam_ph, float* lam_t_2,
                                        float* lam_mu, float* lam_mut, float* lam_t_1, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  if (b.k1 + 1 > g->nz) return PYWRF_E_RANGE;
  AdjPRho a = {lam_al, lam_p, lam_pm1, ph, alt, t_2, t_1, c2a, mu, mut, rdnw,
               lam_al_ori, lam_p_ori, lam_pm1_in, lam_ph, lam_t_2, lam_mu, lam_mut, lam_t_1, F(t0), F(smdiv),
               step};
  PW_LAUNCH_COLS(k_calc_p_rho_adj, whole(g), s, dims_of(g), b, a);
  PW_RET();
}

// ------------------------------------------------------------------ sumflux^T
// sumflux (wrf_dynamics.py:3799-3839) is linear: on its first call the three
// accumulators are zeroed over the tile, each then adds its flux over its box,
// and on the last call the sum is divided by the number of calls (the u_lin
// terms are fixed).  One thread per point.
struct SumLin {
  const float *muu, *muv, *msfuy, *msfvx_inv;
  float *gul, *gvl, *gwl;                               // u_lin, v_lin, ww_lin (null: not wanted)
};

__global__ void k_sumflux_adj(Box all, Dims D, const float* lru_m, const float* lrv_m, const float* lww_m,
                              float* gru, float* grv, float* gww, float* gru_m, float* grv_m, float* gww_m, Box b,
                              Box bu, Box bv, Box bw, int first, int last, float n, SumLin sl) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  const long q = D.ix(k, j, i);
  const bool zeroed = first && b.has(k, j, i);
  const float* lam[3] = {lru_m, lrv_m, lww_m};
  float* gf[3] = {gru, grv, gww};
  float* ga[3] = {gru_m, grv_m, gww_m};
  const Box* bx[3] = {&bu, &bv, &bw};
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const float l = lam[f][q];
    const bool in = bx[f]->has(k, j, i);
    const float s = last ? l / n : l;
    gf[f][q] = in ? s : 0.0f;
    ga[f][q] = zeroed ? 0.0f : (in ? s : l);
  }
  // the last call adds muu u_lin / msfuy, muv v_lin msfvx_inv and ww_lin
  if (sl.gul) {
    const bool iu = last && bu.has(k, j, i), iv = last && bv.has(k, j, i), iw = last && bw.has(k, j, i);
    sl.gul[q] = iu ? lru_m[q] * (AT2(sl.muu, j, i) / AT2(sl.msfuy, j, i)) : 0.0f;
    sl.gvl[q] = iv ? lrv_m[q] * (AT2(sl.muv, j, i) * AT2(sl.msfvx_inv, j, i)) : 0.0f;
    sl.gwl[q] = iw ? lww_m[q] : 0.0f;
  }
}

extern "C" int pywrf_sumflux_adjoint(const pywrf_grid_t* g, const float* lam_ru_m, const float* lam_rv_m,
                                     const float* lam_ww_m, int iteration, int number_of_small_timesteps,
                                     float* lam_ru, float* lam_rv, float* lam_ww, float* lam_ru_m_in,
                                     float* lam_rv_m_in, float* lam_ww_m_in, const float* muu, const float* muv,
                                     const float* msfuy, const float* msfvx_inv, float* lam_u_lin,
                                     float* lam_v_lin, float* lam_ww_lin, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kte > g->nz || number_of_small_timesteps == 0) return PYWRF_E_RANGE;
  int mini = imin(g->ide - 1, g->ite), minj = imin(g->jde - 1, g->jte), mink = imin(g->kde - 1, g->kte);
  Box bu = mkbox(g->kts, mink, g->jts, minj, g->its, g->ite);
  Box bv = mkbox(g->kts, mink, g->jts, g->jte, g->its, mini);
  Box bw = mkbox(g->kts, g->kte, g->jts, minj, g->its, mini);
  Box b = mkbox(g->kts, g->kte, g->jts, g->jte, g->its, g->ite);
  if (lam_u_lin && (!lam_v_lin || !lam_ww_lin || !muu || !muv || !msfuy || !msfvx_inv)) return PYWRF_E_ARG;
  SumLin sl = {muu, muv, msfuy, msfvx_inv, lam_u_lin, lam_v_lin, lam_ww_lin};
  PW_LAUNCH_BOX(k_sumflux_adj, whole(g), s, dims_of(g), lam_ru_m, lam_rv_m, lam_ww_m, lam_ru, lam_rv, lam_ww,
                lam_ru_m_in, lam_rv_m_in, lam_ww_m_in, b, bu, bv, bw, iteration == 1,
                iteration == number_of_small_timesteps, (float)number_of_small_timesteps, sl);
  PW_RET();
}

// ------------------------------------------------------------------ open-boundary fill^T
// The fill of one plane (set_physical_bc3d / bc2d, wrf_dynamics.py:164-278) is
// a run of copies: the x-halos from the first / last interior column (over
// rows that include the y-halos), then the y-halos from the first / last
// interior row.  Its transpose runs the two in the opposite order, each copy
// target handing its sensitivity to its source and keeping none: first the
// y-copies (one thread per column and side, its three targets in turn), a
// barrier, then the x-copies (one thread per row and side).  No two threads
// of one phase share a source or a target, so no atomics are needed; an
// x-copy's source in a y-halo row then receives the sensitivity of the value
// it held before the y-copies, which is what the x-copies read.
__device__ inline void open_bc_plane_adj(float* a, int nx, const BcSpec& c) {
  const int ny_items = 2 * (c.yc1 - c.yc0);
  for (int t = (int)threadIdx.x; t < ny_items; t += (int)blockDim.x) {
    const int col = c.yc0 + t / 2, side = t & 1;
    int srow, d0;
    if (side == 0) { srow = c.jds; d0 = c.jds - 3; }
    else if (c.vstag) { srow = c.jde - 1; d0 = c.jde; }
    else { srow = c.jde - 2; d0 = c.jde - 1; }
    float sum = 0.0f;
    for (int h = 0; h < 3; ++h) {
      float* p = a + (long)(d0 + h) * nx + col;
      sum += *p;
      *p = 0.0f;
    }
    a[(long)srow * nx + col] += sum;
  }
  __syncthreads();
  const int r0 = c.ustag ? c.ur0 : c.xr0, r1 = c.ustag ? c.ur1 : c.xr1;
  const int nstart = c.xr1 - c.xr0, nx_items = nstart + (r1 - r0);
  for (int t = (int)threadIdx.x; t < nx_items; t += (int)blockDim.x) {
    int row, src, d0;
    if (t < nstart) { row = c.xr0 + t; src = c.ids; d0 = c.ids - 3; }
    else { row = r0 + (t - nstart); src = c.ustag ? c.ide - 1 : c.ide - 2; d0 = c.ustag ? c.ide : c.ide - 1; }
    float* base = a + (long)row * nx;
    float sum = 0.0f;
    for (int h = 0; h < 3; ++h) {
      sum += base[d0 + h];
      base[d0 + h] = 0.0f;
    }
    base[src] += sum;
  }
}

#define PW_BC_ADJ_MAX 32
struct BcAdjJobs {
  int n;
  BcJob job[PW_BC_ADJ_MAX];
};

__global__ void k_open_bc_adj_group(long plane, int nx, BcAdjJobs J) {
  if ((int)blockIdx.y >= J.n) return;
  const BcJob& jb = J.job[blockIdx.y];
  if ((int)blockIdx.x >= jb.nk) return;
  open_bc_plane_adj(jb.p + (long)(jb.k0 + (int)blockIdx.x) * plane, nx, jb.c);
}

extern "C" int pywrf_set_physical_bc_group_adjoint(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n,
                                                   void* const* lam_fields, const char* variables,
                                                   const int* is2d, pywrf_stream_t s) {
  if (!pw_size_ok(g3) || !pw_size_ok(g2)) return PYWRF_E_RANGE;
  if (n < 0 || n > PW_BC_ADJ_MAX || g3->ny != g2->ny || g3->nx != g2->nx) return PYWRF_E_ARG;
  BcAdjJobs J;
  J.n = n;
  int maxk = 1;
  for (int f = 0; f < n; ++f) {
    BcJob& jb = J.job[f];
    jb.p = (float*)lam_fields[f];
    int e, k1 = 1;
    if (is2d[f]) {
      e = bc_spec2(g2, variables[f], jb.c);
      jb.k0 = 0;
      if (!e && !bc_spec_ok(g2->ny, g2->nx, 1, 0, 1, jb.c)) e = PYWRF_E_RANGE;
    } else {
      e = bc_spec3(g3, variables[f], jb.c, k1);
      jb.k0 = g3->kts;
      if (!e && !bc_spec_ok(g3->ny, g3->nx, g3->nz, g3->kts, k1, jb.c)) e = PYWRF_E_RANGE;
    }
    if (e) return e;
    // the copies' sources must lie outside their targets (interior rows and columns)
    if (jb.c.jde - 2 <= jb.c.jds || jb.c.ide - 2 <= jb.c.ids) return PYWRF_E_RANGE;
    jb.nk = k1 - jb.k0 > 0 ? k1 - jb.k0 : 0;
    maxk = imax(maxk, jb.nk);
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_open_bc_adj_group, dim3(maxk, n), dim3(BC_THREADS), 0, (hipStream_t)s,
                     (long)g3->ny * g3->nx, g3->nx, J);
  PW_RET();
}

// ------------------------------------------------------------------ small_step_prep^T (past the first stage)
// small_step_prep (wrf_dynamics.py:3152-3266) past the first RK stage, with
// pb, p, alt (so c2a) held fixed: on its boxes it saves u_2 .. ph_2 and ww,
// and forms the acoustic variables
//   u = (muus u_1 - muu u_2) / msfuy,  v = (muvs v_1 - muv v_2) msfvx_inv,
//   t = muts t_1 - mut t_2,  w = (muts w_1 - mut w_2) / msfty,  ph = ph_1 - ph_2,
// with muts = mub + mu_1, muus / muvs the averages of mub + mu_1 onto the u / v
// points, mu_save = mu_2 and mu_2' = mu_1 - mu_2.  Pass 1 (one thread per
// column) gives the 3-D sensitivities and sums down the column those of
// muus, muvs and muts; pass 2 (one thread per point of the plane) gathers
// the averages' and the masses' shares into mu_1 and mu_2.
struct PrepAdj {
  const float *lus, *lvs, *lws, *lts, *lps, *lwws, *luo, *lvo, *lwo, *lto, *lpo;   // 3-D results
  const float *lmu2, *lmuts, *lmuus, *lmuvs, *lmusave;                             // plane results
  const float *u_1, *v_1, *w_1, *t_1, *muus, *muvs, *muts, *muu, *muv, *mut, *msfuy, *msfvx_inv, *msfty;
  float *gu2, *gv2, *gw2, *gt2, *gph2, *gu1, *gv1, *gw1, *gt1, *gph1, *gww;
  float *smuus, *smuvs, *smuts;                                                    // plane scratch
  float *gmu1, *gmu2;
  int k1, kde, j0, j1, i0, i1, i1u, j1v;
};

__global__ void k_ssprep_adj(Box all, Dims D, PrepAdj a) {
  int j, i;
  if (!box_column(all, j, i)) return;
  const bool h = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  const bool ur = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1u;
  const bool vr = j >= a.j0 && j < a.j1v && i >= a.i0 && i < a.i1;
  const bool vsr = j >= a.j0 && j < a.j1v && i >= a.i0 && i < a.i1u;   // the muvs box (i-range i1u)
  const float muus = AT2(a.muus, j, i), muvs = AT2(a.muvs, j, i), muts = AT2(a.muts, j, i);
  const float muu = AT2(a.muu, j, i), muv = AT2(a.muv, j, i), mut = AT2(a.mut, j, i);
  const float my = AT2(a.msfuy, j, i), mvx = AT2(a.msfvx_inv, j, i), mty = AT2(a.msfty, j, i);
  float su = 0.0f, sv = 0.0f, st = 0.0f;
  for (int k = 0; k < D.nz; ++k) {
    const long q = D.ix(k, j, i);
    const bool bu = k < a.k1 && ur, bv = k < a.k1 && vr, bm = k < a.k1 && h, bw = k < a.kde && h;
    const float luo = a.luo[q], lvo = a.lvo[q], lwo = a.lwo[q], lto = a.lto[q], lpo = a.lpo[q];
    if (bu) {
      a.gu2[q] = a.lus[q] - luo * muu / my;
      a.gu1[q] = luo * muus / my;
      su += luo * a.u_1[q] / my;
    } else {
      a.gu2[q] = luo;
      a.gu1[q] = 0.0f;
    }
    if (bv) {
      a.gv2[q] = a.lvs[q] - lvo * muv * mvx;
      a.gv1[q] = lvo * muvs * mvx;
      sv += lvo * a.v_1[q] * mvx;
    } else {
      a.gv2[q] = lvo;
      a.gv1[q] = 0.0f;
    }
    if (bm) {
      a.gt2[q] = a.lts[q] - lto * mut;
      a.gt1[q] = lto * muts;
      st += lto * a.t_1[q];
    } else {
      a.gt2[q] = lto;
      a.gt1[q] = 0.0f;
    }
    if (bw) {
      a.gw2[q] = a.lws[q] - lwo * mut / mty;
      a.gw1[q] = lwo * muts / mty;
      st += lwo * a.w_1[q] / mty;
      a.gph2[q] = a.lps[q] - lpo;
      a.gph1[q] = lpo;
      a.gww[q] = a.lwws[q];
    } else {
      a.gw2[q] = lwo;
      a.gw1[q] = 0.0f;
      a.gph2[q] = lpo;
      a.gph1[q] = 0.0f;
      a.gww[q] = 0.0f;
    }
  }
  AT2(a.smuus, j, i) = ur ? AT2(a.lmuus, j, i) + su : 0.0f;
  AT2(a.smuvs, j, i) = vsr ? AT2(a.lmuvs, j, i) + sv : 0.0f;
  AT2(a.smuts, j, i) = h ? AT2(a.lmuts, j, i) + st : 0.0f;
}

__global__ void k_ssprep_adj_2d(Box all, Dims D, PrepAdj a) {
  int j, i;
  if (!box_column(all, j, i)) return;
  const bool h = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  const float lmu2 = AT2(a.lmu2, j, i);
  // muus[j][i] and muus[j][i+1] average mu_1[j][i]; muvs[j][i] and muvs[j+1][i] too
  float g1 = 0.5f * (AT2(a.smuus, j, i) + (i + 1 < D.nx ? AT2(a.smuus, j, i + 1) : 0.0f)) +
             0.5f * (AT2(a.smuvs, j, i) + (j + 1 < D.ny ? AT2(a.smuvs, j + 1, i) : 0.0f));
  float g2 = lmu2;                                        // outside the box mu_2 passes through
  if (h) {
    g1 += AT2(a.smuts, j, i) + lmu2;                      // muts = mub + mu_1, mu_2' = mu_1 - mu_2
    g2 = AT2(a.lmusave, j, i) - lmu2;
  }
  AT2(a.gmu1, j, i) = g1;
  AT2(a.gmu2, j, i) = g2;
}

extern "C" int pywrf_small_step_prep_adjoint(
    const pywrf_grid_t* g, int rk_step, const float* lam_u_save, const float* lam_v_save, const float* lam_w_save,
    const float* lam_t_save, const float* lam_ph_save, const float* lam_ww_save, const float* lam_u,
    const float* lam_v, const float* lam_w, const float* lam_t, const float* lam_ph, const float* lam_mu_2,
    const float* lam_muts, const float* lam_muus, const float* lam_muvs, const float* lam_mu_save,
    const float* u_1, const float* v_1, const float* w_1, const float* t_1, const float* muus, const float* muvs,
    const float* muts, const float* muu, const float* muv, const float* mut, const float* msfuy,
    const float* msfvx_inv, const float* msfty, float* lam_u_2_in, float* lam_v_2_in, float* lam_w_2_in,
    float* lam_t_2_in, float* lam_ph_2_in, float* lam_u_1, float* lam_v_1, float* lam_w_1, float* lam_t_1,
    float* lam_ph_1, float* lam_ww, float* lam_mu_1, float* lam_mu_2_in, float* work2d, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (rk_step == 1) return PYWRF_E_ARG;            // the first stage copies: not differentiated here
  if (g->kde > g->nz || g->kts != 0) return PYWRF_E_RANGE;
  PrepAdj a = {lam_u_save, lam_v_save, lam_w_save, lam_t_save, lam_ph_save, lam_ww_save, lam_u, lam_v, lam_w,
               lam_t, lam_ph, lam_mu_2, lam_muts, lam_muus, lam_muvs, lam_mu_save, u_1, v_1, w_1, t_1, muus, muvs,
               muts, muu, muv, mut, msfuy, msfvx_inv, msfty, lam_u_2_in, lam_v_2_in, lam_w_2_in, lam_t_2_in,
               lam_ph_2_in, lam_u_1, lam_v_1, lam_w_1, lam_t_1, lam_ph_1, lam_ww, nullptr, nullptr, nullptr,
               lam_mu_1, lam_mu_2_in, imin(g->kte, g->kde - 1), g->kde, g->jts, imin(g->jte, g->jde - 1), g->its,
               imin(g->ite, g->ide - 1), g->ite, g->jte};
  const long n2 = (long)g->ny * g->nx;
  a.smuus = work2d; a.smuvs = work2d + n2; a.smuts = work2d + 2 * n2;
  if (a.j0 < 1 || a.i0 < 1) return PYWRF_E_RANGE;   // the averages read j-1 / i-1
  Box pl = mkbox(0, 1, 0, g->ny, 0, g->nx);
  PW_LAUNCH_COLS(k_ssprep_adj, pl, s, dims_of(g), a);
  PW_LAUNCH_COLS(k_ssprep_adj_2d, pl, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ calc_mu_uv_1^T
// calc_mu_uv_1 (wrf_dynamics.py:3844-3860) sets muu / muv from mu: the
// two-point average inside, mu itself on the first row / column and the
// second last point's value on the last; elsewhere muu / muv keep theirs.
// Transpose: each written point hands its sensitivity to the mu points it
// read (atomic adds onto the zeroed lam_mu) and keeps none.
__global__ void k_calc_mu_uv_1_adj(Box b, Dims D, const float* lmuu, const float* lmuv, float* gmu, float* gmuu,
                                   float* gmuv, int jts, int jtf, int jte, int its, int itf, int ite) {
  int j, i;
  if (!box_column(b, j, i)) return;
  float lu = AT2(lmuu, j, i), lv = AT2(lmuv, j, i);
  bool wu = false, wv = false;
  if (j < jtf) {
    if (i == ite - 1) { atomic_add(&AT2(gmu, j, ite - 2), lu); wu = true; }
    else if (i >= its + 1) { atomic_add(&AT2(gmu, j, i), 0.5f * lu); atomic_add(&AT2(gmu, j, i - 1), 0.5f * lu); wu = true; }
    else if (i == its) { atomic_add(&AT2(gmu, j, its), lu); wu = true; }
  }
  if (i < itf) {
    if (j == jte - 1) { atomic_add(&AT2(gmu, jte - 2, i), lv); wv = true; }
    else if (j >= jts + 1) { atomic_add(&AT2(gmu, j, i), 0.5f * lv); atomic_add(&AT2(gmu, j - 1, i), 0.5f * lv); wv = true; }
    else if (j == jts) { atomic_add(&AT2(gmu, jts, i), lv); wv = true; }
  }
  AT2(gmuu, j, i) = wu ? 0.0f : lu;
  AT2(gmuv, j, i) = wv ? 0.0f : lv;
}

__global__ void k_copy_n(const float* src, float* dst, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = src[t];
}

extern "C" int pywrf_calc_mu_uv_1_adjoint(const pywrf_grid_t* g, const float* lam_muu, const float* lam_muv,
                                          float* lam_mu, float* lam_muu_in, float* lam_muv_in, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  const int jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  if (g->ite - 2 < 0 || g->jte - 2 < 0) return PYWRF_E_RANGE;
  const long n2 = (long)g->ny * g->nx;
  hipLaunchKernelGGL(k_zero_n, dim3(256), dim3(256), 0, (hipStream_t)s, lam_mu, n2);
  // outside the box muu / muv pass through
  hipLaunchKernelGGL(k_copy_n, dim3(256), dim3(256), 0, (hipStream_t)s, lam_muu, lam_muu_in, n2);
  hipLaunchKernelGGL(k_copy_n, dim3(256), dim3(256), 0, (hipStream_t)s, lam_muv, lam_muv_in, n2);
  Box b = mkbox(0, 1, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_COLS(k_calc_mu_uv_1_adj, b, s, dims_of(g), lam_muu, lam_muv, lam_mu, lam_muu_in, lam_muv_in, g->jts, jtf,
                 g->jte, g->its, itf, g->ite);
  PW_RET();
}

// ------------------------------------------------------------------ small_step_finish^T
// small_step_finish (wrf_dynamics.py:3863-3978) couples the acoustic
// variables back on its boxes:
//   u = (msfuy u + u_save muu) / muus,  v = (msfvx v + v_save muv) / muvs,
//   w = (msfty w + w_save mut) / muts,  ph = ph + ph_save,  ww = ww + ww1,
//   t = (t [- dts n mut h_diabatic] + t_save mut) / muts,  mu_2 = mu_2 + mu_save;
// elsewhere the fields pass.  mut, muu, muv, the map factors and h_diabatic
// are held fixed.  One thread per column; the sensitivities of muus, muvs and
// muts are summed down it (muts by division: the results are inputs).
struct FinAdj {
  const float *lu, *lv, *lw, *lph, *lww, *lt, *lmu;                                // results
  const float *uo, *vo, *wo, *to;                                                  // forward results
  const float *muu, *muus, *muv, *muvs, *mut, *muts, *msfuy, *msfvx, *msfty;
  float *gu, *gv, *gw, *gph, *gww, *gt, *gus, *gvs, *gws, *gps, *gww1, *gts;
  float *gmuus, *gmuvs, *gmuts, *gmu2, *gmusave;
  int kde, j0, j1, i0, i1, jte, ite;
};

__global__ void k_ssfinish_adj(Box all, Dims D, FinAdj a) {
  int j, i;
  if (!box_column(all, j, i)) return;
  const bool h = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  const bool vr = j >= a.j0 && j < a.jte && i >= a.i0 && i < a.i1;
  const bool ur = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.ite;
  const float muus = AT2(a.muus, j, i), muvs = AT2(a.muvs, j, i), muts = AT2(a.muts, j, i);
  const float muu = AT2(a.muu, j, i), muv = AT2(a.muv, j, i), mut = AT2(a.mut, j, i);
  const float my = AT2(a.msfuy, j, i), mx = AT2(a.msfvx, j, i), mty = AT2(a.msfty, j, i);
  float su = 0.0f, sv = 0.0f, st = 0.0f;
  for (int k = 0; k < D.nz; ++k) {
    const long q = D.ix(k, j, i);
    const float lv = a.lv[q], lu = a.lu[q], lw = a.lw[q], lph = a.lph[q], lww = a.lww[q], lt = a.lt[q];
    if (k < a.kde - 1 && vr) {
      a.gv[q] = lv * mx / muvs;
      a.gvs[q] = lv * muv / muvs;
      sv -= lv * a.vo[q] / muvs;
    } else {
      a.gv[q] = lv;
      a.gvs[q] = 0.0f;
    }
    if (k < a.kde - 1 && ur) {
      a.gu[q] = lu * my / muus;
      a.gus[q] = lu * muu / muus;
      su -= lu * a.uo[q] / muus;
    } else {
      a.gu[q] = lu;
      a.gus[q] = 0.0f;
    }
    if (k < a.kde && h) {
      a.gw[q] = lw * mty / muts;
      a.gws[q] = lw * mut / muts;
      st -= lw * a.wo[q] / muts;
      a.gps[q] = lph;
      a.gww1[q] = lww;
    } else {
      a.gw[q] = lw;
      a.gws[q] = 0.0f;
      a.gps[q] = 0.0f;
      a.gww1[q] = 0.0f;
    }
    a.gph[q] = lph;
    a.gww[q] = lww;
    if (k < a.kde - 1 && h) {
      a.gt[q] = lt / muts;
      a.gts[q] = lt * mut / muts;
      st -= lt * a.to[q] / muts;
    } else {
      a.gt[q] = lt;
      a.gts[q] = 0.0f;
    }
  }
  AT2(a.gmuus, j, i) = su;
  AT2(a.gmuvs, j, i) = sv;
  AT2(a.gmuts, j, i) = st;
  const float lmu = AT2(a.lmu, j, i);
  AT2(a.gmu2, j, i) = lmu;
  AT2(a.gmusave, j, i) = h ? lmu : 0.0f;
}

extern "C" int pywrf_small_step_finish_adjoint(
    const pywrf_grid_t* g, const float* lam_u, const float* lam_v, const float* lam_w, const float* lam_ph,
    const float* lam_ww, const float* lam_t, const float* lam_mu_2, const float* u_out, const float* v_out,
    const float* w_out, const float* t_out, const float* muu, const float* muus, const float* muv,
    const float* muvs, const float* mut, const float* muts, const float* msfuy, const float* msfvx,
    const float* msfty, float* lam_u_2, float* lam_v_2, float* lam_w_2, float* lam_ph_2, float* lam_ww_in,
    float* lam_t_2, float* lam_u_save, float* lam_v_save, float* lam_w_save, float* lam_ph_save, float* lam_ww1,
    float* lam_t_save, float* lam_muus, float* lam_muvs, float* lam_muts, float* lam_mu_2_in, float* lam_mu_save,
    pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kde > g->nz || g->kds != 0) return PYWRF_E_RANGE;
  FinAdj a = {lam_u, lam_v, lam_w, lam_ph, lam_ww, lam_t, lam_mu_2, u_out, v_out, w_out, t_out, muu, muus, muv,
              muvs, mut, muts, msfuy, msfvx, msfty, lam_u_2, lam_v_2, lam_w_2, lam_ph_2, lam_ww_in, lam_t_2,
              lam_u_save, lam_v_save, lam_w_save, lam_ph_save, lam_ww1, lam_t_save, lam_muus, lam_muvs, lam_muts,
              lam_mu_2_in, lam_mu_save, g->kde, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1),
              g->jte, g->ite};
  Box pl = mkbox(0, 1, 0, g->ny, 0, g->nx);
  PW_LAUNCH_COLS(k_ssfinish_adj, pl, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ advect_u^T
// advect_u (wrf_dynamics.py:1112-1307) with the mass fluxes fixed is linear in
// u except at the two x-faces next to the edges, whose upwind value is chosen
// by the sign of u itself (piecewise linear: the forward's branch is taken).
// Scatter form: each tendency point hands its sensitivity, through the
// coefficients of its six faces' rules, to the u points those faces read
// (atomic adds onto the zeroed output).
struct AdjAdvU {
  const float *lam, *u, *ru, *rv, *rom, *msfux, *fzm, *fzp, *rdzw;
  float* gu;
  float rdx, rdy;
  int ids, ide, jds, jde, ktf;
  Box b, by;
};

// the y-face jf of u column (k, i): weight w times d(face)/d(u at row jf + o)
__device__ inline void u_yface_adj(const AdjAdvU& a, int jf, float rvs, float w, float* acc) {
  auto put = [&](int o, float c) { acc[o] += w * c; };
  if (jf >= a.jds + 3 && jf < a.jde - 3) {
    const float vel = 0.5f * rvs, s = sgn1(vel);
    for (int o = -3; o <= 2; ++o) put(o, vel * c5(o, s));
  } else if (jf == a.jds + 1 || jf == a.jde - 2) {
    put(0, 0.25f * rvs);
    put(-1, 0.25f * rvs);
  } else if (jf == a.jds + 2 || jf == a.jde - 3) {
    const float vel = 0.5f * rvs, s = jf == a.jds + 2 ? sgn1(vel) : sgn1(0.0f);
    for (int o = -2; o <= 1; ++o) put(o, vel * c3(o, s));
  }
}

__device__ inline void u_xface_adj(const AdjAdvU& a, const Dims& D, int k, int j, int f, float rus, float w, float* acc) {
  auto put = [&](int o, float c) { acc[o] += w * c; };
  if (f >= a.ids + 3 && f < a.ide - 2) {
    const float vel = 0.5f * rus, s = sgn1(vel);
    for (int o = -3; o <= 2; ++o) put(o, vel * c5(o, s));
  } else if (f == a.ids + 1) {
    if (AT(a.u, k, j, f) < 0.0f) put(0, 0.5f * rus);
    else { put(0, 0.25f * rus); put(-1, 0.25f * rus); }
  } else if (f == a.ids + 2 || f == a.ide - 2) {
    const float vel = 0.5f * rus, s = sgn1(vel);
    for (int o = -2; o <= 1; ++o) put(o, vel * c3(o, s));
  } else if (f == a.ide - 1) {
    if (AT(a.u, k, j, f - 1) > 0.0f) put(-1, 0.5f * rus);
    else { put(-1, 0.25f * rus); put(0, 0.25f * rus); }
  }
}

__device__ inline void u_zface_adj(const AdjAdvU& a, int kf, float vel, float w, float* acc) {
  if (kf < 1 || kf >= a.ktf) return;
  auto put = [&](int o, float c) { acc[o] += w * c; };
  if (kf >= 3 && kf < a.ktf - 2) {
    const float s = sgn1(-vel);
    for (int o = -3; o <= 2; ++o) put(o, vel * c5(o, s));
  } else if (kf == 1 || kf == a.ktf - 1) {
    put(0, vel * a.fzm[kf]);
    put(-1, vel * a.fzp[kf]);
  } else {
    const float s = sgn1(-vel);
    for (int o = -2; o <= 1; ++o) put(o, vel * c3(o, s));
  }
}

__global__ void k_advect_u_adj(Box all, Dims D, AdjAdvU a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  const float lam = AT(a.lam, k, j, i);
  if (lam == 0.0f) return;
  const float msf = AT2(a.msfux, j, i), rdz = a.rdzw[k];
  // the two faces of each direction read overlapping stencils: their shares
  // are summed per target (offsets -3..3 from the point) before the atomics
  float ay[7] = {}, ax[7] = {}, az[7] = {};
  if (a.by.has(k, j, i)) {
    const float wy = msf * a.rdy * lam;
    u_yface_adj(a, j + 1, AT(a.rv, k, j + 1, i) + AT(a.rv, k, j + 1, i - 1), -wy, ay + 4);
    u_yface_adj(a, j, AT(a.rv, k, j, i) + AT(a.rv, k, j, i - 1), wy, ay + 3);
  }
  const float wx = msf * a.rdx * lam;
  u_xface_adj(a, D, k, j, i + 1, AT(a.ru, k, j, i + 1) + AT(a.ru, k, j, i), -wx, ax + 4);
  u_xface_adj(a, D, k, j, i, AT(a.ru, k, j, i) + AT(a.ru, k, j, i - 1), wx, ax + 3);
  const float wz = rdz * lam;
  u_zface_adj(a, k + 1, 0.5f * (AT(a.rom, k + 1, j, i - 1) + AT(a.rom, k + 1, j, i)), -wz, az + 4);
  u_zface_adj(a, k, 0.5f * (AT(a.rom, k, j, i - 1) + AT(a.rom, k, j, i)), wz, az + 3);
  for (int m = 0; m < 7; ++m) {          // a nonzero share is one a face stencil read: in bounds
    atomic_add(&AT(a.gu, k, j + m - 3, i), ay[m]);
    atomic_add(&AT(a.gu, k, j, i + m - 3), ax[m]);
    atomic_add(&AT(a.gu, k + m - 3, j, i), az[m]);
  }
}

extern "C" int pywrf_advect_u_adjoint(const pywrf_grid_t* g, const float* lam_tendency, const float* u,
                                      const float* ru, const float* rv, const float* rom, const float* msfux,
                                      const float* fzm, const float* fzp, double rdx, double rdy,
                                      const float* rdzw, float* lam_u, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  const int ktf = imin(g->kte, g->kde - 1);
  if (g->kts != 0 || ktf + 1 > g->nz) return PYWRF_E_RANGE;
  const int i0 = imax(g->ids + 1, g->its), i1 = imin(g->ide - 1, g->ite);
  const int j0 = g->jts, j1 = imin(g->jte, g->jde - 1);
  AdjAdvU a;
  a.lam = lam_tendency; a.u = u; a.ru = ru; a.rv = rv; a.rom = rom; a.msfux = msfux; a.fzm = fzm; a.fzp = fzp;
  a.rdzw = rdzw; a.gu = lam_u; a.rdx = F(rdx); a.rdy = F(rdy);
  a.ids = g->ids; a.ide = g->ide; a.jds = g->jds; a.jde = g->jde; a.ktf = ktf;
  a.b = mkbox(g->kts, ktf, j0, j1, i0, i1);
  a.by = mkbox(g->kts, ktf, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), i0, i1);
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_u, (long)g->nz * g->ny * g->nx);
  PW_LAUNCH_BOX(k_advect_u_adj, a.b, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ advect_v^T
// advect_v (wrf_dynamics.py:1332-1532), as advect_u's transpose with the roles
// of x and y exchanged: the upwind choice by the sign of v at the edge
// y-faces, the symmetric rules at the x-faces, the z term scaled by
// msfvy / msfvx, and the reference's +v sign at the z-face ktf-2.
struct AdjAdvV {
  const float *lam, *v, *ru, *rv, *rom, *msfvx, *msfvy, *fzm, *fzp, *rdzw;
  float* gv;
  float rdx, rdy;
  int ids, ide, jds, jde, ktf;
  Box b, bx;
};

__device__ inline void v_yface_adj(const AdjAdvV& a, const Dims& D, int k, int f, int i, float rvs, float w, float* acc) {
  auto put = [&](int o, float c) { acc[o] += w * c; };
  if (f >= a.jds + 3 && f < a.jde - 2) {
    const float vel = 0.5f * rvs, s = sgn1(vel);
    for (int o = -3; o <= 2; ++o) put(o, vel * c5(o, s));
  } else if (f == a.jds + 1) {
    if (AT(a.v, k, f, i) < 0.0f) put(0, 0.5f * rvs);
    else { put(0, 0.25f * rvs); put(-1, 0.25f * rvs); }
  } else if (f == a.jds + 2 || f == a.jde - 2) {
    const float vel = 0.5f * rvs, s = sgn1(vel);
    for (int o = -2; o <= 1; ++o) put(o, vel * c3(o, s));
  } else if (f == a.jde - 1) {
    if (AT(a.v, k, f - 1, i) > 0.0f) put(-1, 0.5f * rvs);
    else { put(-1, 0.25f * rvs); put(0, 0.25f * rvs); }
  }
}

__device__ inline void v_xface_adj(const AdjAdvV& a, const Dims& D, int k, int j, int f, float rus, float w, float* acc) {
  auto put = [&](int o, float c) { acc[o] += w * c; };
  if (f >= a.ids + 3 && f < a.ide - 3) {
    const float vel = 0.5f * rus, s = sgn1(vel);
    for (int o = -3; o <= 2; ++o) put(o, vel * c5(o, s));
  } else if (f == a.ids + 1 || f == a.ide - 2) {
    put(0, 0.25f * rus);
    put(-1, 0.25f * rus);
  } else if (f == a.ids + 2 || f == a.ide - 3) {
    const float vel = 0.5f * rus, s = sgn1(vel);
    for (int o = -2; o <= 1; ++o) put(o, vel * c3(o, s));
  }
}

__device__ inline void v_zface_adj(const AdjAdvV& a, const Dims& D, int kf, int j, int i, float vel, float w, float* acc) {
  if (kf < 1 || kf >= a.ktf) return;
  auto put = [&](int o, float c) { acc[o] += w * c; };
  if (kf >= 3 && kf < a.ktf - 2) {
    const float s = sgn1(-vel);
    for (int o = -3; o <= 2; ++o) put(o, vel * c5(o, s));
  } else if (kf == 1 || kf == a.ktf - 1) {
    put(0, vel * a.fzm[kf]);
    put(-1, vel * a.fzp[kf]);
  } else {
    const float s = kf == 2 ? sgn1(-vel) : sgn1(vel);
    for (int o = -2; o <= 1; ++o) put(o, vel * c3(o, s));
  }
}

__global__ void k_advect_v_adj(Box all, Dims D, AdjAdvV a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  const float lam = AT(a.lam, k, j, i);
  if (lam == 0.0f) return;
  const float my = AT2(a.msfvy, j, i), mvx = AT2(a.msfvx, j, i), rdz = a.rdzw[k];
  // shares summed per target (offsets -3..3 from the point) before the atomics
  float ay[7] = {}, ax[7] = {}, az[7] = {};
  const float wy = my * a.rdy * lam;
  v_yface_adj(a, D, k, j + 1, i, AT(a.rv, k, j + 1, i) + AT(a.rv, k, j, i), -wy, ay + 4);
  v_yface_adj(a, D, k, j, i, AT(a.rv, k, j, i) + AT(a.rv, k, j - 1, i), wy, ay + 3);
  if (a.bx.has(k, j, i)) {
    const float wx = my * a.rdx * lam;
    v_xface_adj(a, D, k, j, i + 1, AT(a.ru, k, j, i + 1) + AT(a.ru, k, j - 1, i + 1), -wx, ax + 4);
    v_xface_adj(a, D, k, j, i, AT(a.ru, k, j, i) + AT(a.ru, k, j - 1, i), wx, ax + 3);
  }
  const float wz = (my / mvx) * rdz * lam;
  v_zface_adj(a, D, k + 1, j, i, 0.5f * (AT(a.rom, k + 1, j, i) + AT(a.rom, k + 1, j - 1, i)), -wz, az + 4);
  v_zface_adj(a, D, k, j, i, 0.5f * (AT(a.rom, k, j, i) + AT(a.rom, k, j - 1, i)), wz, az + 3);
  for (int m = 0; m < 7; ++m) {          // a nonzero share is one a face stencil read: in bounds
    atomic_add(&AT(a.gv, k, j + m - 3, i), ay[m]);
    atomic_add(&AT(a.gv, k, j, i + m - 3), ax[m]);
    atomic_add(&AT(a.gv, k + m - 3, j, i), az[m]);
  }
}

extern "C" int pywrf_advect_v_adjoint(const pywrf_grid_t* g, const float* lam_tendency, const float* v,
                                      const float* ru, const float* rv, const float* rom, const float* msfvx,
                                      const float* msfvy, const float* fzm, const float* fzp, double rdx, double rdy,
                                      const float* rdzw, float* lam_v, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  const int ktf = imin(g->kte, g->kde - 1);
  if (g->kts != 0 || ktf + 1 > g->nz) return PYWRF_E_RANGE;
  const int i0 = g->its, i1 = imin(g->ite, g->ide - 1);
  const int j0 = imax(g->jts, g->jds + 1), j1 = imin(g->jte, g->jde - 1);
  AdjAdvV a;
  a.lam = lam_tendency; a.v = v; a.ru = ru; a.rv = rv; a.rom = rom; a.msfvx = msfvx; a.msfvy = msfvy;
  a.fzm = fzm; a.fzp = fzp; a.rdzw = rdzw; a.gv = lam_v; a.rdx = F(rdx); a.rdy = F(rdy);
  a.ids = g->ids; a.ide = g->ide; a.jds = g->jds; a.jde = g->jde; a.ktf = ktf;
  a.b = mkbox(g->kts, ktf, j0, j1, i0, i1);
  a.bx = mkbox(g->kts, ktf, j0, j1, imax(g->ids + 1, g->its), imin(g->ide - 2, g->ite));
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_v, (long)g->nz * g->ny * g->nx);
  PW_LAUNCH_BOX(k_advect_v_adj, a.b, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ advect_w^T
// advect_w (wrf_dynamics.py:1535-1838) is linear in w with the mass fluxes
// fixed (the horizontal ones interpolated to w levels, extrapolated at the
// top; the vertical faces' rules from 2nd order at the ends to 5th inside,
// with the reference's +w sign at the face ktf-1; at the top level only the
// face below, doubled).  Scatter form as advect_u's.
struct AdjAdvW {
  const float *lam, *ru, *rv, *rom, *msftx, *fzm, *fzp, *rdzu;
  float* gw;
  float rdx, rdy;
  int ids, ide, jds, jde, ktf;
  Box b, by, bx;
};

__device__ inline float w_level_flux_a(const AdjAdvW& a, int k, float rA, float rB) {
  if (k < a.ktf) return a.fzm[k] * rA + a.fzp[k] * rB;
  return (2.0f - a.fzm[a.ktf - 1]) * rA - a.fzp[a.ktf - 1] * rB;
}

// the horizontal face f of a row or column [lo, hi): d(face) / d(w at f + o), times w
template <class Put>
__device__ inline void w_hface_adj(int f, int lo, int hi, float vel, float w, Put put) {
  if (f >= lo + 3 && f < hi - 3) {
    const float s = sgn1(vel);
    for (int o = -3; o <= 2; ++o) put(o, w * vel * c5(o, s));
  } else if (f == lo + 1 || f == hi - 2) {
    put(0, w * 0.5f * vel);
    put(-1, w * 0.5f * vel);
  } else if (f == lo + 2 || f == hi - 3) {
    const float s = sgn1(vel);
    for (int o = -2; o <= 1; ++o) put(o, w * vel * c3(o, s));
  }
}

__device__ inline void w_zface_adj(const AdjAdvW& a, int kf, float rs, float w, float* acc) {
  if (kf < 1 || kf > a.ktf) return;
  auto put = [&](int o, float c) { acc[o] += w * c; };
  if (kf == 1 || kf == a.ktf) {
    put(0, 0.25f * rs);
    put(-1, 0.25f * rs);
    return;
  }
  const float vel = 0.5f * rs;
  if (kf >= 3 && kf < a.ktf - 1) {
    const float s = sgn1(-vel);
    for (int o = -3; o <= 2; ++o) put(o, vel * c5(o, s));
  } else {
    const float s = kf == 2 ? sgn1(-vel) : sgn1(vel);
    for (int o = -2; o <= 1; ++o) put(o, vel * c3(o, s));
  }
}

__global__ void k_advect_w_adj(Box all, Dims D, AdjAdvW a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  const float lam = AT(a.lam, k, j, i);
  if (lam == 0.0f) return;
  const int ktf = a.ktf, top = k >= ktf;
  const int kA = top ? ktf - 1 : k, kB = imax(top ? ktf - 2 : k - 1, 0);
  const float mx = AT2(a.msftx, j, i), rdz = a.rdzu[top ? ktf - 1 : k];
  // shares summed per target (offsets -3..3 from the point) before the atomics
  float ay[7] = {}, ax[7] = {}, az[7] = {};
  auto into = [](float* acc) { return [acc](int o, float c) { acc[o] += c; }; };
  if (a.by.has(k, j, i)) {
    const float wy = mx * a.rdy * lam;
    w_hface_adj(j + 1, a.jds, a.jde, w_level_flux_a(a, k, AT(a.rv, kA, j + 1, i), AT(a.rv, kB, j + 1, i)), -wy,
                into(ay + 4));
    w_hface_adj(j, a.jds, a.jde, w_level_flux_a(a, k, AT(a.rv, kA, j, i), AT(a.rv, kB, j, i)), wy, into(ay + 3));
  }
  if (a.bx.has(k, j, i)) {
    const float wx = mx * a.rdx * lam;
    w_hface_adj(i + 1, a.ids, a.ide, w_level_flux_a(a, k, AT(a.ru, kA, j, i + 1), AT(a.ru, kB, j, i + 1)), -wx,
                into(ax + 4));
    w_hface_adj(i, a.ids, a.ide, w_level_flux_a(a, k, AT(a.ru, kA, j, i), AT(a.ru, kB, j, i)), wx, into(ax + 3));
  }
  const float rm0 = AT(a.rom, k, j, i), rmb = AT(a.rom, imax(k - 1, 0), j, i);
  if (!top) {
    const float rmn = AT(a.rom, imin(k + 1, D.nz - 1), j, i);
    w_zface_adj(a, k + 1, rmn + rm0, -rdz * lam, az + 4);
    w_zface_adj(a, k, rm0 + rmb, rdz * lam, az + 3);
  } else {
    w_zface_adj(a, k, rm0 + rmb, 2.0f * rdz * lam, az + 3);
  }
  for (int m = 0; m < 7; ++m) {          // a nonzero share is one a face stencil read: in bounds
    atomic_add(&AT(a.gw, k, j + m - 3, i), ay[m]);
    atomic_add(&AT(a.gw, k, j, i + m - 3), ax[m]);
    atomic_add(&AT(a.gw, k + m - 3, j, i), az[m]);
  }
}

extern "C" int pywrf_advect_w_adjoint(const pywrf_grid_t* g, const float* lam_tendency, const float* ru,
                                      const float* rv, const float* rom, const float* msftx, const float* fzm,
                                      const float* fzp, double rdx, double rdy, const float* rdzu, float* lam_w,
                                      pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  const int ktf = imin(g->kte, g->kde - 1);
  if (g->kts != 0 || ktf + 1 > g->nz || ktf < 3) return PYWRF_E_RANGE;
  const int i0 = g->its, i1 = imin(g->ite, g->ide - 1);
  const int j0 = g->jts, j1 = imin(g->jte, g->jde - 1);
  AdjAdvW a;
  a.lam = lam_tendency; a.ru = ru; a.rv = rv; a.rom = rom; a.msftx = msftx; a.fzm = fzm; a.fzp = fzp;
  a.rdzu = rdzu; a.gw = lam_w; a.rdx = F(rdx); a.rdy = F(rdy);
  a.ids = g->ids; a.ide = g->ide; a.jds = g->jds; a.jde = g->jde; a.ktf = ktf;
  a.b = mkbox(g->kts + 1, ktf + 1, j0, j1, i0, i1);
  a.by = mkbox(g->kts + 1, ktf + 1, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), i0, i1);
  a.bx = mkbox(g->kts + 1, ktf + 1, j0, j1, imax(g->ids + 1, g->its), imin(g->ide - 2, g->ite));
  hipLaunchKernelGGL(k_zero_n, dim3(1024), dim3(256), 0, (hipStream_t)s, lam_w, (long)g->nz * g->ny * g->nx);
  PW_LAUNCH_BOX(k_advect_w_adj, a.b, s, dims_of(g), a);
  PW_RET();
}
