// Acoustic small-step machinery (pywrf/wrf_dynamics.py:3152-3930):
// small_step_prep, calc_p_rho, calc_coef_w, advance_uv, advance_mu_t,
// advance_w, sumflux, calc_mu_uv_1, small_step_finish.
//
// Where the reference clones an argument and updates a box of the clone, the
// kernel writes the whole output array (box points computed, the rest copied)
// in the same pass.  Column recurrences (eta velocity, Thomas solve) run one
// thread per (j, i) column with k marching; the output buffer doubles as the
// column scratch, so no full-grid temporaries are allocated.
#include "common.h"

#ifndef PW_COL_XCD
#define PW_COL_XCD 1
#endif
#include "consts.h"

#define AT(p, k, j, i) at_b((p), D.ob((k), (j), (i)))
#define AT2(p, j, i) at_b((p), D.ob2((j), (i)))

int pw_check_grid(const pywrf_grid_t* g, int r);
#define CHECK(g, r) do { int _e = pw_check_grid((g), (r)); if (_e) return _e; } while (0)

static Box plane(const pywrf_grid_t* g) { return mkbox(0, 1, 0, g->ny, 0, g->nx); }

// ------------------------------------------------------------------ small_step_prep
struct SsPrep2 {
  const float *mub, *muu, *muv;
  float *mu_1, *mu_2, *muus, *muvs, *muts, *mudf, *mu_save;
  int rk1, j0, j1, i0, i1, i1u, j1v;
};

__global__ void k_ssprep_2d(Box b, Dims D, SsPrep2 a) {
  int j, i;
  if (!box_column(b, j, i)) return;
  bool h = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  bool hu = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1u;
  long q = D.ix2(j, i);
  if (a.rk1) {
    bool hv = j >= a.j0 && j < a.j1v && i >= a.i0 && i < a.i1;
    float m2 = a.mu_2[q];
    if (h) {
      a.mu_1[q] = m2 + 0.0f;
      a.mudf[q] = 0.0f;
      a.muts[q] = a.mub[q] + m2;
      a.mu_save[q] = m2 + 0.0f;
      a.mu_2[q] = 0.0f;
    }
    if (hu) a.muus[q] = a.muu[q] + 0.0f;
    if (hv) a.muvs[q] = a.muv[q] + 0.0f;
  } else {
    bool hv = j >= a.j0 && j < a.j1v && i >= a.i0 && i < a.i1u;  // i-range i1u as in the reference
    if (h) a.muts[q] = a.mub[q] + a.mu_1[q];
    if (hu) a.muus[q] = 0.5f * (a.mub[q] + a.mu_1[q] + AT2(a.mub, j, i - 1) + AT2(a.mu_1, j, i - 1));
    if (hv) a.muvs[q] = 0.5f * (a.mub[q] + a.mu_1[q] + AT2(a.mub, j - 1, i) + AT2(a.mu_1, j - 1, i));
    if (h) {
      float m2 = a.mu_2[q];
      a.mu_save[q] = m2 + 0.0f;
      a.mu_2[q] = a.mu_1[q] - m2;
    }
  }
}

struct SsPrep3 {
  float *u_1, *v_1, *w_1, *t_1, *ph_1;
  const float *u_2, *v_2, *w_2, *t_2, *ph_2;
  float *u_save, *v_save, *w_save, *t_save, *ph_save, *ww_save, *c2a;
  const float *ww, *pb, *p, *alt, *muu, *muus, *muv, *muvs, *mut, *muts, *msfuy, *msfvx_inv, *msfty;
  float *u_out, *v_out, *w_out, *t_out, *ph_out;
  int rk1, k1, kde, j0, j1, i0, i1, i1u, j1v;
};

__global__ void k_ssprep_3d(Box all, Dims D, SsPrep3 a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  long q2 = D.ix2(j, i);
  bool hcol = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  bool bm = k < a.k1 && hcol;
  bool bu = k < a.k1 && j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1u;
  bool bv = k < a.k1 && j >= a.j0 && j < a.j1v && i >= a.i0 && i < a.i1;
  bool bw = k < a.kde && hcol;
  // every value is formed first and stored last, so that no load waits on a
  // store, and every operand (all at this point) is loaded before the
  // branches, so that no branch waits on its loads
  const float u2 = a.u_2[q], v2 = a.v_2[q], t2 = a.t_2[q], w2 = a.w_2[q], p2 = a.ph_2[q];
  // the *_1 fields are read only past the first stage (rk1 is uniform)
  const float u1l = a.rk1 ? 0.0f : a.u_1[q], v1l = a.rk1 ? 0.0f : a.v_1[q], t1l = a.rk1 ? 0.0f : a.t_1[q];
  const float w1l = a.rk1 ? 0.0f : a.w_1[q], p1l = a.rk1 ? 0.0f : a.ph_1[q];
  const float pbq = a.pb[q], pq = a.p[q], altq = a.alt[q], wwq = a.ww[q];
  const float muus = a.muus[q2], muu = a.muu[q2], msfuy = a.msfuy[q2], muvs = a.muvs[q2], muv = a.muv[q2];
  const float msfvx_inv = a.msfvx_inv[q2], muts = a.muts[q2], mut = a.mut[q2], msfty = a.msfty[q2];
  float uo = u2, u1 = 0.0f;
  if (bu) {
    u1 = a.rk1 ? u2 + 0.0f : u1l;
    uo = (muus * u1 - muu * u2) / msfuy;
  }
  float vo = v2, v1 = 0.0f;
  if (bv) {
    v1 = a.rk1 ? v2 + 0.0f : v1l;
    vo = (muvs * v1 - muv * v2) * msfvx_inv;
  }
  float to = t2, t1 = 0.0f, c2 = 0.0f;
  if (bm) {
    t1 = a.rk1 ? t2 + 0.0f : t1l;
    c2 = F(pw::cpovcv) * (pbq + pq) / altq;
    to = muts * t1 - mut * t2;
  }
  float wo = w2, po = p2, w1 = 0.0f, p1 = 0.0f, wws = 0.0f;
  const bool wwz = !bw && hcol && (k == 0 || k == a.kde - 1);
  if (bw) {
    w1 = a.rk1 ? w2 + 0.0f : w1l;
    p1 = a.rk1 ? p2 + 0.0f : p1l;
    wo = (muts * w1 - mut * w2) / msfty;
    po = p1 - p2;
    wws = wwq;
  }
  if (bu) {
    if (a.rk1) a.u_1[q] = u1;
    a.u_save[q] = u2 + 0.0f;
  }
  a.u_out[q] = uo;
  if (bv) {
    if (a.rk1) a.v_1[q] = v1;
    a.v_save[q] = v2 + 0.0f;
  }
  a.v_out[q] = vo;
  if (bm) {
    if (a.rk1) a.t_1[q] = t1;
    a.c2a[q] = c2;
    a.t_save[q] = t2 + 0.0f;
  }
  a.t_out[q] = to;
  if (bw) {
    if (a.rk1) {
      a.w_1[q] = w1;
      a.ph_1[q] = p1;
    }
    a.w_save[q] = w2 + 0.0f;
    a.ph_save[q] = p2 + 0.0f;
    a.ww_save[q] = wws;
  } else if (wwz) {
    a.ww_save[q] = 0.0f;
  }
  a.w_out[q] = wo;
  a.ph_out[q] = po;
}

extern "C" int pywrf_small_step_prep(const pywrf_grid_t* g, float* u_1, const float* u_2, float* v_1,
                                     const float* v_2, float* w_1, const float* w_2, float* t_1, const float* t_2,
                                     float* ph_1, const float* ph_2, const float* mub, float* mu_1, float* mu_2,
                                     const float* muu, float* muus, const float* muv, float* muvs, const float* mut,
                                     float* muts, float* mudf, float* u_save, float* v_save, float* w_save,
                                     float* t_save, float* ph_save, float* mu_save, const float* ww,
                                     float* ww_save, float* c2a, const float* pb, const float* p, const float* alt,
                                     const float* msfuy, const float* msfvx_inv, const float* msfty, int rk_step,
                                     float* u_out, float* v_out, float* w_out, float* t_out, float* ph_out,
                                     pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->kde > g->nz || g->kts != 0) return PYWRF_E_RANGE;
  int j0 = g->jts, j1 = imin(g->jte, g->jde - 1), i0 = g->its, i1 = imin(g->ite, g->ide - 1);
  int k1 = imin(g->kte, g->kde - 1), i1u = g->ite, j1v = g->jte;
  SsPrep2 a2 = {mub, muu, muv, mu_1, mu_2, muus, muvs, muts, mudf, mu_save, rk_step == 1, j0, j1, i0, i1, i1u, j1v};
  Dims D = dims_of(g);
  Box b2 = mkbox(0, 1, j0, imax(j1, j1v), i0, imax(i1, i1u));
  PW_LAUNCH_COLS(k_ssprep_2d, b2, s, D, a2);
  SsPrep3 a3 = {u_1, v_1, w_1, t_1, ph_1, u_2, v_2, w_2, t_2, ph_2, u_save, v_save, w_save, t_save, ph_save,
                ww_save, c2a, ww, pb, p, alt, muu, muus, muv, muvs, mut, muts, msfuy, msfvx_inv, msfty,
                u_out, v_out, w_out, t_out, ph_out, rk_step == 1, k1, g->kde, j0, j1, i0, i1, i1u, j1v};
  PW_LAUNCH_BOX(k_ssprep_3d, whole(g), s, D, a3);
  PW_RET();
}

// ------------------------------------------------------------------ calc_p_rho
struct PRho {
  const float *al_in, *p_in, *ph, *alt, *t_2, *t_1, *c2a, *mu, *mut, *rdnw;
  float *al, *p, *pm1;
  float t0, smdiv;
  int step;
};

__global__ void k_calc_p_rho(Box all, Dims D, Box b, PRho a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  // the box's operands are loaded before the branch (a load under a branch is
  // waited for before the branch ends); the level above is clamped into the
  // array, its value being used only inside the box, where k + 1 is a level.
  // The incoming al and p are read only outside the box (a wave inside it
  // skips them as one), pm1 only past step 0 (uniform)
  const bool in = b.has(k, j, i);
  const float al_in = in ? 0.0f : a.al_in[q], p_in = in ? 0.0f : a.p_in[q];
  const float ph0 = a.ph[q], ph1 = AT(a.ph, imin(k + 1, D.nz - 1), j, i);
  const float altq = a.alt[q], t1 = a.t_1[q], t2 = a.t_2[q], c2 = a.c2a[q], pm = a.step != 0 ? a.pm1[q] : 0.0f;
  const float mut = AT2(a.mut, j, i), mu = AT2(a.mu, j, i), rd = a.rdnw[k];
  float al = al_in, p = p_in, pmo = 0.0f;
  if (in) {
    al = -1.0f / mut * (altq * mu + rd * (ph1 - ph0));
    p = c2 * (altq * (t2 - mu * t1) / (mut * (a.t0 + t1)) - al);
    if (a.step != 0) {
      pmo = p + 0.0f;
      p = p + a.smdiv * (p - pm);
    }
  }
  a.al[q] = al;
  a.p[q] = p;
  if (a.step == 0) a.pm1[q] = p + 0.0f;
  else if (in) a.pm1[q] = pmo;
}

extern "C" int pywrf_calc_p_rho(const pywrf_grid_t* g, const float* al_in, const float* p_in, const float* ph,
                                const float* alt, const float* t_2, const float* t_1, const float* c2a, float* pm1,
                                const float* mu, const float* mut, double t0, const float* rdnw, double smdiv,
                                int step, float* al_out, float* p_out, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  if (b.k1 + 1 > g->nz) return PYWRF_E_RANGE;
  PRho a = {al_in, p_in, ph, alt, t_2, t_1, c2a, mu, mut, rdnw, al_out, p_out, pm1, F(t0), F(smdiv), step};
  PW_LAUNCH_BOX(k_calc_p_rho, whole(g), s, dims_of(g), b, a);
  PW_RET();
}

// ------------------------------------------------------------------ calc_coef_w
// Column block (a 512-thread workgroup owns 64 i-columns of one row j): the
// tridiagonal coefficients of every level are formed level-parallel (a written
// out, the diagonal and upper coefficients staged in LDS); one wave then runs
// the forward elimination alpha/gamma per column from LDS.
#define CW_TI 64
#define CW_TB 512
#define CW_TG (CW_TB / CW_TI)
struct CoefW {
  float *a, *alpha, *gamma;
  const float *mut, *cqw, *rdn, *rdnw, *c2a;
  float C, m2C, p2C, lid;
  int kde;
};

__global__ void __launch_bounds__(CW_TB) k_calc_coef_w(Box b, Dims D, CoefW p) {
  extern __shared__ float sm[];
  float* s_a = sm;                  // [nz][CW_TI] sub-diagonal
  float* s_b = sm + D.nz * CW_TI;   // diagonal
  float* s_c = sm + 2 * D.nz * CW_TI;   // super-diagonal
  const int lane = threadIdx.x % CW_TI, grp = threadIdx.x / CW_TI;
  int bx, by;
  xcd_block_2d(bx, by);   // rows of one XCD contiguous (neighbouring rows share an L2)
  const int j = b.j0 + by, i = b.i0 + bx * CW_TI + lane;
  const bool ok = i < b.i1;
  const int kde = p.kde;
  float m = 0.0f;
  if (ok) {
    m = AT2(p.mut, j, i);
    for (int k = 1 + grp; k < kde - 1; k += CW_TG) {
      float cq = AT(p.cqw, k, j, i), c2 = AT(p.c2a, k, j, i), c2m = AT(p.c2a, k - 1, j, i);
      float ak = 0.0f;
      if (k >= 2) {
        ak = -cq * p.C * p.rdn[k] * p.rdnw[k - 1] * c2m / (m * m);
        AT(p.a, k, j, i) = ak;
      }
      s_a[k * CW_TI + lane] = ak;
      s_b[k * CW_TI + lane] = 1.0f + cq * p.C * p.rdn[k] * (p.rdnw[k] * c2 / (m * m) + p.rdnw[k - 1] * c2m / (m * m));
      s_c[k * CW_TI + lane] = -cq * p.C * p.rdn[k] * p.rdnw[k] * c2 / (m * m);
    }
  }
  __syncthreads();
  if (!ok || grp != 0) return;
  float r2 = p.rdnw[kde - 2] * p.rdnw[kde - 2];
  AT(p.a, 1, j, i) = 0.0f;
  float atop = p.m2C * r2 * AT(p.c2a, kde - 2, j, i) * p.lid / (m * m);
  AT(p.a, kde - 1, j, i) = atop;
  AT(p.gamma, 0, j, i) = 0.0f;
  float gm = 0.0f;
  for (int k = 1; k < kde - 1; ++k) {
    float al = 1.0f / (s_b[k * CW_TI + lane] - s_a[k * CW_TI + lane] * gm);
    gm = s_c[k * CW_TI + lane] * al;
    AT(p.alpha, k, j, i) = al;
    AT(p.gamma, k, j, i) = gm;
  }
  float bb = 1.0f + p.p2C * r2 * AT(p.c2a, kde - 2, j, i) / (m * m);
  AT(p.alpha, kde - 1, j, i) = 1.0f / (bb - atop * gm);
  AT(p.gamma, kde - 1, j, i) = 0.0f;
}

extern "C" int pywrf_calc_coef_w(const pywrf_grid_t* g, float* a, float* alpha, float* gamma, const float* mut,
                                 const float* cqw, const float* rdn, const float* rdnw, const float* c2a,
                                 double dts, double g_, double epssm, int top_lid, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kde > g->nz || g->kde < 3) return PYWRF_E_RANGE;
  double C = (0.5 * dts * g_ * (1. + epssm));
  C = C * C;
  Box b = mkbox(0, 1, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  const size_t lds = 3 * (size_t)g->nz * CW_TI * sizeof(float);
  if (lds > 64 * 1024) return PYWRF_E_RANGE;
  if (b.nj() <= 0 || b.ni() <= 0) PW_RET();
  CoefW p = {a, alpha, gamma, mut, cqw, rdn, rdnw, c2a, F(C), F(-2. * C), F(2. * C), top_lid ? 0.0f : 1.0f, g->kde};
  hipLaunchKernelGGL(k_calc_coef_w, dim3((unsigned)((b.ni() + CW_TI - 1) / CW_TI), (unsigned)b.nj()), dim3(CW_TB),
                     lds, (hipStream_t)s, b, dims_of(g), p);
  PW_RET();
}

// ------------------------------------------------------------------ advance_uv
struct AdvUV {
  const float *u_in, *ru_tend, *v_in, *rv_tend, *p, *pb, *ph, *php, *alt, *al, *mu, *muu, *cqu, *muv, *cqv, *mudf;
  const float *msfux, *msfuy, *msfvx, *msfvx_inv, *msfvy, *cf, *fnm, *fnp, *rdnw;
  float *u, *v;
  float rdx, rdy, dts, dfx, dfy;
  int kde;
};

__device__ inline float uv_dpn(const Dims& D, const AdvUV& a, int k, int j, int i, int dj, int di) {
  auto P = [&](int kk, int jj, int ii) { return AT(a.p, kk, jj, ii); };
  if (k == 0)
    return 0.5f * (a.cf[0] * (P(0, j, i) + P(0, j + dj, i + di)) + a.cf[1] * (P(1, j, i) + P(1, j + dj, i + di)) +
                   a.cf[2] * (P(2, j, i) + P(2, j + dj, i + di)));
  if (k == a.kde - 1) return 0.0f;
  return 0.5f * (a.fnm[k] * (P(k, j, i) + P(k, j + dj, i + di)) + a.fnp[k] * (P(k - 1, j, i) + P(k - 1, j + dj, i + di)));
}

__device__ inline float uv_dpxy(const Dims& D, const AdvUV& a, int k, int j, int i, int dj, int di, const float* mv,
                                const float* mx, const float* my, float rd) {
  int jd = j + dj, id = i + di;
  float r = AT2(mx, j, i) / AT2(my, j, i);
  float d = r * 0.5f * rd * AT2(mv, j, i) *
            (((AT(a.ph, k + 1, j, i) - AT(a.ph, k + 1, jd, id)) + (AT(a.ph, k, j, i) - AT(a.ph, k, jd, id))) +
             (AT(a.alt, k, j, i) + AT(a.alt, k, jd, id)) * (AT(a.p, k, j, i) - AT(a.p, k, jd, id)) +
             (AT(a.al, k, j, i) + AT(a.al, k, jd, id)) * (AT(a.pb, k, j, i) - AT(a.pb, k, jd, id)));
  return d + r * rd * (AT(a.php, k, j, i) - AT(a.php, k, jd, id)) *
                 (a.rdnw[k] * (uv_dpn(D, a, k + 1, j, i, dj, di) - uv_dpn(D, a, k, j, i, dj, di)) -
                  0.5f * (AT2(a.mu, jd, id) + AT2(a.mu, j, i)));
}

__global__ void k_advance_uv(Box all, Dims D, AdvUV a, Box bu, Box bv) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  float u = a.u_in[q];
  if (bu.has(k, j, i)) {
    u = u + a.dts * a.ru_tend[q];
    float mxy = a.dfx * (AT2(a.mudf, j, i) - AT2(a.mudf, j, i - 1)) / AT2(a.msfuy, j, i);
    float d = uv_dpxy(D, a, k, j, i, 0, -1, a.muu, a.msfux, a.msfuy, a.rdx);
    u = u - a.dts * a.cqu[q] * d + mxy;
  }
  float v = a.v_in[q];
  if (bv.has(k, j, i)) {
    v = v + a.dts * a.rv_tend[q];
    float mxy = a.dfy * (AT2(a.mudf, j, i) - AT2(a.mudf, j - 1, i)) * AT2(a.msfvx_inv, j, i);
    float d = uv_dpxy(D, a, k, j, i, -1, 0, a.muv, a.msfvy, a.msfvx, a.rdy);
    v = v - a.dts * a.cqv[q] * d + mxy;
  }
  a.u[q] = u;   // stores last: no load waits on them
  a.v[q] = v;
}

extern "C" int pywrf_advance_uv(const pywrf_grid_t* g, const float* u_in, const float* ru_tend, const float* v_in,
                                const float* rv_tend, const float* p, const float* pb, const float* ph,
                                const float* php, const float* alt, const float* al, const float* mu,
                                const float* muu, const float* cqu, const float* muv, const float* cqv,
                                const float* mudf, const float* msfux, const float* msfuy, const float* msfvx,
                                const float* msfvx_inv, const float* msfvy, double rdx, double rdy, double dts,
                                const float* cf, const float* fnm, const float* fnp, double emdiv,
                                const float* rdnw, int spec_zone, float* u_out, float* v_out, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  int k1 = imin(g->kte, g->kde - 1);
  if (k1 + 1 > g->nz || g->kde > g->nz) return PYWRF_E_RANGE;
  int i0 = imax(g->its, g->ids + spec_zone), i1 = imin(g->ite, g->ide - spec_zone - 1);
  int j0 = imax(g->jts, g->jds + spec_zone), j1 = imin(g->jte, g->jde - spec_zone - 1);
  int i1u = imin(g->ite, g->ide - spec_zone), j1v = imin(g->jte, g->jde - spec_zone);
  double dx = 1. / rdx, dy = 1. / rdy;
  AdvUV a = {u_in, ru_tend, v_in, rv_tend, p, pb, ph, php, alt, al, mu, muu, cqu, muv, cqv, mudf, msfux, msfuy,
             msfvx, msfvx_inv, msfvy, cf, fnm, fnp, rdnw, u_out, v_out, F(rdx), F(rdy), F(dts), F(-emdiv * dx),
             F(-emdiv * dy), g->kde};
  Box bu = mkbox(g->kts, k1, j0, j1, i0, i1u);
  Box bv = mkbox(g->kts, k1, j0, j1v, i0, i1);
  PW_LAUNCH_BOX(k_advance_uv, whole(g), s, dims_of(g), a, bu, bv);
  PW_RET();
}

// ------------------------------------------------------------------ advance_mu_t
struct AdvMuT {
  const float *ww_in, *ww_1, *u, *u_1, *v, *v_1, *mu_in, *mut, *muave_in, *muts_in, *muu, *muv, *mudf_in;
  const float *t_in, *t_1, *t_ave_in, *ft, *mu_tend, *dnw, *fnm, *fnp, *rdnw;
  const float *msfuy, *msfvx_inv, *msftx, *msfty;
  float *ww, *t, *t_ave, *muave, *muts, *mudf, *mu;
  float rdx, rdy, dts, ep1, em1, hrdx, hrdy;
  int k1, kde, j0, j1, i0, i1;
};

// Operands are loaded into registers before the arithmetic so that the loads
// of one level are in flight together.
__device__ inline float mut_dvdxi(const Dims& D, const AdvMuT& a, int k, int j, int i) {
  const float vn = AT(a.v, k, j + 1, i), mvn = AT2(a.muv, j + 1, i), v1n = AT(a.v_1, k, j + 1, i);
  const float xn = AT2(a.msfvx_inv, j + 1, i);
  const float vs = AT(a.v, k, j, i), mvs = AT2(a.muv, j, i), v1s = AT(a.v_1, k, j, i), xs = AT2(a.msfvx_inv, j, i);
  const float ue = AT(a.u, k, j, i + 1), mue = AT2(a.muu, j, i + 1), u1e = AT(a.u_1, k, j, i + 1);
  const float ye = AT2(a.msfuy, j, i + 1);
  const float uw = AT(a.u, k, j, i), muw = AT2(a.muu, j, i), u1w = AT(a.u_1, k, j, i), yw = AT2(a.msfuy, j, i);
  const float mx = AT2(a.msftx, j, i), my = AT2(a.msfty, j, i);
  float vy = (vn + mvn * v1n * xn) - (vs + mvs * v1s * xs);
  float ux = (ue + mue * u1e / ye) - (uw + muw * u1w / yw);
  return mx * my * (a.rdy * vy + a.rdx * ux);
}

// Column block: a 512-thread workgroup owns TI consecutive i-columns of one
// row j and all levels.  The divergence dvdxi of every (k, i) is computed by
// all threads and staged in LDS; one thread per column then forms the column
// sum, the column-mass update and the eta-velocity recurrence from LDS; all
// threads finally apply the theta update level by level (coalesced in i).
// Columns outside the update box copy their inputs (clone semantics).
#define TI 64
#define TB 512
#define TG (TB / TI)

#define CB_MAXNZ 128   // column blocks hold at most 64 KB of per-level LDS, so nz <= 128

__global__ void __launch_bounds__(TB) k_advance_mu_t(Dims D, AdvMuT a) {
  extern __shared__ float sm[];
  __shared__ float s_dnw[CB_MAXNZ], s_fnm[CB_MAXNZ], s_fnp[CB_MAXNZ], s_rdnw[CB_MAXNZ];
  __shared__ float s_part[(CB_MAXNZ / 16 + 1) * TI];
  const int nz = D.nz;
  for (int t = threadIdx.x; t < nz; t += TB) {
    s_dnw[t] = a.dnw[t]; s_fnm[t] = a.fnm[t]; s_fnp[t] = a.fnp[t]; s_rdnw[t] = a.rdnw[t];
  }
  float* s_dv = sm;             // [nz][TI] dvdxi
  float* s_ww = sm + nz * TI;   // [nz][TI] eta velocity before subtracting ww_1
  const int lane = threadIdx.x % TI, grp = threadIdx.x / TI;
  int bx, by;
#if PW_COL_XCD
  xcd_block_2d(bx, by);
#else
  bx = (int)blockIdx.x; by = (int)blockIdx.y;
#endif
  const int j = by, i = bx * TI + lane;
  const bool vi = i < D.nx;
  const bool h = vi && j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  const long q2 = vi ? D.ix2(j, i) : 0;
  for (int k = grp; k < nz; k += TG) {
    if (!vi) continue;
    long q = D.ix(k, j, i);
    if (!h) {
      a.ww[q] = a.ww_in[q];
      a.t[q] = a.t_in[q];
      a.t_ave[q] = a.t_ave_in[q];
      continue;
    }
    s_dv[k * TI + lane] = k < a.k1 ? mut_dvdxi(D, a, k, j, i) : 0.0f;
  }
  if (vi && !h && grp == 0) {
    a.muave[q2] = a.muave_in[q2];
    a.muts[q2] = a.muts_in[q2];
    a.mudf[q2] = a.mudf_in[q2];
    a.mu[q2] = a.mu_in[q2];
  }
  __syncthreads();
  // The column sum dmdt in torch's order (torch_outer_sum with nz <= 128:
  // blocks of 16 levels summed from +0 one after the other, the tail from +0,
  // tail + (sum of the block sums) + 0 + 0): each block's chain and the
  // tail's are independent, so the groups form them side by side.
  const int nb = nz >> 4;
  for (int c = grp; c <= nb; c += TG) {
    float acc = 0.0f;
    if (h) {
      const int k0 = c << 4, k1 = imin(k0 + 16, nz);
      for (int k = k0; k < k1; ++k) acc += k < a.k1 ? s_dnw[k] * s_dv[k * TI + lane] : 0.0f;
    }
    s_part[c * TI + lane] = acc;
  }
  __syncthreads();
  if (grp == 0 && h) {
    float blocks = 0.0f;
    for (int c = 0; c < nb; ++c) blocks += s_part[c * TI + lane];
    float dmdt = s_part[nb * TI + lane] + blocks;
    dmdt = dmdt + 0.0f;
    dmdt = dmdt + 0.0f;
    float mt = a.mu_tend[q2];
    float mu_old = a.mu_in[q2];
    float mu = mu_old + a.dts * (dmdt + mt);
    a.mu[q2] = mu;
    a.mudf[q2] = dmdt + mt;
    a.muts[q2] = a.mut[q2] + mu;
    a.muave[q2] = 0.5f * (a.ep1 * mu + a.em1 * (mu_old + 0.0f));
    s_part[lane] = dmdt;
  }
  __syncthreads();
  // the eta-velocity decrements dnw(k-1) (dmdt + dvdxi(k-1) + mu_tend) / msfty
  // of every level at once (in place of dvdxi(k-1), which nothing reads again)
  if (h) {
    const float dmdt = s_part[lane], mt = a.mu_tend[q2], my = a.msfty[q2];
    for (int k = 1 + grp; k < a.k1; k += TG)
      s_dv[(k - 1) * TI + lane] = s_dnw[k - 1] * (dmdt + s_dv[(k - 1) * TI + lane] + mt) / my;
  }
  __syncthreads();
  if (grp == 0 && h) {
    // the decrements 8 levels at a time into registers first (s_dv and s_ww
    // share one LDS buffer, so a read after a store to s_ww would wait for it)
    float wk = a.ww_in[D.ix(0, j, i)];
    s_ww[lane] = wk;
    int k = 1;
    for (; k + 7 < a.k1; k += 8) {
      float dv[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) dv[m] = s_dv[(k - 1 + m) * TI + lane];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        wk = wk - dv[m];
        s_ww[(k + m) * TI + lane] = wk;
      }
    }
    for (; k < a.k1; ++k) {
      wk = wk - s_dv[(k - 1) * TI + lane];
      s_ww[k * TI + lane] = wk;
    }
  }
  __syncthreads();
  if (!h) return;
  const float my = a.msfty[q2], mx = a.msftx[q2];
  // theta and eta velocity of level k < k1 from loads only (indices clamped;
  // the values taken at clamped indices are not used)
  auto level = [&](int k, float& o_ww, float& o_t) {
    const long q = D.ix(k, j, i);
    const int ku = imin(k + 1, nz - 1), kd = imax(k - 1, 0);
    const float tin = a.t_in[q], ftq = a.ft[q], t1 = a.t_1[q], w1k = a.ww_1[q];
    const float vn = AT(a.v, k, j + 1, i), vs = AT(a.v, k, j, i), ue = AT(a.u, k, j, i + 1), uw = AT(a.u, k, j, i);
    const float tn = AT(a.t_1, k, j + 1, i), ts = AT(a.t_1, k, j - 1, i);
    const float te = AT(a.t_1, k, j, i + 1), tw = AT(a.t_1, k, j, i - 1);
    const bool up = k + 1 < a.k1, dn = k > 0;   // where the eta fluxes wdtn(k + 1), wdtn(k) are nonzero
    const float w1u = AT(a.ww_1, ku, j, i), t1u = AT(a.t_1, ku, j, i), t1d = AT(a.t_1, kd, j, i);
    const float wwk = s_ww[k * TI + lane] - w1k;
    o_ww = wwk;
    float t = tin + my * a.dts * ftq;
    const float wdu = up ? (s_ww[ku * TI + lane] - w1u) * (s_fnm[ku] * t1u + s_fnp[ku] * t1) : 0.0f;
    const float wdk = dn ? wwk * (s_fnm[k] * t1 + s_fnp[k] * t1d) : 0.0f;
    float yterm = a.hrdy * (vn * (tn + t1) - vs * (t1 + ts));
    float xterm = a.hrdx * (ue * (te + t1) - uw * (t1 + tw));
    o_t = t - a.dts * my * (mx * (yterm + xterm) + s_rdnw[k] * (wdu - wdk));
  };
  // two of the thread's levels at a time, every load before the stores (the
  // second level clamped to the last one below k1 and stored only if real)
  for (int k = grp; k < a.k1; k += 2 * TG) {
    const int k2 = imin(k + TG, a.k1 - 1);
    float w0, t0, w1, t1;
    level(k, w0, t0);
    level(k2, w1, t1);
    const long q = D.ix(k, j, i), qb = D.ix(k2, j, i);
    const float ta0 = a.t_in[q] + 0.0f, ta1 = a.t_in[qb] + 0.0f;
    a.ww[q] = w0;
    a.t_ave[q] = ta0;
    a.t[q] = t0;
    if (k + TG < a.k1) {
      a.ww[qb] = w1;
      a.t_ave[qb] = ta1;
      a.t[qb] = t1;
    }
  }
  for (int k = a.k1 + grp; k < nz; k += TG) {   // levels at and above k1 copy their inputs
    const long q = D.ix(k, j, i);
    a.ww[q] = a.ww_in[q];
    a.t[q] = a.t_in[q];
    a.t_ave[q] = a.t_ave_in[q];
  }
}

static int column_block_launch_check(const pywrf_grid_t* g, int narrays) {
  size_t bytes = (size_t)narrays * g->nz * TI * sizeof(float);
  return bytes <= 64 * 1024 ? 0 : PYWRF_E_RANGE;
}

extern "C" int pywrf_advance_mu_t(const pywrf_grid_t* g, const float* ww_in, const float* ww_1, const float* u,
                                  const float* u_1, const float* v, const float* v_1, const float* mu_in,
                                  const float* mut, const float* muave_in, const float* muts_in, const float* muu,
                                  const float* muv, const float* mudf_in, const float* t_in, const float* t_1,
                                  const float* t_ave_in, const float* ft, const float* mu_tend, double rdx,
                                  double rdy, double dts, double epssm, const float* dnw, const float* fnm,
                                  const float* fnp, const float* rdnw, const float* msfuy, const float* msfvx_inv,
                                  const float* msftx, const float* msfty, float* ww_out, float* t_out,
                                  float* t_ave_out, float* muave_out, float* muts_out, float* mudf_out,
                                  float* mu_out, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->kte > g->nz || g->kte < 2) return PYWRF_E_RANGE;
  AdvMuT a = {ww_in, ww_1, u, u_1, v, v_1, mu_in, mut, muave_in, muts_in, muu, muv, mudf_in, t_in, t_1, t_ave_in,
              ft, mu_tend, dnw, fnm, fnp, rdnw, msfuy, msfvx_inv, msftx, msfty, ww_out, t_out, t_ave_out,
              muave_out, muts_out, mudf_out, mu_out, F(rdx), F(rdy), F(dts), F(1. + epssm), F(1. - epssm),
              F(0.5 * rdx), F(0.5 * rdy), g->kte - 1, g->kde, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2),
              imax(g->its, g->ids + 1), imin(g->ite, g->ide - 2)};
  if (column_block_launch_check(g, 2)) return PYWRF_E_RANGE;
  dim3 grid((g->nx + TI - 1) / TI, g->ny);
  hipLaunchKernelGGL(k_advance_mu_t, grid, dim3(TB), 2 * g->nz * TI * sizeof(float), (hipStream_t)s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ advance_w
struct AdvW {
  const float *w_in, *rw_tend, *ww, *w_save, *u, *v, *mut, *muave, *muts, *t_2ave_in, *t_2, *t_1, *ph_in, *ph_1;
  const float *phb, *ph_tend, *ht, *c2a, *cqw, *alt, *a, *alpha, *gamma, *fnm, *fnp, *rdnw, *rdn, *cf;
  const float *msftx, *msfty;
  float *t_2ave, *w, *ph;
  float dts, t0, ep1, em1, hgem1, rdx, rdy, hdtsg, dtsg, mhdtsg, dampmag, hpi, hdepth;
  int k1, j0, j1, i0, i1;
};

// Column block as in advance_mu_t.  Phase 1 (all threads): t_2ave and the
// phi right-hand side rhs for every (k, i), staged in LDS with the
// tridiagonal factors a, alpha, gamma.  Phase 2 (all threads): surface w and
// the explicit w predictor from the LDS rhs / t_2ave.  Phase 3 (one thread
// per column): Thomas forward sweep and back substitution in LDS.  Phase 4
// (all threads): Rayleigh damping and the new phi, written back coalesced.
#ifndef AW_UNROLL
#define AW_UNROLL 2
#endif
__global__ void __launch_bounds__(TB) k_advance_w(Dims D, AdvW a) {
  extern __shared__ float sm[];
  __shared__ float s_fnm[CB_MAXNZ], s_fnp[CB_MAXNZ], s_rdnw[CB_MAXNZ], s_rdn[CB_MAXNZ];
  const int nz = D.nz, k1 = a.k1;
  for (int t = threadIdx.x; t < nz; t += TB) {
    s_fnm[t] = a.fnm[t]; s_fnp[t] = a.fnp[t]; s_rdnw[t] = a.rdnw[t]; s_rdn[t] = a.rdn[t];
  }
  float* s_t2a = sm;
  float* s_rhs = sm + 1 * nz * TI;
  float* s_w = sm + 2 * nz * TI;
  float* s_a = sm + 3 * nz * TI;
  float* s_al = sm + 4 * nz * TI;
  float* s_ga = sm + 5 * nz * TI;
  const int lane = threadIdx.x % TI, grp = threadIdx.x / TI;
  int bx, by;
#if PW_COL_XCD
  xcd_block_2d(bx, by);
#else
  bx = (int)blockIdx.x; by = (int)blockIdx.y;
#endif
  const int j = by, i = bx * TI + lane;
  const bool vi = i < D.nx;
  const bool h = vi && j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  const long q2 = vi ? D.ix2(j, i) : 0;
  const float G = F(pw::g);
  float mut = 0.0f, muts = 0.0f, muave = 0.0f, my = 1.0f, mx = 1.0f;
  if (h) {
    mut = a.mut[q2]; muts = a.muts[q2]; muave = a.muave[q2]; my = a.msfty[q2]; mx = a.msftx[q2];
  }
  const float mi = 1.0f / my;
  __syncthreads();   // level arrays staged
  // phase 1 (operands loaded before the arithmetic, so one level's loads are in
  // flight together; no global store in the loop, so the levels can overlap)
#pragma unroll AW_UNROLL
  for (int k = grp; k < nz; k += TG) {
    if (!vi) continue;
    long q = D.ix(k, j, i);
    if (!h) {
      a.t_2ave[q] = a.t_2ave_in[q];
      a.w[q] = a.w_in[q];
      a.ph[q] = a.ph_in[q];
      continue;
    }
    const bool lo = k < k1, rk = k >= 1 && k <= k1, wd = rk && k < k1;
    const float t2i = a.t_2ave_in[q], t2 = lo ? a.t_2[q] : 0.0f, t1 = lo ? a.t_1[q] : 0.0f;
    const float pht = rk ? a.ph_tend[q] : 0.0f, win = rk ? a.w_in[q] : 0.0f, phi = rk ? a.ph_in[q] : 0.0f;
    const float ca = a.a[q], cal = a.alpha[q], cga = a.gamma[q];
    float wwp = 0.0f, ww0 = 0.0f, wwm = 0.0f, p1p = 0.0f, p10 = 0.0f, p1m = 0.0f, pbp = 0.0f, pb0 = 0.0f, pbm = 0.0f;
    if (wd) {
      wwp = AT(a.ww, k + 1, j, i); ww0 = AT(a.ww, k, j, i); wwm = AT(a.ww, k - 1, j, i);
      p1p = AT(a.ph_1, k + 1, j, i); p10 = AT(a.ph_1, k, j, i); p1m = AT(a.ph_1, k - 1, j, i);
      pbp = AT(a.phb, k + 1, j, i); pb0 = AT(a.phb, k, j, i); pbm = AT(a.phb, k - 1, j, i);
    }
    float t = t2i;
    if (lo) {
      t = 0.5f * (a.ep1 * t2 + a.em1 * t);
      t = (t + muave * a.t0) / (muts * (a.t0 + t1));
    }
    s_t2a[k * TI + lane] = t;
    float r = 0.0f;
    if (rk) {
      r = a.dts * (pht + a.hgem1 * win);
      if (wd) {
        // wdwn(kk) = 0.5 (ww[kk] + ww[kk-1]) rdnw[kk-1] (ph_1[kk] - ph_1[kk-1] + phb[kk] - phb[kk-1])
        const float wup = 0.5f * (wwp + ww0) * s_rdnw[k] * (p1p - p10 + pbp - pb0);
        const float wdn = 0.5f * (ww0 + wwm) * s_rdnw[k - 1] * (p10 - p1m + pb0 - pbm);
        r = r - a.dts * (s_fnm[k] * wup + s_fnp[k] * wdn);
      }
      r = phi + my * r / mut;
    }
    s_rhs[k * TI + lane] = r;
    s_a[k * TI + lane] = ca;
    s_al[k * TI + lane] = cal;
    s_ga[k * TI + lane] = cga;
  }
  __syncthreads();
  auto RHS = [&](int k) -> float { return s_rhs[k * TI + lane]; };   // rhs[0] is 0
  // phase 2
#pragma unroll AW_UNROLL
  for (int k = grp; k < nz; k += TG) {
    if (!h) continue;
    long q = D.ix(k, j, i);
    float w;
    if (k == 0) {
      float cf1 = a.cf[0], cf2 = a.cf[1], cf3 = a.cf[2];
      auto V3 = [&](int jj, int ii) {
        return cf1 * AT(a.v, 0, jj, ii) + cf2 * AT(a.v, 1, jj, ii) + cf3 * AT(a.v, 2, jj, ii);
      };
      auto U3 = [&](int jj, int ii) {
        return cf1 * AT(a.u, 0, jj, ii) + cf2 * AT(a.u, 1, jj, ii) + cf3 * AT(a.u, 2, jj, ii);
      };
      float htc = AT2(a.ht, j, i);
      float ty = (AT2(a.ht, j + 1, i) - htc) * V3(j + 1, i) + (htc - AT2(a.ht, j - 1, i)) * V3(j, i);
      float tx = (AT2(a.ht, j, i + 1) - htc) * U3(j, i + 1) + (htc - AT2(a.ht, j, i - 1)) * U3(j, i);
      w = my * 0.5f * a.rdy * ty + mx * 0.5f * a.rdx * tx;
    } else if (k < k1) {
      const float c2 = a.c2a[q], c2m = AT(a.c2a, k - 1, j, i), al0 = a.alt[q], alm = AT(a.alt, k - 1, j, i);
      const float win = a.w_in[q], rwt = a.rw_tend[q], cq = a.cqw[q];
      const float php = AT(a.ph_in, k + 1, j, i), ph0 = a.ph_in[q], phm = AT(a.ph_in, k - 1, j, i);
      float X = a.hdtsg * s_rdn[k] *
                (c2 * s_rdnw[k] / mut * (a.ep1 * (RHS(k + 1) - RHS(k)) + a.em1 * (php - ph0)) -
                 c2m * s_rdnw[k - 1] / mut * (a.ep1 * (RHS(k) - RHS(k - 1)) + a.em1 * (ph0 - phm)));
      float Y = s_rdn[k] * (c2 * al0 * s_t2a[k * TI + lane] - c2m * alm * s_t2a[(k - 1) * TI + lane]) - muave;
      w = win + a.dts * rwt + mi * cq * X + a.dtsg * mi * Y;
    } else if (k == k1) {
      const float c2m = AT(a.c2a, k1 - 1, j, i), alm = AT(a.alt, k1 - 1, j, i);
      const float win = a.w_in[q], rwt = a.rw_tend[q];
      const float ph0 = a.ph_in[q], phm = AT(a.ph_in, k1 - 1, j, i);
      float rn = s_rdnw[k1 - 1];
      float Z = a.mhdtsg / mut * (rn * rn) * 2.0f * c2m * (a.ep1 * (RHS(k1) - RHS(k1 - 1)) + a.em1 * (ph0 - phm)) -
                a.dtsg * (2.0f * rn * c2m * alm * s_t2a[(k1 - 1) * TI + lane] + muave);
      w = win + a.dts * rwt + mi * Z;
    } else {
      w = a.w_in[q];
    }
    s_w[k * TI + lane] = w;
  }
  __syncthreads();
  // phase 3
  // The sweeps take their operands 8 levels at a time into registers before
  // the dependent arithmetic: s_w, s_a, s_al, s_ga are runtime offsets into
  // one LDS buffer, so a read placed after a store to s_w would be made to
  // wait for it.
  if (grp == 0 && h) {
    float wm = s_w[lane];
    int k = 1;
    for (; k + 7 <= k1; k += 8) {
      float wv[8], av[8], lv[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        wv[m] = s_w[(k + m) * TI + lane]; av[m] = s_a[(k + m) * TI + lane]; lv[m] = s_al[(k + m) * TI + lane];
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        wm = (wv[m] - av[m] * wm) * lv[m];
        s_w[(k + m) * TI + lane] = wm;
      }
    }
    for (; k <= k1; ++k) {
      wm = (s_w[k * TI + lane] - s_a[k * TI + lane] * wm) * s_al[k * TI + lane];
      s_w[k * TI + lane] = wm;
    }
    float wp = wm;
    k = k1 - 1;
    for (; k - 7 >= 1; k -= 8) {
      float wv[8], gv[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) { wv[m] = s_w[(k - m) * TI + lane]; gv[m] = s_ga[(k - m) * TI + lane]; }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        wp = wv[m] - gv[m] * wp;
        s_w[(k - m) * TI + lane] = wp;
      }
    }
    for (; k >= 1; --k) {
      wp = s_w[k * TI + lane] - s_ga[k * TI + lane] * wp;
      s_w[k * TI + lane] = wp;
    }
  }
  __syncthreads();
  if (!h) return;
  // phase 4: the new w and phi of this thread's levels into LDS (s_w, s_rhs),
  // then written out with t_2ave in a loop of stores only
  float htop = (AT(a.ph_1, k1, j, i) + AT(a.phb, k1, j, i)) / G;
  float hbot = htop - a.hdepth;
#pragma unroll AW_UNROLL
  for (int k = grp; k < nz; k += TG) {
    long q = D.ix(k, j, i);
    float w = s_w[k * TI + lane];
    if (k == 0 || k > k1) {
      s_rhs[k * TI + lane] = a.ph_in[q];
    } else {
      const float p1 = a.ph_1[q], pb = a.phb[q], ws = a.w_save[q];
      float hk = (p1 + pb) / G;
      float dw = 0.0f;
      if (hk >= hbot) {
        float sn = sinf(a.hpi * (hk - hbot) / a.hdepth);
        dw = a.dampmag * (sn * sn);
      }
      w = (w - dw * mut * ws) / (1.0f + dw);
      s_w[k * TI + lane] = w;
      s_rhs[k * TI + lane] = s_rhs[k * TI + lane] + my * 0.5f * a.dts * G * a.ep1 * w / muts;
    }
  }
  for (int k = grp; k < nz; k += TG) {
    long q = D.ix(k, j, i);
    a.w[q] = s_w[k * TI + lane];
    a.ph[q] = s_rhs[k * TI + lane];
    a.t_2ave[q] = s_t2a[k * TI + lane];
  }
}

extern "C" int pywrf_advance_w(const pywrf_grid_t* g, const float* w_in, const float* rw_tend, const float* ww,
                               const float* w_save, const float* u, const float* v, const float* mut,
                               const float* muave, const float* muts, const float* t_2ave_in, const float* t_2,
                               const float* t_1, const float* ph_in, const float* ph_1, const float* phb,
                               const float* ph_tend, const float* ht, const float* c2a, const float* cqw,
                               const float* alt, const float* a, const float* alpha, const float* gamma,
                               double rdx, double rdy, double dts, double t0, double epssm, const float* fnm,
                               const float* fnp, const float* rdnw, const float* rdn, const float* cf,
                               const float* msftx, const float* msfty, float* t_2ave_out, float* w_out,
                               float* ph_out, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->kte > g->nz || g->kte < 3) return PYWRF_E_RANGE;
  const double pi_ = 3.14159265359, G = pw::g;
  AdvW p = {w_in, rw_tend, ww, w_save, u, v, mut, muave, muts, t_2ave_in, t_2, t_1, ph_in, ph_1, phb, ph_tend, ht,
            c2a, cqw, alt, a, alpha, gamma, fnm, fnp, rdnw, rdn, cf, msftx, msfty, t_2ave_out, w_out, ph_out,
            F(dts), F(t0), F(1. + epssm), F(1. - epssm), F(0.5 * G * (1. - epssm)), F(rdx), F(rdy),
            F(0.5 * dts * G), F(dts * G), F(-0.5 * dts * G), F(dts * 0.2), F(0.5 * pi_), F(5000.), g->kte - 1,
            imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), imax(g->its, g->ids + 1), imin(g->ite, g->ide - 2)};
  if (column_block_launch_check(g, 6)) return PYWRF_E_RANGE;
  dim3 grid((g->nx + TI - 1) / TI, g->ny);
  hipLaunchKernelGGL(k_advance_w, grid, dim3(TB), 6 * g->nz * TI * sizeof(float), (hipStream_t)s, dims_of(g), p);
  PW_RET();
}

// ------------------------------------------------------------------ sumflux
__global__ void k_sumflux(Box b, Dims D, const float* ru, const float* rv, const float* ww, const float* u_lin,
                          const float* v_lin, const float* ww_lin, const float* muu, const float* muv, float* ru_m,
                          float* rv_m, float* ww_m, const float* msfuy, const float* msfvx_inv, Box bu, Box bv,
                          Box bw, int first, int last, float n) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  // every operand is at this point, so all loads are issued before the
  // branches (a load inside a divergent branch is waited for before the
  // branch ends); `first` and `last` are uniform
  float rm = first ? 0.0f : ru_m[q], vm = first ? 0.0f : rv_m[q], wm = first ? 0.0f : ww_m[q];
  const float rq = ru[q], vq = rv[q], wq = ww[q];
  float ul = 0.0f, vl = 0.0f, wl = 0.0f, mu = 0.0f, my = 0.0f, mv = 0.0f, mx = 0.0f;
  if (last) {
    ul = u_lin[q]; vl = v_lin[q]; wl = ww_lin[q];
    mu = AT2(muu, j, i); my = AT2(msfuy, j, i); mv = AT2(muv, j, i); mx = AT2(msfvx_inv, j, i);
  }
  if (bu.has(k, j, i)) {
    rm = rm + rq;
    if (last) rm = rm / n + mu * ul / my;
  }
  if (bv.has(k, j, i)) {
    vm = vm + vq;
    if (last) vm = vm / n + mv * vl * mx;
  }
  if (bw.has(k, j, i)) {
    wm = wm + wq;
    if (last) wm = wm / n + wl;
  }
  ru_m[q] = rm;
  rv_m[q] = vm;
  ww_m[q] = wm;
}

extern "C" int pywrf_sumflux(const pywrf_grid_t* g, const float* ru, const float* rv, const float* ww,
                             const float* u_lin, const float* v_lin, const float* ww_lin, const float* muu,
                             const float* muv, float* ru_m, float* rv_m, float* ww_m, const float* msfuy,
                             const float* msfvx_inv, int iteration, int number_of_small_timesteps,
                             pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kte > g->nz || number_of_small_timesteps == 0) return PYWRF_E_RANGE;
  int mini = imin(g->ide - 1, g->ite), minj = imin(g->jde - 1, g->jte), mink = imin(g->kde - 1, g->kte);
  Box bu = mkbox(g->kts, mink, g->jts, minj, g->its, g->ite);
  Box bv = mkbox(g->kts, mink, g->jts, g->jte, g->its, mini);
  Box bw = mkbox(g->kts, g->kte, g->jts, minj, g->its, mini);
  Box b = mkbox(g->kts, g->kte, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_BOX(k_sumflux, b, s, dims_of(g), ru, rv, ww, u_lin, v_lin, ww_lin, muu, muv, ru_m, rv_m, ww_m, msfuy,
                msfvx_inv, bu, bv, bw, iteration == 1, iteration == number_of_small_timesteps,
                (float)number_of_small_timesteps);
  PW_RET();
}

// ------------------------------------------------------------------ calc_mu_uv_1
__global__ void k_calc_mu_uv_1(Box b, Dims D, const float* mu, float* muu, float* muv, int jts, int jtf, int jte,
                               int its, int itf, int ite) {
  int j, i;
  if (!box_column(b, j, i)) return;
  if (j < jtf) {
    if (i >= its + 1 && i < ite - 1) AT2(muu, j, i) = 0.5f * (AT2(mu, j, i) + AT2(mu, j, i - 1));
    else if (i == its) AT2(muu, j, i) = AT2(mu, j, i);
    if (i == ite - 1) AT2(muu, j, i) = AT2(mu, j, ite - 2);
  }
  if (i < itf) {
    if (j >= jts + 1 && j < jte - 1) AT2(muv, j, i) = 0.5f * (AT2(mu, j, i) + AT2(mu, j - 1, i));
    else if (j == jts) AT2(muv, j, i) = AT2(mu, j, i);
    if (j == jte - 1) AT2(muv, j, i) = AT2(mu, jte - 2, i);
  }
}

extern "C" int pywrf_calc_mu_uv_1(const pywrf_grid_t* g, const float* mu, float* muu, float* muv, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  int jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  Box b = mkbox(0, 1, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_COLS(k_calc_mu_uv_1, b, s, dims_of(g), mu, muu, muv, g->jts, jtf, g->jte, g->its, itf, g->ite);
  PW_RET();
}

// ------------------------------------------------------------------ small_step_finish
struct SsFin {
  const float *u_2, *v_2, *w_2, *t_2, *ph_2, *ww1, *mut, *muts, *muu, *muus, *muv, *muvs;
  const float *u_save, *v_save, *w_save, *t_save, *ph_save, *mu_save, *msfuy, *msfvx, *msfty, *h_diabatic;
  float *ww, *mu_2, *u, *v, *w, *ph, *t;
  float dtsn;
  int last, kde, j0, j1, i0, i1, jte, ite;
};

__global__ void k_ssfinish(Box all, Dims D, SsFin a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i), q2 = D.ix2(j, i);
  bool hcol = j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.i1;
  // every value is formed first and stored last, so that no load waits on a
  // store, and every operand (all at this point) is loaded before the
  // branches, so that no branch waits on its loads
  const float v2 = a.v_2[q], vs = a.v_save[q], u2 = a.u_2[q], us = a.u_save[q];
  const float w2 = a.w_2[q], ws = a.w_save[q], p2 = a.ph_2[q], ps = a.ph_save[q], wwq = a.ww[q], ww1 = a.ww1[q];
  const float t2 = a.t_2[q], ts = a.t_save[q], hd = a.last ? a.h_diabatic[q] : 0.0f;
  const float msfvx = a.msfvx[q2], muv = a.muv[q2], muvs = a.muvs[q2], msfuy = a.msfuy[q2], muu = a.muu[q2];
  const float muus = a.muus[q2], msfty = a.msfty[q2], mut = a.mut[q2], muts = a.muts[q2];
  const float mu2 = a.mu_2[q2], musv = a.mu_save[q2];
  float v = v2;
  if (k < a.kde - 1 && j >= a.j0 && j < a.jte && i >= a.i0 && i < a.i1) v = (msfvx * v + vs * muv) / muvs;
  float u = u2;
  if (k < a.kde - 1 && j >= a.j0 && j < a.j1 && i >= a.i0 && i < a.ite) u = (msfuy * u + us * muu) / muus;
  float w = w2, ph = p2, wwn = 0.0f;
  const bool hw = k < a.kde && hcol;
  if (hw) {
    w = (msfty * w + ws * mut) / muts;
    ph = ph + ps;
    wwn = wwq + ww1;
  }
  float t = t2;
  if (k < a.kde - 1 && hcol) {
    if (!a.last) t = (t + ts * mut) / muts;
    else t = (t - a.dtsn * mut * hd + ts * mut) / muts;
  }
  const bool hm = k == 0 && hcol;
  const float mu = hm ? mu2 + musv : 0.0f;
  a.v[q] = v;
  a.u[q] = u;
  if (hw) a.ww[q] = wwn;
  a.w[q] = w;
  a.ph[q] = ph;
  a.t[q] = t;
  if (hm) a.mu_2[q2] = mu;
}

extern "C" int pywrf_small_step_finish(const pywrf_grid_t* g, const float* u_2, const float* v_2, const float* w_2,
                                       const float* t_2, const float* ph_2, float* ww, const float* ww1,
                                       float* mu_2, const float* mut, const float* muts, const float* muu,
                                       const float* muus, const float* muv, const float* muvs,
                                       const float* u_save, const float* v_save, const float* w_save,
                                       const float* t_save, const float* ph_save, const float* mu_save,
                                       const float* msfuy, const float* msfvx, const float* msfty,
                                       const float* h_diabatic, int number_of_small_timesteps, double dts,
                                       int rk_step, int rk_order, float* u_out, float* v_out, float* w_out,
                                       float* ph_out, float* t_out, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kde > g->nz || g->kds != 0) return PYWRF_E_RANGE;
  SsFin a = {u_2, v_2, w_2, t_2, ph_2, ww1, mut, muts, muu, muus, muv, muvs, u_save, v_save, w_save, t_save,
             ph_save, mu_save, msfuy, msfvx, msfty, h_diabatic, ww, mu_2, u_out, v_out, w_out, ph_out, t_out,
             F(dts * number_of_small_timesteps), rk_step >= rk_order, g->kde, g->jts, imin(g->jte, g->jde - 1),
             g->its, imin(g->ite, g->ide - 1), g->jte, g->ite};
  PW_LAUNCH_BOX(k_ssfinish, whole(g), s, dims_of(g), a);
  PW_RET();
}
