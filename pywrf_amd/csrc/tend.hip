// Large-step tendencies other than advection: geopotential RHS, horizontal
// pressure gradient, vertical PGF and buoyancy, Coriolis, curvature,
// horizontal and vertical diffusion, and the first-stage tendency fold
// (pywrf/wrf_dynamics.py:2031-3148).  One launch per call; a thread owns one
// output point and applies the reference's statements for that point in order.
#include "common.h"
#include "consts.h"
#include "diff.h"


int pw_check_grid(const pywrf_grid_t* g, int r);
#define CHECK(g, r) do { int _e = pw_check_grid((g), (r)); if (_e) return _e; } while (0)

// ------------------------------------------------------------------ rhs_ph
struct RhsPh {
  const float *u, *v, *ww, *ph, *phb, *w, *mut, *muu, *muv, *fnm, *fnp, *rdnw, *cfn, *cfn1;
  const float *msfux, *msfvy, *msfty;
  float* tend;
  float rdx, rdy, c4x, c4y, c2x, c2y;  // 0.25*rd and 0.5*rd rounded once
  int kte, ids, ide, jds, jde, jts, jtf, its, itf;
};

// One direction of the horizontal phi advection, order 6, 4 or 2
// (wrf_dynamics.py:2082-2295), from operand values: P[0..6] and B[0..6] are ph
// and phb at -3..3 steps from the point; vA / vB the normal velocity at the two
// levels the face value uses (k and k-1, or k-1 and k-2 at the top level, where
// they are extrapolated with cfn, cfn1) at the point (0) and one step on (1);
// mu0 / mu1 and m0 / m1 the mass and map factor there.
__device__ inline float ph_hflux_v(const RhsPh& a, bool top, int order, float c, float vA0, float vB0,
                                   float vA1, float vB1, float mu0, float mu1, float m0, float m1, const float* P,
                                   const float* B) {
  const float cfn = *a.cfn, cfn1 = *a.cfn1;
  const float V0 = top ? cfn * vA0 + cfn1 * vB0 : vA0 + vB0;
  const float V1 = top ? cfn * vA1 + cfn1 * vB1 : vA1 + vB1;
  float F1 = mu1 * V1 * m1;
  float F0 = mu0 * V0 * m0;
  if (order == 6) {
    float s = 45.0f * (P[4] - P[2]) - 9.0f * (P[5] - P[1]) + (P[6] - P[0]) + 45.0f * (B[4] - B[2]) -
              9.0f * (B[5] - B[1]) + (B[6] - B[0]);
    return c * ((F1 + F0) * F(1. / 60.) * s);
  }
  if (order == 4) {
    float s = 8.0f * (P[4] - P[2]) - (P[5] - P[1]) + 8.0f * (B[4] - B[2]) - (B[5] - B[1]);
    return c * ((F1 + F0) * F(1. / 12.) * s);
  }
  float d1 = B[4] - B[3] + P[4] - P[3];
  float d0 = B[3] - B[2] + P[3] - P[2];
  return c * (F1 * d1 + F0 * d0);
}

// fresh: the tendency array is taken as zero (the caller's zeroing is folded
// in): the launch covers the whole array, points outside the box get 0, and the
// box's points start from 0 instead of the stored value.
__global__ void k_rhs_ph(Box all, Dims D, RhsPh a, Box b, int fresh) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  if (!b.has(k, j, i)) {
    if (fresh) a.tend[q] = 0.0f;
    return;
  }
  // every operand is loaded first, outside the branches, at indices clamped
  // into the array (k is one per workgroup; the halo check puts every value a
  // flux formula uses in range, so clamping changes only unused values)
  const bool top = k == a.kte - 1;
  const int ku = imin(k + 1, D.nz - 1), kA = top ? k - 1 : k, kB = imax(top ? k - 2 : k - 1, 0);
  const int jn = imin(j + 1, D.ny - 1), ie = imin(i + 1, D.nx - 1);
  float py[7], by_[7], px[7], bx_[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    const int jj = imin(imax(j - 3 + o, 0), D.ny - 1), ii = imin(imax(i - 3 + o, 0), D.nx - 1);
    py[o] = AT(a.ph, k, jj, i); by_[o] = AT(a.phb, k, jj, i);
    px[o] = AT(a.ph, k, j, ii); bx_[o] = AT(a.phb, k, j, ii);
  }
  const float wwu = AT(a.ww, ku, j, i), ww0 = AT(a.ww, k, j, i), wwd = AT(a.ww, k - 1, j, i);
  const float phu = AT(a.ph, ku, j, i), phd = AT(a.ph, k - 1, j, i);
  const float pbu = AT(a.phb, ku, j, i), pbd = AT(a.phb, k - 1, j, i);
  const float vA0 = AT(a.v, kA, j, i), vB0 = AT(a.v, kB, j, i), vA1 = AT(a.v, kA, jn, i), vB1 = AT(a.v, kB, jn, i);
  const float uA0 = AT(a.u, kA, j, i), uB0 = AT(a.u, kB, j, i), uA1 = AT(a.u, kA, j, ie), uB1 = AT(a.u, kB, j, ie);
  const float mv0 = AT2(a.muv, j, i), mv1 = AT2(a.muv, jn, i), sv0 = AT2(a.msfvy, j, i), sv1 = AT2(a.msfvy, jn, i);
  const float mu0 = AT2(a.muu, j, i), mu1 = AT2(a.muu, j, ie), su0 = AT2(a.msfux, j, i), su1 = AT2(a.msfux, j, ie);
  const float mut = AT2(a.mut, j, i), my = AT2(a.msfty, j, i), wq = a.w[q];
  float t = fresh ? 0.0f : a.tend[q];
  // the vertical phi advection terms at w levels k + 1 and k
  const float w1 = 0.5f * (wwu + ww0) * a.rdnw[top ? k - 1 : k] * (phu - py[3] + pbu - by_[3]);
  const float w0 = 0.5f * (ww0 + wwd) * a.rdnw[k - 1] * (py[3] - phd + by_[3] - pbd);
  if (!top) t = t - (a.fnm[k] * w1 + a.fnp[k] * w0);
  else t = 0.0f;
  t = t + mut * F(pw::g) * wq / my;
  // y direction: 6th order inside, 4th order at jds+2 and jde-4, 2nd order at jds+1 and jde-3
  int oy = 0;
  if (j >= imax(a.jts, a.jds + 3) && j < imin(a.jtf, a.jde - 4)) oy = 6;
  else if (j == a.jds + 2 || j == a.jde - 4) oy = 4;
  else if (j == a.jds + 1 || j == a.jde - 3) oy = 2;
  const float cy = (top ? a.c2y : a.c4y) / my, cx = (top ? a.c2x : a.c4x) / my;
  const float fy = ph_hflux_v(a, top, oy, cy, vA0, vB0, vA1, vB1, mv0, mv1, sv0, sv1, py, by_);
  if (oy) t = t - fy;
  // x direction: 6th order inside, 2nd order at ids+1 and ide-3 (no 4th-order columns)
  int ox = 0;
  if (i >= imax(a.its, a.ids + 3) && i < imin(a.itf, a.ide - 4)) ox = 6;
  else if (i == a.ids + 1 || i == a.ide - 3) ox = 2;
  const float fx = ph_hflux_v(a, top, ox, cx, uA0, uB0, uA1, uB1, mu0, mu1, su0, su1, px, bx_);
  if (ox) t = t - fx;
  a.tend[q] = t;
}

static int rhs_ph(int fresh, const pywrf_grid_t* g, float* ph_tend, const float* u, const float* v,
                  const float* ww, const float* ph, const float* phb, const float* w, const float* mut,
                  const float* muuf, const float* muvf, const float* fnm, const float* fnp, const float* rdnw,
                  const float* cfn, const float* cfn1, double rdx, double rdy, const float* msfux,
                  const float* msfvy, const float* msfty, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  if (g->kte > g->nz || g->kte < 3) return PYWRF_E_RANGE;
  RhsPh a;
  a.u = u; a.v = v; a.ww = ww; a.ph = ph; a.phb = phb; a.w = w; a.mut = mut; a.muu = muuf; a.muv = muvf;
  a.fnm = fnm; a.fnp = fnp; a.rdnw = rdnw; a.cfn = cfn; a.cfn1 = cfn1; a.msfux = msfux; a.msfvy = msfvy;
  a.msfty = msfty; a.tend = ph_tend;
  a.rdx = F(rdx); a.rdy = F(rdy);
  a.c4x = F(0.25 * rdx); a.c4y = F(0.25 * rdy); a.c2x = F(0.5 * rdx); a.c2y = F(0.5 * rdy);
  a.kte = g->kte; a.ids = g->ids; a.ide = g->ide; a.jds = g->jds; a.jde = g->jde;
  a.jts = g->jts; a.jtf = imin(g->jte, g->jde - 1); a.its = g->its; a.itf = imin(g->ite, g->ide - 1);
  Box b = mkbox(1, g->kte, a.jts, a.jtf, a.its, a.itf);
  PW_LAUNCH_BOX(k_rhs_ph, fresh ? whole(g) : b, s, dims_of(g), a, b, fresh);
  PW_RET();
}

extern "C" int pywrf_rhs_ph(const pywrf_grid_t* g, float* ph_tend, const float* u, const float* v,
                            const float* ww, const float* ph, const float* phb, const float* w, const float* mut,
                            const float* muuf, const float* muvf, const float* fnm, const float* fnp,
                            const float* rdnw, const float* cfn, const float* cfn1, double rdx, double rdy,
                            const float* msfux, const float* msfvy, const float* msfty, pywrf_stream_t s) {
  return rhs_ph(0, g, ph_tend, u, v, ww, ph, phb, w, mut, muuf, muvf, fnm, fnp, rdnw, cfn, cfn1, rdx, rdy, msfux,
                msfvy, msfty, s);
}

extern "C" int pywrf_rhs_ph_fresh(const pywrf_grid_t* g, float* ph_tend, const float* u, const float* v,
                                  const float* ww, const float* ph, const float* phb, const float* w,
                                  const float* mut, const float* muuf, const float* muvf, const float* fnm,
                                  const float* fnp, const float* rdnw, const float* cfn, const float* cfn1,
                                  double rdx, double rdy, const float* msfux, const float* msfvy,
                                  const float* msfty, pywrf_stream_t s) {
  return rhs_ph(1, g, ph_tend, u, v, ww, ph, phb, w, mut, muuf, muvf, fnm, fnp, rdnw, cfn, cfn1, rdx, rdy, msfux,
                msfvy, msfty, s);
}

// ------------------------------------------------------------------ horizontal_pressure_gradient
struct Hpg {
  const float *ph, *alt, *p, *pb, *al, *php, *cqu, *cqv, *muu, *muv, *mu, *fnm, *fnp, *rdnw, *cf;
  const float *msfux, *msfuy, *msfvx, *msfvy;
  float *ru_tend, *rv_tend;
  float rdx, rdy;
  int ktf, kde;
};

// pressure on w level k averaged over the face between (j, i) and (j+dj, i+di)
__device__ inline float hpg_dpn(const Dims& D, const Hpg& a, int k, int j, int i, int dj, int di) {
  auto P = [&](int kk, int jj, int ii) { return AT(a.p, kk, jj, ii); };
  if (k == 0)
    return 0.5f * (a.cf[0] * (P(0, j + dj, i + di) + P(0, j, i)) + a.cf[1] * (P(1, j + dj, i + di) + P(1, j, i)) +
                   a.cf[2] * (P(2, j + dj, i + di) + P(2, j, i)));
  if (k == a.kde - 1) return 0.0f;
  return 0.5f * (a.fnm[k] * (P(k, j + dj, i + di) + P(k, j, i)) + a.fnp[k] * (P(k - 1, j + dj, i + di) + P(k - 1, j, i)));
}

__device__ inline float hpg_dpx(const Dims& D, const Hpg& a, int k, int j, int i, int dj, int di, const float* mv,
                                const float* mx, const float* my, float rd) {
  int jd = j + dj, id = i + di;
  float r = AT2(mx, j, i) / AT2(my, j, i);
  float dpx = r * 0.5f * rd * AT2(mv, j, i) *
              ((AT(a.ph, k + 1, j, i) - AT(a.ph, k + 1, jd, id) + AT(a.ph, k, j, i) - AT(a.ph, k, jd, id)) +
               (AT(a.alt, k, j, i) + AT(a.alt, k, jd, id)) * (AT(a.p, k, j, i) - AT(a.p, k, jd, id)) +
               (AT(a.al, k, j, i) + AT(a.al, k, jd, id)) * (AT(a.pb, k, j, i) - AT(a.pb, k, jd, id)));
  dpx = dpx + r * rd * (AT(a.php, k, j, i) - AT(a.php, k, jd, id)) *
                  (a.rdnw[k] * (hpg_dpn(D, a, k + 1, j, i, dj, di) - hpg_dpn(D, a, k, j, i, dj, di)) -
                   0.5f * (AT2(a.mu, jd, id) + AT2(a.mu, j, i)));
  return dpx;
}

__global__ void k_hpg(Box b, Dims D, Hpg a, Box bv, Box bu) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  if (bv.has(k, j, i))
    a.rv_tend[q] = a.rv_tend[q] - a.cqv[q] * hpg_dpx(D, a, k, j, i, -1, 0, a.muv, a.msfvy, a.msfvx, a.rdy);
  if (bu.has(k, j, i))
    a.ru_tend[q] = a.ru_tend[q] - a.cqu[q] * hpg_dpx(D, a, k, j, i, 0, -1, a.muu, a.msfux, a.msfuy, a.rdx);
}

extern "C" int pywrf_horizontal_pressure_gradient(const pywrf_grid_t* g, float* ru_tend, float* rv_tend,
                                                  const float* ph, const float* alt, const float* p,
                                                  const float* pb, const float* al, const float* php,
                                                  const float* cqu, const float* cqv, const float* muu,
                                                  const float* muv, const float* mu, const float* fnm,
                                                  const float* fnp, const float* rdnw, const float* cf, double rdx,
                                                  double rdy, const float* msfux, const float* msfuy,
                                                  const float* msfvx, const float* msfvy, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  Hpg a;
  a.ph = ph; a.alt = alt; a.p = p; a.pb = pb; a.al = al; a.php = php; a.cqu = cqu; a.cqv = cqv; a.muu = muu;
  a.muv = muv; a.mu = mu; a.fnm = fnm; a.fnp = fnp; a.rdnw = rdnw; a.cf = cf; a.msfux = msfux; a.msfuy = msfuy;
  a.msfvx = msfvx; a.msfvy = msfvy; a.ru_tend = ru_tend; a.rv_tend = rv_tend; a.rdx = F(rdx); a.rdy = F(rdy);
  a.ktf = imin(g->kte, g->kde - 1);
  a.kde = g->kde;
  if (a.ktf + 1 > g->nz || g->kde > g->nz) return PYWRF_E_RANGE;
  Box bv = mkbox(0, a.ktf, g->jts + 1, g->jte - 1, g->its, imin(g->ite, g->ide - 1));
  Box bu = mkbox(0, a.ktf, g->jts, imin(g->jte, g->jde - 1), g->its + 1, g->ite - 1);
  Box b = mkbox(0, a.ktf, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_BOX(k_hpg, b, s, dims_of(g), a, bv, bu);
  PW_RET();
}

// ------------------------------------------------------------------ pg_buoy_w
__global__ void k_pg_buoy_w(Box b, Dims D, float* rw_tend, const float* p, float* cqw, const float* muf,
                            const float* mubf, const float* rdnw, const float* rdn, float g_, const float* msfty,
                            int kde) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  float cqo = cqw[q];
  float rmsf = 1.0f / AT2(msfty, j, i);
  if (k == kde - 2) {  // the top level reads cqw[kde-2] before it is overwritten
    float cq1 = 1.0f / (1.0f + cqo);
    float cq2 = cqo * cq1;
    long t = D.ix(kde - 1, j, i);
    rw_tend[t] = rw_tend[t] + rmsf * g_ * (cq1 * 2.0f * rdnw[kde - 2] * (-p[q]) - AT2(muf, j, i) -
                                           cq2 * AT2(mubf, j, i));
  }
  float cq1 = 1.0f / (1.0f + cqo);
  float cq2 = cqo * cq1;
  cqw[q] = cq1;
  rw_tend[q] = rw_tend[q] + rmsf * g_ * (cq1 * rdn[k] * (p[q] - AT(p, k - 1, j, i)) - AT2(muf, j, i) -
                                         cq2 * AT2(mubf, j, i));
}

extern "C" int pywrf_pg_buoy_w(const pywrf_grid_t* g, float* rw_tend, const float* p, float* cqw, const float* muf,
                               const float* mubf, const float* rdnw, const float* rdn, double g_,
                               const float* msfty, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kde > g->nz || g->kde < 3) return PYWRF_E_RANGE;
  Box b = mkbox(1, g->kde - 1, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  PW_LAUNCH_BOX(k_pg_buoy_w, b, s, dims_of(g), rw_tend, p, cqw, muf, mubf, rdnw, rdn, F(g_), msfty, g->kde);
  PW_RET();
}

// ------------------------------------------------------------------ coriolis
struct Cor {
  const float *ru, *rv, *rw, *msftx, *msfty, *msfux, *msfuy, *msfvx, *msfvy, *f, *e, *sina, *cosa, *fzm, *fzp;
  float *ru_tend, *rv_tend, *rw_tend;
};

__global__ void k_coriolis(Box b, Dims D, Cor a, Box bu, Box bv, Box bw) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  if (bu.has(k, j, i)) {
    float x = (AT2(a.msfux, j, i) / AT2(a.msfuy, j, i)) * 0.5f * (AT2(a.f, j, i) + AT2(a.f, j, i - 1)) * 0.25f *
              (AT(a.rv, k, j + 1, i - 1) + AT(a.rv, k, j + 1, i) + AT(a.rv, k, j, i - 1) + AT(a.rv, k, j, i));
    float y = 0.5f * (AT2(a.e, j, i) + AT2(a.e, j, i - 1)) * 0.5f * (AT2(a.cosa, j, i) + AT2(a.cosa, j, i - 1)) *
              0.25f * (AT(a.rw, k + 1, j, i - 1) + AT(a.rw, k, j, i - 1) + AT(a.rw, k + 1, j, i) + AT(a.rw, k, j, i));
    a.ru_tend[q] = a.ru_tend[q] + x - y;
  }
  if (bv.has(k, j, i)) {
    float r = AT2(a.msfvy, j, i) / AT2(a.msfvx, j, i);
    float x = r * 0.5f * (AT2(a.f, j, i) + AT2(a.f, j - 1, i)) * 0.25f *
              (AT(a.ru, k, j, i) + AT(a.ru, k, j, i + 1) + AT(a.ru, k, j - 1, i) + AT(a.ru, k, j - 1, i + 1));
    float y = r * 0.5f * (AT2(a.e, j, i) + AT2(a.e, j - 1, i)) * 0.5f * (AT2(a.sina, j, i) + AT2(a.sina, j - 1, i)) *
              0.25f * (AT(a.rw, k + 1, j - 1, i) + AT(a.rw, k, j - 1, i) + AT(a.rw, k + 1, j, i) + AT(a.rw, k, j, i));
    a.rv_tend[q] = a.rv_tend[q] - x + y;
  }
  if (bw.has(k, j, i)) {
    float fzm = a.fzm[k], fzp = a.fzp[k];
    float cu = AT2(a.cosa, j, i) * 0.5f *
               (fzm * (AT(a.ru, k, j, i) + AT(a.ru, k, j, i + 1)) + fzp * (AT(a.ru, k - 1, j, i) + AT(a.ru, k - 1, j, i + 1)));
    float sv = (AT2(a.msftx, j, i) / AT2(a.msfty, j, i)) * AT2(a.sina, j, i) * 0.5f *
               (fzm * (AT(a.rv, k, j, i) + AT(a.rv, k, j + 1, i)) + fzp * (AT(a.rv, k - 1, j, i) + AT(a.rv, k - 1, j + 1, i)));
    a.rw_tend[q] = a.rw_tend[q] + AT2(a.e, j, i) * (cu - sv);
  }
}

extern "C" int pywrf_coriolis(const pywrf_grid_t* g, const float* ru, const float* rv, const float* rw,
                              float* ru_tend, float* rv_tend, float* rw_tend, const float* msftx, const float* msfty,
                              const float* msfux, const float* msfuy, const float* msfvx, const float* msfvy,
                              const float* f, const float* e, const float* sina, const float* cosa,
                              const float* fzm, const float* fzp, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  int ktf = imin(g->kte, g->kde - 1);
  if (ktf + 1 > g->nz) return PYWRF_E_RANGE;
  Cor a = {ru, rv, rw, msftx, msfty, msfux, msfuy, msfvx, msfvy, f, e, sina, cosa, fzm, fzp, ru_tend, rv_tend,
           rw_tend};
  Box bu = mkbox(g->kts, ktf, g->jts, g->jte - 1, imax(g->ids + 1, g->its), imin(g->ide - 1, g->ite));
  Box bv = mkbox(g->kts, ktf, imax(g->jds + 1, g->jts), imin(g->jde - 1, g->jte), g->its, g->ite - 1);
  Box bw = mkbox(g->kts + 1, ktf, g->jts, g->jte - 1, g->its, g->ite - 1);
  Box b = mkbox(g->kts, ktf, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_BOX(k_coriolis, b, s, dims_of(g), a, bu, bv, bw);
  PW_RET();
}

// ------------------------------------------------------------------ curvature
struct Curv {
  const float *ru, *rv, *rw, *u, *v, *msfuy, *msfvx, *msfvy, *msftx, *msfty, *fzm, *fzp;
  float *ru_tend, *rv_tend, *rw_tend;
  float rdx, rdy;
  int jts, jte, its, ite;
};

// vxgm on [jts-1, jte) x [its-1, ite): the interior value on [jts, jte-1) x
// [its, ite-1), copied outward to the edge ring (columns first, then rows);
// zero elsewhere.
__device__ inline float curv_vxgm(const Dims& D, const Curv& a, int k, int j, int i) {
  if (j < a.jts - 1 || j >= a.jte || i < a.its - 1 || i >= a.ite) return 0.0f;
  j = j < a.jts ? a.jts : (j > a.jte - 2 ? a.jte - 2 : j);
  i = i < a.its ? a.its : (i > a.ite - 2 ? a.ite - 2 : i);
  return 0.5f * (AT(a.u, k, j, i) + AT(a.u, k, j, i + 1)) * (AT2(a.msfvx, j + 1, i) - AT2(a.msfvx, j, i)) * a.rdy -
         0.5f * (AT(a.v, k, j, i) + AT(a.v, k, j + 1, i)) * (AT2(a.msfuy, j, i + 1) - AT2(a.msfuy, j, i)) * a.rdx;
}

__global__ void k_curvature(Box b, Dims D, Curv a, Box bu, Box bv, Box bw) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  const float rr = F(pw::reradius);
  if (bu.has(k, j, i)) {
    float x = 0.5f * (curv_vxgm(D, a, k, j, i) + curv_vxgm(D, a, k, j, i - 1)) * 0.25f *
              (AT(a.rv, k, j + 1, i - 1) + AT(a.rv, k, j + 1, i) + AT(a.rv, k, j, i - 1) + AT(a.rv, k, j, i));
    float y = AT(a.u, k, j, i) * rr * 0.25f *
              (AT(a.rw, k + 1, j, i - 1) + AT(a.rw, k, j, i - 1) + AT(a.rw, k + 1, j, i) + AT(a.rw, k, j, i));
    a.ru_tend[q] = a.ru_tend[q] + x - y;
  }
  if (bv.has(k, j, i)) {
    float x = 0.5f * (curv_vxgm(D, a, k, j, i) + curv_vxgm(D, a, k, j - 1, i)) * 0.25f *
              (AT(a.ru, k, j, i) + AT(a.ru, k, j, i + 1) + AT(a.ru, k, j - 1, i) + AT(a.ru, k, j - 1, i + 1));
    float y = (AT2(a.msfvy, j, i) / AT2(a.msfvx, j, i)) * AT(a.v, k, j, i) * rr * 0.25f *
              (AT(a.rw, k + 1, j - 1, i) + AT(a.rw, k, j - 1, i) + AT(a.rw, k + 1, j, i) + AT(a.rw, k, j, i));
    a.rv_tend[q] = a.rv_tend[q] - x - y;
  }
  if (bw.has(k, j, i)) {
    float fzm = a.fzm[k], fzp = a.fzp[k];
    auto fz = [&](const float* f, int dj, int di) {
      return fzm * (AT(f, k, j, i) + AT(f, k, j + dj, i + di)) + fzp * (AT(f, k - 1, j, i) + AT(f, k - 1, j + dj, i + di));
    };
    a.rw_tend[q] = a.rw_tend[q] + rr * (0.5f * fz(a.ru, 0, 1) * 0.5f * fz(a.u, 0, 1) +
                                        (AT2(a.msftx, j, i) / AT2(a.msfty, j, i)) * 0.5f * fz(a.rv, 1, 0) * 0.5f *
                                            fz(a.v, 1, 0));
  }
}

extern "C" int pywrf_curvature(const pywrf_grid_t* g, const float* ru, const float* rv, const float* rw,
                               const float* u, const float* v, float* ru_tend, float* rv_tend, float* rw_tend,
                               const float* msfuy, const float* msfvx, const float* msfvy, const float* msftx,
                               const float* msfty, const float* fzm, const float* fzp, double rdx, double rdy,
                               pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 2);
  int ktf = imin(g->kte, g->kde - 1);
  if (ktf + 1 > g->nz) return PYWRF_E_RANGE;
  Curv a = {ru, rv, rw, u, v, msfuy, msfvx, msfvy, msftx, msfty, fzm, fzp, ru_tend, rv_tend, rw_tend,
            F(rdx), F(rdy), g->jts, g->jte, g->its, g->ite};
  Box bu = mkbox(g->kts, ktf, g->jts, g->jte - 1, imax(g->ids + 1, g->its), imin(g->ide - 1, g->ite));
  Box bv = mkbox(g->kts, ktf, imax(g->jds + 1, g->jts), imin(g->jde - 1, g->jte), g->its, g->ite - 1);
  Box bw = mkbox(g->kts + 1, ktf, g->jts, g->jte - 1, g->its, g->ite - 1);
  Box b = mkbox(g->kts, ktf, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_BOX(k_curvature, b, s, dims_of(g), a, bu, bv, bw);
  PW_RET();
}

// ------------------------------------------------------------------ coriolis, then curvature
// The two calls the driver makes back to back (solver.py:1063-1079) in one
// pass: same boxes, each tendency point updated by the Coriolis terms and then
// by the curvature terms, in that order (the same floats as the two kernels).
// Coriolis then curvature of the three momentum tendencies at (k, j, i)
// (the fused coriolis + curvature), from their incoming values u0, v0, w0;
// each result is meaningful where its box flag is set.
__device__ inline void cor_curv_point(const Dims& D, const Cor& a, const Curv& c, bool hu, bool hv, bool hw, int k,
                                      int j, int i, float u0, float v0, float w0, float& tu, float& tv, float& tw) {
  const float rr = F(pw::reradius);
  tu = 0.0f; tv = 0.0f; tw = 0.0f;
  if (hu) {
    const float rv4 = AT(a.rv, k, j + 1, i - 1) + AT(a.rv, k, j + 1, i) + AT(a.rv, k, j, i - 1) + AT(a.rv, k, j, i);
    const float rw4 = AT(a.rw, k + 1, j, i - 1) + AT(a.rw, k, j, i - 1) + AT(a.rw, k + 1, j, i) + AT(a.rw, k, j, i);
    float x = (AT2(a.msfux, j, i) / AT2(a.msfuy, j, i)) * 0.5f * (AT2(a.f, j, i) + AT2(a.f, j, i - 1)) * 0.25f * rv4;
    float y = 0.5f * (AT2(a.e, j, i) + AT2(a.e, j, i - 1)) * 0.5f * (AT2(a.cosa, j, i) + AT2(a.cosa, j, i - 1)) *
              0.25f * rw4;
    float t = u0 + x - y;
    x = 0.5f * (curv_vxgm(D, c, k, j, i) + curv_vxgm(D, c, k, j, i - 1)) * 0.25f * rv4;
    y = AT(c.u, k, j, i) * rr * 0.25f * rw4;
    tu = t + x - y;
  }
  if (hv) {
    const float ru4 = AT(a.ru, k, j, i) + AT(a.ru, k, j, i + 1) + AT(a.ru, k, j - 1, i) + AT(a.ru, k, j - 1, i + 1);
    const float rw4 = AT(a.rw, k + 1, j - 1, i) + AT(a.rw, k, j - 1, i) + AT(a.rw, k + 1, j, i) + AT(a.rw, k, j, i);
    float r = AT2(a.msfvy, j, i) / AT2(a.msfvx, j, i);
    float x = r * 0.5f * (AT2(a.f, j, i) + AT2(a.f, j - 1, i)) * 0.25f * ru4;
    float y = r * 0.5f * (AT2(a.e, j, i) + AT2(a.e, j - 1, i)) * 0.5f * (AT2(a.sina, j, i) + AT2(a.sina, j - 1, i)) *
              0.25f * rw4;
    float t = v0 - x + y;
    x = 0.5f * (curv_vxgm(D, c, k, j, i) + curv_vxgm(D, c, k, j - 1, i)) * 0.25f * ru4;
    y = r * AT(c.v, k, j, i) * rr * 0.25f * rw4;
    tv = t - x - y;
  }
  if (hw) {
    const float fzm = a.fzm[k], fzp = a.fzp[k];
    auto fz = [&](const float* f, int dj, int di) {
      return fzm * (AT(f, k, j, i) + AT(f, k, j + dj, i + di)) + fzp * (AT(f, k - 1, j, i) + AT(f, k - 1, j + dj, i + di));
    };
    const float mr = AT2(a.msftx, j, i) / AT2(a.msfty, j, i);
    float cu = AT2(a.cosa, j, i) * 0.5f * fz(a.ru, 0, 1);
    float sv = mr * AT2(a.sina, j, i) * 0.5f * fz(a.rv, 1, 0);
    float t = w0 + AT2(a.e, j, i) * (cu - sv);
    tw = t + rr * (0.5f * fz(a.ru, 0, 1) * 0.5f * fz(c.u, 0, 1) + mr * 0.5f * fz(a.rv, 1, 0) * 0.5f * fz(c.v, 1, 0));
  }
}

__global__ void k_cor_curv(Box b, Dims D, Cor a, Curv c, Box bu, Box bv, Box bw) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  const int q = D.ix(k, j, i);
  // the three tendencies are formed first and stored last, so that no load waits on a store
  const bool hu = bu.has(k, j, i), hv = bv.has(k, j, i), hw = bw.has(k, j, i);
  const float u0 = a.ru_tend[q], v0 = a.rv_tend[q], w0 = a.rw_tend[q];
  float tu, tv, tw;
  cor_curv_point(D, a, c, hu, hv, hw, k, j, i, u0, v0, w0, tu, tv, tw);
  if (hu) a.ru_tend[q] = tu;
  if (hv) a.rv_tend[q] = tv;
  if (hw) a.rw_tend[q] = tw;
}

// horizontal_pressure_gradient, pg_buoy_w and the fused coriolis + curvature
// of one RK stage in one pass (each point in the calls' order: the pressure
// gradient on ru, rv, the buoyancy on rw and cqw, then Coriolis and curvature
// on the three).  The calls' boxes nest in the launch box; only the buoyancy
// thread at kde-2 touches level kde-1 (the top rw tendency, from the cqw it is
// about to overwrite), which no other part updates.
struct PgBuoy {
  float *rw_tend, *cqw;
  const float *p, *muf, *mubf, *rdnw, *rdn, *msfty;
  float g;
  int kde;
};

__global__ void k_pgf_buoy_cor(Box b, Dims D, Hpg h, Box hbv, Box hbu, PgBuoy pg, Box pbox, Cor a, Curv c, Box bu,
                               Box bv, Box bw) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  const int q = D.ix(k, j, i);
  float u0 = a.ru_tend[q], v0 = a.rv_tend[q], w0 = a.rw_tend[q];
  const bool gv = hbv.has(k, j, i), gu = hbu.has(k, j, i), gb = pbox.has(k, j, i);
  if (gv) v0 = v0 - h.cqv[q] * hpg_dpx(D, h, k, j, i, -1, 0, h.muv, h.msfvy, h.msfvx, h.rdy);
  if (gu) u0 = u0 - h.cqu[q] * hpg_dpx(D, h, k, j, i, 0, -1, h.muu, h.msfux, h.msfuy, h.rdx);
  float cq1 = 0.0f, wtop = 0.0f;
  const bool top = gb && k == pg.kde - 2;
  if (gb) {
    const float cqo = pg.cqw[q];
    const float rmsf = 1.0f / AT2(pg.msfty, j, i);
    if (top) {
      const float c1 = 1.0f / (1.0f + cqo), c2 = cqo * c1;
      wtop = AT(pg.rw_tend, pg.kde - 1, j, i) + rmsf * pg.g * (c1 * 2.0f * pg.rdnw[pg.kde - 2] * (-pg.p[q]) -
                                                               AT2(pg.muf, j, i) - c2 * AT2(pg.mubf, j, i));
    }
    cq1 = 1.0f / (1.0f + cqo);
    const float cq2 = cqo * cq1;
    w0 = w0 + rmsf * pg.g * (cq1 * pg.rdn[k] * (pg.p[q] - AT(pg.p, k - 1, j, i)) - AT2(pg.muf, j, i) -
                             cq2 * AT2(pg.mubf, j, i));
  }
  const bool hu = bu.has(k, j, i), hv = bv.has(k, j, i), hw = bw.has(k, j, i);
  float tu, tv, tw;
  cor_curv_point(D, a, c, hu, hv, hw, k, j, i, u0, v0, w0, tu, tv, tw);
  if (hu || gu) a.ru_tend[q] = hu ? tu : u0;
  if (hv || gv) a.rv_tend[q] = hv ? tv : v0;
  if (hw || gb) a.rw_tend[q] = hw ? tw : w0;
  if (gb) pg.cqw[q] = cq1;
  if (top) AT(pg.rw_tend, pg.kde - 1, j, i) = wtop;
}

extern "C" int pywrf_coriolis_curvature(const pywrf_grid_t* g, const float* ru, const float* rv, const float* rw,
                                        const float* u, const float* v, float* ru_tend, float* rv_tend,
                                        float* rw_tend, const float* msftx, const float* msfty, const float* msfux,
                                        const float* msfuy, const float* msfvx, const float* msfvy, const float* f,
                                        const float* e, const float* sina, const float* cosa, const float* fzm,
                                        const float* fzp, double rdx, double rdy, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 2);
  int ktf = imin(g->kte, g->kde - 1);
  if (ktf + 1 > g->nz) return PYWRF_E_RANGE;
  Cor a = {ru, rv, rw, msftx, msfty, msfux, msfuy, msfvx, msfvy, f, e, sina, cosa, fzm, fzp, ru_tend, rv_tend,
           rw_tend};
  Curv c = {ru, rv, rw, u, v, msfuy, msfvx, msfvy, msftx, msfty, fzm, fzp, ru_tend, rv_tend, rw_tend,
            F(rdx), F(rdy), g->jts, g->jte, g->its, g->ite};
  Box bu = mkbox(g->kts, ktf, g->jts, g->jte - 1, imax(g->ids + 1, g->its), imin(g->ide - 1, g->ite));
  Box bv = mkbox(g->kts, ktf, imax(g->jds + 1, g->jts), imin(g->jde - 1, g->jte), g->its, g->ite - 1);
  Box bw = mkbox(g->kts + 1, ktf, g->jts, g->jte - 1, g->its, g->ite - 1);
  Box b = mkbox(g->kts, ktf, g->jts, g->jte, g->its, g->ite);
  PW_LAUNCH_BOX(k_cor_curv, b, s, dims_of(g), a, c, bu, bv, bw);
  PW_RET();
}

// ------------------------------------------------------------------ horizontal_diffusion
__global__ void k_hdiff(Box b, Dims D, HDiff a) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  a.tend[q] = hdiff_apply(D, a, k, j, i, a.tend[q]);
}

// the struct and box of one horizontal_diffusion call
static int hdiff_setup(const pywrf_grid_t* g, char name, const float* field, float* tendency, const float* base,
                       const float* mut, const float* msfux, const float* msfuy, const float* msfvx,
                       const float* msfvx_inv, const float* msfvy, const float* msftx, const float* msfty,
                       const float* xkmhd, float rdx, float rdy, HDiff& a, Box& b) {
  CHECK(g, 2);
  int ktf = imin(g->kte, g->kde - 1);
  a = {field, base, mut, xkmhd, msfux, msfuy, msfvx, msfvx_inv, msfvy, msftx, msfty, tendency, rdx, rdy, name};
  int j0 = imax(g->jds + 1, g->jts), i0 = imax(g->ids + 1, g->its);
  if (name == 'u') b = mkbox(g->kts, ktf, j0, imin(g->jde - 2, g->jte), i0, imin(g->ide - 1, g->ite));
  else if (name == 'v') b = mkbox(g->kts, ktf, j0, imin(g->jde - 1, g->jte), i0, imin(g->ide - 2, g->ite));
  else if (name == 'w') b = mkbox(g->kts + 1, ktf, j0, imin(g->jde - 2, g->jte), i0, imin(g->ide - 2, g->ite));
  else b = mkbox(g->kts, ktf, j0, imin(g->jde - 2, g->jte), i0, imin(g->ide - 2, g->ite));
  return 0;
}

int pw_horizontal_diffusion(const pywrf_grid_t* g, char name, const float* field, float* tendency,
                            const float* base, const float* mut, const float* msfux, const float* msfuy,
                            const float* msfvx, const float* msfvx_inv, const float* msfvy, const float* msftx,
                            const float* msfty, const float* xkmhd, float rdx, float rdy, pywrf_stream_t s) {
  HDiff a;
  Box b;
  int e = hdiff_setup(g, name, field, tendency, base, mut, msfux, msfuy, msfvx, msfvx_inv, msfvy, msftx, msfty,
                      xkmhd, rdx, rdy, a, b);
  if (e) return e;
  PW_LAUNCH_BOX(k_hdiff, b, s, dims_of(g), a);
  PW_RET();
}

extern "C" int pywrf_horizontal_diffusion(const pywrf_grid_t* g, char name, const float* field, float* tendency,
                                          const float* mut, const float* msfux, const float* msfuy,
                                          const float* msfvx, const float* msfvx_inv, const float* msfvy,
                                          const float* msftx, const float* msfty, const float* xkmhd, double rdx,
                                          double rdy, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return pw_horizontal_diffusion(g, name, field, tendency, nullptr, mut, msfux, msfuy, msfvx, msfvx_inv, msfvy,
                                 msftx, msfty, xkmhd, F(rdx), F(rdy), s);
}

extern "C" int pywrf_horizontal_diffusion_3dmp(const pywrf_grid_t* g, const float* field, float* tendency,
                                               const float* mut, const float* base_3d, const float* msfux,
                                               const float* msfuy, const float* msfvx, const float* msfvx_inv,
                                               const float* msfvy, const float* msftx, const float* msfty,
                                               const float* xkmhd, double rdx, double rdy, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  // the 3dmp form always uses the scalar coefficients
  return pw_horizontal_diffusion(g, 'm', field, tendency, base_3d, mut, msfux, msfuy, msfvx, msfvx_inv, msfvy,
                                 msftx, msfty, xkmhd, F(rdx), F(rdy), s);
}

// ------------------------------------------------------------------ vertical diffusion
// kind: 'w' (vertical_diffusion 'w'), 'm' (vertical_diffusion 'm'), 'u'/'v'
// (vertical_diffusion_u/_v, 1-D base), 'p' (vertical_diffusion_mp, 1-D base),
// '3' (vertical_diffusion_3dmp, 3-D base).
__global__ void k_vdiff(Box b, Dims D, VDiff a) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  a.tend[q] = vdiff_apply(D, a, k, j, i, a.tend[q]);
}

// the struct and box of one vertical diffusion call
static int vdiff_setup(const pywrf_grid_t* g, char kind, const float* field, float* tendency, const float* base1,
                       const float* base3, const float* alt, const float* mus, const float* rdn, const float* rdnw,
                       double kvdif, VDiff& a, Box& b) {
  CHECK(g, 1);
  int ktf = imin(g->kte, g->kde - 1);
  if (ktf + 1 > g->nz) return PYWRF_E_RANGE;
  a = {field, alt, mus, rdn, rdnw, base1, base3, tendency, F(kvdif), ktf, g->kts, kind};
  int jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  if (kind == 'u') b = mkbox(g->kts + 1, ktf - 1, g->jts, jtf, imax(g->ids + 1, g->its), imin(g->ide - 1, g->ite));
  else if (kind == 'v') b = mkbox(g->kts + 1, ktf - 1, g->jts, jtf, g->its, itf);
  else b = mkbox(g->kts + 1, ktf, g->jts, jtf, g->its, itf);
  return 0;
}

static int vdiff_launch(const pywrf_grid_t* g, char kind, const float* field, float* tendency, const float* base1,
                        const float* base3, const float* alt, const float* mus, const float* rdn,
                        const float* rdnw, double kvdif, pywrf_stream_t s) {
  VDiff a;
  Box b;
  int e = vdiff_setup(g, kind, field, tendency, base1, base3, alt, mus, rdn, rdnw, kvdif, a, b);
  if (e) return e;
  PW_LAUNCH_BOX(k_vdiff, b, s, dims_of(g), a);
  PW_RET();
}

extern "C" int pywrf_vertical_diffusion(const pywrf_grid_t* g, char name, const float* field, float* tendency,
                                        const float* alt, const float* mut, const float* rdn, const float* rdnw,
                                        double kvdif, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  if (name != 'w' && name != 'm') return 0;  // the reference leaves other names untouched
  return vdiff_launch(g, name, field, tendency, nullptr, nullptr, alt, mut, rdn, rdnw, kvdif, s);
}

extern "C" int pywrf_vertical_diffusion_u(const pywrf_grid_t* g, const float* field, float* tendency,
                                          const float* u_base, const float* alt, const float* muu, const float* rdn,
                                          const float* rdnw, double kvdif, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return vdiff_launch(g, 'u', field, tendency, u_base, nullptr, alt, muu, rdn, rdnw, kvdif, s);
}

extern "C" int pywrf_vertical_diffusion_v(const pywrf_grid_t* g, const float* field, float* tendency,
                                          const float* v_base, const float* alt, const float* muv, const float* rdn,
                                          const float* rdnw, double kvdif, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return vdiff_launch(g, 'v', field, tendency, v_base, nullptr, alt, muv, rdn, rdnw, kvdif, s);
}

extern "C" int pywrf_vertical_diffusion_mp(const pywrf_grid_t* g, const float* field, float* tendency,
                                           const float* base, const float* alt, const float* mut, const float* rdn,
                                           const float* rdnw, double kvdif, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return vdiff_launch(g, 'p', field, tendency, base, nullptr, alt, mut, rdn, rdnw, kvdif, s);
}

extern "C" int pywrf_vertical_diffusion_3dmp(const pywrf_grid_t* g, const float* field, float* tendency,
                                             const float* base_3d, const float* alt, const float* mut,
                                             const float* rdn, const float* rdnw, double kvdif, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return vdiff_launch(g, '3', field, tendency, nullptr, base_3d, alt, mut, rdn, rdnw, kvdif, s);
}

// ------------------------------------------------------------------ the first stage's diffusion
// Array A of the four (u, v, w, theta): its tendency at the point gains the
// horizontal then the vertical term, as the two calls add them one after the
// other; neither term reads a tendency.
struct DiffDry {
  HDiff h[4];
  VDiff v[4];
  Box hb[4], vb[4];
  int nk;
};

template <int A>
__device__ inline void diff_dry_point(const Dims& D, const DiffDry& a, int k, int j, int i) {
  const bool hh = a.hb[A].has(k, j, i), hv = a.vb[A].has(k, j, i);
  if (!hh && !hv) return;
  const long q = D.ix(k, j, i);
  float t = a.h[A].tend[q];
  if (hh) t = hdiff_apply(D, a.h[A], k, j, i, t);
  if (hv) t = vdiff_apply(D, a.v[A], k, j, i, t);
  a.h[A].tend[q] = t;
}

// blockIdx.z = k - b.k0 + nk * array, so that every block works on one array
__global__ void k_diff_dry(Box b, Dims D, DiffDry a) {
  const int i = b.i0 + (int)(blockIdx.x * PW_TX + threadIdx.x), j = b.j0 + (int)(blockIdx.y * PW_TY + threadIdx.y);
  const int arr = (int)blockIdx.z / a.nk, k = b.k0 + (int)blockIdx.z - arr * a.nk;
  if (i >= b.i1 || j >= b.j1) return;
  switch (arr) {
    case 0: diff_dry_point<0>(D, a, k, j, i); break;
    case 1: diff_dry_point<1>(D, a, k, j, i); break;
    case 2: diff_dry_point<2>(D, a, k, j, i); break;
    default: diff_dry_point<3>(D, a, k, j, i); break;
  }
}

extern "C" int pywrf_diffusion_dry(const pywrf_grid_t* g, const float* u, float* ru_tendf, const float* v,
                                   float* rv_tendf, const float* w, float* rw_tendf, const float* t,
                                   float* t_tendf, const float* t_init, const float* mut, const float* msfux,
                                   const float* msfuy, const float* msfvx, const float* msfvx_inv,
                                   const float* msfvy, const float* msftx, const float* msfty, const float* xkmh,
                                   const float* xkhh, const float* u_base, const float* v_base, const float* alt,
                                   const float* muu, const float* muv, const float* rdn, const float* rdnw,
                                   double kvdif_m, double kvdif_t, double rdx, double rdy, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  DiffDry a;
  int e = 0;
  const float* fld[4] = {u, v, w, t};
  float* tend[4] = {ru_tendf, rv_tendf, rw_tendf, t_tendf};
  const char hn[4] = {'u', 'v', 'w', 'm'};
  for (int n = 0; n < 4 && !e; ++n)
    e = hdiff_setup(g, hn[n], fld[n], tend[n], n == 3 ? t_init : nullptr, mut, msfux, msfuy, msfvx, msfvx_inv,
                    msfvy, msftx, msfty, n == 3 ? xkhh : xkmh, F(rdx), F(rdy), a.h[n], a.hb[n]);
  if (!e) e = vdiff_setup(g, 'u', u, ru_tendf, u_base, nullptr, alt, muu, rdn, rdnw, kvdif_m, a.v[0], a.vb[0]);
  if (!e) e = vdiff_setup(g, 'v', v, rv_tendf, v_base, nullptr, alt, muv, rdn, rdnw, kvdif_m, a.v[1], a.vb[1]);
  if (!e) e = vdiff_setup(g, 'w', w, rw_tendf, nullptr, nullptr, alt, mut, rdn, rdnw, kvdif_m, a.v[2], a.vb[2]);
  if (!e) e = vdiff_setup(g, '3', t, t_tendf, nullptr, t_init, alt, mut, rdn, rdnw, kvdif_t, a.v[3], a.vb[3]);
  if (e) return e;
  // the launch covers the union of the eight boxes (empty boxes add nothing)
  Box b = mkbox(g->nz, 0, g->ny, 0, g->nx, 0);
  for (int n = 0; n < 8; ++n) {
    const Box& c = n < 4 ? a.hb[n] : a.vb[n - 4];
    if (c.size() == 0) continue;
    b.k0 = imin(b.k0, c.k0); b.k1 = imax(b.k1, c.k1);
    b.j0 = imin(b.j0, c.j0); b.j1 = imax(b.j1, c.j1);
    b.i0 = imin(b.i0, c.i0); b.i1 = imax(b.i1, c.i1);
  }
  if (b.size() == 0) return 0;
  a.nk = b.nk();
  if ((long)a.nk * 4 > 65535) return PYWRF_E_RANGE;
  dim3 grd((unsigned)((b.ni() + PW_TX - 1) / PW_TX), (unsigned)((b.nj() + PW_TY - 1) / PW_TY), (unsigned)(4 * a.nk));
  hipLaunchKernelGGL(k_diff_dry, grd, dim3(PW_TX, PW_TY), 0, (hipStream_t)s, b, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ rk_addtend_dry
struct AddT {
  float *ru_tend, *rv_tend, *rw_tend, *ph_tend, *t_tend, *ru_tendf, *rv_tendf, *rw_tendf, *ph_tendf, *t_tendf;
  const float *u_save, *v_save, *w_save, *ph_save, *t_save, *h_diabatic, *mut, *msfty, *msfuy, *msfvx,
      *msfvx_inv;
  float *mu_tend;
  const float* mu_tendf;
  int rk1;
};

__global__ void k_addtend(Box b, Dims D, AddT a, Box bu, Box bv, Box bw, Box bt) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  // every value is formed first and stored last, so that no load waits on a store
  const bool hu = bu.has(k, j, i), hv = bv.has(k, j, i), hw = bw.has(k, j, i), ht = bt.has(k, j, i);
  // the point's operands are loaded before the branches (a load under a
  // divergent branch is waited for before the branch ends); rk1 is uniform
  const float rut = a.ru_tend[q], rutf = a.ru_tendf[q], rvt = a.rv_tend[q], rvtf = a.rv_tendf[q];
  const float rwt = a.rw_tend[q], rwtf = a.rw_tendf[q], pht = a.ph_tend[q], phtf = a.ph_tendf[q];
  const float tt0 = a.t_tend[q], ttf = a.t_tendf[q], hd = a.h_diabatic[q];
  float us = 0.0f, vs = 0.0f, ws = 0.0f, ps = 0.0f, tsv = 0.0f;
  if (a.rk1) {
    us = a.u_save[q]; vs = a.v_save[q]; ws = a.w_save[q]; ps = a.ph_save[q]; tsv = a.t_save[q];
  }
  const float msfuy = AT2(a.msfuy, j, i), msfvx = AT2(a.msfvx, j, i), msfvx_inv = AT2(a.msfvx_inv, j, i);
  const float msfty = AT2(a.msfty, j, i), mut = AT2(a.mut, j, i);
  const float mt0 = AT2(a.mu_tend, j, i), mtf = AT2(a.mu_tendf, j, i);
  float fu = 0.0f, tu = 0.0f, fv = 0.0f, tv = 0.0f, fw = 0.0f, tw = 0.0f, fp = 0.0f, tp = 0.0f, ft = 0.0f, tt = 0.0f;
  if (hu) {
    fu = rutf;
    if (a.rk1) fu = fu + us * msfuy;
    tu = rut + fu / msfuy;
  }
  if (hv) {
    fv = rvtf;
    if (a.rk1) fv = fv + vs * msfvx;
    tv = rvt + fv * msfvx_inv;
  }
  if (hw) {
    fw = rwtf;
    if (a.rk1) fw = fw + ws * msfty;
    tw = rwt + fw / msfty;
    fp = phtf;
    if (a.rk1) fp = fp + ps;
    tp = pht + fp / msfty;
  }
  if (ht) {
    ft = ttf;
    if (a.rk1) ft = ft + tsv;
    tt = tt0 + ft / msfty + mut * hd / msfty;
  }
  const bool hm = k == bt.k0 && ht;
  const float tm = hm ? mt0 + mtf : 0.0f;
  if (hu) {
    if (a.rk1) a.ru_tendf[q] = fu;
    a.ru_tend[q] = tu;
  }
  if (hv) {
    if (a.rk1) a.rv_tendf[q] = fv;
    a.rv_tend[q] = tv;
  }
  if (hw) {
    if (a.rk1) {
      a.rw_tendf[q] = fw;
      a.ph_tendf[q] = fp;
    }
    a.rw_tend[q] = tw;
    a.ph_tend[q] = tp;
  }
  if (ht) {
    if (a.rk1) a.t_tendf[q] = ft;
    a.t_tend[q] = tt;
  }
  if (hm) AT2(a.mu_tend, j, i) = tm;
}

extern "C" int pywrf_rk_addtend_dry(const pywrf_grid_t* g, float* ru_tend, float* rv_tend, float* rw_tend,
                                    float* ph_tend, float* t_tend, float* ru_tendf, float* rv_tendf,
                                    float* rw_tendf, float* ph_tendf, float* t_tendf, const float* u_save,
                                    const float* v_save, const float* w_save, const float* ph_save,
                                    const float* t_save, float* mu_tend, const float* mu_tendf, int rk_step,
                                    const float* h_diabatic, const float* mut, const float* msfty,
                                    const float* msfuy, const float* msfvx, const float* msfvx_inv,
                                    pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kte > g->nz || g->kte < 2) return PYWRF_E_RANGE;
  AddT a = {ru_tend, rv_tend, rw_tend, ph_tend, t_tend, ru_tendf, rv_tendf, rw_tendf, ph_tendf, t_tendf,
            u_save, v_save, w_save, ph_save, t_save, h_diabatic, mut, msfty, msfuy, msfvx, msfvx_inv,
            mu_tend, mu_tendf, rk_step == 1};
  Box bu = mkbox(g->kts, g->kte - 1, g->jts, g->jde - 1, g->its, g->ite);
  Box bv = mkbox(g->kts, g->kte - 1, g->jts, g->jte, g->its, g->ide - 1);
  Box bw = mkbox(g->kts, g->kte, g->jts, g->jde - 1, g->its, g->ide - 1);
  Box bt = mkbox(g->kts, g->kte - 1, g->jts, g->jde - 1, g->its, g->ide - 1);
  Box b = mkbox(g->kts, g->kte, g->jts, imax(g->jde - 1, g->jte), g->its, imax(g->ite, g->ide - 1));
  PW_LAUNCH_BOX(k_addtend, b, s, dims_of(g), a, bu, bv, bw, bt);
  PW_RET();
}

extern "C" int pywrf_pgf_buoy_coriolis_curvature(
    const pywrf_grid_t* g, float* ru_tend, float* rv_tend, float* rw_tend, const float* ph, const float* alt,
    const float* p, const float* pb, const float* al, const float* php, const float* cqu, const float* cqv,
    float* cqw, const float* muu, const float* muv, const float* mu, const float* mub, const float* fnm,
    const float* fnp, const float* rdnw, const float* rdn, const float* cf, double g_, const float* ru,
    const float* rv, const float* rw, const float* u, const float* v, const float* msftx, const float* msfty,
    const float* msfux, const float* msfuy, const float* msfvx, const float* msfvy, const float* f, const float* e,
    const float* sina, const float* cosa, double rdx, double rdy, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 2);
  Hpg h;
  h.ph = ph; h.alt = alt; h.p = p; h.pb = pb; h.al = al; h.php = php; h.cqu = cqu; h.cqv = cqv; h.muu = muu;
  h.muv = muv; h.mu = mu; h.fnm = fnm; h.fnp = fnp; h.rdnw = rdnw; h.cf = cf; h.msfux = msfux; h.msfuy = msfuy;
  h.msfvx = msfvx; h.msfvy = msfvy; h.ru_tend = ru_tend; h.rv_tend = rv_tend; h.rdx = F(rdx); h.rdy = F(rdy);
  h.ktf = imin(g->kte, g->kde - 1);
  h.kde = g->kde;
  if (h.ktf + 1 > g->nz || g->kde > g->nz || g->kde < 3) return PYWRF_E_RANGE;
  Box hbv = mkbox(0, h.ktf, g->jts + 1, g->jte - 1, g->its, imin(g->ite, g->ide - 1));
  Box hbu = mkbox(0, h.ktf, g->jts, imin(g->jte, g->jde - 1), g->its + 1, g->ite - 1);
  Box hb = mkbox(0, h.ktf, g->jts, g->jte, g->its, g->ite);
  PgBuoy pg = {rw_tend, cqw, p, mu, mub, rdnw, rdn, msfty, F(g_), g->kde};
  Box pbox = mkbox(1, g->kde - 1, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  Cor a = {ru, rv, rw, msftx, msfty, msfux, msfuy, msfvx, msfvy, f, e, sina, cosa, fnm, fnp, ru_tend, rv_tend,
           rw_tend};
  Curv c = {ru, rv, rw, u, v, msfuy, msfvx, msfvy, msftx, msfty, fnm, fnp, ru_tend, rv_tend, rw_tend,
            F(rdx), F(rdy), g->jts, g->jte, g->its, g->ite};
  const int ktf = h.ktf;
  Box bu = mkbox(g->kts, ktf, g->jts, g->jte - 1, imax(g->ids + 1, g->its), imin(g->ide - 1, g->ite));
  Box bv = mkbox(g->kts, ktf, imax(g->jds + 1, g->jts), imin(g->jde - 1, g->jte), g->its, g->ite - 1);
  Box bw = mkbox(g->kts + 1, ktf, g->jts, g->jte - 1, g->its, g->ite - 1);
  Box cb = mkbox(g->kts, ktf, g->jts, g->jte, g->its, g->ite);
  // the launch box holds the three calls' boxes (pg_buoy_w's top-level write is made by its kde-2 thread)
  Box b = mkbox(imin(hb.k0, imin(pbox.k0, cb.k0)), imax(hb.k1, imax(pbox.k1, cb.k1)),
                imin(hb.j0, imin(pbox.j0, cb.j0)), imax(hb.j1, imax(pbox.j1, cb.j1)),
                imin(hb.i0, imin(pbox.i0, cb.i0)), imax(hb.i1, imax(pbox.i1, cb.i1)));
  PW_LAUNCH_BOX(k_pgf_buoy_cor, b, s, dims_of(g), h, hbv, hbu, pg, pbox, a, c, bu, bv, bw);
  PW_RET();
}
