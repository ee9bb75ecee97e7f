// WSM6 single-moment 6-class microphysics with semi-Lagrangian
// sedimentation, and the moist-physics prep / finish wrappers
// (pywrf/wrf_physics.py:27-1584, 1988-2137).
//
// Layout: one 64-lane wavefront per (j, i) column, lane k = level k.  Every
// per-level quantity of the scheme lives in the lane's registers; the column
// operations of the sedimentation (cumulative heights, the backward CFL
// limiter, the interval-overlap remap and the surface precipitation sum) go
// through LDS in the reference's summation order.  No full-grid scratch.
//
// Quirks kept because they fix the reference's numbers: the rain slope caps
// are used for snow and graupel, the ice number uses 5.387e7 before and
// 5.38e7 after sedimentation, paacw enters the cloud-water sink twice, pgacs
// is identically zero, the sedimentation fall speed update always uses the
// rain slope, the remap drops the top source cell and caps both top cells at
// 3800 m, and the saturation-adjustment latent heat is the one from the start
// of the call.
#include "common.h"
#include "consts.h"

#define AT(p, k, j, i) at_b((p), D.ob((k), (j), (i)))
#define AT2(p, j, i) at_b((p), D.ob2((j), (i)))

int pw_check_grid(const pywrf_grid_t* g, int r);

using namespace pw;

// torch.maximum / torch.minimum (NaN-propagating)
__device__ inline float tmax(float a, float b) { return (a != a) ? a : ((b != b) ? b : (a > b ? a : b)); }
__device__ inline float tmin(float a, float b) { return (a != a) ? a : ((b != b) ? b : (a < b ? a : b)); }

#define WAVE 64

// Fractional powers of the scheme.  x^(1/4) is two square roots and x^y is
// exp2(y log2 x): a few instructions instead of the general powf, with a
// relative error of a few 1e-7 for the magnitudes here (x^y and the
// reference's pow agree to far below the parity tolerance).  Zero, infinity,
// NaN and negative x behave as pow does for these positive, non-integer y.
__device__ inline float qroot(float x) { return sqrtf(sqrtf(x)); }
__device__ inline float fpow(float x, float y) { return exp2f(y * log2f(x)); }

struct Slope3 {
  float rs[3], rb[3], r2[3], r3[3], vt[3];
};

// slope_wsm6 (wrf_physics.py:2048-2103): rain caps for all three species
__device__ inline void slope3(float q0, float q1, float q2, float den, float denfac, float t, Slope3& s) {
  float supcol = F(t0c) - t;
  float n0sfac = tmax(tmin(expf(F(alpha) * supcol), F(n0smax / n0s)), 1.0f);
  float q[3] = {q0, q1, q2};
  const float bexp[3] = {F(bvtr), F(bvts), F(bvtg)}, pv[3] = {F(pvtr), F(pvts), F(pvtg)};
  for (int n = 0; n < 3; ++n) {
    bool small = q[n] <= F(qcrmin);
    float lam;
    if (n == 0) lam = qroot(F(pidn0r) / (q[n] * den));
    else if (n == 1) lam = qroot(F(pidn0s) * n0sfac / (q[n] * den));
    else lam = qroot(F(pidn0g) / (q[n] * den));
    float r = small ? F(rslopermax) : 1.0f / lam;
    s.rs[n] = r;
    s.rb[n] = small ? F(rsloperbmax) : fpow(r, bexp[n]);
    s.r2[n] = small ? F(rsloper2max) : r * r;
    s.r3[n] = small ? F(rsloper3max) : r * r * r;
    float v = pv[n] * s.rb[n] * denfac;
    s.vt[n] = q[n] <= 0.0f ? 0.0f : v;
  }
}

// slope_rain / slope_snow / slope_graup fall speed (wrf_physics.py:1988-2046)
__device__ inline float slope1(int kind, float qr, float den, float denfac, float t) {
  bool small = qr <= F(qcrmin);
  float r, rb, v;
  if (kind == 0) {
    r = small ? F(rslopermax) : 1.0f / qroot(F(pidn0r) / (qr * den));
    rb = small ? F(rsloperbmax) : fpow(r, F(bvtr));
    v = F(pvtr) * rb * denfac;
  } else if (kind == 1) {
    float supcol = F(t0c) - t;
    float n0sfac = tmax(tmin(expf(F(alpha) * supcol), F(n0smax / n0s)), 1.0f);
    r = small ? F(rslopermax) : 1.0f / qroot(F(pidn0s) * n0sfac / (qr * den));
    rb = small ? F(rsloperbmax) : fpow(r, F(bvts));
    v = F(pvts) * rb * denfac;
  } else {
    r = small ? F(rslopermax) : 1.0f / qroot(F(pidn0g) / (qr * den));
    rb = small ? F(rsloperbmax) : fpow(r, F(bvtg));
    v = F(pvtg) * rb * denfac;
  }
  return qr <= 0.0f ? 0.0f : v;
}

struct SedLds {
  float dz[WAVE + 1], ww[WAVE + 1], wi[WAVE + 1], za[WAVE + 1], dza[WAVE + 1], qa[WAVE + 1], qa2[WAVE + 1],
      zi[WAVE + 1];
  float red[2];
};

// Ordering point for an exchange through LDS among the lanes of one wave:
// a wave's LDS operations execute in issue order, so only the compiler must
// be kept from moving them across this point.  (A workgroup barrier would
// pair with the other waves' barriers when several waves each work on their
// own column.)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline float lane_value(float v, int m) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), m));
}

// nislfv_rain_plm (two == false) and nislfv_rain_plm6 (two == true) for one
// column; lane k holds level k, km levels are active.  Inputs dz, den,
// denfac, tk, ww (fall speed), q1 (and q2) are per-lane values (zero for
// k >= km); returns the new q1 (q2) in place and the surface precipitation
// on lane 0 in prec1 (prec2).  wrf_physics.py:1216-1584.
//
// The serial parts run as wave-wide operations: the sign test of the column
// sums is a ballot (the inputs are non-negative or NaN), the interface heights
// are each lane's in-order sum of the layers below it, the backward CFL
// limiter walks the lanes top-down with readlane, and the overlap remap sums,
// in the reference's summation order, only the source cells that can overlap
// the target cell when the departure heights are increasing (otherwise, or
// with non-finite values, every cell, as the reference does).
__device__ void sediment(SedLds& L, int k, int km, float dz, float den, float denfac, float tk, float ww, float& q1,
                         float& q2, bool two, float dt, int iter, float& prec1, float& prec2) {
  // sign of the all-level column sums: > 0 iff no NaN and some value > 0
  const bool anynan = __ballot((q1 != q1) || (two && q2 != q2)) != 0;
  const bool anypos = __ballot((q1 > 0.0f) || (two && q2 > 0.0f)) != 0;
  const bool cond = !anynan && anypos;
  L.dz[k] = dz;
  wave_sync();
  // heights of the level interfaces, summed bottom-up in order
  float zi = 0.0f;
  if (k <= km) {
#pragma unroll
    for (int m = 0; m < WAVE; ++m)
      if (m < k) zi = zi + L.dz[m];
  }
  L.zi[k] = zi;
  const float wd = ww;
  float was = 0.0f;
  float za = 0.0f, dza = 0.0f, qa = 0.0f, qa2 = 0.0f;
  int n = iter == 0 ? 0 : 1;
  while (n <= iter + 1) {
    // arrival-interface fall speeds
    wave_sync();
    L.ww[k] = ww;
    wave_sync();
    float wi = 0.0f;
    if (k <= km) {
      if (k == 0) wi = L.ww[0] + 0.0f;
      else if (k == 1) wi = 0.5f * (L.ww[1] + L.ww[0]);
      else if (k <= km - 2) wi = 0.5625f * (L.ww[k] + L.ww[k - 1]) - 0.0625f * (L.ww[k + 1] + L.ww[k - 2]);
      else if (k == km - 1) wi = 0.5f * (L.ww[km - 1] + L.ww[km - 2]);
      else wi = L.ww[km - 1] + 0.0f;
      if (k >= 1 && k < km && L.ww[k] == 0.0f) wi = L.ww[k - 1];
    }
    // backward CFL limiter from the top: level m takes the already limited
    // speed of level m + 1 (every lane runs the same uniform chain).  Levels
    // above the highest level that the limiter changes keep their speeds, and
    // that level is the highest one whose test already fires with the
    // unchanged speeds; so the test is first made on all levels at once and
    // the chain runs from there down (not at all in columns it never touches).
    const float wup = __shfl_down(wi, 1, WAVE);
    const unsigned long long fire = __ballot(k < km && (wup - wi) * dt / dz > 0.05f);
    if (fire) {
      const int mtop = 63 - __clzll(fire);
      float wn = lane_value(wi, mtop + 1), wnew = wi;
      for (int m = mtop; m >= 0; --m) {
        float wm = lane_value(wi, m);
        const float dzm = lane_value(dz, m);
        float decfl = (wn - wm) * dt / dzm;
        if (decfl > 0.05f) wm = wn - 0.05f * dzm / dt;
        wnew = k == m ? wm : wnew;
        wn = wm;
      }
      wi = wnew;
    }
    za = zi - wi * dt;
    wave_sync();
    L.za[k] = za;
    wave_sync();
    if (k < km) dza = L.za[k + 1] - za;
    else if (k == km) dza = zi - za;
    float qr = 0.0f, qr2 = 0.0f;
    if (k < km) {
      qa = q1 * dz / dza;
      qr = qa / den;
      if (two) {
        qa2 = q2 * dz / dza;
        qr2 = qa2 / den;
      }
    } else {
      qa = 0.0f;
      qa2 = 0.0f;
    }
    if (n <= iter) {
      float wa = 0.0f;
      if (!two) {
        if (k < km) wa = slope1(0, qr, den, denfac, tk);
      } else {
        float w1 = 0.0f, w2 = 0.0f;
        if (k < km) {
          w1 = slope1(1, qr, den, denfac, tk);
          w2 = slope1(2, qr2, den, denfac, tk);
        }
        float tmp = tmax(qr + qr2, 1.e-15f);
        wa = tmp > 1.e-15f ? (w1 * qr + w2 * qr2) / tmp : 0.0f;
      }
      if (n >= 2 && k < km) wa = 0.5f * (wa + was);
      ww = 0.5f * (wd + wa);
      was = wa + 0.0f;
    }
    n = n + 1;
    if (iter == 0) n = 2;
  }
  // remap onto the fixed grid; the top source cell is dropped
  wave_sync();
  L.dza[k] = dza;
  const float qs1 = (k == km - 1) ? 0.0f : qa, qs2 = (k == km - 1) ? 0.0f : qa2;
  L.qa[k] = qs1;
  L.qa2[k] = qs2;
  wave_sync();
  // departure heights increasing and all values finite: the overlaps of a
  // target cell are a contiguous run of source cells, the rest exactly +0
  const bool bad = (k < km - 1 && !(L.za[k + 1] > za)) || !isfinite(za) || !isfinite(qs1) || !isfinite(qs2);
  const bool ordered = __ballot(bad) == 0;
  auto zah_of = [&](int m) { return (m < km - 1) ? L.za[m + 1] : L.za[km - 1] + 3800.0f; };
  if (k < km) {
    const float zlo = L.zi[k];
    const float zhi = (k < km - 1) ? L.zi[k + 1] : L.zi[km - 1] + 3800.0f;
    auto ov = [&](int m) { return tmax(0.0f, tmin(zhi, zah_of(m)) - tmax(zlo, L.za[m])); };
    int lo = 0, hi = km - 1;
    if (ordered) {
      // first source cell whose top is above zlo, last whose bottom is below zhi
      int a = 0, b = km;
      while (a < b) { int c = (a + b) >> 1; if (zah_of(c) > zlo) b = c; else a = c + 1; }
      lo = a;
      a = 0; b = km;
      while (a < b) { int c = (a + b) >> 1; if (L.za[c] < zhi) a = c + 1; else b = c; }
      hi = a - 1;
    }
    float s1 = torch_outer_sum_range(km, lo, hi, [&](int m) { return L.qa[m] * ov(m); });
    float s2 = two ? torch_outer_sum_range(km, lo, hi, [&](int m) { return L.qa2[m] * ov(m); }) : 0.0f;
    q1 = cond ? s1 / dz : 0.0f;
    if (two) q2 = cond ? s2 / dz : 0.0f;
  }
  if (k == 0) {
    float p1 = prec1, p2 = prec2;
    for (int m = 0; m < km; ++m) {
      float z0 = L.za[m], z1 = L.za[m + 1];
      if (ordered && !(z0 < 0.0f)) break;   // increasing heights: no later cell reaches the ground
      if (z0 < 0.0f && z1 < 0.0f) {
        p1 = p1 + L.qa[m] * L.dza[m];
        if (two) p2 = p2 + L.qa2[m] * L.dza[m];
      }
      if (z0 < 0.0f && z1 > 0.0f) {
        p1 = p1 + L.qa[m] * (0.0f - z0);
        if (two) p2 = p2 + L.qa2[m] * (0.0f - z0);
      }
    }
    prec1 = cond ? p1 : 0.0f;
    prec2 = cond ? p2 : 0.0f;
  }
  wave_sync();
}

struct Wsm6 {
  float *th, *q, *qc, *qr, *qi, *qs, *qg;
  const float *den, *pii, *p, *delz;
  float *rain, *rainncv, *snow, *snowncv, *graupel, *graupelncv, *sr;
  float delt;
  int km, j0, j1, i0, i1, ncol, per;
  const int* cols;   // non-null: run only the columns cols[0 .. *ncols) (indices into the tile's column order)
  const int* ncols;
};

// rates limited so that a sink cannot remove more than value
#define LIMIT(value, source, cond, ...)                                    \
  do {                                                                     \
    float _v = (value), _s = (source);                                     \
    if (_s > _v && (cond)) {                                               \
      float _f = _v / _s;                                                  \
      float* _r[] = {__VA_ARGS__};                                         \
      for (unsigned _n = 0; _n < sizeof(_r) / sizeof(_r[0]); ++_n) *_r[_n] = *_r[_n] * _f; \
    }                                                                      \
  } while (0)

__device__ inline void fpvs(float t, float pp, float& s0, float& s1) {
  const double ttp = t0c + 0.01, xa = -(Cpv - cliq) / r_v, xb = xa + XLV / (r_v * ttp);
  const double xai = -(Cpv - cice) / r_v, xbi = xai + XLS / (r_v * ttp);
  float tr = F(ttp) / t;
  float liq = F(psat) * expf(logf(tr) * F(xa)) * expf(F(xb) * (1.0f - tr));
  float a = tmin(liq, 0.99f * pp);
  a = F(ep_2) * a / (pp - a);
  s0 = tmax(a, F(epsilon));
  float ice = F(psat) * expf(logf(tr) * F(xai)) * expf(F(xbi) * (1.0f - tr));
  float b = t < F(ttp) ? ice : liq;
  b = tmin(b, 0.99f * pp);
  b = F(ep_2) * b / (pp - b);
  s1 = tmax(b, F(epsilon));
}

__device__ inline float viscos(float x, float y) { return F(1.496e-6) * (x * sqrtf(x)) / (x + 120.0f) / y; }
__device__ inline float diffus(float x, float y) { return F(8.794e-5) * expf(logf(x) * 1.81f) / y; }
__device__ inline float xka(float x, float y) { return 1414.0f * viscos(x, y) * y; }
__device__ inline float diffac(float a, float b, float c, float dd, float e) {
  return dd * a * a / (xka(c, dd) * F(r_v) * c * c) + 1.0f / (e * diffus(c, b));
}
__device__ inline float venfac(float a, float b, float c) {
  return expf(logf(viscos(b, c) / diffus(b, a)) * 0.3333333f) / sqrtf(viscos(b, c)) * sqrtf(sqrtf(F(rhoair0) / c));
}

// One column of the scheme: lane k = level k (km levels active), the
// column's inputs per lane in `in` (the fields th, pii, den, delz, p, q, qc,
// qi, qr, qs, qg), its outputs per lane in `out` (th, q, qc, qi, qr, qs, qg;
// out.done is set on the lanes that produced them).  The surface
// precipitation of the column is written by lane 0.  L: the wave's LDS for
// the sedimentation.  Every lane of the wave must enter (the sedimentation
// uses wave-wide operations).
enum { CI_TH = 0, CI_PII, CI_DEN, CI_DZ, CI_P, CI_QV, CI_QC, CI_QI, CI_QR, CI_QS, CI_QG, CI_N };
struct ColVals {
  float v[CI_N];
  bool done;
};

__device__ __forceinline__ void wsm6_column(SedLds& L, const Dims& D, const Wsm6& a, int j, int i, int k,
                                            const ColVals& in, ColVals& out) {
  const int km = a.km;
  const bool act = k < km;
  out.done = false;
  const float qmin = F(epsilon), dtcld = a.delt;
  const float cpd = F(Cp), cpv = F(Cpv), xls = F(XLS), xlv0 = F(XLV), xlf0 = F(XLF), denr = F(rhowater);
  // column inputs (inactive lanes carry zeros, like the reference's zero-filled tiles)
  float th = act ? in.v[CI_TH] : 0.0f, pii = act ? in.v[CI_PII] : 1.0f;
  float d = act ? in.v[CI_DEN] : 0.0f, dz = act ? in.v[CI_DZ] : 0.0f, pp = act ? in.v[CI_P] : 0.0f;
  float qv = act ? in.v[CI_QV] : 0.0f;
  float t = th * pii;
  float qci0 = act ? tmax(in.v[CI_QC] + 0.0f, 0.0f) : 0.0f, qci1 = act ? tmax(in.v[CI_QI] + 0.0f, 0.0f) : 0.0f;
  float qrs0 = act ? tmax(in.v[CI_QR] + 0.0f, 0.0f) : 0.0f, qrs1 = act ? tmax(in.v[CI_QS] + 0.0f, 0.0f) : 0.0f;
  float qrs2 = act ? tmax(in.v[CI_QG] + 0.0f, 0.0f) : 0.0f;
  float cpm = qv >= qmin ? cpv * qv + (1.0f - qv) * cpd : F(Cp * (1 - epsilon) + Cpv * epsilon);
  float xl = xlv0 - F(xlv1) * (t - F(t0c));
  float denfac = sqrtf(1.0f / d * F(rhoair0));
  float qs0_, qs1_;
  fpvs(t, pp, qs0_, qs1_);
  float rh0 = tmax(qv / qs0_, qmin), rh1 = tmax(qv / qs1_, qmin);
  float xni = tmin(tmax(F(5.387e7) * fpow(d * tmax(qci1, qmin), 0.75f), 1.e3f), 1.e6f);
  (void)xni;
  Slope3 s;
  slope3(qrs0, qrs1, qrs2, d, denfac, t, s);
  float workr = s.vt[0];
  float qsum = tmax(qrs1 + qrs2, 1.e-15f);
  float worka = qsum > 1.e-15f ? (s.vt[1] * qrs1 + s.vt[2] * qrs2) / qsum : 0.0f;
  workr = qrs0 <= 0.0f ? 0.0f : workr;
  if (!act) { workr = 0.0f; worka = 0.0f; denfac = 0.0f; }
  // rain, then snow + graupel sedimentation
  float dq1 = d * qrs0, dq2 = d * qrs1, dq3 = d * qrs2;
  if (!act) { dq1 = dq2 = dq3 = 0.0f; }
  float tk = act ? t : 0.0f;
  float delqrs1 = 0.0f, unused = 0.0f, delqrs2 = 0.0f, delqrs3 = 0.0f;
  float z0 = 0.0f;
  sediment(L, k, km, dz, d, denfac, tk, workr, dq1, z0, false, dtcld, 1, delqrs1, unused);
  sediment(L, k, km, dz, d, denfac, tk, worka, dq2, dq3, true, dtcld, 1, delqrs2, delqrs3);
  qrs0 = tmax(dq1 / d, 0.0f);
  qrs1 = tmax(dq2 / d, 0.0f);
  qrs2 = tmax(dq3 / d, 0.0f);
  slope3(qrs0, qrs1, qrs2, d, denfac, t, s);
  // melting of snow and graupel
  const float xlf = xlf0 + 0.0f;
  float supcol = F(t0c) - t;
  float n0sfac = tmax(tmin(expf(F(alpha) * supcol), F(n0smax / n0s)), 1.0f);
  bool warm = t > F(t0c);
  float work2 = warm ? venfac(pp, t, d) : 0.0f;
  if (qrs1 > 0.0f && warm) {
    float coeres = s.r2[1] * sqrtf(s.rs[1] * s.rb[1]);
    float psmlt = xka(t, d) / xlf * (F(t0c) - t) * F(pi) / 2.0f * n0sfac *
                  (F(precs1) * s.r2[1] + F(precs2) * work2 * coeres) / d;
    psmlt = tmin(tmax(psmlt * dtcld / 1.0f, -qrs1 / 1.0f), 0.0f);
    qrs1 = qrs1 + psmlt;
    qrs0 = qrs0 - psmlt;
    t = t + xlf / cpm * psmlt;
  }
  if (qrs2 > 0.0f && warm) {
    float coeres = s.r2[2] * sqrtf(s.rs[2] * s.rb[2]);
    float pgmlt = xka(t, d) / xlf * (F(t0c) - t) * (F(precg1) * s.r2[2] + F(precg2) * work2 * coeres) / d;
    pgmlt = tmin(tmax(pgmlt * dtcld / 1.0f, -qrs2 / 1.0f), 0.0f);
    qrs2 = qrs2 + pgmlt;
    qrs0 = qrs0 - pgmlt;
    t = t + xlf / cpm * pgmlt;
  }
  // cloud-ice fall (one pass with the ice fall speed)
  {
    float xni0 = tmin(tmax(F(5.387e7) * fpow(d * tmax(qci1, qmin), 0.75f), 1.e3f), 1.e6f);
    float wt = tmax(tmin(F(dicon) * sqrtf(d * qci1 / xni0), F(dimax)), 1.e-25f);
    wt = 1.49e4f * expf(logf(wt) * 1.31f);
    float work1c = qci1 <= 0.0f ? 0.0f : wt;
    if (!act) work1c = 0.0f;
    float dqi = act ? d * qci1 : 0.0f, z1 = 0.0f, delqi = 0.0f, u2 = 0.0f;
    sediment(L, k, km, dz, d, denfac, act ? t : 0.0f, work1c, dqi, z1, false, dtcld, 0, delqi, u2);
    qci1 = tmax(dqi / d, 0.0f);
    if (k == 0) {
      long q2 = D.ix2(j, i);
      float fall0 = delqrs1 / dz / dtcld, fall1 = delqrs2 / dz / dtcld, fall2 = delqrs3 / dz / dtcld;
      float fallc0 = delqi / dz / dtcld;
      float fallsum = fall0 + fall1 + fall2 + fallc0;
      float fallsum_qsi = fall1 + fallc0, fallsum_qg = fall2 + 0.0f;
      auto acc = [&](float fs) { return fs * dz / denr * dtcld * 1000.0f; };
      float rainncv = 0.0f, snowncv = 0.0f, graupelncv = 0.0f, sr = 0.0f;
      if (fallsum > 0.0f) {
        rainncv = acc(fallsum) + rainncv;
        a.rain[q2] = acc(fallsum) + a.rain[q2];
      }
      if (fallsum_qsi > 0.0f) {
        snowncv = acc(fallsum_qsi) + snowncv;
        a.snow[q2] = acc(fallsum_qsi) + a.snow[q2];
      }
      if (fallsum_qg > 0.0f) {
        graupelncv = acc(fallsum_qg) + graupelncv;
        a.graupel[q2] = acc(fallsum_qg) + a.graupel[q2];
      }
      if (fallsum > 0.0f) sr = (snowncv + graupelncv) / (rainncv + 1.e-12f);
      a.rainncv[q2] = rainncv;
      a.snowncv[q2] = snowncv;
      a.graupelncv[q2] = graupelncv;
      a.sr[q2] = sr;
    }
  }
  if (!act) return;  // no further column operations
  // instantaneous melting and freezing
  supcol = F(t0c) - t;
  float xlf3 = xls - xl;
  if (supcol < 0.0f && qci1 > 0.0f) {
    qci0 = qci0 + qci1;
    t = t - xlf3 / cpm * qci1;
    qci1 = 0.0f;
  }
  if (supcol > 40.0f && qci0 > 0.0f) {
    qci1 = qci1 + qci0;
    t = t + xlf3 / cpm * qci0;
    qci0 = 0.0f;
  }
  {
    float pfrzdtc = tmin(F(pfrz1) * (expf(F(pfrz2) * tmin(supcol, 50.0f)) - 1.0f) * d / denr / F(xncr) * qci0 *
                             qci0 * dtcld,
                         qci0);
    if (supcol > 0.0f && qci0 > qmin) {
      qci1 = qci1 + pfrzdtc;
      t = t + xlf3 / cpm * pfrzdtc;
      qci0 = qci0 - pfrzdtc;
    }
  }
  {
    float pfrzdtr = tmin(F(20. * pi * pi * pfrz1 * n0r * rhowater) / d *
                             (expf(F(pfrz2) * tmin(supcol, 50.0f)) - 1.0f) * s.r3[0] * s.r3[0] * s.rs[0] * dtcld,
                         qrs0);
    if (supcol > 0.0f && qrs0 > 0.0f) {
      qrs2 = qrs2 + pfrzdtr;
      t = t + xlf3 / cpm * pfrzdtr;
      qrs0 = qrs0 - pfrzdtr;
    }
  }
  slope3(qrs0, qrs1, qrs2, d, denfac, t, s);
  float work1_0 = diffac(xl, pp, t, d, qs0_);
  float work1_1 = diffac(xls, pp, t, d, qs1_);
  work2 = venfac(pp, t, d);
  // warm rain
  float supsat = tmax(qv, qmin) - qs0_;
  float praut = F(qck1) * fpow(qci0, F(7. / 3.));
  praut = qci0 > F(qc0) ? tmin(praut, qci0 / dtcld) : 0.0f;
  float pracw = 0.0f;
  if (qrs0 > F(qcrmin) && qci0 > qmin) pracw = tmin(F(pacrr) * s.r3[0] * s.rb[0] * qci0 * denfac, qci0 / dtcld);
  float prevp = 0.0f;
  {
    bool c = qrs0 > 0.0f;
    float coeres = s.r2[0] * sqrtf(s.rs[0] * s.rb[0]);
    if (c) prevp = (rh0 - 1.0f) * (F(precr1) * s.r2[0] + F(precr2) * work2 * coeres) / work1_0;
    bool c2 = prevp < 0.0f && c, c3 = prevp >= 0.0f && c;
    if (c2) prevp = tmax(prevp, -qrs0 / dtcld);
    if (c2) prevp = tmax(prevp, supsat / dtcld / 2.0f);
    if (c3) prevp = tmin(prevp, supsat / dtcld / 2.0f);
  }
  // cold rain
  supcol = F(t0c) - t;
  n0sfac = tmax(tmin(expf(F(alpha) * supcol), F(n0smax / n0s)), 1.0f);
  supsat = tmax(qv, qmin) - qs1_;
  float satdt = supsat / dtcld;
  int ifsat = 0;
  xni = tmin(tmax(F(5.38e7) * fpow(d * tmax(qci1, qmin), 0.75f), 1.e3f), 1.e6f);
  float eacrs = expf(0.07f * (-supcol));
  float xmi = d * qci1 / xni;
  float diameter = tmin(F(dicon) * sqrtf(xmi), F(dimax));
  float vt2i = 1.49e4f * fpow(diameter, 1.31f);
  float vt2r = F(pvtr) * s.rb[0] * denfac, vt2s = F(pvts) * s.rb[1] * denfac, vt2g = F(pvtg) * s.rb[2] * denfac;
  qsum = tmax(qrs1 + qrs2, 1.e-15f);
  float vt2ave = qsum > 1.e-15f ? (vt2s * qrs1 + vt2g * qrs2) / qsum : 0.0f;
  bool cc = supcol > 0.0f && qci1 > qmin;
  float praci = 0.0f, piacr = 0.0f, psaci = 0.0f, pgaci = 0.0f;
  {
    bool c1 = cc && qrs0 > F(qcrmin);
    float acr = 2.0f * s.r3[0] + 2.0f * diameter * s.r2[0] + diameter * diameter * s.rs[0];
    float v = F(pi) * qci1 * F(n0r) * fabsf(vt2r - vt2i) * acr;
    praci = c1 ? tmin(v, qci1 / dtcld) : 0.0f;
    v = F(pi * pi * avtr * n0r * rhowater) * xni * denfac * F(g6pbr) * (s.r3[0] * s.r3[0]) * s.rb[0] / 24.0f / d;
    piacr = c1 ? tmin(v, qrs0 / dtcld) : 0.0f;
    c1 = qrs1 > F(qcrmin) && cc;
    acr = 2.0f * s.r3[1] + 2.0f * diameter * s.r2[1] + diameter * diameter * s.rs[1];
    v = F(pi) * qci1 * eacrs * F(n0s) * n0sfac * fabsf(vt2ave - vt2i) * acr / 4.0f;
    psaci = c1 ? tmin(v, qci1 / dtcld) : 0.0f;
    c1 = qrs2 > F(qcrmin) && cc;
    float egi = expf(-0.07f * supcol);
    acr = 2.0f * s.r3[2] + 2.0f * diameter * s.r2[2] + diameter * diameter * s.rs[2];
    v = F(pi) * egi * qci1 * F(n0g) * fabsf(vt2ave - vt2i) * acr / 4.0f;
    pgaci = c1 ? tmin(v, qci1 / dtcld) : 0.0f;
  }
  float psacw = 0.0f, pgacw = 0.0f;
  if (qrs1 > F(qcrmin) && qci0 > qmin)
    psacw = tmin(F(pacrc) * n0sfac * s.r3[1] * s.rb[1] * qci0 * denfac, qci0 / dtcld);
  if (qrs2 > F(qcrmin) && qci0 > qmin) pgacw = tmin(F(pacrc) * s.r3[2] * s.rb[2] * qci0 * denfac, qci0 / dtcld);
  float paacw = qsum > 1.e-15f ? (qrs1 * psacw + qrs2 * pgacw) / qsum : 0.0f;
  float pracs = 0.0f, psacr = 0.0f, pgacr = 0.0f, pgacs = 0.0f;
  {
    bool c = qrs1 > F(qcrmin) && qrs0 > F(qcrmin);
    bool c1 = supcol > 0.0f && c;
    float acr = 5.0f * (s.r3[1] * s.r3[1]) * s.rs[0] + 2.0f * s.r3[1] * s.r2[1] * s.r2[0] +
                0.5f * (s.r2[1] * s.r2[1]) * s.r3[0];
    float v = F(pi * pi * n0r * n0s) * n0sfac * fabsf(vt2r - vt2ave) * (F(dens) / d) * acr;
    pracs = c1 ? tmin(v, qrs1 / dtcld) : 0.0f;
    acr = 5.0f * (s.r3[0] * s.r3[0]) * s.rs[1] + 2.0f * s.r3[0] * s.r2[0] * s.r2[1] +
          0.5f * (s.r2[0] * s.r2[0]) * s.r3[1];
    v = F(pi * pi * n0r * n0s) * n0sfac * fabsf(vt2r - vt2ave) * (denr / d) * acr;
    psacr = c ? tmin(v, qrs0 / dtcld) : 0.0f;
    c = qrs2 > F(qcrmin) && qrs0 > F(qcrmin);
    acr = 5.0f * (s.r3[0] * s.r3[0]) * s.rs[2] + 2.0f * s.r3[0] * s.r2[0] * s.r2[2] +
          0.5f * (s.r2[0] * s.r2[0]) * s.r3[2];
    v = F(pi * pi * n0r * n0g) * fabsf(vt2r - vt2ave) * (denr / d) * acr;
    pgacr = c ? tmin(v, qrs0 / dtcld) : 0.0f;
  }
  float pseml = 0.0f, pgeml = 0.0f;
  if (supcol <= 0.0f) {
    if (qrs1 > 0.0f) pseml = tmin(tmax(F(cliq) * supcol * (paacw + psacr) / xlf0, -qrs1 / dtcld), 0.0f);
    if (qrs2 > 0.0f) pgeml = tmin(tmax(F(cliq) * supcol * (paacw + psacr) / xlf0, -qrs2 / dtcld), 0.0f);
  }
  // deposition / sublimation, with the ifsat chain
  float pidep = 0.0f, psdep = 0.0f, pgdep = 0.0f, pigen = 0.0f, psaut = 0.0f, pgaut = 0.0f;
  cc = supcol > 0.0f;
  {
    bool c1 = qci1 > 0.0f && ifsat != 1 && cc;
    if (c1) {
      pidep = 4.0f * diameter * xni * (rh1 - 1.0f) / work1_1;
      float supice = satdt - prevp;
      if (pidep < 0.0f) pidep = tmax(tmax(tmax(pidep, satdt / 2.0f), supice), -qci1 / dtcld);
      else pidep = tmin(tmin(pidep, satdt / 2.0f), supice);
      if (fabsf(prevp + pidep) >= fabsf(satdt)) ifsat = 1;
    }
    c1 = qrs1 > 0.0f && ifsat != 1 && cc;
    if (c1) {
      float coeres = s.r2[1] * sqrtf(s.rs[1] * s.rb[1]);
      psdep = (rh1 - 1.0f) * n0sfac * (F(precs1) * s.r2[1] + F(precs2) * work2 * coeres) / work1_1;
      float supice = satdt - prevp - pidep;
      if (psdep < 0.0f) psdep = tmax(tmax(tmax(psdep, -qrs1 / dtcld), satdt / 2.0f), supice);
      else psdep = tmin(tmin(psdep, satdt / 2.0f), supice);
      if (fabsf(prevp + pidep + psdep) >= fabsf(satdt)) ifsat = 1;
    }
    c1 = qrs2 > 0.0f && ifsat != 1 && cc;
    if (c1) {
      float coeres = s.r2[2] * sqrtf(s.rs[2] * s.rb[2]);
      pgdep = (rh1 - 1.0f) * (F(precg1) * s.r2[2] + F(precg2) * work2 * coeres) / work1_1;
      float supice = satdt - prevp - pidep - psdep;
      if (pgdep < 0.0f) pgdep = tmax(tmax(tmax(pgdep, -qrs2 / dtcld), satdt / 2.0f), supice);
      else pgdep = tmin(tmin(pgdep, satdt / 2.0f), supice);
      if (fabsf(prevp + pidep + psdep + pgdep) >= fabsf(satdt)) ifsat = 1;
    }
    c1 = supsat > 0.0f && ifsat != 1 && cc;
    float supice = satdt - prevp - pidep - psdep - pgdep;
    float roqi0 = F(4.92e-11) * fpow(1.e3f * expf(0.1f * supcol), 1.33f);
    float pg = tmax(0.0f, (roqi0 / d - tmax(qci1, 0.0f)) / dtcld);
    pigen = c1 ? tmin(tmin(pg, satdt), supice) : 0.0f;
    if (qci1 > 0.0f && cc) psaut = tmax(0.0f, (qci1 - F(roqimax) / d) / dtcld);
    if (qrs1 > 0.0f && cc) pgaut = tmin(tmax(0.0f, 1.e-3f * expf(-0.09f * supcol) * (qrs1 - F(qs0))), qrs1 / dtcld);
  }
  float psevp = 0.0f, pgevp = 0.0f;
  cc = supcol < 0.0f;
  {
    float coeres = s.r2[1] * sqrtf(s.rs[1] * s.rb[1]);
    float v = (rh0 - 1.0f) * n0sfac * (F(precs1) * s.r2[1] + F(precs2) * work2 * coeres) / work1_0;
    psevp = (qrs1 > 0.0f && rh0 < 1.0f && cc) ? tmin(tmax(v, -qrs1 / dtcld), 0.0f) : 0.0f;
    coeres = s.r2[2] * sqrtf(s.rs[2] * s.rb[2]);
    v = (rh0 - 1.0f) * (F(precg1) * s.r2[2] + F(precg2) * work2 * coeres) / work1_0;
    pgevp = (qrs2 > 0.0f && rh0 < 1.0f && cc) ? tmin(tmax(v, -qrs2 / dtcld), 0.0f) : 0.0f;
  }
  // mass-conservation limiters and the update, T <= T0
  float delta2 = (qrs0 < 1.e-4f && qrs1 < 1.e-4f) ? 1.0f : 0.0f;
  float delta3 = qrs0 < 1.e-4f ? 1.0f : 0.0f;
  cc = t <= F(t0c);
  LIMIT(tmax(qmin, qci0), (praut + pracw + paacw + paacw) * dtcld, cc, &praut, &pracw, &paacw);
  LIMIT(tmax(qmin, qci1), (psaut - pigen - pidep + praci + psaci + pgaci) * dtcld, cc, &psaut, &pigen, &pidep,
        &praci, &psaci, &pgaci);
  LIMIT(tmax(qmin, qrs0), (-praut - prevp - pracw + piacr + psacr + pgacr) * dtcld, cc, &praut, &prevp, &pracw,
        &piacr, &psacr, &pgacr);
  LIMIT(tmax(qmin, qrs1),
        -(psdep + psaut - pgaut + paacw + piacr * delta3 + praci * delta3 - pracs * (1.0f - delta2) +
          psacr * delta2 + psaci - pgacs) *
            dtcld,
        cc, &psdep, &psaut, &pgaut, &paacw, &piacr, &praci, &psaci, &pracs, &psacr, &pgacs);
  LIMIT(tmax(qmin, qrs2),
        -(pgdep + pgaut + piacr * (1.0f - delta3) + praci * (1.0f - delta3) + psacr * (1.0f - delta2) +
          pracs * (1.0f - delta2) + pgaci + paacw + pgacr + pgacs) *
            dtcld,
        cc, &pgdep, &pgaut, &piacr, &praci, &psacr, &pracs, &paacw, &pgaci, &pgacr, &pgacs);
  float qvn = qv;
  if (cc) {
    work2 = -(prevp + psdep + pgdep + pigen + pidep);
    qvn = qv + work2 * dtcld;
    qci0 = tmax(qci0 - (praut + pracw + paacw + paacw) * dtcld, 0.0f);
    qrs0 = tmax(qrs0 + (praut + pracw + prevp - piacr - pgacr - psacr) * dtcld, 0.0f);
    qci1 = tmax(qci1 - (psaut + praci + psaci + pgaci - pigen - pidep) * dtcld, 0.0f);
    qrs1 = tmax(qrs1 + (psdep + psaut + paacw - pgaut + piacr * delta3 + praci * delta3 + psaci - pgacs -
                        pracs * (1.0f - delta2) + psacr * delta2) *
                           dtcld,
                0.0f);
    qrs2 = tmax(qrs2 + (pgdep + pgaut + piacr * (1.0f - delta3) + praci * (1.0f - delta3) + psacr * (1.0f - delta2) +
                        pracs * (1.0f - delta2) + pgaci + paacw + pgacr + pgacs) *
                           dtcld,
                0.0f);
    xlf3 = xls - xl;
    float xlwork2 = F(-XLS) * (psdep + pgdep + pidep + pigen) - xl * prevp -
                    xlf3 * (piacr + paacw + paacw + pgacr + psacr);
    t = t - xlwork2 / cpm * dtcld;
  }
  // T > T0
  cc = t > F(t0c);
  LIMIT(tmax(qmin, qci0), (praut + pracw + paacw + paacw) * dtcld, cc, &praut, &pracw, &paacw);
  LIMIT(tmax(qmin, qrs0), (-paacw - praut + pseml + pgeml - pracw - paacw - prevp) * dtcld, cc, &praut, &prevp,
        &pracw, &paacw, &pseml, &pgeml);
  LIMIT(tmax(F(qcrmin), qrs1), (pgacs - pseml - psevp) * dtcld, cc, &pgacs, &psevp, &pseml);
  LIMIT(tmax(F(qcrmin), qrs2), -(pgacs + pgevp + pgeml) * dtcld, cc, &pgacs, &pgevp, &pgeml);
  if (cc) {
    work2 = -(prevp + psevp + pgevp);
    qvn = qvn + work2 * dtcld;
    qci0 = tmax(qci0 - (praut + pracw + paacw + paacw) * dtcld, 0.0f);
    qrs0 = tmax(qrs0 + (praut + pracw + prevp + paacw + paacw - pseml - pgeml) * dtcld, 0.0f);
    qrs1 = tmax(qrs1 + (psevp - pgacs + pseml) * dtcld, 0.0f);
    qrs2 = tmax(qrs2 + (pgacs + pgevp + pgeml) * dtcld, 0.0f);
    xlf3 = xls - xl;
    float xlwork2 = -xl * (prevp + psevp + pgevp) - xlf3 * (pseml + pgeml);
    t = t - xlwork2 / cpm * dtcld;
  }
  // saturation adjustment
  fpvs(t, pp, qs0_, qs1_);
  float w1 = (tmax(qvn, qmin) - qs0_) / (1.0f + xl * xl / (F(r_v) * cpm) * qs0_ / (t * t));
  float pcond = tmin(tmax(w1 / dtcld, 0.0f), tmax(qvn, 0.0f) / dtcld);
  if (qci0 > 0.0f && w1 < 0.0f) pcond = tmax(w1, -qci0) / dtcld;
  qvn = qvn - pcond * dtcld;
  qci0 = tmax(qci0 + pcond * dtcld, 0.0f);
  t = t + pcond * xl / cpm * dtcld;
  qci0 = qci0 <= qmin ? 0.0f : qci0;
  qci1 = qci1 <= qmin ? 0.0f : qci1;
  out.v[CI_QV] = qvn;
  out.v[CI_TH] = t / pii;
  out.v[CI_QC] = qci0;
  out.v[CI_QI] = qci1;
  out.v[CI_QR] = qrs0;
  out.v[CI_QS] = qrs1;
  out.v[CI_QG] = qrs2;
  out.done = true;
}


// One 64-lane wave per column (lane = level), reading and writing the fields
// in place.  Columns are dealt to the XCDs in contiguous runs: XCD x takes
// [x * per, (x + 1) * per) of the tile's order -- or of the device list, with
// per formed from the list's length on the device, so that a short list
// (the surrogate's fallback columns) is spread over all XCDs, not left to the
// first -- and the columns sharing a cache line are served by one L2.  The
// launch is sized for the whole tile; waves beyond the run return at once.
// (A wave walking several columns in a loop needs 107 VGPRs instead of 64
// and halves the occupancy: 408 against 283 us per call at 240 x 240.)
__global__ void __launch_bounds__(WAVE) k_wsm6(Dims D, Wsm6 a) {
  __shared__ SedLds L;
  const int k = (int)threadIdx.x;
  const int total = a.cols ? imin(*a.ncols, a.ncol) : a.ncol;
  const int per = a.cols ? (total + PW_XCDS - 1) / PW_XCDS : a.per;
  const int x = (int)(blockIdx.x % PW_XCDS);
  const int cc = x * per + (int)(blockIdx.x / PW_XCDS);
  if (cc >= imin(x * per + per, total)) return;
  int c = cc;
  if (a.cols) {
    c = a.cols[cc];
    if (c < 0 || c >= a.ncol) return;
  }
  const int ni = a.i1 - a.i0;
  const int j = a.j0 + c / ni, i = a.i0 + c % ni;
  const bool act = k < a.km;
  const long q = act ? D.ix(k, j, i) : D.ix(0, j, i);
  const float* src[CI_N] = {a.th, a.pii, a.den, a.delz, a.p, a.q, a.qc, a.qi, a.qr, a.qs, a.qg};
  ColVals in, out;
#pragma unroll
  for (int f = 0; f < CI_N; ++f) in.v[f] = act ? src[f][q] : 0.0f;
  wsm6_column(L, D, a, j, i, k, in, out);
  if (out.done) {
    a.q[q] = out.v[CI_QV];
    a.th[q] = out.v[CI_TH];
    a.qc[q] = out.v[CI_QC];
    a.qi[q] = out.v[CI_QI];
    a.qr[q] = out.v[CI_QR];
    a.qs[q] = out.v[CI_QS];
    a.qg[q] = out.v[CI_QG];
  }
}

static int wsm6_launch(const pywrf_grid_t* g, float* th, float* q, float* qc, float* qr, float* qi, float* qs,
                       float* qg, const float* den, const float* pii, const float* p, const float* delz, double delt,
                       float* rain, float* rainncv, float* snow, float* snowncv, float* sr, float* graupel,
                       float* graupelncv, const int* cols, const int* ncols, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  int e = pw_check_grid(g, 0);
  if (e) return e;
  Wsm6 a;
  a.th = th; a.q = q; a.qc = qc; a.qr = qr; a.qi = qi; a.qs = qs; a.qg = qg;
  a.den = den; a.pii = pii; a.p = p; a.delz = delz;
  a.rain = rain; a.rainncv = rainncv; a.snow = snow; a.snowncv = snowncv; a.graupel = graupel;
  a.graupelncv = graupelncv; a.sr = sr;
  double dtcld = delt <= dtcldcr ? delt + 0.0 : 1.0 * delt / (delt / dtcldcr > 1 ? delt / dtcldcr : 1);
  a.delt = F(dtcld);
  // the scheme's tile: [kts, kte-1) x [jts, jte-1) x [its, ite-1)
  a.km = (g->kte - 1) - g->kts;
  a.j0 = g->jts; a.j1 = g->jte - 1; a.i0 = g->its; a.i1 = g->ite - 1;
  if (g->kts != 0 || a.km < 4 || a.km + 1 > WAVE || g->kte > g->nz) return PYWRF_E_RANGE;
  long ncol = (long)(a.j1 - a.j0) * (a.i1 - a.i0);
  if (ncol <= 0) return 0;
  a.ncol = (int)ncol;
  a.per = (int)((ncol + PW_XCDS - 1) / PW_XCDS);
  a.cols = cols;
  a.ncols = ncols;
  hipLaunchKernelGGL(k_wsm6, dim3((unsigned)(a.per * PW_XCDS)), dim3(WAVE), 0, (hipStream_t)s, dims_of(g), a);
  return (int)hipGetLastError();
}

extern "C" int pywrf_wsm6(const pywrf_grid_t* g, float* th, float* q, float* qc, float* qr, float* qi, float* qs,
                          float* qg, const float* den, const float* pii, const float* p, const float* delz,
                          double delt, float* rain, float* rainncv, float* snow, float* snowncv, float* sr,
                          float* graupel, float* graupelncv, pywrf_stream_t s) {
  return wsm6_launch(g, th, q, qc, qr, qi, qs, qg, den, pii, p, delz, delt, rain, rainncv, snow, snowncv, sr, graupel,
                     graupelncv, nullptr, nullptr, s);
}

// WSM6 on a device list of the tile's columns (c = (j - jts) * (ite - 1 - its) +
// (i - its)), cols[0 .. *ncols): the surrogate's fallback.  Columns outside
// the list are untouched; each listed column gets exactly the plain scheme.
extern "C" int pywrf_wsm6_columns(const pywrf_grid_t* g, float* th, float* q, float* qc, float* qr, float* qi,
                                  float* qs, float* qg, const float* den, const float* pii, const float* p,
                                  const float* delz, double delt, float* rain, float* rainncv, float* snow,
                                  float* snowncv, float* sr, float* graupel, float* graupelncv, const int* cols,
                                  const int* ncols, pywrf_stream_t s) {
  if (!cols || !ncols) return PYWRF_E_ARG;
  return wsm6_launch(g, th, q, qc, qr, qi, qs, qg, den, pii, p, delz, delt, rain, rainncv, snow, snowncv, sr, graupel,
                     graupelncv, cols, ncols, s);
}

extern "C" int pywrf_microphysics_driver(const pywrf_grid_t* g, float* th, const float* rho, const float* pi_phy,
                                         const float* p, const float* dz8w, double dt, float* qv, float* qc,
                                         float* qr, float* qi, float* qs, float* qg, float* rainnc, float* rainncv,
                                         float* snownc, float* snowncv, float* sr, float* graupelnc,
                                         float* graupelncv, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return pywrf_wsm6(g, th, qv, qc, qr, qi, qs, qg, rho, pi_phy, p, dz8w, dt, rainnc, rainncv, snownc, snowncv, sr,
                    graupelnc, graupelncv, s);
}

// ------------------------------------------------------------------ moist_physics_prep_em
struct MpPrep {
  const float *t_new, *al, *alb, *p, *pb, *ph, *phb, *qv, *qc, *fzm, *fzp;
  float *rho, *th_phy, *pii, *pf, *z, *z_at_w, *dz8w, *p8w, *h_diabatic, *qv_diabatic, *qc_diabatic;
  float t0, p0;
  int kte, kde, k1;
};

// One thread per point of levels [0, kte): every output is local in k except
// the neighbours' z_at_w and full pressure, which the thread recomputes with
// the same operations (so the values are those the column sequence of the
// reference produces).  The k = 0 thread also writes the extrapolated surface
// and top rows of p8w.
__global__ void k_mp_prep(Box b, Dims D, MpPrep a) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  const float G = F(pw::g);
  // levels this call writes are recomputed; others (only for unusual index
  // sets, kde > kte + 1) are read as stored, as the reference reads them
  auto ZW = [&](int kk) { return kk < a.kte ? (AT(a.ph, kk, j, i) + AT(a.phb, kk, j, i)) / G : AT(a.z_at_w, kk, j, i); };
  auto PF = [&](int kk) { return kk < a.k1 ? AT(a.p, kk, j, i) + AT(a.pb, kk, j, i) : AT(a.pf, kk, j, i); };
  auto ZM = [&](int kk) { return 0.5f * (ZW(kk) + ZW(kk + 1)); };
  const long q = D.ix(k, j, i);
  const float zw = ZW(k);
  a.z_at_w[q] = zw;
  a.dz8w[q] = k < a.kte - 1 ? ZW(k + 1) - zw : 0.0f;
  if (k < a.k1) {
    float th = a.t_new[q] + a.t0;
    a.th_phy[q] = th;
    a.h_diabatic[q] = th + 0.0f;
    a.qv_diabatic[q] = a.qv[q] + 0.0f;
    a.qc_diabatic[q] = a.qc[q] + 0.0f;
    a.rho[q] = 1.0f / (a.al[q] + a.alb[q]);
    float pf = PF(k);
    a.pii[q] = fpow(pf / a.p0, F(pw::rcp));
    a.z[q] = 0.5f * (zw + ZW(k + 1));
    a.pf[q] = pf;
    if (k >= 1) a.p8w[q] = a.fzm[k] * pf + a.fzp[k] * PF(k - 1);
  }
  if (k == 0) {
    float w1 = (zw - ZM(1)) / (ZM(0) - ZM(1));
    float w2 = 1.0f - w1;
    AT(a.p8w, 0, j, i) = w1 * PF(0) + w2 * PF(1);
    float u1 = (ZW(a.kte - 1) - ZM(a.k1 - 2)) / (ZM(a.k1 - 1) - ZM(a.k1 - 2));
    float u2 = 1.0f - u1;
    AT(a.p8w, a.kde - 1, j, i) = expf(u1 * logf(PF(a.kde - 2)) + u2 * logf(PF(a.kde - 3)));
  }
}

extern "C" int pywrf_moist_physics_prep_em(const pywrf_grid_t* g, const float* t_new, double t0, float* rho,
                                           const float* al, const float* alb, const float* p, float* p8w, double p0,
                                           const float* pb, const float* ph, const float* phb, float* th_phy,
                                           float* pii, float* pf, float* z, float* z_at_w, float* dz8w,
                                           float* h_diabatic, const float* qv, float* qv_diabatic, const float* qc,
                                           float* qc_diabatic, const float* fzm, const float* fzp,
                                           pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  int e = pw_check_grid(g, 0);
  if (e) return e;
  if (g->kts != 0 || g->kte > g->nz || g->kde > g->nz || g->kte < 4) return PYWRF_E_RANGE;
  MpPrep a = {t_new, al, alb, p, pb, ph, phb, qv, qc, fzm, fzp, rho, th_phy, pii, pf, z, z_at_w, dz8w, p8w,
              h_diabatic, qv_diabatic, qc_diabatic, F(t0), F(p0), g->kte, g->kde, imin(g->kte, g->kde - 1)};
  Box b = mkbox(0, g->kte, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  PW_LAUNCH_BOX(k_mp_prep, b, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ moist_physics_finish_em
__global__ void k_mp_finish(Box b, Dims D, float* t_new, const float* th_phy, float* h_diabatic, const float* qv,
                            float* qv_diabatic, const float* qc, float* qc_diabatic, float dt) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  float mpten = th_phy[q] - h_diabatic[q];
  float qvten = qv[q] - qv_diabatic[q];
  float qcten = qc[q] - qc_diabatic[q];
  t_new[q] = t_new[q] + mpten;
  h_diabatic[q] = mpten / dt;
  qv_diabatic[q] = qvten / dt;
  qc_diabatic[q] = qcten / dt;
}

extern "C" int pywrf_moist_physics_finish_em(const pywrf_grid_t* g, float* t_new, const float* th_phy,
                                             float* h_diabatic, double dt, const float* qv, float* qv_diabatic,
                                             const float* qc, float* qc_diabatic, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  int e = pw_check_grid(g, 0);
  if (e) return e;
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  PW_LAUNCH_BOX(k_mp_finish, b, s, dims_of(g), t_new, th_phy, h_diabatic, qv, qv_diabatic, qc, qc_diabatic, F(dt));
  PW_RET();
}
