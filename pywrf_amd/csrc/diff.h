// Horizontal and vertical diffusion at one point, shared by the diffusion
// entry points and the fused scalar-transport kernel
// (pywrf/wrf_dynamics.py:2623-3086).
#pragma once
#include "common.h"
#include "consts.h"

#ifndef AT
#define AT(p, k, j, i) at_b((p), D.ob((k), (j), (i)))
#define AT2(p, j, i) at_b((p), D.ob2((j), (i)))
#endif

struct HDiff {
  const float *f, *base, *mut, *xk, *msfux, *msfuy, *msfvx, *msfvx_inv, *msfvy, *msftx, *msfty;
  float* tend;
  float rdx, rdy;
  char name;
};

struct HCoef {
  float mkxm, mkxp, mrdx, mkym, mkyp, mrdy;
};

__device__ inline HCoef hdiff_coefs(const Dims& D, const HDiff& a, int k, int j, int i) {
  HCoef c;
  auto M = [&](const float* f, int dj, int di) { return AT2(f, j + dj, i + di); };
  auto X = [&](int dk, int dj, int di) { return AT(a.xk, k + dk, j + dj, i + di); };
  const float rdx = a.rdx, rdy = a.rdy;
  if (a.name == 'u') {
    c.mkxm = (M(a.msftx, 0, -1) / M(a.msfty, 0, -1)) * M(a.mut, 0, -1) * X(0, 0, -1) * rdx;
    c.mkxp = (M(a.msftx, 0, 0) / M(a.msfty, 0, 0)) * M(a.mut, 0, 0) * X(0, 0, 0) * rdx;
    c.mrdx = M(a.msfux, 0, 0) * M(a.msfuy, 0, 0) * rdx;
    c.mkym = ((M(a.msfuy, 0, 0) + M(a.msfuy, -1, 0)) / (M(a.msfux, 0, 0) + M(a.msfux, -1, 0))) * 0.25f *
             (M(a.mut, 0, 0) + M(a.mut, -1, 0) + M(a.mut, -1, -1) + M(a.mut, 0, -1)) * 0.25f *
             (X(0, 0, 0) + X(0, -1, 0) + X(0, -1, -1) + X(0, 0, -1)) * rdy;
    c.mkyp = ((M(a.msfuy, 0, 0) + M(a.msfuy, 1, 0)) / (M(a.msfux, 0, 0) + M(a.msfux, 1, 0))) * 0.25f *
             (M(a.mut, 0, 0) + M(a.mut, 1, 0) + M(a.mut, 1, -1) + M(a.mut, 0, -1)) * 0.25f *
             (X(0, 0, 0) + X(0, 1, 0) + X(0, 1, -1) + X(0, 0, -1)) * rdy;
    c.mrdy = M(a.msfux, 0, 0) * M(a.msfuy, 0, 0) * rdy;
  } else if (a.name == 'v') {
    c.mkxm = ((M(a.msfvx, 0, 0) + M(a.msfvx, 0, -1)) / (M(a.msfvy, 0, 0) + M(a.msfvy, 0, -1))) * 0.25f *
             (M(a.mut, 0, 0) + M(a.mut, -1, 0) + M(a.mut, -1, -1) + M(a.mut, 0, -1)) * 0.25f *
             (X(0, 0, 0) + X(0, -1, 0) + X(0, -1, -1) + X(0, 0, -1)) * rdx;
    c.mkxp = ((M(a.msfvx, 0, 0) + M(a.msfvx, 0, 1)) / (M(a.msfvy, 0, 0) + M(a.msfvy, 0, 1))) * 0.25f *
             (M(a.mut, 0, 0) + M(a.mut, -1, 0) + M(a.mut, -1, 1) + M(a.mut, 0, 1)) * 0.25f *
             (X(0, 0, 0) + X(0, -1, 0) + X(0, -1, 1) + X(0, 0, 1)) * rdx;
    c.mrdx = M(a.msfvx, 0, 0) * M(a.msfvy, 0, 0) * rdx;
    c.mkym = (M(a.msfty, -1, 0) / M(a.msftx, -1, 0)) * X(0, -1, 0) * rdy;
    c.mkyp = (M(a.msfty, 0, 0) / M(a.msftx, 0, 0)) * X(0, 0, 0) * rdy;
    c.mrdy = M(a.msfvx, 0, 0) * M(a.msfvy, 0, 0) * rdy;
  } else if (a.name == 'w') {
    c.mkxm = (M(a.msfux, 0, 0) / M(a.msfuy, 0, 0)) * 0.25f *
             (M(a.mut, 0, 0) + M(a.mut, 0, -1) + M(a.mut, 0, 0) + M(a.mut, 0, -1)) * 0.25f *
             (X(0, 0, 0) + X(0, 0, -1) + X(-1, 0, 0) + X(-1, 0, -1)) * rdx;
    c.mkxp = (M(a.msfux, 0, 1) / M(a.msfuy, 0, 1)) * 0.25f *
             (M(a.mut, 0, 1) + M(a.mut, 0, 0) + M(a.mut, 0, 1) + M(a.mut, 0, 0)) * 0.25f *
             (X(0, 0, 1) + X(0, 0, 0) + X(-1, 0, 1) + X(-1, 0, 0)) * rdx;
    c.mrdx = M(a.msftx, 0, 0) * M(a.msfty, 0, 0) * rdx;
    c.mkym = (M(a.msfvy, 0, 0) / M(a.msfvx, 0, 0)) * 0.25f *
             (M(a.mut, 0, 0) + M(a.mut, -1, 0) + M(a.mut, 0, 0) + M(a.mut, -1, 0)) * 0.25f *
             (X(0, 0, 0) + X(0, -1, 0) + X(-1, 0, 0) + X(-1, -1, 0)) * rdy;
    c.mkyp = (M(a.msfvy, 1, 0) / M(a.msfvx, 1, 0)) * 0.25f *
             (M(a.mut, 1, 0) + M(a.mut, 0, 0) + M(a.mut, 1, 0) + M(a.mut, 0, 0)) * 0.25f *
             (X(0, 1, 0) + X(0, 0, 0) + X(-1, 1, 0) + X(-1, 0, 0)) * rdy;
    c.mrdy = M(a.msftx, 0, 0) * M(a.msfty, 0, 0) * rdy;
  } else {
    c.mkxm = (M(a.msfux, 0, 0) / M(a.msfuy, 0, 0)) * 0.5f * (X(0, 0, 0) + X(0, 0, -1)) * 0.5f *
             (M(a.mut, 0, 0) + M(a.mut, 0, -1)) * rdx;
    c.mkxp = (M(a.msfux, 0, 1) / M(a.msfuy, 0, 1)) * 0.5f * (X(0, 0, 1) + X(0, 0, 0)) * 0.5f *
             (M(a.mut, 0, 1) + M(a.mut, 0, 0)) * rdx;
    c.mrdx = M(a.msftx, 0, 0) * M(a.msfty, 0, 0) * rdx;
    c.mkym = (M(a.msfvy, 0, 0) * M(a.msfvx_inv, 0, 0)) * 0.5f * (X(0, 0, 0) + X(0, -1, 0)) * 0.5f *
             (M(a.mut, 0, 0) + M(a.mut, -1, 0)) * rdy;
    c.mkyp = (M(a.msfvy, 1, 0) * M(a.msfvx_inv, 1, 0)) * 0.5f * (X(0, 1, 0) + X(0, 0, 0)) * 0.5f *
             (M(a.mut, 1, 0) + M(a.mut, 0, 0)) * rdy;
    c.mrdy = M(a.msftx, 0, 0) * M(a.msfty, 0, 0) * rdy;
  }
  return c;
}

// tendency + horizontal diffusion of f (or f - base) at (k, j, i)
__device__ inline float hdiff_apply(const Dims& D, const HDiff& a, int k, int j, int i, float t) {
  HCoef c = hdiff_coefs(D, a, k, j, i);
  auto Fv = [&](int dj, int di) { return AT(a.f, k, j + dj, i + di); };
  if (!a.base)
    return t + (c.mrdx * (c.mkxp * (Fv(0, 1) - Fv(0, 0)) - c.mkxm * (Fv(0, 0) - Fv(0, -1))) +
                c.mrdy * (c.mkyp * (Fv(1, 0) - Fv(0, 0)) - c.mkym * (Fv(0, 0) - Fv(-1, 0))));
  auto S = [&](int dj, int di) { return AT(a.base, k, j + dj, i + di); };
  return t + (c.mrdx * (c.mkxp * (Fv(0, 1) - Fv(0, 0) - S(0, 1) + S(0, 0)) -
                        c.mkxm * (Fv(0, 0) - Fv(0, -1) - S(0, 0) + S(0, -1))) +
              c.mrdy * (c.mkyp * (Fv(1, 0) - Fv(0, 0) - S(1, 0) + S(0, 0)) -
                        c.mkym * (Fv(0, 0) - Fv(-1, 0) - S(0, 0) + S(-1, 0))));
}

// kind: 'w' (vertical_diffusion 'w'), 'm' (vertical_diffusion 'm'), 'u'/'v'
// (vertical_diffusion_u/_v, 1-D base), 'p' (vertical_diffusion_mp, 1-D base),
// '3' (vertical_diffusion_3dmp, 3-D base).
struct VDiff {
  const float *f, *alt, *mus, *rdn, *rdnw, *base1, *base3;
  float* tend;
  float kvdif;
  int ktf, kts;
  char kind;
};

__device__ inline float vdiff_flux(const Dims& D, const VDiff& a, int k, int j, int i) {
  auto A = [&](int dk, int dj, int di) { return AT(a.alt, k + dk, j + dj, i + di); };
  auto Fv = [&](int dk) { return AT(a.f, k + dk, j, i); };
  if (k < a.kts || k >= a.ktf - 1) return 0.0f;
  switch (a.kind) {
    case 'w':
      return (a.kvdif / A(0, 0, 0)) * a.rdnw[k] * (Fv(1) - Fv(0));
    case 'm':
      if (k == a.kts) k = a.kts + 1;  // vflux[kts] = vflux[kts+1]
      if (k >= a.ktf - 1) return 0.0f;
      return a.kvdif * a.rdn[k + 1] / (0.5f * (AT(a.alt, k, j, i) + AT(a.alt, k + 1, j, i))) *
             (AT(a.f, k + 1, j, i) - AT(a.f, k, j, i));
    case 'u':
    case 'v': {
      int dj = a.kind == 'v' ? -1 : 0, di = a.kind == 'u' ? -1 : 0;
      return a.kvdif * a.rdn[k + 1] / (0.25f * (A(0, 0, 0) + A(0, dj, di) + A(1, 0, 0) + A(1, dj, di))) *
             (Fv(1) - Fv(0) - a.base1[k + 1] + a.base1[k]);
    }
    case 'p':
      return a.kvdif * a.rdn[k + 1] / (0.5f * (A(0, 0, 0) + A(1, 0, 0))) *
             (Fv(1) - Fv(0) - a.base1[k + 1] + a.base1[k]);
    default:
      return a.kvdif * a.rdn[k + 1] / (0.5f * (A(0, 0, 0) + A(1, 0, 0))) *
             (Fv(1) - Fv(0) - AT(a.base3, k + 1, j, i) + AT(a.base3, k, j, i));
  }
}

// tendency + vertical diffusion at (k, j, i)
__device__ inline float vdiff_apply(const Dims& D, const VDiff& a, int k, int j, int i, float t) {
  long q = D.ix(k, j, i);
  float dv = vdiff_flux(D, a, k, j, i) - vdiff_flux(D, a, k - 1, j, i);
  const float gg = F(pw::g * pw::g), G = F(pw::g);
  float m = AT2(a.mus, j, i);
  switch (a.kind) {
    case 'w':
      return t + a.rdn[k] * G * G / m / (0.5f * (a.alt[q] + AT(a.alt, k - 1, j, i))) * dv;
    case 'm':
      return t + gg / m / a.alt[q] * a.rdnw[k] * dv;
    case 'u':
      return t + gg * a.rdnw[k] / m / (0.5f * (AT(a.alt, k, j, i - 1) + a.alt[q])) * dv;
    case 'v':
      return t + gg * a.rdnw[k] / m / (0.5f * (AT(a.alt, k, j - 1, i) + a.alt[q])) * dv;
    default:
      return t + gg * m / a.alt[q] * a.rdnw[k] * dv;
  }
}

