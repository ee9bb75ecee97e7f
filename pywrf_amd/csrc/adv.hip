// Flux-form 5th/3rd-order advection of u, v, w and mass-point scalars
// (pywrf/wrf_dynamics.py:1112-2028).
//
// One launch per call.  For u, v and w a thread owns one tendency point and
// evaluates the face fluxes on both sides of it in y, x and z (the reference
// stores each face-flux plane in a full-grid scratch array and differences
// it; here the face flux is a pure function of the inputs, so neighbouring
// threads recompute shared faces from cached loads).  Scalars share x-faces
// between neighbouring lanes and carry z-faces up the column.  The three
// directional updates are applied to the tendency in the reference's order,
// y then x then z.
//
// Face rules, including the order reductions on the rows and columns next
// to the domain edge and the sign-source quirks of some 3rd-order faces, are
// those of the reference; a face the reference never sets carries zero flux.
// The upwind sign is sign(vel) * sign(time_step) with the reference's
// module-level time_step = 60 > 0.
#include "common.h"
#include "consts.h"
#include "diff.h"


int pw_check_grid(const pywrf_grid_t* g, int r);
#define CHECK(g, r) do { int _e = pw_check_grid((g), (r)); if (_e) return _e; } while (0)

struct AdvIdx {
  int ids, ide, jds, jde, ktf;
};

// ------------------------------------------------------------------ advect_u
struct AdvU {
  const float *u, *ru, *rv, *rom, *msfux, *fzm, *fzp, *rdzw;
  float *tend;
  float rdx, rdy;
};

// The u faces from operand values: Q[0..5] are u at offsets -3..2 from the
// face, rvs / rus the sum of the two mass fluxes beside the face, vel the
// z-face velocity.
__device__ inline float u_yface_v(const AdvIdx& x, int jf, float rvs, const float* Q) {
  if (jf >= x.jds + 3 && jf < x.jde - 3) {
    float vel = 0.5f * rvs;
    return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], vel);
  }
  if (jf == x.jds + 1 || jf == x.jde - 2) return 0.25f * rvs * (Q[3] + Q[2]);
  if (jf == x.jds + 2) {
    float vel = 0.5f * rvs;
    return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);
  }
  if (jf == x.jde - 3) {
    float vel = 0.5f * rvs;
    return vel * flux3(Q[1], Q[2], Q[3], Q[4], 0.0f);
  }
  return 0.0f;
}

__device__ inline float u_xface_v(const AdvIdx& x, int f, float rus, const float* Q) {
  if (f >= x.ids + 3 && f < x.ide - 2) {
    float vel = 0.5f * rus;
    return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], vel);
  }
  if (f == x.ids + 1) {
    float ub = Q[3] < 0.0f ? Q[3] : Q[2];
    return 0.25f * rus * (Q[3] + ub);
  }
  if (f == x.ids + 2 || f == x.ide - 2) {
    float vel = 0.5f * rus;
    return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);
  }
  if (f == x.ide - 1) {
    float ub = Q[2] > 0.0f ? Q[2] : Q[3];
    return 0.25f * rus * (Q[2] + ub);
  }
  return 0.0f;
}

__device__ inline float u_zface_v(const AdvU& a, const AdvIdx& x, int kf, float vel, const float* Q) {
  if (kf < 1 || kf >= x.ktf) return 0.0f;
  if (kf >= 3 && kf < x.ktf - 2) return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], -vel);
  if (kf == 1 || kf == x.ktf - 1) return vel * (a.fzm[kf] * Q[3] + a.fzp[kf] * Q[2]);
  return vel * flux3(Q[1], Q[2], Q[3], Q[4], -vel);  // kf == 2 or ktf-2
}

// fresh: the tendency array is taken as zero (the caller's zeroing is folded
// in): the launch covers the whole array, points outside the box get 0, and the
// box's points start from 0 instead of the stored value.
__global__ void k_advect_u(Box all, Dims D, AdvU a, AdvIdx x, Box by, Box bz, Box b, int fresh) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  if (!b.has(k, j, i)) {
    if (fresh) a.tend[q] = 0.0f;
    return;
  }
  // every operand is loaded first, outside the branches, at indices clamped
  // into the array (the halo check puts every value a face formula uses in
  // range, so clamping changes only values that are never used)
  const int jn = imin(j + 1, D.ny - 1), iw = imax(i - 1, 0), ie = imin(i + 1, D.nx - 1), kn = imin(k + 1, D.nz - 1);
  float uy[7], ux[7], uz[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    uy[o] = AT(a.u, k, imin(imax(j - 3 + o, 0), D.ny - 1), i);
    ux[o] = AT(a.u, k, j, imin(imax(i - 3 + o, 0), D.nx - 1));
    uz[o] = AT(a.u, imin(imax(k - 3 + o, 0), D.nz - 1), j, i);
  }
  const float rv0 = AT(a.rv, k, j, i), rv0w = AT(a.rv, k, j, iw), rv1 = AT(a.rv, k, jn, i), rv1w = AT(a.rv, k, jn, iw);
  const float ruw = AT(a.ru, k, j, iw), ru0 = AT(a.ru, k, j, i), rue = AT(a.ru, k, j, ie);
  const float rmw0 = AT(a.rom, k, j, iw), rm0 = AT(a.rom, k, j, i), rmw1 = AT(a.rom, kn, j, iw), rm1 = AT(a.rom, kn, j, i);
  const float msf = AT2(a.msfux, j, i), rdz = a.rdzw[k];
  float t = fresh ? 0.0f : a.tend[q];
  if (by.has(k, j, i)) t = t - (msf * a.rdy) * (u_yface_v(x, j + 1, rv1 + rv1w, uy + 1) - u_yface_v(x, j, rv0 + rv0w, uy));
  t = t - (msf * a.rdx) * (u_xface_v(x, i + 1, rue + ru0, ux + 1) - u_xface_v(x, i, ru0 + ruw, ux));
  t = t - rdz * (u_zface_v(a, x, k + 1, 0.5f * (rmw1 + rm1), uz + 1) - u_zface_v(a, x, k, 0.5f * (rmw0 + rm0), uz));
  a.tend[q] = t;
}

static int advect_u(int fresh, const pywrf_grid_t* g, const float* u, float* tendency, const float* ru,
                    const float* rv, const float* rom, const float* msfux, const float* fzm, const float* fzp,
                    double rdx, double rdy, const float* rdzw, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  AdvIdx x = {g->ids, g->ide, g->jds, g->jde, imin(g->kte, g->kde - 1)};
  if (g->kts != 0 || x.ktf + 1 > g->nz) return PYWRF_E_RANGE;
  AdvU a = {u, ru, rv, rom, msfux, fzm, fzp, rdzw, tendency, F(rdx), F(rdy)};
  int i0 = imax(g->ids + 1, g->its), i1 = imin(g->ide - 1, g->ite);
  int j0 = g->jts, j1 = imin(g->jte, g->jde - 1);
  Box b = mkbox(g->kts, x.ktf, j0, j1, i0, i1);
  Box by = mkbox(g->kts, x.ktf, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), i0, i1);
  Box bz = b;
  PW_LAUNCH_BOX(k_advect_u, fresh ? whole(g) : b, s, dims_of(g), a, x, by, bz, b, fresh);
  PW_RET();
}

extern "C" int pywrf_advect_u(const pywrf_grid_t* g, const float* u, float* tendency, const float* ru,
                              const float* rv, const float* rom, const float* msfux, const float* fzm,
                              const float* fzp, double rdx, double rdy, const float* rdzw, pywrf_stream_t s) {
  return advect_u(0, g, u, tendency, ru, rv, rom, msfux, fzm, fzp, rdx, rdy, rdzw, s);
}

extern "C" int pywrf_advect_u_fresh(const pywrf_grid_t* g, const float* u, float* tendency, const float* ru,
                                    const float* rv, const float* rom, const float* msfux, const float* fzm,
                                    const float* fzp, double rdx, double rdy, const float* rdzw, pywrf_stream_t s) {
  return advect_u(1, g, u, tendency, ru, rv, rom, msfux, fzm, fzp, rdx, rdy, rdzw, s);
}

// ------------------------------------------------------------------ advect_v
struct AdvV {
  const float *v, *ru, *rv, *rom, *msfvx, *msfvy, *fzm, *fzp, *rdzw;
  float *tend;
  float rdx, rdy;
};

// The v faces from operand values (as the u faces above).
__device__ inline float v_yface_v(const AdvIdx& x, int f, float rvs, const float* Q) {
  if (f >= x.jds + 3 && f < x.jde - 2) {
    float vel = 0.5f * rvs;
    return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], vel);
  }
  if (f == x.jds + 1) {
    float vb = Q[3] < 0.0f ? Q[3] : Q[2];
    return 0.25f * rvs * (Q[3] + vb);
  }
  if (f == x.jds + 2 || f == x.jde - 2) {
    float vel = 0.5f * rvs;
    return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);
  }
  if (f == x.jde - 1) {
    float vb = Q[2] > 0.0f ? Q[2] : Q[3];
    return 0.25f * rvs * (Q[2] + vb);
  }
  return 0.0f;
}

__device__ inline float v_xface_v(const AdvIdx& x, int f, float rus, const float* Q) {
  if (f >= x.ids + 3 && f < x.ide - 3) {
    float vel = 0.5f * rus;
    return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], vel);
  }
  if (f == x.ids + 1 || f == x.ide - 2) return 0.25f * rus * (Q[3] + Q[2]);
  if (f == x.ids + 2 || f == x.ide - 3) {
    float vel = 0.5f * rus;
    return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);
  }
  return 0.0f;
}

__device__ inline float v_zface_v(const AdvV& a, const AdvIdx& x, int kf, float vel, const float* Q) {
  if (kf < 1 || kf >= x.ktf) return 0.0f;
  if (kf >= 3 && kf < x.ktf - 2) return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], -vel);
  if (kf == 1 || kf == x.ktf - 1) return vel * (a.fzm[kf] * Q[3] + a.fzp[kf] * Q[2]);
  if (kf == 2) return vel * flux3(Q[1], Q[2], Q[3], Q[4], -vel);
  return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);  // kf == ktf-2: +vel in the reference
}

__global__ void k_advect_v(Box all, Dims D, AdvV a, AdvIdx x, Box bx, Box b, int fresh) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  if (!b.has(k, j, i)) {
    if (fresh) a.tend[q] = 0.0f;
    return;
  }
  // every operand is loaded first, outside the branches, at clamped indices
  // (as in k_advect_u)
  const int jn = imin(j + 1, D.ny - 1), js = imax(j - 1, 0), ie = imin(i + 1, D.nx - 1), kn = imin(k + 1, D.nz - 1);
  float vy[7], vx[7], vz[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    vy[o] = AT(a.v, k, imin(imax(j - 3 + o, 0), D.ny - 1), i);
    vx[o] = AT(a.v, k, j, imin(imax(i - 3 + o, 0), D.nx - 1));
    vz[o] = AT(a.v, imin(imax(k - 3 + o, 0), D.nz - 1), j, i);
  }
  const float rvs_ = AT(a.rv, k, js, i), rv0 = AT(a.rv, k, j, i), rvn = AT(a.rv, k, jn, i);
  const float ru0 = AT(a.ru, k, j, i), ru0s = AT(a.ru, k, js, i), rue = AT(a.ru, k, j, ie), rues = AT(a.ru, k, js, ie);
  const float rm0 = AT(a.rom, k, j, i), rm0s = AT(a.rom, k, js, i), rm1 = AT(a.rom, kn, j, i), rm1s = AT(a.rom, kn, js, i);
  const float my = AT2(a.msfvy, j, i), mvx = AT2(a.msfvx, j, i), rdz = a.rdzw[k];
  float t = fresh ? 0.0f : a.tend[q];
  t = t - (my * a.rdy) * (v_yface_v(x, j + 1, rvn + rv0, vy + 1) - v_yface_v(x, j, rv0 + rvs_, vy));
  if (bx.has(k, j, i)) t = t - (my * a.rdx) * (v_xface_v(x, i + 1, rue + rues, vx + 1) - v_xface_v(x, i, ru0 + ru0s, vx));
  t = t - (my / mvx) * rdz * (v_zface_v(a, x, k + 1, 0.5f * (rm1 + rm1s), vz + 1) - v_zface_v(a, x, k, 0.5f * (rm0 + rm0s), vz));
  a.tend[q] = t;
}

static int advect_v(int fresh, const pywrf_grid_t* g, const float* v, float* tendency, const float* ru,
                    const float* rv, const float* rom, const float* msfvx, const float* msfvy, const float* fzm,
                    const float* fzp, double rdx, double rdy, const float* rdzw, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  AdvIdx x = {g->ids, g->ide, g->jds, g->jde, imin(g->kte, g->kde - 1)};
  if (g->kts != 0 || x.ktf + 1 > g->nz) return PYWRF_E_RANGE;
  AdvV a = {v, ru, rv, rom, msfvx, msfvy, fzm, fzp, rdzw, tendency, F(rdx), F(rdy)};
  int i0 = g->its, i1 = imin(g->ite, g->ide - 1);
  int j0 = imax(g->jts, g->jds + 1), j1 = imin(g->jte, g->jde - 1);
  Box b = mkbox(g->kts, x.ktf, j0, j1, i0, i1);
  Box bx = mkbox(g->kts, x.ktf, j0, j1, imax(g->ids + 1, g->its), imin(g->ide - 2, g->ite));
  PW_LAUNCH_BOX(k_advect_v, fresh ? whole(g) : b, s, dims_of(g), a, x, bx, b, fresh);
  PW_RET();
}

extern "C" int pywrf_advect_v(const pywrf_grid_t* g, const float* v, float* tendency, const float* ru,
                              const float* rv, const float* rom, const float* msfvx, const float* msfvy,
                              const float* fzm, const float* fzp, double rdx, double rdy, const float* rdzw,
                              pywrf_stream_t s) {
  return advect_v(0, g, v, tendency, ru, rv, rom, msfvx, msfvy, fzm, fzp, rdx, rdy, rdzw, s);
}

extern "C" int pywrf_advect_v_fresh(const pywrf_grid_t* g, const float* v, float* tendency, const float* ru,
                                    const float* rv, const float* rom, const float* msfvx, const float* msfvy,
                                    const float* fzm, const float* fzp, double rdx, double rdy, const float* rdzw,
                                    pywrf_stream_t s) {
  return advect_v(1, g, v, tendency, ru, rv, rom, msfvx, msfvy, fzm, fzp, rdx, rdy, rdzw, s);
}

// ------------------------------------------------------------------ advect_w
struct AdvW {
  const float *w, *ru, *rv, *rom, *msftx, *fzm, *fzp, *rdzu;
  float *tend;
  float rdx, rdy;
};

// Horizontal mass flux interpolated to w level k (k < ktf) from its values rA
// at level k and rB at k-1, or extrapolated to the model top (k == ktf) from
// rA at ktf-1 and rB at ktf-2.
__device__ inline float w_level_flux_v(const AdvW& a, int ktf, int k, float rA, float rB) {
  if (k < ktf) return a.fzm[k] * rA + a.fzp[k] * rB;
  return (2.0f - a.fzm[ktf - 1]) * rA - a.fzp[ktf - 1] * rB;
}

// The w faces from operand values: Q[0..5] are w at offsets -3..2 from the
// face, vel the face's mass flux (rs for z: the sum of the two rom values).
__device__ inline float w_hface_v(int f, int lo, int hi, float vel, const float* Q) {
  if (f >= lo + 3 && f < hi - 3) return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], vel);
  if (f == lo + 1 || f == hi - 2) return 0.5f * vel * (Q[3] + Q[2]);
  if (f == lo + 2 || f == hi - 3) return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);
  return 0.0f;
}

// Vertical flux at mass level kf (between w[kf-1] and w[kf]), kf in [1, ktf].
__device__ inline float w_zface_v(const AdvIdx& x, int kf, float rs, const float* Q) {
  if (kf < 1 || kf > x.ktf) return 0.0f;
  if (kf == 1 || kf == x.ktf) return 0.25f * rs * (Q[3] + Q[2]);
  float vel = 0.5f * rs;
  if (kf >= 3 && kf < x.ktf - 1) return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], -vel);
  if (kf == 2) return vel * flux3(Q[1], Q[2], Q[3], Q[4], -vel);
  return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);  // kf == ktf-1: +vel in the reference
}

__global__ void k_advect_w(Box all, Dims D, AdvW a, AdvIdx x, Box by, Box bx, Box b, int fresh) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i);
  if (!b.has(k, j, i)) {
    if (fresh) a.tend[q] = 0.0f;
    return;
  }
  // every operand is loaded first, outside the branches, at clamped indices
  // (as in k_advect_u); the two levels of the horizontal mass fluxes are k
  // and k-1 below the top, ktf-1 and ktf-2 at it (k is one per workgroup)
  const int ktf = x.ktf, top = k >= ktf;
  const int kA = top ? ktf - 1 : k, kB = imax(top ? ktf - 2 : k - 1, 0);
  const int jn = imin(j + 1, D.ny - 1), ie = imin(i + 1, D.nx - 1);
  float wy[7], wx[7], wz[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    wy[o] = AT(a.w, k, imin(imax(j - 3 + o, 0), D.ny - 1), i);
    wx[o] = AT(a.w, k, j, imin(imax(i - 3 + o, 0), D.nx - 1));
    wz[o] = AT(a.w, imin(imax(k - 3 + o, 0), D.nz - 1), j, i);
  }
  const float rvA0 = AT(a.rv, kA, j, i), rvB0 = AT(a.rv, kB, j, i), rvA1 = AT(a.rv, kA, jn, i), rvB1 = AT(a.rv, kB, jn, i);
  const float ruA0 = AT(a.ru, kA, j, i), ruB0 = AT(a.ru, kB, j, i), ruA1 = AT(a.ru, kA, j, ie), ruB1 = AT(a.ru, kB, j, ie);
  const float rmn = AT(a.rom, imin(k + 1, D.nz - 1), j, i), rm0 = AT(a.rom, k, j, i), rmb = AT(a.rom, imax(k - 1, 0), j, i);
  const float mx = AT2(a.msftx, j, i), rdz = a.rdzu[top ? ktf - 1 : k];
  float t = fresh ? 0.0f : a.tend[q];
  if (by.has(k, j, i))
    t = t - (mx * a.rdy) * (w_hface_v(j + 1, x.jds, x.jde, w_level_flux_v(a, ktf, k, rvA1, rvB1), wy + 1) -
                            w_hface_v(j, x.jds, x.jde, w_level_flux_v(a, ktf, k, rvA0, rvB0), wy));
  if (bx.has(k, j, i))
    t = t - (mx * a.rdx) * (w_hface_v(i + 1, x.ids, x.ide, w_level_flux_v(a, ktf, k, ruA1, ruB1), wx + 1) -
                            w_hface_v(i, x.ids, x.ide, w_level_flux_v(a, ktf, k, ruA0, ruB0), wx));
  if (!top) t = t - rdz * (w_zface_v(x, k + 1, rmn + rm0, wz + 1) - w_zface_v(x, k, rm0 + rmb, wz));
  else t = t + 2.0f * rdz * w_zface_v(x, k, rm0 + rmb, wz);
  a.tend[q] = t;
}

static int advect_w(int fresh, const pywrf_grid_t* g, const float* w, float* tendency, const float* ru,
                    const float* rv, const float* rom, const float* msftx, const float* fzm, const float* fzp,
                    double rdx, double rdy, const float* rdzu, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 4);
  AdvIdx x = {g->ids, g->ide, g->jds, g->jde, imin(g->kte, g->kde - 1)};
  if (g->kts != 0 || x.ktf + 1 > g->nz) return PYWRF_E_RANGE;
  AdvW a = {w, ru, rv, rom, msftx, fzm, fzp, rdzu, tendency, F(rdx), F(rdy)};
  int i0 = g->its, i1 = imin(g->ite, g->ide - 1);
  int j0 = g->jts, j1 = imin(g->jte, g->jde - 1);
  Box b = mkbox(g->kts + 1, x.ktf + 1, j0, j1, i0, i1);
  Box by = mkbox(g->kts + 1, x.ktf + 1, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), i0, i1);
  Box bx = mkbox(g->kts + 1, x.ktf + 1, j0, j1, imax(g->ids + 1, g->its), imin(g->ide - 2, g->ite));
  PW_LAUNCH_BOX(k_advect_w, fresh ? whole(g) : b, s, dims_of(g), a, x, by, bx, b, fresh);
  PW_RET();
}

extern "C" int pywrf_advect_w(const pywrf_grid_t* g, const float* w, float* tendency, const float* ru,
                              const float* rv, const float* rom, const float* msftx, const float* fzm,
                              const float* fzp, double rdx, double rdy, const float* rdzu, pywrf_stream_t s) {
  return advect_w(0, g, w, tendency, ru, rv, rom, msftx, fzm, fzp, rdx, rdy, rdzu, s);
}

extern "C" int pywrf_advect_w_fresh(const pywrf_grid_t* g, const float* w, float* tendency, const float* ru,
                                    const float* rv, const float* rom, const float* msftx, const float* fzm,
                                    const float* fzp, double rdx, double rdy, const float* rdzu, pywrf_stream_t s) {
  return advect_w(1, g, w, tendency, ru, rv, rom, msftx, fzm, fzp, rdx, rdy, rdzu, s);
}

// ------------------------------------------------------------------ advect_scalar
struct AdvS {
  const float *f, *ru, *rv, *rom, *msftx, *fzm, *fzp, *rdzw;
  float *tend;
  float rdx, rdy;
};

__device__ inline float s_yface(const Dims& D, const AdvS& a, const AdvIdx& x, int k, int jf, int i) {
  auto Q = [&](int o) { return AT(a.f, k, jf + o, i); };
  float vel = AT(a.rv, k, jf, i);
  if (jf >= x.jds + 3 && jf < x.jde - 3) return vel * flux5(Q(-3), Q(-2), Q(-1), Q(0), Q(1), Q(2), vel);
  if (jf == x.jds + 1 || jf == x.jde - 2) return 0.5f * vel * (Q(0) + Q(-1));
  if (jf == x.jds + 2 || jf == x.jde - 3) return vel * flux3(Q(-2), Q(-1), Q(0), Q(1), vel);
  return 0.0f;
}

__device__ inline float s_xface(const Dims& D, const AdvS& a, const AdvIdx& x, int k, int j, int f) {
  auto Q = [&](int o) { return AT(a.f, k, j, f + o); };
  float vel = AT(a.ru, k, j, f);
  if (f >= x.ids + 3 && f < x.ide - 3) return vel * flux5(Q(-3), Q(-2), Q(-1), Q(0), Q(1), Q(2), vel);
  if (f == x.ids + 1 || f == x.ide - 2) return 0.5f * vel * (Q(0) + Q(-1));
  if (f == x.ids + 2) return vel * flux3(Q(-2), Q(-1), Q(0), Q(1), vel);
  if (f == x.ide - 3) return vel * flux3(Q(-2), Q(-1), Q(0), Q(1), AT(a.ru, k, j, f + 1));  // sign from ru[ide-2]
  return 0.0f;
}

__device__ inline float s_zface(const Dims& D, const AdvS& a, const AdvIdx& x, int kf, int j, int i) {
  auto Q = [&](int o) { return AT(a.f, kf + o, j, i); };
  if (kf < 1 || kf >= x.ktf) return 0.0f;
  float vel = AT(a.rom, kf, j, i);
  if (kf >= 3 && kf < x.ktf - 2) return vel * flux5(Q(-3), Q(-2), Q(-1), Q(0), Q(1), Q(2), -vel);
  if (kf == 1 || kf == x.ktf - 1) return vel * (a.fzm[kf] * Q(0) + a.fzp[kf] * Q(-1));
  return vel * flux3(Q(-2), Q(-1), Q(0), Q(1), -vel);  // kf == 2 or ktf-2
}


// The same faces from operand values: Q[0..5] are the field at offsets -3..2
// from the face (rows for y, levels for z), vel the face velocity.
__device__ inline float s_yface_v(const AdvIdx& x, int jf, float vel, const float* Q) {
  if (jf >= x.jds + 3 && jf < x.jde - 3) return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], vel);
  if (jf == x.jds + 1 || jf == x.jde - 2) return 0.5f * vel * (Q[3] + Q[2]);
  if (jf == x.jds + 2 || jf == x.jde - 3) return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);
  return 0.0f;
}

__device__ inline float s_xface_v(const AdvIdx& x, int f, float vel, float vel1, const float* Q) {
  if (f >= x.ids + 3 && f < x.ide - 3) return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], vel);
  if (f == x.ids + 1 || f == x.ide - 2) return 0.5f * vel * (Q[3] + Q[2]);
  if (f == x.ids + 2) return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel);
  if (f == x.ide - 3) return vel * flux3(Q[1], Q[2], Q[3], Q[4], vel1);  // sign from ru[ide-2]
  return 0.0f;
}

__device__ inline float s_zface_v(const AdvS& a, const AdvIdx& x, int kf, float vel, const float* Q) {
  if (kf < 1 || kf >= x.ktf) return 0.0f;
  if (kf >= 3 && kf < x.ktf - 2) return vel * flux5(Q[0], Q[1], Q[2], Q[3], Q[4], Q[5], -vel);
  if (kf == 1 || kf == x.ktf - 1) return vel * (a.fzm[kf] * Q[3] + a.fzp[kf] * Q[2]);
  return vel * flux3(Q[1], Q[2], Q[3], Q[4], -vel);  // kf == 2 or ktf-2
}

// ------------------------------------------------------------------ rk_scalar_tend, all species in one pass
// Zeroes h_tendency / z_tendency (whole array, in the same launch); for every species of
// [scs, sce] writes its advection tendency on the tile (zero outside the
// advection box) and, on the first RK stage, adds the horizontal and vertical
// diffusion into scalar_tends -- the per-species sequence of rk_scalar_tend
// (wrf_dynamics.py:3982-4051) -- with the mass fluxes read once for all species.
struct ScalarTend {
  AdvS adv;
  AdvIdx x;
  HDiff hd;
  VDiff vd;
  const float* scalar;
  float *adv_tend, *sc_tend, *h_t, *z_t;
  long n3;
  int scs, sce, rk1, moist_step, accumulate;
  Box bzero, badv, by, bx, bh, bv, span;
  // a second species array over the same fluxes and index set (scalar2 null:
  // none); species groups [0, nsg1) are the first array's, the rest the second's
  const float* scalar2;
  float *adv_tend2, *sc_tend2;
  int scs2, sce2, moist_step2, nsg1;
};

// Tile of 63 i x 4 j points per workgroup (lane l of a row evaluates the x-face
// at i0 + l; the point takes its right-hand face from lane l + 1 by a lane
// shuffle), marching SA_KC levels up the column and carrying each z-face to
// the level above.  Face values and the tendency arithmetic are those of the
// one-thread-per-point form (s_yface, s_xface, s_zface in the reference's
// y, x, z order), so the results are the same floats; each x- and z-face is
// evaluated about once instead of by both neighbours.
#define SA_OX 63
#ifndef SA_KC
#define SA_KC 8
#endif
#ifndef SA_SPW
#define SA_SPW 1   // species per wave
#endif
__global__ void __launch_bounds__(PW_TX * PW_TY) k_scalar_adv(Dims D, ScalarTend a) {
  const Box& b = a.bzero;   // points that receive a tendency
  const Box& sp = a.span;   // launch region: b, or the whole array when h_t / z_t are zeroed as well
  const int i = sp.i0 + (int)(blockIdx.x * SA_OX + threadIdx.x);
  const int j = sp.j0 + (int)(blockIdx.y * PW_TY + threadIdx.y);
  if (j >= sp.j1) return;                     // a whole wave (one row) leaves together
  const int nkc = (sp.nk() + SA_KC - 1) / SA_KC;
  const int kc = (int)blockIdx.z % nkc, sg = (int)blockIdx.z / nkc;   // level chunk, species group
  const int k0 = sp.k0 + kc * SA_KC, k1 = imin(k0 + SA_KC, sp.k1);
  const bool second = sg >= a.nsg1;                                   // the second species array
  const int scs = second ? a.scs2 : a.scs, sce = second ? a.sce2 : a.sce;
  const int ms = scs + (second ? sg - a.nsg1 : sg) * SA_SPW, me = imin(sce, ms + SA_SPW - 1);
  const float* const sbase = second ? a.scalar2 : a.scalar;
  float* const tbase = second ? a.adv_tend2 : a.adv_tend;
  const bool mine = threadIdx.x < SA_OX && i < sp.i1;
  if (a.h_t && mine && sg == 0)
    for (int k = k0; k < k1; ++k) {
      const int q = D.ix(k, j, i);
      a.h_t[q] = 0.0f;
      a.z_t[q] = 0.0f;
    }
  if (j < b.j0 || j >= b.j1) return;          // whole wave: no tendency points in this row
  const bool live = i < D.nx;                 // may evaluate faces
  const bool own = mine && i >= b.i0 && i < b.i1;
  const float mx = (own && a.badv.has2(j, i)) ? AT2(a.adv.msftx, j, i) : 0.0f;
#pragma unroll 1
  for (int im = ms; im <= me; ++im) {
    AdvS ad = a.adv;
    ad.f = sbase + im * a.n3;
    // the column's values at levels kf-3 .. kf+2 around the current z-face kf,
    // slid up one level per step (one new load per level); levels outside the
    // array are never used by the face formulas and are not loaded
    // (indices clamped into the array so that every load is unconditional; the
    // clamped values are never used)
    const int ic = imin(i, D.nx - 1);
    auto ZQ = [&](int kk) { return AT(ad.f, imin(imax(kk, 0), D.nz - 1), j, ic); };
    float wz[6];
#pragma unroll
    for (int o = 0; o < 6; ++o) wz[o] = ZQ(k0 - 3 + o);
    float zlo = live ? s_zface_v(ad, a.x, k0, AT(ad.rom, imin(imax(k0, 0), D.nz - 1), j, ic), wz) : 0.0f;
    // every operand of a level is loaded before its arithmetic and outside
    // the branches (indices clamped into the array; values at clamped indices
    // are never used), so that a level waits on memory once
    const int jn = imin(j + 1, D.ny - 1);
    int xi[6], yj[7];
#pragma unroll
    for (int o = 0; o < 6; ++o) xi[o] = imin(imax(ic - 3 + o, 0), D.nx - 1);
#pragma unroll
    for (int o = 0; o < 7; ++o) yj[o] = imin(imax(j - 3 + o, 0), D.ny - 1);
    const int ic1 = imin(ic + 1, D.nx - 1);
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
      float xq[6], r[7];
#pragma unroll
      for (int o = 0; o < 6; ++o) xq[o] = AT(ad.f, k, j, xi[o]);
#pragma unroll
      for (int o = 0; o < 7; ++o) r[o] = AT(ad.f, k, yj[o], ic);
      const float xvel = AT(ad.ru, k, j, ic), xvel1 = AT(ad.ru, k, j, ic1);
      const float rvl = AT(ad.rv, k, j, ic), rvh = AT(ad.rv, k, jn, ic);
      const int kf = k + 1;
      const float romv = AT(ad.rom, imin(kf, D.nz - 1), j, ic);
      const float wnew = ZQ(k + 3);
      const float told = a.accumulate ? tbase[im * a.n3 + D.ix(k, j, ic)] : 0.0f;
      const float rdz = k < b.k1 ? ad.rdzw[k] : 0.0f;
      const float xf = live ? s_xface_v(a.x, i, xvel, xvel1, xq) : 0.0f;
      const float xr = __shfl_down(xf, 1, PW_TX);
#pragma unroll
      for (int o = 0; o < 5; ++o) wz[o] = wz[o + 1];
      wz[5] = wnew;
      const float zhi = live ? s_zface_v(ad, a.x, kf, romv, wz) : 0.0f;
      if (own && k >= b.k0 && k < b.k1) {
        const int q = im * a.n3 + D.ix(k, j, i);
        float t = told;
        if (a.badv.has(k, j, i)) {
          if (a.by.has(k, j, i)) {
            // rows j-3 .. j+3 serve the faces at j (rows j-3 .. j+2) and j+1 (rows j-2 .. j+3)
            const float yhi = s_yface_v(a.x, j + 1, rvh, r + 1);
            const float ylo = s_yface_v(a.x, j, rvl, r);
            t = t - (mx * ad.rdy) * (yhi - ylo);
          }
          if (a.bx.has(k, j, i)) t = t - (mx * ad.rdx) * (xr - xf);
          t = t - rdz * (zhi - zlo);
        }
        tbase[q] = t;
      }
      zlo = zhi;
    }
  }
}

static void pw_launch_scalar_adv(const pywrf_grid_t* g, const ScalarTend& a, pywrf_stream_t s) {
  const Box& sp = a.span;
  if (sp.size() <= 0) return;
  dim3 grid((unsigned)((sp.ni() + SA_OX - 1) / SA_OX), (unsigned)((sp.nj() + PW_TY - 1) / PW_TY),
            (unsigned)((sp.nk() + SA_KC - 1) / SA_KC *
                       (a.nsg1 + (a.scalar2 ? (a.sce2 - a.scs2 + SA_SPW) / SA_SPW : 0))));
  hipLaunchKernelGGL(k_scalar_adv, grid, dim3(PW_TX, PW_TY), 0, (hipStream_t)s, dims_of(g), a);
}

// zero_box 1: first zero the tendency on [kts, kte) x [jts, jte) x [its, ite)
// (rk_scalar_tend, wrf_dynamics.py:3995) and accumulate into it; 2: the whole
// tendency array is taken as zero (the caller's zeroing folded in).
int pw_advect_scalar_box(const pywrf_grid_t* g, const float* field, float* tendency, const float* ru,
                         const float* rv, const float* rom, const float* msftx, const float* fzm, const float* fzp,
                         float rdx, float rdy, const float* rdzw, int zero_box, pywrf_stream_t s) {
  CHECK(g, 4);
  AdvIdx x = {g->ids, g->ide, g->jds, g->jde, imin(g->kte, g->kde - 1)};
  if (g->kts != 0 || x.ktf + 1 > g->nz || g->kte > g->nz) return PYWRF_E_RANGE;
  AdvS a = {field, ru, rv, rom, msftx, fzm, fzp, rdzw, tendency, rdx, rdy};
  int i0 = g->its, i1 = imin(g->ite, g->ide - 1);
  int j0 = g->jts, j1 = imin(g->jte, g->jde - 1);
  Box badv = mkbox(g->kts, x.ktf, j0, j1, i0, i1);
  Box by = mkbox(g->kts, x.ktf, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), i0, i1);
  Box bx = mkbox(g->kts, x.ktf, j0, j1, imax(g->ids + 1, g->its), imin(g->ide - 2, g->ite));
  ScalarTend t;
  t.adv = a; t.x = x;
  t.scalar = field; t.adv_tend = tendency; t.n3 = 0; t.scs = 0; t.sce = 0; t.accumulate = zero_box ? 0 : 1;
  t.scalar2 = nullptr; t.adv_tend2 = t.sc_tend2 = nullptr; t.scs2 = 0; t.sce2 = -1; t.moist_step2 = 0;
  t.nsg1 = 1;
  t.bzero = zero_box == 2 ? whole(g) : zero_box ? mkbox(g->kts, g->kte, g->jts, g->jte, g->its, g->ite) : badv;
  t.badv = badv; t.by = by; t.bx = bx;
  t.span = t.bzero;
  t.h_t = nullptr;
  t.z_t = nullptr;
  pw_launch_scalar_adv(g, t, s);
  PW_RET();
}

extern "C" int pywrf_advect_scalar(const pywrf_grid_t* g, const float* field, float* tendency, const float* ru,
                                   const float* rv, const float* rom, const float* msftx, const float* fzm,
                                   const float* fzp, double rdx, double rdy, const float* rdzw, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return pw_advect_scalar_box(g, field, tendency, ru, rv, rom, msftx, fzm, fzp, F(rdx), F(rdy), rdzw, 0, s);
}

extern "C" int pywrf_advect_scalar_fresh(const pywrf_grid_t* g, const float* field, float* tendency,
                                         const float* ru, const float* rv, const float* rom, const float* msftx,
                                         const float* fzm, const float* fzp, double rdx, double rdy,
                                         const float* rdzw, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return pw_advect_scalar_box(g, field, tendency, ru, rv, rom, msftx, fzm, fzp, F(rdx), F(rdy), rdzw, 2, s);
}

// First RK stage only: horizontal ('m') and vertical diffusion of every
// species into scalar_tends (wrf_dynamics.py:4018-4046), one thread per point
// for all species of the call.  The diffusion coefficients depend only on the
// point: the horizontal ones (hdiff_coefs), the vertical flux factor
// kvdif rdn[k+1] / (0.5 (alt[k] + alt[k+1])) of the two faces and the factor
// the flux difference is applied with are formed once, each with the
// reference's operation order up to the product with the species' own
// difference, and applied to every species: the floats are those of the
// per-species form (hdiff_apply, vdiff_apply with kind 'm', or 'p' for the
// water vapour of the moist step).
struct VFace {
  float c;     // flux factor, 0 where the flux is zero
  int k;       // level of the face's lower point (after the kts rule of kind 'm')
  bool live;
};

// Face below level kf+1 (the reference's vflux[kf]) of kind 'm' or 'p'.
__device__ inline VFace vface(const Dims& D, const VDiff& v, char kind, int kf, int j, int i) {
  VFace r = {0.0f, kf, false};
  if (kf < v.kts || kf >= v.ktf - 1) return r;
  if (kind == 'm' && kf == v.kts) kf = v.kts + 1;   // vflux[kts] = vflux[kts+1]
  if (kf >= v.ktf - 1) return r;
  r.k = kf;
  r.live = true;
  r.c = v.kvdif * v.rdn[kf + 1] / (0.5f * (AT(v.alt, kf, j, i) + AT(v.alt, kf + 1, j, i)));
  return r;
}

__device__ inline float vface_flux(const Dims& D, const VFace& fc, const float* f, const float* base1, int j,
                                   int i) {
  if (!fc.live) return 0.0f;
  const float df = AT(f, fc.k + 1, j, i) - AT(f, fc.k, j, i);
  return base1 ? fc.c * (df - base1[fc.k + 1] + base1[fc.k]) : fc.c * df;
}

__global__ void k_scalar_diff(Box b, Dims D, ScalarTend a) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  const bool dh = a.bh.has(k, j, i), dv = a.bv.has(k, j, i);
  if (!dh && !dv) return;
  const int q0 = D.ix(k, j, i);
  HCoef c = {};
  if (dh) c = hdiff_coefs(D, a.hd, k, j, i);
  // kinds needed: 'p' for the water vapour of a moist array, 'm' for the rest
  const bool p1 = a.moist_step && a.scs <= pw::P_QV && pw::P_QV <= a.sce;
  const bool p2 = a.scalar2 && a.moist_step2 && a.scs2 <= pw::P_QV && pw::P_QV <= a.sce2;
  const bool has_p = p1 || p2;
  const bool has_m = !p1 || a.sce > a.scs || (a.scalar2 && (!p2 || a.sce2 > a.scs2));
  VFace um = {}, dm = {}, up = {}, dp = {};
  float am = 0.0f, ap = 0.0f;
  if (dv) {
    const VDiff& v = a.vd;
    const float gg = F(pw::g * pw::g), m = AT2(v.mus, j, i), al = v.alt[q0];
    if (has_m) {
      um = vface(D, v, 'm', k, j, i);
      dm = vface(D, v, 'm', k - 1, j, i);
      am = gg / m / al * v.rdnw[k];
    }
    if (has_p) {
      up = vface(D, v, 'p', k, j, i);
      dp = vface(D, v, 'p', k - 1, j, i);
      ap = gg * m / al * v.rdnw[k];
    }
  }
  for (int arr = 0; arr < (a.scalar2 ? 2 : 1); ++arr) {
    const float* const sbase = arr ? a.scalar2 : a.scalar;
    float* const st_base = arr ? a.sc_tend2 : a.sc_tend;
    const int scs = arr ? a.scs2 : a.scs, sce = arr ? a.sce2 : a.sce, moist = arr ? a.moist_step2 : a.moist_step;
#pragma unroll 1
    for (int im = scs; im <= sce; ++im) {
      const float* f = sbase + im * a.n3;
      const int q = im * (int)a.n3 + q0;
      float st = st_base[q];
      if (dh) {
        auto Fv = [&](int dj, int di) { return AT(f, k, j + dj, i + di); };
        const float f0 = Fv(0, 0);
        st = st + (c.mrdx * (c.mkxp * (Fv(0, 1) - f0) - c.mkxm * (f0 - Fv(0, -1))) +
                   c.mrdy * (c.mkyp * (Fv(1, 0) - f0) - c.mkym * (f0 - Fv(-1, 0))));
      }
      if (dv) {
        if (moist && im == pw::P_QV)
          st = st + ap * (vface_flux(D, up, f, a.vd.base1, j, i) - vface_flux(D, dp, f, a.vd.base1, j, i));
        else
          st = st + am * (vface_flux(D, um, f, nullptr, j, i) - vface_flux(D, dm, f, nullptr, j, i));
      }
      st_base[q] = st;
    }
  }
}

// scalar2 non-null: a second species array [scs2, sce2] (its own tendencies
// and moist flag) over the same fluxes, index set and diffusion, in the same
// launches.
int pw_rk_scalar_tend_fused(const pywrf_grid_t* g, int scs, int sce, int rk_step, const float* ru, const float* rv,
                            const float* ww, const float* mut, const float* alt, const float* scalar,
                            float* scalar_tends, float* advect_tend, float* h_tendency, float* z_tendency,
                            const float* base, int moist_step, const float* fnm, const float* fnp,
                            const float* msfux, const float* msfuy, const float* msfvx, const float* msfvx_inv,
                            const float* msfvy, const float* msftx, const float* msfty, float rdx, float rdy,
                            const float* rdn, const float* rdnw, float kvdq, const float* xkmhd, pywrf_stream_t s,
                            const float* scalar2, float* scalar_tends2, float* advect_tend2, int scs2, int sce2,
                            int moist_step2) {
  CHECK(g, 4);
  ScalarTend a;
  a.x = {g->ids, g->ide, g->jds, g->jde, imin(g->kte, g->kde - 1)};
  if (g->kts != 0 || a.x.ktf + 1 > g->nz || g->kte > g->nz || scs < 0 || sce < scs) return PYWRF_E_RANGE;
  a.adv = {nullptr, ru, rv, ww, msftx, fnm, fnp, rdnw, nullptr, rdx, rdy};
  a.hd = {nullptr, nullptr, mut, xkmhd, msfux, msfuy, msfvx, msfvx_inv, msfvy, msftx, msfty, nullptr, rdx, rdy, 'm'};
  a.vd = {nullptr, alt, mut, rdn, rdnw, base, nullptr, nullptr, kvdq, a.x.ktf, g->kts, 'm'};
  a.scalar = scalar; a.adv_tend = advect_tend; a.sc_tend = scalar_tends; a.h_t = h_tendency; a.z_t = z_tendency;
  a.n3 = (long)g->nz * g->ny * g->nx;
  if ((long)(sce + 1) * a.n3 >= (1L << 31)) return PYWRF_E_RANGE;   // species offsets are 32-bit
  a.scs = scs; a.sce = sce; a.rk1 = rk_step == 1; a.moist_step = moist_step; a.accumulate = 0;
  a.nsg1 = (sce - scs + SA_SPW) / SA_SPW;
  a.scalar2 = scalar2; a.sc_tend2 = scalar_tends2; a.adv_tend2 = advect_tend2;
  a.scs2 = scs2; a.sce2 = sce2; a.moist_step2 = moist_step2;
  if (scalar2 && (scs2 < 0 || sce2 < scs2 || (long)(sce2 + 1) * a.n3 >= (1L << 31) || !scalar_tends2 ||
                  !advect_tend2))
    return PYWRF_E_RANGE;
  int ktf = a.x.ktf, jtf = imin(g->jte, g->jde - 1), itf = imin(g->ite, g->ide - 1);
  a.bzero = mkbox(g->kts, g->kte, g->jts, g->jte, g->its, g->ite);
  a.badv = mkbox(g->kts, ktf, g->jts, jtf, g->its, itf);
  a.by = mkbox(g->kts, ktf, imax(g->jts, g->jds + 1), imin(g->jte, g->jde - 2), g->its, itf);
  a.bx = mkbox(g->kts, ktf, g->jts, jtf, imax(g->ids + 1, g->its), imin(g->ide - 2, g->ite));
  a.bh = mkbox(g->kts, ktf, imax(g->jds + 1, g->jts), imin(g->jde - 2, g->jte), imax(g->ids + 1, g->its),
               imin(g->ide - 2, g->ite));
  a.bv = mkbox(g->kts + 1, ktf, g->jts, jtf, g->its, itf);
  a.span = mkbox(0, g->nz, 0, g->ny, 0, g->nx);   // h_tendency and z_tendency are zeroed everywhere
  pw_launch_scalar_adv(g, a, s);
  if (a.rk1 && a.bzero.size() > 0) {
    const Box& bz = a.bzero;
    dim3 grid((unsigned)((bz.ni() + PW_TX - 1) / PW_TX), (unsigned)((bz.nj() + PW_TY - 1) / PW_TY),
              (unsigned)bz.nk());
    hipLaunchKernelGGL(k_scalar_diff, grid, dim3(PW_TX, PW_TY), 0, (hipStream_t)s, bz, dims_of(g), a);
  }
  PW_RET();
}
