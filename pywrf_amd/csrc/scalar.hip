// Scalar transport and update, the moist equation of state, PBL decoupling and
// the kinematic surface w (pywrf/wrf_dynamics.py:3982-4324).
#include "common.h"
#include "consts.h"

#define AT(p, k, j, i) at_b((p), D.ob((k), (j), (i)))
#define AT2(p, j, i) at_b((p), D.ob2((j), (i)))

int pw_check_grid(const pywrf_grid_t* g, int r);
#define CHECK(g, r) do { int _e = pw_check_grid((g), (r)); if (_e) return _e; } while (0)

int pw_rk_scalar_tend_fused(const pywrf_grid_t* g, int scs, int sce, int rk_step, const float* ru, const float* rv,
                            const float* ww, const float* mut, const float* alt, const float* scalar,
                            float* scalar_tends, float* advect_tend, float* h_tendency, float* z_tendency,
                            const float* base, int moist_step, const float* fnm, const float* fnp,
                            const float* msfux, const float* msfuy, const float* msfvx, const float* msfvx_inv,
                            const float* msfvy, const float* msftx, const float* msfty, float rdx, float rdy,
                            const float* rdn, const float* rdnw, float kvdq, const float* xkmhd, pywrf_stream_t s,
                            const float* scalar2 = nullptr, float* scalar_tends2 = nullptr,
                            float* advect_tend2 = nullptr, int scs2 = 0, int sce2 = -1, int moist_step2 = 0);

// Fill kernel used instead of hipMemsetAsync so that the calls stay safe under
// stream capture.
__global__ void k_fill(float* a, long n, float v) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t < n; t += (long)gridDim.x * blockDim.x) a[t] = v;
}

int pw_fill(float* a, long n, float v, pywrf_stream_t s) {
  if (n <= 0) return 0;
  long blocks = (n + PW_BLOCK - 1) / PW_BLOCK;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(PW_BLOCK), 0, (hipStream_t)s, a, n, v);
  return (int)hipGetLastError();
}

// Grouped zero fill: blockIdx.y selects the array, a grid-stride loop over x
// covers it.
#define PW_ZERO_MAX 32
struct ZeroJobs {
  float* p[PW_ZERO_MAX];
  long n[PW_ZERO_MAX];
};

__global__ void k_zero_group(ZeroJobs J) {
  float* p = J.p[blockIdx.y];
  const long n = J.n[blockIdx.y];
  const long stride = (long)gridDim.x * blockDim.x, t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long done = 0;
  if (((uintptr_t)p & 15) == 0) {   // 16-byte stores over the aligned body
    const long n4 = n / 4;
    float4* p4 = (float4*)p;
    for (long t = t0; t < n4; t += stride) p4[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    done = n4 * 4;
  }
  for (long t = done + t0; t < n; t += stride) p[t] = 0.0f;
}

extern "C" int pywrf_zero_group(int n, void* const* fields, const long long* counts, pywrf_stream_t s) {
  if (n < 0 || n > PW_ZERO_MAX || (n > 0 && (!fields || !counts))) return PYWRF_E_ARG;
  ZeroJobs J;
  long most = 0;
  for (int f = 0; f < n; ++f) {
    if (!fields[f] || counts[f] < 0) return PYWRF_E_ARG;
    J.p[f] = (float*)fields[f];
    J.n[f] = (long)counts[f];
    most = J.n[f] > most ? J.n[f] : most;
  }
  if (n == 0 || most == 0) return 0;
  long blocks = (most / 4 + PW_BLOCK - 1) / PW_BLOCK;
  if (blocks > 2048) blocks = 2048;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_zero_group, dim3((unsigned)blocks, (unsigned)n), dim3(PW_BLOCK), 0, (hipStream_t)s, J);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------------ rk_scalar_tend
extern "C" int pywrf_rk_scalar_tend(const pywrf_grid_t* g, int scs, int sce, int rk_step, const float* ru,
                                    const float* rv, const float* ww, const float* mut, const float* alt,
                                    const float* scalar, float* scalar_tends, float* advect_tend,
                                    float* h_tendency, float* z_tendency, const float* base, int moist_step,
                                    const float* fnm, const float* fnp, const float* msfux, const float* msfuy,
                                    const float* msfvx, const float* msfvx_inv, const float* msfvy,
                                    const float* msftx, const float* msfty, double rdx, double rdy,
                                    const float* rdn, const float* rdnw, double kvdif, const float* xkmhd,
                                    pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  // kvdif / prandtl is the reference's kvdq (formed in double, rounded once)
  return pw_rk_scalar_tend_fused(g, scs, sce, rk_step, ru, rv, ww, mut, alt, scalar, scalar_tends, advect_tend,
                                 h_tendency, z_tendency, base, moist_step, fnm, fnp, msfux, msfuy, msfvx, msfvx_inv,
                                 msfvy, msftx, msfty, F(rdx), F(rdy), rdn, rdnw, F(kvdif / pw::prandtl), xkmhd, s);
}

extern "C" int pywrf_rk_scalar_tend_pair(const pywrf_grid_t* g, int rk_step, int scs, int sce, const float* scalar,
                                         float* scalar_tends, float* advect_tend, int moist_step, int scs2,
                                         int sce2, const float* scalar2, float* scalar_tends2, float* advect_tend2,
                                         int moist_step2, const float* ru, const float* rv, const float* ww,
                                         const float* mut, const float* alt, float* h_tendency, float* z_tendency,
                                         const float* base, const float* fnm, const float* fnp, const float* msfux,
                                         const float* msfuy, const float* msfvx, const float* msfvx_inv,
                                         const float* msfvy, const float* msftx, const float* msfty, double rdx,
                                         double rdy, const float* rdn, const float* rdnw, double kvdif,
                                         const float* xkmhd, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  if (!scalar2 || scalar2 == scalar || scalar_tends2 == scalar_tends || advect_tend2 == advect_tend)
    return PYWRF_E_ARG;
  return pw_rk_scalar_tend_fused(g, scs, sce, rk_step, ru, rv, ww, mut, alt, scalar, scalar_tends, advect_tend,
                                 h_tendency, z_tendency, base, moist_step, fnm, fnp, msfux, msfuy, msfvx, msfvx_inv,
                                 msfvy, msftx, msfty, F(rdx), F(rdy), rdn, rdnw, F(kvdif / pw::prandtl), xkmhd, s,
                                 scalar2, scalar_tends2, advect_tend2, scs2, sce2, moist_step2);
}

// ------------------------------------------------------------------ rk_update_scalar
// All species of [scs, sce] in one pass; muold / munew formed once per point.
struct UpdS {
  float *s1, *s2;
  const float *sc_tend, *adv, *msfty, *mu_old, *mu_new, *mub;
  long n3;
  float dt;
  int rk1, scs, sce;
  // a second species array over the same mass and index set (s1b null: none)
  float *s1b, *s2b;
  const float *sc_tend_b, *adv_b;
  int scs2, sce2;
  Box bs, bi, bf;
  int jf0, jf1, if0, if1;  // rows / columns where muold and munew are set
};

#define UPD_CHUNK 8
__global__ void k_rk_update_scalar(Box all, Dims D, UpdS a) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  long q = D.ix(k, j, i), q2 = D.ix2(j, i);
  const bool inf = a.bf.has(k, j, i), ini = a.bi.has(k, j, i), ins = a.bs.has(k, j, i);
  bool hm = j >= a.jf0 && j < a.jf1 && i >= a.if0 && i < a.if1;
  // the point's operands are loaded whatever the masks say (a load under a
  // divergent branch is waited for before the branch ends); the masks select
  const float mo = a.mu_old[q2], mn = a.mu_new[q2], mb = a.mub[q2], msf = a.msfty[q2];
  float muold = hm ? mo + mb : 0.0f;
  float munew = hm ? mn + mb : 0.0f;
  const float my = (inf && ini) ? msf : 0.0f;
  // species in chunks of UPD_CHUNK: every load of a chunk is issued before its
  // stores (the arrays are distinct, which the compiler cannot assume)
  for (int arr = 0; arr < (a.s1b ? 2 : 1); ++arr) {
    float* const s1 = arr ? a.s1b : a.s1;
    float* const s2 = arr ? a.s2b : a.s2;
    const float* const adv = arr ? a.adv_b : a.adv;
    const float* const sct = arr ? a.sc_tend_b : a.sc_tend;
    const int scs = arr ? a.scs2 : a.scs, sce = arr ? a.sce2 : a.sce;
    for (int m0 = scs; m0 <= sce; m0 += UPD_CHUNK) {
      float v1[UPD_CHUNK], va[UPD_CHUNK], vt[UPD_CHUNK];
#pragma unroll
      for (int m = 0; m < UPD_CHUNK; ++m) {
        if (m0 + m > sce) break;
        const long qq = q + (m0 + m) * a.n3;
        v1[m] = a.rk1 ? s2[qq] + 0.0f : s1[qq];   // rk1 is uniform
        const float av = adv[qq], tv = sct[qq];
        va[m] = (inf && ini) ? av : 0.0f;
        vt[m] = (inf && ins) ? tv : 0.0f;
      }
#pragma unroll
      for (int m = 0; m < UPD_CHUNK; ++m) {
        if (m0 + m > sce) break;
        const long qq = q + (m0 + m) * a.n3;
        if (a.rk1) s1[qq] = v1[m];
        if (!inf) continue;
        float t = 0.0f;
        if (ini) t = va[m] * my;
        if (ins) t = t + vt[m];
        s2[qq] = (muold * v1[m] + a.dt * t) / munew;
      }
    }
  }
}

// scalar_1b non-null: a second species array [scs2, sce2] with its own
// tendencies over the same mass, map factors and index set, in the same launch.
static int rk_update_scalar(const pywrf_grid_t* g, int scs, int sce, float* scalar_1, float* scalar_2,
                            const float* sc_tend, const float* advect_tend, const float* msfty, const float* mu_old,
                            const float* mu_new, const float* mu_base, int rk_step, double dt, int spec_zone,
                            pywrf_stream_t s, float* scalar_1b, float* scalar_2b, const float* sc_tend_b,
                            const float* advect_tend_b, int scs2, int sce2) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  if (g->kte > g->nz || g->kde > g->nz || scs < 0 || sce < scs) return PYWRF_E_RANGE;
  UpdS a;
  a.s1 = scalar_1; a.s2 = scalar_2; a.sc_tend = sc_tend; a.adv = advect_tend;
  a.mu_old = mu_old; a.mu_new = mu_new; a.mub = mu_base; a.msfty = msfty; a.dt = F(dt); a.rk1 = rk_step == 1;
  a.n3 = (long)g->nz * g->ny * g->nx; a.scs = scs; a.sce = sce;
  a.s1b = scalar_1b; a.s2b = scalar_2b; a.sc_tend_b = sc_tend_b; a.adv_b = advect_tend_b;
  a.scs2 = scs2; a.sce2 = sce2;
  if (scalar_1b && (scs2 < 0 || sce2 < scs2 || !scalar_2b || !sc_tend_b || !advect_tend_b)) return PYWRF_E_RANGE;
  a.bs = mkbox(g->kts, g->kte - 1, g->jts, imin(g->jte, g->jde - 1), g->its, imin(g->ite, g->ide - 1));
  a.bi = mkbox(g->kts, imin(g->kte, g->kde - 1), imax(g->jts, g->jds + spec_zone), imin(g->jte, g->jde - spec_zone - 1),
               imax(g->its, g->ids + spec_zone), imin(g->ite, g->ide - spec_zone - 1));
  a.bf = mkbox(g->kts, g->kde - 1, g->jts, g->jde - 1, g->its, g->ide - 1);
  a.jf0 = g->jts; a.jf1 = g->jde - 1; a.if0 = g->its; a.if1 = g->ide - 1;
  Box b = a.rk1 ? mkbox(0, g->nz, 0, g->ny, 0, g->nx) : a.bf;
  PW_LAUNCH_BOX(k_rk_update_scalar, b, s, dims_of(g), a);
  PW_RET();
}

extern "C" int pywrf_rk_update_scalar(const pywrf_grid_t* g, int scs, int sce, float* scalar_1, float* scalar_2,
                                      const float* sc_tend, const float* advect_tend, const float* msfty,
                                      const float* mu_old, const float* mu_new, const float* mu_base, int rk_step,
                                      double dt, int spec_zone, pywrf_stream_t s) {
  return rk_update_scalar(g, scs, sce, scalar_1, scalar_2, sc_tend, advect_tend, msfty, mu_old, mu_new, mu_base,
                          rk_step, dt, spec_zone, s, nullptr, nullptr, nullptr, nullptr, 0, -1);
}

extern "C" int pywrf_rk_update_scalar_pair(const pywrf_grid_t* g, int rk_step, int scs, int sce, float* scalar_1,
                                           float* scalar_2, const float* sc_tend, const float* advect_tend,
                                           int scs2, int sce2, float* scalar_1b, float* scalar_2b,
                                           const float* sc_tend_b, const float* advect_tend_b, const float* msfty,
                                           const float* mu_old, const float* mu_new, const float* mu_base,
                                           double dt, int spec_zone, pywrf_stream_t s) {
  if (!scalar_1b || scalar_1b == scalar_1 || scalar_2b == scalar_2) return PYWRF_E_ARG;
  return rk_update_scalar(g, scs, sce, scalar_1, scalar_2, sc_tend, advect_tend, msfty, mu_old, mu_new, mu_base,
                          rk_step, dt, spec_zone, s, scalar_1b, scalar_2b, sc_tend_b, advect_tend_b, scs2, sce2);
}

// ------------------------------------------------------------------ calc_p_rho_phi
__global__ void k_calc_p_rho_phi(Box b, Dims D, const float* qv, float* al, const float* alb, const float* mu,
                                 const float* muts, const float* ph, float* p, const float* pb, const float* t,
                                 float p0, float t0, const float* rdnw) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  float a = -1.0f / AT2(muts, j, i) * (alb[q] * AT2(mu, j, i) + rdnw[k] * (AT(ph, k + 1, j, i) - ph[q]));
  al[q] = a;
  float qvf = 1.0f + F(pw::rvovrd) * qv[q];
  float temp = (F(pw::r_d) * (t0 + t[q]) * qvf) / (p0 * (a + alb[q]));
  float pp = powf(temp, F(pw::cpovcv));
  p[q] = pp * p0 - pb[q];
}

extern "C" int pywrf_calc_p_rho_phi(const pywrf_grid_t* g, const float* moist, float* al, const float* alb,
                                    const float* mu, const float* muts, const float* ph, float* p, const float* pb,
                                    const float* t, double p0, double t0, const float* rdnw, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  if (b.k1 + 1 > g->nz) return PYWRF_E_RANGE;
  long n3 = (long)g->nz * g->ny * g->nx;
  PW_LAUNCH_BOX(k_calc_p_rho_phi, b, s, dims_of(g), moist + pw::P_QV * n3, al, alb, mu, muts, ph, p, pb, t, F(p0),
                F(t0), rdnw);
  PW_RET();
}

// ------------------------------------------------------------------ phy_prep_part2
__global__ void k_div_by_mut(Box b, Dims D, const float* mut, float* a0, float* a1, float* a2, float* a3,
                             float* a4, float* a5) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  float m = AT2(mut, j, i);
  const float v0 = a0[q], v1 = a1[q], v2 = a2[q], v3 = a3[q], v4 = a4[q], v5 = a5[q];   // loads before stores
  a0[q] = v0 / m;
  a1[q] = v1 / m;
  a2[q] = v2 / m;
  a3[q] = v3 / m;
  a4[q] = v4 / m;
  a5[q] = v5 / m;
}

extern "C" int pywrf_phy_prep_part2(const pywrf_grid_t* g, const float* mut, float* rthblten, float* rublten,
                                    float* rvblten, float* rqvblten, float* rqcblten, float* rqiblten,
                                    pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 0);
  Box b = mkbox(g->kts, imin(g->kte, g->kde - 1), g->jts, imin(g->jte, g->jde - 1), g->its,
                imin(g->ite, g->ide - 1));
  PW_LAUNCH_BOX(k_div_by_mut, b, s, dims_of(g), mut, rublten, rvblten, rthblten, rqvblten, rqcblten, rqiblten);
  PW_RET();
}

// ------------------------------------------------------------------ set_w_surface
// One-sided terrain-slope stencils on the outer rows / columns of [jts, jte-1)
// x [its, ite-1); at the (jts, ite-2) corner the y slope is the reference's
// ht[jts+1] - ht[jts+1], and at the (jte-2, ite-2) corner the x term samples
// msftx at (jts, its).
__global__ void k_set_w_surface(Box b, Dims D, float* w, const float* ht, const float* u, const float* v,
                                const float* cf, float rdx, float rdy, const float* msftx, const float* msfty,
                                int jts, int jte, int its, int ite) {
  int j, i;
  if (!box_column(b, j, i)) return;
  float cf1 = cf[0], cf2 = cf[1], cf3 = cf[2];
  auto cv = [&](int jj) { return cf1 * AT(v, 0, jj, i) + cf2 * AT(v, 1, jj, i) + cf3 * AT(v, 2, jj, i); };
  auto cu = [&](int ii) { return cf1 * AT(u, 0, j, ii) + cf2 * AT(u, 1, j, ii) + cf3 * AT(u, 2, j, ii); };
  auto H = [&](int jj, int ii) { return AT2(ht, jj, ii); };
  bool jlo = (j == jts), jhi = (j == jte - 2), ilo = (i == its), ihi = (i == ite - 2);
  float ty, tx;
  if (jlo && ihi) ty = (H(j + 1, i) - H(j + 1, i)) * cv(j + 1);
  else if (jlo) ty = (H(j + 1, i) - H(j, i)) * cv(j + 1);
  else if (jhi) ty = (H(j, i) - H(j - 1, i)) * cv(j);
  else ty = (H(j + 1, i) - H(j, i)) * cv(j + 1) + (H(j, i) - H(j - 1, i)) * cv(j);
  if (ilo) tx = (H(j, i + 1) - H(j, i)) * cu(i + 1);
  else if (ihi) tx = (H(j, i) - H(j, i - 1)) * cu(i);
  else tx = (H(j, i + 1) - H(j, i)) * cu(i + 1) + (H(j, i) - H(j, i - 1)) * cu(i);
  float msx = (jhi && ihi) ? AT2(msftx, jts, its) : AT2(msftx, j, i);
  AT(w, 0, j, i) = AT2(msfty, j, i) * 0.5f * rdy * ty + msx * 0.5f * rdx * tx;
}

extern "C" int pywrf_set_w_surface(const pywrf_grid_t* g, float* w, const float* ht, const float* u, const float* v,
                                   const float* cf, double rdx, double rdy, const float* msftx, const float* msfty,
                                   pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (g->nz < 3) return PYWRF_E_RANGE;
  Box b = mkbox(0, 1, g->jts, g->jte - 1, g->its, g->ite - 1);
  PW_LAUNCH_COLS(k_set_w_surface, b, s, dims_of(g), w, ht, u, v, cf, F(rdx), F(rdy), msftx, msfty, g->jts, g->jte,
                 g->its, g->ite);
  PW_RET();
}

// ------------------------------------------------------------------ device NaN flag
// The reference stops when p holds a NaN after an RK stage or after the
// physics (solver.py:1648-1654, 1768-1774: torch.isnan(p).any(), the indices
// printed, sys.exit).  Here the check stays on the device: a scan of the
// field records, in a 4-int flag array, the code of the first check that saw
// a NaN (flag[0], 0 while none has), the linear index of one NaN point of
// that check (flag[1]) and the number of NaN points over all checks
// (flag[2]).  The host reads the flag when it chooses (after a graph replay,
// a run), so the step needs no synchronisation.
__global__ void k_nan_scan(const float* f, long n, int code, int* flag) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  int cnt = 0;
  long first = -1;
  for (long q = i; q < n; q += stride) {
    const float v = f[q];
    if (v != v) {
      ++cnt;
      if (first < 0) first = q;
    }
  }
  if (cnt) {
    atomicAdd(flag + 2, cnt);
    if (atomicCAS(flag, 0, code) == 0) atomicExch(flag + 1, (int)first);
  }
}

__global__ void k_nan_flag_clear(int* flag) {
  if (threadIdx.x < 4) flag[threadIdx.x] = 0;
}

extern "C" int pywrf_nan_scan(const pywrf_grid_t* g, const float* field, int code, int* flag, pywrf_stream_t s) {
  if (!pw_size_ok(g) || !flag || code <= 0) return PYWRF_E_ARG;
  const long n = (long)g->nz * g->ny * g->nx;
  const int blocks = (int)((n + 256L * 8 - 1) / (256L * 8));
  hipLaunchKernelGGL(k_nan_scan, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(256), 0, (hipStream_t)s, field, n,
                     code, flag);
  PW_RET();
}

extern "C" int pywrf_nan_flag_clear(int* flag, pywrf_stream_t s) {
  if (!flag) return PYWRF_E_ARG;
  hipLaunchKernelGGL(k_nan_flag_clear, dim3(1), dim3(64), 0, (hipStream_t)s, flag);
  PW_RET();
}

// ------------------------------------------------------------------ boundary clock
// The driver's step counter, boundary record index and boundary time
// (solver.py:457-458, 713) advanced on the device at the start of a step, so
// that a captured graph of steps forms each replay's dtbc itself: step += 1;
// it_bdy = floor(step / bdy_interval) when step mod bdy_interval == 1 (mod and
// division rounding towards minus infinity, as the host's Python integers);
// dtbc = (step - it_bdy bdy_interval) dt formed in double and rounded to
// float once, as the kernels take the host's Python float.
__global__ void k_clock_advance(long long* step, long long* it_bdy, float* dtbc, long long n, double dt) {
  if (threadIdx.x != 0) return;
  const long long s = *step + 1;
  long long q = s / n, r = s % n;
  if (r != 0 && ((r < 0) != (n < 0))) {
    r += n;
    --q;
  }
  const long long it = r == 1 ? q : *it_bdy;
  *step = s;
  *it_bdy = it;
  *dtbc = (float)((double)(s - it * n) * dt);
}

extern "C" int pywrf_clock_advance(void* step, void* it_bdy, float* dtbc, long long bdy_interval, double dt,
                                   pywrf_stream_t s) {
  if (!step || !it_bdy || !dtbc || bdy_interval <= 0) return PYWRF_E_ARG;
  hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, (hipStream_t)s, (long long*)step, (long long*)it_bdy,
                     dtbc, bdy_interval, dt);
  PW_RET();
}
