// Lateral-boundary zone kernels (pywrf/wrf_boundary.py:52-587): relaxation
// and specified-zone tendencies, specified-zone updates, zero-gradient and
// flow-dependent edges, mass weighting and the end-of-step specified values.
//
// Boundary arrays have the reference driver's layout (solver.py:468-703):
// an x-boundary array is the (kz, ny, 5) view permute(1, 2, 0) of a
// contiguous (5, kz, ny) array, a y-boundary array the (kz, 5, nx) view
// permute(1, 0, 2) of a contiguous (5, kz, nx) array, so element [k, j, c] of
// an x-boundary lies at (c kz + k) ny + j and element [k, r, i] of a
// y-boundary at (r kz + k) nx + i; c / r count inward from the domain edge and
// kz is the level count of the field the array belongs to.  Each kernel is one launch over the ring of points it writes;
// where the reference writes a point in several sequential statements the
// kernel reproduces the last writer (or repeats the update as often as the
// statements cover the point).
#include "common.h"
#include "consts.h"
#include "openbc.h"

#define AT(p, k, j, i) at_b((p), D.ob((k), (j), (i)))
#define AT2(p, j, i) at_b((p), D.ob2((j), (i)))

int pw_check_grid(const pywrf_grid_t* g, int r);
#define CHECK(g, r) do { int _e = pw_check_grid((g), (r)); if (_e) return _e; } while (0)

struct Bdy {
  const float *xs, *xe, *ys, *ye, *txs, *txe, *tys, *tye;
};

__device__ inline float bx(const Dims& D, const float* a, int k, int j, int c) { return a[((long)c * D.nz + k) * D.ny + j]; }
__device__ inline float by(const Dims& D, const float* a, int k, int r, int i) { return a[((long)r * D.nz + k) * D.nx + i]; }

// staggering of the relax/spec routines (wrf_boundary.py:_bounds logic)
static void bounds_of(char v, const pywrf_grid_t* g, int& ibe, int& jbe, int& ktf) {
  ibe = g->ide - 1;
  jbe = g->jde - 1;
  ktf = g->kde - 1;
  if (v == 'u') ibe = g->ide;
  if (v == 'v') jbe = g->jde;
  if (v == 'm' || v == 'h') ktf = g->kte;
}

// ------------------------------------------------------------------ relax_bdytend
struct Relax {
  const float* field;
  const float* mut;  // non-null: field is the unweighted scalar, weighted by mut on mbox (relax_bdy_scalar)
  Box mbox;
  float* tend;
  Bdy b;
  const float *fcx, *gcx;
  float dtbc;
  const float* dtbc_dev;  // non-null: the boundary time is read from device memory (graph replays)
  int ibs, ibe, jbs, jbe;
};

__device__ inline float relax_field(const Dims& D, const Relax& a, int k, int j, int i) {
  if (!a.mut) return AT(a.field, k, j, i);
  return a.mbox.has(k, j, i) ? AT(a.field, k, j, i) * AT2(a.mut, j, i) : 0.0f;
}

__global__ void k_relax(Box all, int w, Dims D, Relax a) {
  int k, j, i;
  if (!edge_point(all, w, k, j, i)) return;
  // which strip (y-start, y-end, x-start, x-end) and b_dist owns this point
  int side = -1, bd = 0;
  for (int b = 1; b <= 3 && side < 0; ++b) {
    if (j == a.jbs + b && i >= a.ibs + b && i < a.ibe - b) { side = 0; bd = b; }
    else if (j == a.jbe - 1 - b && i >= a.ibs + b && i < a.ibe - b) { side = 1; bd = b; }
    else if (i == a.ibs + b && j >= a.jbs + b + 1 && j < a.jbe - b - 1) { side = 2; bd = b; }
    else if (i == a.ibe - 1 - b && j >= a.jbs + b + 1 && j < a.jbe - b - 1) { side = 3; bd = b; }
  }
  if (side < 0) return;
  const float dt = a.dtbc_dev ? *a.dtbc_dev : a.dtbc;
  auto FY = [&](const float* bv, const float* bt, int r, int jj, int ii) {
    return by(D, bv, k, r, ii) + dt * by(D, bt, k, r, ii) - relax_field(D, a, k, jj, ii);
  };
  auto FX = [&](const float* bv, const float* bt, int c, int jj, int ii) {
    return bx(D, bv, k, jj, c) + dt * bx(D, bt, k, jj, c) - relax_field(D, a, k, jj, ii);
  };
  float f0, f1, f2, f3, f4;
  if (side == 0) {
    f0 = FY(a.b.ys, a.b.tys, bd, j, i); f1 = FY(a.b.ys, a.b.tys, bd, j, i - 1); f2 = FY(a.b.ys, a.b.tys, bd, j, i + 1);
    f3 = FY(a.b.ys, a.b.tys, bd - 1, j - 1, i); f4 = FY(a.b.ys, a.b.tys, bd + 1, j + 1, i);
  } else if (side == 1) {
    f0 = FY(a.b.ye, a.b.tye, bd, j, i); f1 = FY(a.b.ye, a.b.tye, bd, j, i - 1); f2 = FY(a.b.ye, a.b.tye, bd, j, i + 1);
    f3 = FY(a.b.ye, a.b.tye, bd - 1, j + 1, i); f4 = FY(a.b.ye, a.b.tye, bd + 1, j - 1, i);
  } else if (side == 2) {
    f0 = FX(a.b.xs, a.b.txs, bd, j, i); f1 = FX(a.b.xs, a.b.txs, bd, j - 1, i); f2 = FX(a.b.xs, a.b.txs, bd, j + 1, i);
    f3 = FX(a.b.xs, a.b.txs, bd - 1, j, i - 1); f4 = FX(a.b.xs, a.b.txs, bd + 1, j, i + 1);
  } else {
    f0 = FX(a.b.xe, a.b.txe, bd, j, i); f1 = FX(a.b.xe, a.b.txe, bd, j - 1, i); f2 = FX(a.b.xe, a.b.txe, bd, j + 1, i);
    f3 = FX(a.b.xe, a.b.txe, bd - 1, j, i + 1); f4 = FX(a.b.xe, a.b.txe, bd + 1, j, i - 1);
  }
  long q = D.ix(k, j, i);
  a.tend[q] = a.tend[q] + a.fcx[bd] * f0 - a.gcx[bd] * (f1 + f2 + f3 + f4 - 4.0f * f0);
}

static int relax_launch(const pywrf_grid_t* g, Relax& a, char v, pywrf_stream_t s) {
  int ibe, jbe, ktf;
  bounds_of(v, g, ibe, jbe, ktf);
  a.ibs = g->ids; a.jbs = g->jds; a.ibe = ibe; a.jbe = jbe;
  if (g->ids < 0 || ibe + 2 > g->nx || g->jds < 0 || jbe + 2 > g->ny || ktf > g->nz) return PYWRF_E_RANGE;
  Box all = mkbox(g->kts, ktf, g->jds, jbe, g->ids, ibe);
  pw_launch_edges(k_relax, all, 4, s, dims_of(g), a);  // strips b_dist = 1..3 lie in a band of 4
  PW_RET();
}

extern "C" int pywrf_relax_bdytend(const pywrf_grid_t* g, const float* field, float* field_tend, const float* bxs,
                                   const float* bxe, const float* bys, const float* bye, const float* btxs,
                                   const float* btxe, const float* btys, const float* btye, char variable_in,
                                   double dtbc, const float* fcx, const float* gcx, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  Relax a;
  a.field = field; a.mut = nullptr; a.mbox = mkbox(0, 0, 0, 0, 0, 0); a.tend = field_tend;
  a.b = {bxs, bxe, bys, bye, btxs, btxe, btys, btye};
  a.fcx = fcx; a.gcx = gcx; a.dtbc = F(dtbc); a.dtbc_dev = nullptr;
  return relax_launch(g, a, variable_in, s);
}

extern "C" int pywrf_relax_bdytend_dtbc_dev(const pywrf_grid_t* g, const float* field, float* field_tend,
                                            const float* bxs, const float* bxe, const float* bys, const float* bye,
                                            const float* btxs, const float* btxe, const float* btys,
                                            const float* btye, char variable_in, const float* dtbc, const float* fcx,
                                            const float* gcx, pywrf_stream_t s) {
  if (!pw_size_ok(g) || !dtbc) return PYWRF_E_RANGE;
  Relax a;
  a.field = field; a.mut = nullptr; a.mbox = mkbox(0, 0, 0, 0, 0, 0); a.tend = field_tend;
  a.b = {bxs, bxe, bys, bye, btxs, btxe, btys, btye};
  a.fcx = fcx; a.gcx = gcx; a.dtbc = 0.0f; a.dtbc_dev = dtbc;
  return relax_launch(g, a, variable_in, s);
}

extern "C" int pywrf_relax_bdy_scalar(const pywrf_grid_t* g, float* scalar_tend, const float* scalar,
                                      const float* mu, const float* bxs, const float* bxe, const float* bys,
                                      const float* bye, const float* btxs, const float* btxe, const float* btys,
                                      const float* btye, double dtbc, const float* fcx, const float* gcx,
                                      pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  Relax a;
  a.field = scalar; a.mut = mu; a.tend = scalar_tend;
  a.mbox = mkbox(g->kts, g->kte - 1, imax(g->jts - 1, g->jds), imin(g->jte + 1, g->jde - 1), imax(g->its - 1, g->ids),
                 imin(g->ite + 1, g->ide - 1));
  a.b = {bxs, bxe, bys, bye, btxs, btxe, btys, btye};
  a.fcx = fcx; a.gcx = gcx; a.dtbc = F(dtbc); a.dtbc_dev = nullptr;
  return relax_launch(g, a, 'q', s);
}

extern "C" int pywrf_relax_bdy_scalar_dtbc_dev(const pywrf_grid_t* g, float* scalar_tend, const float* scalar,
                                               const float* mu, const float* bxs, const float* bxe, const float* bys,
                                               const float* bye, const float* btxs, const float* btxe,
                                               const float* btys, const float* btye, const float* dtbc,
                                               const float* fcx, const float* gcx, pywrf_stream_t s) {
  if (!pw_size_ok(g) || !dtbc) return PYWRF_E_RANGE;
  Relax a;
  a.field = scalar; a.mut = mu; a.tend = scalar_tend;
  a.mbox = mkbox(g->kts, g->kte - 1, imax(g->jts - 1, g->jds), imin(g->jte + 1, g->jde - 1), imax(g->its - 1, g->ids),
                 imin(g->ite + 1, g->ide - 1));
  a.b = {bxs, bxe, bys, bye, btxs, btxe, btys, btye};
  a.fcx = fcx; a.gcx = gcx; a.dtbc = 0.0f; a.dtbc_dev = dtbc;
  return relax_launch(g, a, 'q', s);
}

// ------------------------------------------------------------------ spec_bdytend
struct TendJob {
  Box all;
  Dims D;
  float* tend;
  Bdy b;
  int ibs, ibe, jbs, jbe;
};

__device__ inline void spec_tend_point(const TendJob& a, int k, int j, int i) {
  const Dims& D = a.D;
  float v;
  bool xr = j >= a.jbs && j < a.jbe;
  if (xr && i == a.ibe - 1) v = bx(D, a.b.txe, k, j, 0);
  else if (xr && i == a.ibs) v = bx(D, a.b.txs, k, j, 0);
  else if (j == a.jbe - 1) v = by(D, a.b.tye, k, 0, i);
  else if (j == a.jbs) v = by(D, a.b.tys, k, 0, i);
  else return;
  AT(a.tend, k, j, i) = v;
}

__global__ void k_spec_bdytend(Box all, int w, TendJob a) {
  int k, j, i;
  if (!edge_point(all, w, k, j, i)) return;
  spec_tend_point(a, k, j, i);
}

static int spec_tend_job(const pywrf_grid_t* g, float* field_tend, const float* btxs, const float* btxe,
                         const float* btys, const float* btye, char variable_in, TendJob& a) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  int ibe, jbe, ktf;
  bounds_of(variable_in, g, ibe, jbe, ktf);
  if (g->ids < 0 || ibe > g->nx || g->jds < 0 || jbe > g->ny || ktf > g->nz) return PYWRF_E_RANGE;
  a.all = mkbox(g->kts, ktf, g->jds, jbe, g->ids, ibe);
  a.D = dims_of(g);
  a.tend = field_tend;
  a.b = {nullptr, nullptr, nullptr, nullptr, btxs, btxe, btys, btye};
  a.ibs = g->ids; a.ibe = ibe; a.jbs = g->jds; a.jbe = jbe;
  return 0;
}

extern "C" int pywrf_spec_bdytend(const pywrf_grid_t* g, float* field_tend, const float* btxs, const float* btxe,
                                  const float* btys, const float* btye, char variable_in, pywrf_stream_t s) {
  TendJob a;
  int e = spec_tend_job(g, field_tend, btxs, btxe, btys, btye, variable_in, a);
  if (e) return e;
  pw_launch_edges(k_spec_bdytend, a.all, 1, s, a);
  PW_RET();
}

// Several spec_bdytend calls on distinct tendencies in one launch
// (blockIdx.y selects the job).
#define PW_TEND_GROUP_MAX 8
struct TendJobs {
  TendJob j[PW_TEND_GROUP_MAX];
};

__global__ void k_spec_bdytend_group(TendJobs J) {
  const TendJob& a = J.j[blockIdx.y];
  if ((int)blockIdx.z >= a.all.nk()) return;
  int k, j, i;
  if (!edge_point(a.all, 1, k, j, i)) return;
  spec_tend_point(a, k, j, i);
}

extern "C" int pywrf_spec_bdytend_group(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* tends,
                                        void* const* bdy_tends, const char* variables, const int* is2d,
                                        pywrf_stream_t s) {
  if (!pw_size_ok(g3) || !pw_size_ok(g2)) return PYWRF_E_RANGE;
  if (n < 0 || n > PW_TEND_GROUP_MAX) return PYWRF_E_ARG;
  TendJobs J;
  int m = 0, maxk = 0;
  long most = 0;
  for (int f = 0; f < n; ++f) {
    TendJob a;
    int e = spec_tend_job(is2d[f] ? g2 : g3, (float*)tends[f], (const float*)bdy_tends[4 * f],
                          (const float*)bdy_tends[4 * f + 1], (const float*)bdy_tends[4 * f + 2],
                          (const float*)bdy_tends[4 * f + 3], variables[f], a);
    if (e) return e;
    if (a.all.size() <= 0) continue;
    if (a.all.nj() < 2 || a.all.ni() < 2) {
      pw_launch_edges(k_spec_bdytend, a.all, 1, s, a);
      continue;
    }
    J.j[m++] = a;
    most = edge_count(a.all, 1) > most ? edge_count(a.all, 1) : most;
    maxk = a.all.nk() > maxk ? a.all.nk() : maxk;
  }
  if (m > 0)
    hipLaunchKernelGGL(k_spec_bdytend_group, dim3((unsigned)((most + PW_EDGE_BLOCK - 1) / PW_EDGE_BLOCK), (unsigned)m,
                       (unsigned)maxk), dim3(PW_EDGE_BLOCK), 0, (hipStream_t)s, J);
  PW_RET();
}

extern "C" int pywrf_spec_bdy_scalar(const pywrf_grid_t* g, float* scalar_tend, const float* btxs,
                                     const float* btxe, const float* btys, const float* btye, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  return pywrf_spec_bdytend(g, scalar_tend, btxs, btxe, btys, btye, 'q', s);
}

// ------------------------------------------------------------------ spec_bdyupdate
struct UpdJob {
  Box all;
  Dims D;
  float* f;
  const float* t;
  float dt;
  int jts, jtf, its, itf, jbs, jbe;
};

__device__ inline void spec_upd_point(const UpdJob& b, int k, int j, int i) {
  const Dims& D = b.D;
  bool ic = i >= b.its && i < b.itf, jc = j >= b.jbs + 1 && j < b.jbe - 1;
  int n = (j == b.jts && ic) + (j == b.jtf - 1 && ic) + (i == b.its && jc) + (i == b.itf - 1 && jc);
  if (!n) return;
  long q = D.ix(k, j, i);
  float v = b.f[q], tv = b.t[q];
  for (int m = 0; m < n; ++m) v = v + b.dt * tv;
  b.f[q] = v;
}

__global__ void k_spec_bdyupdate(Box all, int w, UpdJob b) {
  int k, j, i;
  if (!edge_point(all, w, k, j, i)) return;
  spec_upd_point(b, k, j, i);
}

// Host side of spec_bdyupdate: the job and its edge-band width (1 when the
// updated rows jts, jtf-1 and columns its, itf-1 are the box's sides, else 0).
static int spec_upd_job(const pywrf_grid_t* g, float* field, const float* field_tend, double dt, char variable,
                        UpdJob& b, int& w) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  char v = variable;
  if (v == 'U') v = 'u'; else if (v == 'V') v = 'v'; else if (v == 'M') v = 'm'; else if (v == 'H') v = 'h';
  int itf = imin(g->ite, g->ide - 1), jtf = imin(g->jte, g->jde - 1), ktf = g->kde - 1, jbe = g->jde - 1;
  if (v == 'u') itf = imin(g->ite, g->ide);
  if (v == 'v') { jbe = g->jde; jtf = imin(g->jte, g->jde); }
  if (v == 'm' || v == 'h') ktf = g->kte;
  int nz = g->nz;
  if (g->kds == 0 && g->kde == 0) { ktf = 1; nz = 1; }  // 2-D field (the reference's unsqueeze path)
  if (itf > g->nx || jbe > g->ny || jtf > g->ny || ktf > nz || g->its < 0 || g->jts < 0) return PYWRF_E_RANGE;
  pywrf_grid_t g2 = *g;
  g2.nz = nz;
  b.all = mkbox(g->kts, ktf, imin(g->jts, g->jds), imax(jtf, jbe), g->its, itf);
  b.D = dims_of(&g2);
  b.f = field; b.t = field_tend; b.dt = F(dt);
  b.jts = g->jts; b.jtf = jtf; b.its = g->its; b.itf = itf; b.jbs = g->jds; b.jbe = jbe;
  w = (g->jts == b.all.j0 && jtf == b.all.j1) ? 1 : 0;
  return 0;
}

extern "C" int pywrf_spec_bdyupdate(const pywrf_grid_t* g, float* field, const float* field_tend, double dt,
                                    char variable, pywrf_stream_t s) {
  UpdJob b;
  int w, e = spec_upd_job(g, field, field_tend, dt, variable, b, w);
  if (e) return e;
  pw_launch_edges(k_spec_bdyupdate, b.all, w, s, b);
  PW_RET();
}

// Several spec_bdyupdate calls on distinct fields in one launch (blockIdx.y
// selects the job); a job whose update rows are not its box's sides runs on
// its own as in pywrf_spec_bdyupdate.
#define PW_UPD_GROUP_MAX 8
struct UpdJobs {
  int n;
  UpdJob j[PW_UPD_GROUP_MAX];
};

__global__ void k_spec_bdyupdate_group(UpdJobs J) {
  const UpdJob& b = J.j[blockIdx.y];
  if ((int)blockIdx.z >= b.all.nk()) return;
  int k, j, i;
  if (!edge_point(b.all, 1, k, j, i)) return;
  spec_upd_point(b, k, j, i);
}

extern "C" int pywrf_spec_bdyupdate_group(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                                          void* const* tends, const char* variables, const int* is2d, double dt,
                                          pywrf_stream_t s) {
  if (!pw_size_ok(g3) || !pw_size_ok(g2)) return PYWRF_E_RANGE;
  if (n < 0 || n > PW_UPD_GROUP_MAX) return PYWRF_E_ARG;
  UpdJobs J;
  J.n = 0;
  long most = 0;
  int maxk = 0;
  for (int f = 0; f < n; ++f) {
    UpdJob b;
    int w, e = spec_upd_job(is2d[f] ? g2 : g3, (float*)fields[f], (const float*)tends[f], dt, variables[f], b, w);
    if (e) return e;
    if (b.all.size() <= 0) continue;
    if (w == 1 && 2 <= b.all.nj() && 2 <= b.all.ni()) {
      J.j[J.n++] = b;
      most = edge_count(b.all, 1) > most ? edge_count(b.all, 1) : most;
      maxk = b.all.nk() > maxk ? b.all.nk() : maxk;
    } else {
      pw_launch_edges(k_spec_bdyupdate, b.all, w, s, b);
    }
  }
  if (J.n > 0)
    hipLaunchKernelGGL(k_spec_bdyupdate_group, dim3((unsigned)((most + PW_EDGE_BLOCK - 1) / PW_EDGE_BLOCK),
                       (unsigned)J.n, (unsigned)maxk), dim3(PW_EDGE_BLOCK), 0, (hipStream_t)s, J);
  PW_RET();
}

// ------------------------------------------------------------------ spec_bdyupdate_ph
struct PhStrip {
  Box all;
  int w;
  const float *ph_save, *ft, *mu_tend, *muts;
  float* f;
  float dt;
  int jts, ibs, ibe, jbs, jbe;
};

__device__ inline void ph_strip_point(const Dims& D, const PhStrip& a, int k, int j, int i) {
  bool ir = i >= a.ibs && i < a.ibe, jc = j >= a.jbs + 1 && j < a.jbe - 1;
  int n = (j == a.jts && ir) + (j == a.jbe - 1 && ir) + (i == a.ibs && jc) + (i == a.ibe - 1 && jc);
  if (!n) return;
  long q = D.ix(k, j, i);
  float mt = AT2(a.muts, j, i);
  float mo = mt - a.dt * AT2(a.mu_tend, j, i);
  float v = a.f[q];
  for (int m = 0; m < n; ++m) v = v * mo / mt + a.dt * a.ft[q] / mt + a.ph_save[q] * (mo / mt - 1.0f);
  a.f[q] = v;
}

__global__ void k_spec_bdyupdate_ph(Box all, int w, Dims D, PhStrip a) {
  int k, j, i;
  if (!edge_point(all, w, k, j, i)) return;
  ph_strip_point(D, a, k, j, i);
}

static int ph_strip_setup(const pywrf_grid_t* g, const float* ph_save, float* field, const float* field_tend,
                          const float* mu_tend, const float* muts, double dt, char variable_in, PhStrip& a) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  char v = variable_in;
  if (v == 'U') v = 'u'; else if (v == 'V') v = 'v'; else if (v == 'M') v = 'm'; else if (v == 'H') v = 'h';
  int ibe, jbe, ktf;
  bounds_of(v, g, ibe, jbe, ktf);
  if (ibe > g->nx || jbe > g->ny || ktf > g->nz || g->ids < 0 || g->jds < 0) return PYWRF_E_RANGE;
  a.all = mkbox(g->kts, ktf, imin(g->jts, g->jds), imax(jbe, g->jts + 1), g->ids, ibe);
  a.w = (g->jts == a.all.j0 && jbe == a.all.j1) ? 1 : 0;
  a.ph_save = ph_save; a.f = field; a.ft = field_tend; a.mu_tend = mu_tend; a.muts = muts; a.dt = F(dt);
  a.jts = g->jts; a.ibs = g->ids; a.ibe = ibe; a.jbs = g->jds; a.jbe = jbe;
  return 0;
}

extern "C" int pywrf_spec_bdyupdate_ph(const pywrf_grid_t* g, const float* ph_save, float* field,
                                       const float* field_tend, const float* mu_tend, const float* muts, double dt,
                                       char variable_in, pywrf_stream_t s) {
  PhStrip a;
  if (int e = ph_strip_setup(g, ph_save, field, field_tend, mu_tend, muts, dt, variable_in, a)) return e;
  pw_launch_edges(k_spec_bdyupdate_ph, a.all, a.w, s, dims_of(g), a);
  PW_RET();
}

// ------------------------------------------------------------------ zero_grad_bdy
struct ZgStrip {
  Box all;
  float* f;
  int ibs, ibe, jbs, jbe;
};

__device__ inline void zero_grad_point(const Dims& D, const ZgStrip& a, int k, int j, int i) {
  bool ring = (j == a.jbs || j == a.jbe - 1) || (i == a.ibs || i == a.ibe - 1);
  if (!ring) return;
  int sj = j < a.jbs + 1 ? a.jbs + 1 : (j > a.jbe - 2 ? a.jbe - 2 : j);
  int si = i < a.ibs + 1 ? a.ibs + 1 : (i > a.ibe - 2 ? a.ibe - 2 : i);
  AT(a.f, k, j, i) = AT(a.f, k, sj, si);
}

__global__ void k_zero_grad(Box all, int w, Dims D, ZgStrip a) {
  int k, j, i;
  if (!edge_point(all, w, k, j, i)) return;
  zero_grad_point(D, a, k, j, i);
}

static int zero_grad_setup(const pywrf_grid_t* g, float* field, char variable_in, ZgStrip& a) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  char v = variable_in;
  if (v == 'U') v = 'u'; else if (v == 'V') v = 'v';
  int ibe = g->ide - 1, jbe = g->jde - 1, ktf = g->kde - 1;
  if (v == 'u') ibe = g->ide;
  if (v == 'v') jbe = g->jde;
  if (v == 'w') ktf = g->kde;
  if (ibe > g->nx || jbe > g->ny || ktf > g->nz || g->ids < 0 || g->jds < 0) return PYWRF_E_RANGE;
  a.all = mkbox(g->kts, ktf, g->jds, jbe, g->ids, ibe);
  a.f = field; a.ibs = g->ids; a.ibe = ibe; a.jbs = g->jds; a.jbe = jbe;
  return 0;
}

extern "C" int pywrf_zero_grad_bdy(const pywrf_grid_t* g, float* field, char variable_in, pywrf_stream_t s) {
  ZgStrip a;
  if (int e = zero_grad_setup(g, field, variable_in, a)) return e;
  pw_launch_edges(k_zero_grad, a.all, 1, s, dims_of(g), a);
  PW_RET();
}

// spec_bdyupdate_ph and zero_grad_bdy (on another array) in one launch: grid
// row y = 0 takes the first's band, y = 1 the second's (both one cell wide).
__global__ void k_ph_zero_grad(Dims D, PhStrip p, ZgStrip z) {
  int k, j, i;
  if (blockIdx.y == 0) {
    if ((int)blockIdx.z >= p.all.nk() || !edge_point(p.all, 1, k, j, i)) return;
    ph_strip_point(D, p, k, j, i);
  } else {
    if ((int)blockIdx.z >= z.all.nk() || !edge_point(z.all, 1, k, j, i)) return;
    zero_grad_point(D, z, k, j, i);
  }
}

extern "C" int pywrf_spec_bdyupdate_ph_zero_grad(const pywrf_grid_t* g, const float* ph_save, float* field,
                                                 const float* field_tend, const float* mu_tend, const float* muts,
                                                 double dt, char variable_in, float* zg_field, char zg_variable,
                                                 pywrf_stream_t s) {
  if (zg_field == field) return PYWRF_E_ARG;
  PhStrip p;
  ZgStrip z;
  if (int e = ph_strip_setup(g, ph_save, field, field_tend, mu_tend, muts, dt, variable_in, p)) return e;
  if (int e = zero_grad_setup(g, zg_field, zg_variable, z)) return e;
  const bool band = p.w == 1 && 2 <= p.all.nj() && 2 <= p.all.ni() && 2 <= z.all.nj() && 2 <= z.all.ni();
  if (!band || p.all.size() <= 0 || z.all.size() <= 0) {          // as the two calls
    pw_launch_edges(k_spec_bdyupdate_ph, p.all, p.w, s, dims_of(g), p);
    pw_launch_edges(k_zero_grad, z.all, 1, s, dims_of(g), z);
    PW_RET();
  }
  const long n = imax((int)edge_count(p.all, 1), (int)edge_count(z.all, 1));
  dim3 grid((unsigned)((n + PW_EDGE_BLOCK - 1) / PW_EDGE_BLOCK), 2, (unsigned)imax(p.all.nk(), z.all.nk()));
  hipLaunchKernelGGL(k_ph_zero_grad, grid, dim3(PW_EDGE_BLOCK), 0, (hipStream_t)s, dims_of(g), p, z);
  PW_RET();
}

// ------------------------------------------------------------------ spec_bdyupdate_ph^T, zero_grad_bdy^T
// On its strips spec_bdyupdate_ph sets f' = f r + dt ft / muts + ph_save (r - 1)
// with r = (muts - dt mu_tend) / muts (applied once per strip holding the
// point).  Its transpose with ph_save, ft and mu_tend fixed: f's sensitivity
// is scaled by r per application, and muts's gathers, down the column, the
// derivative of f' with respect to muts, carried through the applications in
// forward mode, and so do those of mu_tend (through r) and of ft (dt / muts
// per application).  One thread per column of the band; lam holds f''s
// sensitivity on entry and f's on return; lam_muts and lam_mu_tend are
// written on the band's columns only, lam_ft on the strips' points only.
__global__ void k_spec_bdyupdate_ph_adj(Box cols, Dims D, PhStrip a, const float* f_in, float* lam,
                                        float* lam_muts, float* lam_ft, float* lam_mu_tend, float* lam_ph_save) {
  int j, i;
  if (!box_column(cols, j, i)) return;
  bool ir = i >= a.ibs && i < a.ibe, jc = j >= a.jbs + 1 && j < a.jbe - 1;
  int n = (j == a.jts && ir) + (j == a.jbe - 1 && ir) + (i == a.ibs && jc) + (i == a.ibe - 1 && jc);
  if (!n) return;
  const float mt = AT2(a.muts, j, i), mtend = AT2(a.mu_tend, j, i);
  const float mo = mt - a.dt * mtend, r = mo / mt, dr = a.dt * mtend / (mt * mt), drt = -a.dt / mt;
  float gm = 0.0f, gt = 0.0f;
  for (int k = a.all.k0; k < a.all.k1; ++k) {
    const long q = D.ix(k, j, i);
    const float ft = a.ft[q], ps = a.ph_save[q];
    const float dc = -a.dt * ft / (mt * mt) + ps * dr;
    float v = f_in[q], dv = 0.0f, dft = 0.0f, dtn = 0.0f, dps = 0.0f, s = 1.0f;
    for (int m = 0; m < n; ++m) {
      dv = dv * r + v * dr + dc;
      dft = dft * r + a.dt / mt;
      dps = dps * r + (mo / mt - 1.0f);
      dtn = dtn * r + (v + ps) * drt;
      v = v * mo / mt + a.dt * ft / mt + ps * (mo / mt - 1.0f);
      s *= r;
    }
    const float l = lam[q];
    gm += l * dv;
    gt += l * dtn;
    lam_ft[q] = l * dft;
    if (lam_ph_save) lam_ph_save[q] = l * dps;
    lam[q] = l * s;
  }
  AT2(lam_muts, j, i) = gm;
  AT2(lam_mu_tend, j, i) = gt;
}

extern "C" int pywrf_spec_bdyupdate_ph_adjoint(const pywrf_grid_t* g, const float* ph_save, const float* field_in,
                                               const float* field_tend, const float* mu_tend, const float* muts,
                                               double dt, char variable_in, float* lam_field, float* lam_muts,
                                               float* lam_field_tend, float* lam_mu_tend, float* lam_ph_save,
                                               pywrf_stream_t s) {
  PhStrip a;
  if (int e = ph_strip_setup(g, ph_save, nullptr, field_tend, mu_tend, muts, dt, variable_in, a)) return e;
  Box cols = mkbox(0, 1, a.all.j0, a.all.j1, a.all.i0, a.all.i1);
  PW_LAUNCH_COLS(k_spec_bdyupdate_ph_adj, cols, s, dims_of(g), a, field_in, lam_field, lam_muts, lam_field_tend,
                 lam_mu_tend, lam_ph_save);
  PW_RET();
}

// zero_grad_bdy copies, on every ring point of its box, the nearest point of
// the next ring in.  Its transpose, out of place over the whole array: a ring
// point keeps no sensitivity, a point of the next ring in gathers those of
// the ring points that copy it (its outward neighbours in j, in i and on the
// diagonal), every other point passes its own through.
__global__ void k_zero_grad_adj(Box all, Dims D, ZgStrip a, const float* lam, float* out) {
  int k, j, i;
  if (!box_point(all, k, j, i)) return;
  const long q = D.ix(k, j, i);
  float v = lam[q];
  if (a.all.has(k, j, i)) {
    const bool ring = (j == a.jbs || j == a.jbe - 1) || (i == a.ibs || i == a.ibe - 1);
    if (ring) {
      v = 0.0f;
    } else {
      // ring rows / columns whose clamped source row / column is j / i
      int rj[3] = {j, -1, -1}, ri[3] = {i, -1, -1};
      if (j == a.jbs + 1) rj[1] = a.jbs;
      if (j == a.jbe - 2) rj[2] = a.jbe - 1;
      if (i == a.ibs + 1) ri[1] = a.ibs;
      if (i == a.ibe - 2) ri[2] = a.ibe - 1;
      for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y)
          if ((x || y) && rj[x] >= 0 && ri[y] >= 0) v += AT(lam, k, rj[x], ri[y]);
    }
  }
  out[q] = v;
}

extern "C" int pywrf_zero_grad_bdy_adjoint(const pywrf_grid_t* g, const float* lam, char variable_in, float* lam_in,
                                           pywrf_stream_t s) {
  ZgStrip a;
  if (int e = zero_grad_setup(g, nullptr, variable_in, a)) return e;
  // every ring point's source must be an interior point
  if (a.jbe - a.jbs < 3 || a.ibe - a.ibs < 3) return PYWRF_E_RANGE;
  if (lam == lam_in) return PYWRF_E_ARG;
  PW_LAUNCH_BOX(k_zero_grad_adj, whole(g), s, dims_of(g), a, lam, lam_in);
  PW_RET();
}

// ------------------------------------------------------------------ mass_weight
__global__ void k_mass_weight(Box b, Dims D, const float* field, const float* mut, float* rfield) {
  int k, j, i;
  if (!box_point(b, k, j, i)) return;
  long q = D.ix(k, j, i);
  rfield[q] = field[q] * AT2(mut, j, i);
}

extern "C" int pywrf_mass_weight(const pywrf_grid_t* g, const float* field, const float* mut, float* rfield,
                                 pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  Box b = mkbox(g->kts, g->kte, g->jts, g->jte, g->its, g->ite);
  if (b.k0 < 0 || b.k1 > g->nz || b.j0 < 0 || b.j1 > g->ny || b.i0 < 0 || b.i1 > g->nx) return PYWRF_E_RANGE;
  PW_LAUNCH_BOX(k_mass_weight, b, s, dims_of(g), field, mut, rfield);
  PW_RET();
}

// ------------------------------------------------------------------ flow_dep_bdy
__device__ inline void flow_point(const Dims& D, float* f, const float* u, const float* v, int jts, int jtf, int its,
                                  int itf, int k, int j, int i) {
  bool top = j == jts, bot = j == jtf - 1, lft = i == its, rgt = i == itf - 1;
  if (!(top || bot || lft || rgt)) return;
  float r;
  if ((top || bot) && (lft || rgt)) {
    r = AT(f, k, top ? jts + 1 : jtf - 2, lft ? its + 1 : itf - 2) + 0.0f;
  } else if (top) {
    r = AT(v, k, jts, i) < 0.0f ? AT(f, k, jts + 1, i) : 0.0f;
  } else if (bot) {
    r = AT(v, k, jtf, i) > 0.0f ? AT(f, k, jtf - 2, i) : 0.0f;
  } else if (lft) {
    r = AT(u, k, j, its) < 0.0f ? AT(f, k, j, its + 1) : 0.0f;
  } else {
    r = AT(u, k, j, itf) > 0.0f ? AT(f, k, j, itf - 2) : 0.0f;
  }
  AT(f, k, j, i) = r;
}

__global__ void k_flow_dep(Box all, int w, Dims D, float* f, const float* u, const float* v, int jts, int jtf, int its,
                           int itf) {
  int k, j, i;
  if (!edge_point(all, w, k, j, i)) return;
  flow_point(D, f, u, v, jts, jtf, its, itf, k, j, i);
}

extern "C" int pywrf_flow_dep_bdy(const pywrf_grid_t* g, float* field, const float* u, const float* v,
                                  pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  int itf = imin(g->ite, g->ide - 1), jtf = imin(g->jte, g->jde - 1), ktf = g->kde - 1;
  if (ktf > g->nz) return PYWRF_E_RANGE;
  Box all = mkbox(g->kts, ktf, g->jts, jtf, g->its, itf);
  pw_launch_edges(k_flow_dep, all, 1, s, dims_of(g), field, u, v, g->jts, jtf, g->its, itf);
  PW_RET();
}

// flow_dep_bdy on up to 8 distinct fields of one grid with the same u, v, in
// one launch: blockIdx.y selects the field.  The fields are independent (each
// point reads only its own field), so this equals the calls in sequence.
#define PW_FLOW_GROUP_MAX 8
struct FlowFields {
  float* f[PW_FLOW_GROUP_MAX];
};

__global__ void k_flow_dep_group(Box all, Dims D, FlowFields ff, const float* u, const float* v, int jts, int jtf,
                                 int its, int itf) {
  int k, j, i;
  if (!edge_point(all, 1, k, j, i)) return;
  flow_point(D, ff.f[blockIdx.y], u, v, jts, jtf, its, itf, k, j, i);
}

extern "C" int pywrf_flow_dep_bdy_group(const pywrf_grid_t* g, int n, void* const* fields, const float* u,
                                        const float* v, pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  CHECK(g, 1);
  if (n < 0 || n > PW_FLOW_GROUP_MAX) return PYWRF_E_ARG;
  int itf = imin(g->ite, g->ide - 1), jtf = imin(g->jte, g->jde - 1), ktf = g->kde - 1;
  if (ktf > g->nz) return PYWRF_E_RANGE;
  Box all = mkbox(g->kts, ktf, g->jts, jtf, g->its, itf);
  if (n == 0 || all.size() <= 0) return 0;
  FlowFields ff;
  for (int f = 0; f < n; ++f) {
    if (!fields[f]) return PYWRF_E_ARG;
    ff.f[f] = (float*)fields[f];
  }
  if (2 > all.nj() || 2 > all.ni()) {
    for (int f = 0; f < n; ++f)
      pw_launch_edges(k_flow_dep, all, 1, s, dims_of(g), ff.f[f], u, v, g->jts, jtf, g->its, itf);
    PW_RET();
  }
  const long cnt = edge_count(all, 1);
  hipLaunchKernelGGL(k_flow_dep_group, dim3((unsigned)((cnt + PW_EDGE_BLOCK - 1) / PW_EDGE_BLOCK), (unsigned)n,
                     (unsigned)all.nk()), dim3(PW_EDGE_BLOCK), 0, (hipStream_t)s, all, dims_of(g), ff, u, v, g->jts, jtf,
                     g->its, itf);
  PW_RET();
}

// ------------------------------------------------------------------ spec_bdy_final
struct FinalJob {
  Box all;
  int w;
  Dims D;
  float* f;
  const float *mu, *msf;
  Bdy b;
  float dtbc;
  const float* dtbc_dev;  // non-null: the boundary time is read from device memory (graph replays)
  int jts, jtf, its, itf, ibe, jbe;
};

__device__ inline void spec_final_point(const FinalJob& a, int k, int j, int i) {
  const Dims& D = a.D;
  bool xr = j >= a.jts && j < a.jbe, ir = i >= a.its && i < a.ibe;
  float bv, bt;
  if (xr && i == a.itf - 1) { bv = bx(D, a.b.xe, k, j, 0); bt = bx(D, a.b.txe, k, j, 0); }
  else if (xr && i == a.its) { bv = bx(D, a.b.xs, k, j, 0); bt = bx(D, a.b.txs, k, j, 0); }
  else if (ir && j == a.jtf - 1) { bv = by(D, a.b.ye, k, 0, i); bt = by(D, a.b.tye, k, 0, i); }
  else if (ir && j == a.jts) { bv = by(D, a.b.ys, k, 0, i); bt = by(D, a.b.tys, k, 0, i); }
  else return;
  float sm = a.msf ? AT2(a.msf, j, i) : 1.0f;
  float su = a.mu ? AT2(a.mu, j, i) : 1.0f;
  const float dtbc = a.dtbc_dev ? *a.dtbc_dev : a.dtbc;
  AT(a.f, k, j, i) = sm * (bv + dtbc * bt) / su;
}

__global__ void k_spec_final(Box all, int w, FinalJob a) {
  int k, j, i;
  if (!edge_point(all, w, k, j, i)) return;
  spec_final_point(a, k, j, i);
}

static int spec_final_job(const pywrf_grid_t* g, float* field, const float* mu, const float* msf, const float* bxs,
                          const float* bxe, const float* bys, const float* bye, const float* btxs,
                          const float* btxe, const float* btys, const float* btye, char variable_in, double dtbc,
                          FinalJob& a) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  char v = variable_in;
  switch (v) {
    case 'U': v = 'u'; break;
    case 'V': v = 'v'; break;
    case 'W': v = 'w'; break;
    case 'M': v = 'm'; break;
    case 'T': v = 't'; break;
    case 'H': v = 'h'; break;
    default: break;
  }
  int ibe = g->ide - 1, jbe = g->jde - 1;
  int itf = imin(g->ite, g->ide - 1), jtf = imin(g->jte, g->jde - 1), ktf = g->kde - 1;
  if (v == 'u') { ibe = g->ide; itf = imin(g->ite, g->ide); }
  if (v == 'v') { jbe = g->jde; jtf = imin(g->jte, g->jde); }
  if (v == 'm' || v == 'h' || v == 'w') ktf = g->kde;
  if (ibe > g->nx || jbe > g->ny || jtf > g->ny || ktf > g->nz || g->its < 0 || g->jts < 0) return PYWRF_E_RANGE;
  a.msf = (v == 'u' || v == 'v' || v == 'w') ? msf : nullptr;
  a.mu = (v != 'm') ? mu : nullptr;
  a.b = {bxs, bxe, bys, bye, btxs, btxe, btys, btye};
  a.all = mkbox(g->kts, ktf, g->jts, imax(jtf, jbe), g->its, imax(ibe, itf));
  a.w = (itf >= ibe && jtf >= jbe) ? 1 : 0;  // rows jts, jtf-1 and columns its, itf-1 are the box's sides
  a.D = dims_of(g);
  a.f = field; a.dtbc = F(dtbc); a.dtbc_dev = nullptr;
  a.jts = g->jts; a.jtf = jtf; a.its = g->its; a.itf = itf; a.ibe = ibe; a.jbe = jbe;
  return 0;
}

extern "C" int pywrf_spec_bdy_final(const pywrf_grid_t* g, float* field, const float* mu, const float* msf,
                                    const float* bxs, const float* bxe, const float* bys, const float* bye,
                                    const float* btxs, const float* btxe, const float* btys, const float* btye,
                                    char variable_in, double dtbc, pywrf_stream_t s) {
  FinalJob a;
  int e = spec_final_job(g, field, mu, msf, bxs, bxe, bys, bye, btxs, btxe, btys, btye, variable_in, dtbc, a);
  if (e) return e;
  pw_launch_edges(k_spec_final, a.all, a.w, s, a);
  PW_RET();
}

// Several spec_bdy_final calls on distinct fields in one launch (blockIdx.y
// selects the job); a job whose rows are not its box's sides runs on its own.
#define PW_FINAL_GROUP_MAX 8
struct FinalJobs {
  FinalJob j[PW_FINAL_GROUP_MAX];
};

__global__ void k_spec_final_group(FinalJobs J) {
  const FinalJob& a = J.j[blockIdx.y];
  if ((int)blockIdx.z >= a.all.nk()) return;
  int k, j, i;
  if (!edge_point(a.all, 1, k, j, i)) return;
  spec_final_point(a, k, j, i);
}

static int spec_final_group(const pywrf_grid_t* g, int n, void* const* fields, void* const* mus, void* const* msfs,
                            void* const* bdys, const char* variables, double dtbc, const float* dtbc_dev,
                            pywrf_stream_t s) {
  if (!pw_size_ok(g)) return PYWRF_E_RANGE;
  if (n < 0 || n > PW_FINAL_GROUP_MAX) return PYWRF_E_ARG;
  FinalJobs J;
  int m = 0, maxk = 0;
  long most = 0;
  for (int f = 0; f < n; ++f) {
    FinalJob a;
    const float* const* bd = (const float* const*)bdys + 8 * f;
    int e = spec_final_job(g, (float*)fields[f], (const float*)mus[f], (const float*)msfs[f], bd[0], bd[1], bd[2],
                           bd[3], bd[4], bd[5], bd[6], bd[7], variables[f], dtbc, a);
    if (e) return e;
    a.dtbc_dev = dtbc_dev;
    if (a.all.size() <= 0) continue;
    if (a.w != 1 || a.all.nj() < 2 || a.all.ni() < 2) {
      pw_launch_edges(k_spec_final, a.all, a.w, s, a);
      continue;
    }
    J.j[m++] = a;
    most = edge_count(a.all, 1) > most ? edge_count(a.all, 1) : most;
    maxk = a.all.nk() > maxk ? a.all.nk() : maxk;
  }
  if (m > 0)
    hipLaunchKernelGGL(k_spec_final_group, dim3((unsigned)((most + PW_EDGE_BLOCK - 1) / PW_EDGE_BLOCK), (unsigned)m,
                       (unsigned)maxk), dim3(PW_EDGE_BLOCK), 0, (hipStream_t)s, J);
  PW_RET();
}

extern "C" int pywrf_spec_bdy_final_group(const pywrf_grid_t* g, int n, void* const* fields, void* const* mus,
                                          void* const* msfs, void* const* bdys, const char* variables, double dtbc,
                                          pywrf_stream_t s) {
  return spec_final_group(g, n, fields, mus, msfs, bdys, variables, dtbc, nullptr, s);
}

extern "C" int pywrf_spec_bdy_final_group_dtbc_dev(const pywrf_grid_t* g, int n, void* const* fields,
                                                   void* const* mus, void* const* msfs, void* const* bdys,
                                                   const char* variables, const float* dtbc, pywrf_stream_t s) {
  if (!dtbc) return PYWRF_E_RANGE;
  return spec_final_group(g, n, fields, mus, msfs, bdys, variables, 0.0, dtbc, s);
}

// ------------------------------------------------------------------ end-of-step boundary epilogue
// The end of a step (solver.py:1776-1910) fills the open lateral halos of a
// group of fields (set_physical_bc3d / bc2d) and then sets the specified
// boundary rows of five of them (spec_bdy_final).  Both act plane by plane on
// their own field, and spec_bdy_final reads only the boundary arrays, mu and
// the map factor, so one workgroup per (field, level) does the plane's halo
// fill, waits for its own writes, and then sets the plane's boundary rows:
// the order of the calls, in one launch.
#define PW_EPI_MAX 24
#define PW_EPI_FINAL_MAX 6
struct EpiJobs {
  int n;
  BcJob bc[PW_EPI_MAX];
  int fin[PW_EPI_MAX];           // index into F of the field's spec_bdy_final, or -1
  FinalJob F[PW_EPI_FINAL_MAX];
};

__global__ void __launch_bounds__(BC_THREADS) k_bdy_epilogue(long plane, int nx, EpiJobs J) {
  const int f = (int)blockIdx.y;
  if (f >= J.n) return;
  const BcJob& jb = J.bc[f];
  const int lev = (int)blockIdx.x;
  if (lev < jb.nk) open_bc_plane(jb.p + (long)(jb.k0 + lev) * plane, nx, jb.c);
  if (J.fin[f] < 0) return;                       // uniform across the workgroup
  __syncthreads();                                // the halo fill's stores before the rows it read change
  const FinalJob& a = J.F[J.fin[f]];
  const int k = a.all.k0 + lev;
  if (k >= a.all.k1) return;
  const int ni = a.all.ni(), nj = a.all.nj();
  const long per_level = edge_count(a.all, 1);
  for (long t = threadIdx.x; t < per_level; t += blockDim.x) {
    int j, i;
    const int rows = ni;                          // rows j0 and j1 - 1, then columns i0 and i1 - 1
    if (t < rows) { j = a.all.j0; i = a.all.i0 + (int)t; }
    else if (t < 2 * rows) { j = a.all.j1 - 1; i = a.all.i0 + (int)(t - rows); }
    else if (t < 2 * rows + (nj - 2)) { j = a.all.j0 + 1 + (int)(t - 2 * rows); i = a.all.i0; }
    else { j = a.all.j0 + 1 + (int)(t - 2 * rows - (nj - 2)); i = a.all.i1 - 1; }
    spec_final_point(a, k, j, i);
  }
}

static int bdy_epilogue(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                        const char* variables, const int* is2d, int nfinal, const int* final_of, void* const* mus,
                        void* const* msfs, void* const* bdys, const char* final_variables, double dtbc,
                        const float* dtbc_dev, pywrf_stream_t s) {
  if (!pw_size_ok(g3) || !pw_size_ok(g2)) return PYWRF_E_RANGE;
  if (n < 0 || n > PW_EPI_MAX || nfinal < 0 || nfinal > PW_EPI_FINAL_MAX || g3->ny != g2->ny || g3->nx != g2->nx)
    return PYWRF_E_ARG;
  EpiJobs J;
  J.n = n;
  int maxk = 1;
  for (int f = 0; f < n; ++f) {
    BcJob& jb = J.bc[f];
    jb.p = (float*)fields[f];
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
    jb.nk = k1 - jb.k0 > 0 ? k1 - jb.k0 : 0;
    maxk = imax(maxk, jb.nk);
    J.fin[f] = -1;
  }
  for (int m = 0; m < nfinal; ++m) {
    const int f = final_of[m];
    if (f < 0 || f >= n || is2d[f] || J.fin[f] >= 0) return PYWRF_E_ARG;
    const float* const* bd = (const float* const*)bdys + 8 * m;
    FinalJob& a = J.F[m];
    int e = spec_final_job(g3, (float*)fields[f], (const float*)mus[m], (const float*)msfs[m], bd[0], bd[1], bd[2],
                           bd[3], bd[4], bd[5], bd[6], bd[7], final_variables[m], dtbc, a);
    if (e) return e;
    a.dtbc_dev = dtbc_dev;
    // the rows the workgroup sets must be its box's sides (as in the grouped launch)
    if (a.all.size() > 0 && (a.w != 1 || a.all.nj() < 2 || a.all.ni() < 2)) return PYWRF_E_RANGE;
    J.fin[f] = m;
    maxk = imax(maxk, a.all.nk());
  }
  if (n == 0) return 0;
  static_assert(sizeof(EpiJobs) <= 4096, "kernel arguments of the epilogue exceed 4 KiB");
  hipLaunchKernelGGL(k_bdy_epilogue, dim3(maxk, n), dim3(BC_THREADS), 0, (hipStream_t)s, (long)g3->ny * g3->nx,
                     g3->nx, J);
  PW_RET();
}

extern "C" int pywrf_boundary_epilogue(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                                       const char* variables, const int* is2d, int nfinal, const int* final_of,
                                       void* const* mus, void* const* msfs, void* const* bdys,
                                       const char* final_variables, double dtbc, pywrf_stream_t s) {
  return bdy_epilogue(g3, g2, n, fields, variables, is2d, nfinal, final_of, mus, msfs, bdys, final_variables, dtbc,
                      nullptr, s);
}

extern "C" int pywrf_boundary_epilogue_dtbc_dev(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n,
                                                void* const* fields, const char* variables, const int* is2d,
                                                int nfinal, const int* final_of, void* const* mus, void* const* msfs,
                                                void* const* bdys, const char* final_variables, const float* dtbc,
                                                pywrf_stream_t s) {
  if (!dtbc) return PYWRF_E_RANGE;
  return bdy_epilogue(g3, g2, n, fields, variables, is2d, nfinal, final_of, mus, msfs, bdys, final_variables, 0.0,
                      dtbc, s);
}
