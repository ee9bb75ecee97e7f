// Open lateral-boundary fill of one plane (set_physical_bc3d / bc2d,
// wrf_dynamics.py:164-312), shared by the single, grouped and end-of-step
// launches, and the host-side description of a field's fill.
#pragma once
#include "common.h"

// One workgroup per level: the x-direction copies, a barrier, then the
// y-direction copies (which read cells the x-copies may have written).
struct BcSpec {
  int xr0, xr1;      // rows of the x-start copies and of the non-u x-end copies
  int ur0, ur1;      // rows of the u-staggered x-end copies
  int ustag;         // x-end copies ide..ide+2 from ide-1 (else ide-1..ide+1 from ide-2)
  int yc0, yc1;      // columns of the y copies
  int vstag;         // y-end copies jde..jde+2 from jde-1 (else jde-1..jde+1 from jde-2)
  int ids, ide, jds, jde;
};

// One plane's open-boundary fill by one workgroup of BC_THREADS threads.  The
// reference copies the x-halos and then the y-halos (whose sources, the edge
// rows, include cells the x-copies wrote).  Here every value is first loaded
// -- a y-copy reads the cell the x-copies take its value from -- and, after
// one barrier, stored; an x-copy whose target a y-copy overwrites later is not
// stored.  The result equals the two passes in order, with one round of loads.
// Planes with more copies than BC_THREADS * BC_ITEMS take the two passes.
#define BC_THREADS 512
#define BC_ITEMS 8
__device__ inline void open_bc_plane(float* a, int nx, const BcSpec& c) {
  const int r0 = c.ustag ? c.ur0 : c.xr0, r1 = c.ustag ? c.ur1 : c.xr1;
  const int nstart = c.xr1 - c.xr0, nend = r1 - r0;
  const int nxi = 3 * (nstart + nend), nyi = 6 * (c.yc1 - c.yc0);
  const int yl0 = c.jds - 3, yh0 = c.vstag ? c.jde : c.jde - 1;   // y-copy target rows [yl0, jds), [yh0, yh0 + 3)
  auto xitem = [&](int t, int& src, int& dst) {
    const int h = t % 3, r = t / 3;
    if (r < nstart) {
      const int row = c.xr0 + r;
      src = row * nx + c.ids;
      dst = row * nx + c.ids - 1 - h;
    } else {
      const int row = r0 + (r - nstart);
      src = c.ustag ? row * nx + c.ide - 1 : row * nx + c.ide - 2;
      dst = c.ustag ? row * nx + c.ide + h : row * nx + c.ide - 1 + h;
    }
  };
  auto yitem = [&](int t, int& srow, int& col, int& dst) {
    const int h = t % 3, side = (t / 3) & 1;
    col = c.yc0 + t / 6;
    if (side == 0) { srow = c.jds; dst = (c.jds - 1 - h) * nx + col; }
    else if (c.vstag) { srow = c.jde - 1; dst = (c.jde + h) * nx + col; }
    else { srow = c.jde - 2; dst = (c.jde - 1 + h) * nx + col; }
  };
  const int chunk = BC_ITEMS * (int)blockDim.x;
  if (nxi + nyi <= chunk) {
    float v[BC_ITEMS];
    int dst[BC_ITEMS];
#pragma unroll
    for (int u = 0; u < BC_ITEMS; ++u) {
      const int t = u * (int)blockDim.x + (int)threadIdx.x;
      dst[u] = -1;
      if (t < nxi) {
        int src, d;
        xitem(t, src, d);
        const int row = d / nx, col = d - row * nx;
        const bool ylater = ((row >= yl0 && row < c.jds) || (row >= yh0 && row < yh0 + 3)) && col >= c.yc0 &&
                            col < c.yc1;
        if (!ylater) {
          v[u] = a[src];
          dst[u] = d;
        }
      } else if (t < nxi + nyi) {
        int srow, col, d;
        yitem(t - nxi, srow, col, d);
        // the value of (srow, col) once the x-copies are done
        int src = srow * nx + col;
        if (srow >= c.xr0 && srow < c.xr1 && col >= c.ids - 3 && col <= c.ids - 1) src = srow * nx + c.ids;
        else if (srow >= r0 && srow < r1 && c.ustag && col >= c.ide && col <= c.ide + 2) src = srow * nx + c.ide - 1;
        else if (srow >= r0 && srow < r1 && !c.ustag && col >= c.ide - 1 && col <= c.ide + 1)
          src = srow * nx + c.ide - 2;
        v[u] = a[src];
        dst[u] = d;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BC_ITEMS; ++u)
      if (dst[u] >= 0) a[dst[u]] = v[u];
    return;
  }
  // two passes: x-copies, then y-copies
  for (int base = 0; base < nxi; base += chunk) {
    float v[BC_ITEMS];
    int dst[BC_ITEMS];
#pragma unroll
    for (int u = 0; u < BC_ITEMS; ++u) {
      const int t = base + u * (int)blockDim.x + (int)threadIdx.x;
      dst[u] = -1;
      if (t < nxi) {
        int src;
        xitem(t, src, dst[u]);
        v[u] = a[src];
      }
    }
#pragma unroll
    for (int u = 0; u < BC_ITEMS; ++u)
      if (dst[u] >= 0) a[dst[u]] = v[u];
  }
  __syncthreads();
  for (int base = 0; base < nyi; base += chunk) {
    float v[BC_ITEMS];
    int dst[BC_ITEMS];
#pragma unroll
    for (int u = 0; u < BC_ITEMS; ++u) {
      const int t = base + u * (int)blockDim.x + (int)threadIdx.x;
      dst[u] = -1;
      if (t < nyi) {
        int srow, col;
        yitem(t, srow, col, dst[u]);
        v[u] = a[srow * nx + col];
      }
    }
#pragma unroll
    for (int u = 0; u < BC_ITEMS; ++u)
      if (dst[u] >= 0) a[dst[u]] = v[u];
  }
}

// One field of a grouped fill: its base pointer, first level and level count.
struct BcJob {
  float* p;
  int k0, nk;
  BcSpec c;
};

static inline char lower(char v) { return (v >= 'A' && v <= 'Z') ? (char)(v - 'A' + 'a') : v; }
static inline bool in(char v, const char* set) { for (; *set; ++set) if (*set == v) return true; return false; }

static inline int bc_spec3(const pywrf_grid_t* g, char variable, BcSpec& c, int& k1) {
  char v = variable;
  if (v == 'U') v = 'u';
  else if (v == 'V') v = 'v';
  if (!in(v, "uvwtdexyfrp")) return PYWRF_E_ARG;
  int istag = in(v, "uxde") ? 0 : -1, jstag = in(v, "vydf") ? 0 : -1;
  k1 = in(v, "efw") ? imin(g->kde, g->kte) : imin(g->kde - 1, g->kte);
  c.ids = g->ids; c.ide = g->ide; c.jds = g->jds; c.jde = g->jde;
  c.xr0 = g->jts - 4; c.xr1 = g->jde + jstag + 4;
  c.ur0 = g->jts - 5; c.ur1 = g->jde + jstag + 4;
  c.ustag = in(v, "ux"); c.vstag = in(v, "vy");
  c.yc0 = g->ids; c.yc1 = g->ide + istag;
  return 0;
}

static inline int bc_spec2(const pywrf_grid_t* g, char variable, BcSpec& c) {
  char v = lower(variable);
  if (!in(v, "uvwtxyrp")) return PYWRF_E_ARG;
  int istag = in(v, "uxde") ? 0 : -1, jstag = in(v, "vydf") ? 0 : -1;
  c.ids = g->ids; c.ide = g->ide; c.jds = g->jds; c.jde = g->jde;
  c.xr0 = c.ur0 = g->jds; c.xr1 = c.ur1 = g->jde + jstag;
  c.ustag = in(v, "ux"); c.vstag = in(v, "vy");
  c.yc0 = imax(g->ids, g->its - 1); c.yc1 = imin(g->ite + 1, g->ide + istag);
  return 0;
}

static inline int bc_spec_ok(int ny, int nx, int nz, int k0, int k1, const BcSpec& c) {
  int rlo = imin(c.xr0, c.ustag ? c.ur0 : c.xr0), rhi = imax(c.xr1, c.ustag ? c.ur1 : c.xr1);
  if (rlo < 0 || rhi > ny || c.ids - 3 < 0 || c.ide + 3 > nx) return 0;
  if (c.yc0 < 0 || c.yc1 > nx || c.jds - 3 < 0 || c.jde + 3 > ny) return 0;
  return k0 >= 0 && k1 <= nz;
}

