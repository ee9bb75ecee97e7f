// Shared device helpers for the PyWRF hot-path kernels (gfx950).
//
// Layout: every 3-D field is a contiguous (nz, ny, nx) FP32 array, i fastest,
// with the reference's 5-point halo; 2-D fields are (ny, nx); level arrays
// (nz).  Index sets follow WRF (ids..kte); boxes are half-open [lo, hi).
//
// All kernels are built with -ffp-contract=off so that every multiply and add
// rounds exactly as the reference's element-wise tensor ops do; scalar
// prefactors that the reference forms in Python double precision are formed in
// double on the host (or as constexpr double) and rounded once to float.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/pywrf_amd.h"

#define PW_BLOCK 256

struct Box {
  int k0, k1, j0, j1, i0, i1;
  __host__ __device__ int nk() const { return k1 > k0 ? k1 - k0 : 0; }
  __host__ __device__ int nj() const { return j1 > j0 ? j1 - j0 : 0; }
  __host__ __device__ int ni() const { return i1 > i0 ? i1 - i0 : 0; }
  __host__ __device__ long size() const { return (long)nk() * nj() * ni(); }
  __host__ __device__ bool has(int k, int j, int i) const {
    return k >= k0 && k < k1 && j >= j0 && j < j1 && i >= i0 && i < i1;
  }
  __host__ __device__ bool has2(int j, int i) const { return j >= j0 && j < j1 && i >= i0 && i < i1; }
};

__host__ __device__ inline Box mkbox(int k0, int k1, int j0, int j1, int i0, int i1) {
  Box b; b.k0 = k0; b.k1 = k1; b.j0 = j0; b.j1 = j1; b.i0 = i0; b.i1 = i1; return b;
}

// Launch helper: one thread per point of a box, i fastest.
#define PW_LAUNCH_BOX(kern, box, stream, ...)                                              \
  do {                                                                                     \
    long _n = (box).size();                                                                \
    if (_n > 0) {                                                                          \
      unsigned _g = (unsigned)((_n + PW_BLOCK - 1) / PW_BLOCK);                            \
      hipLaunchKernelGGL(kern, dim3(_g), dim3(PW_BLOCK), 0, (hipStream_t)(stream), (box), __VA_ARGS__); \
    }                                                                                      \
  } while (0)

// Decode the flat thread id of a PW_LAUNCH_BOX launch; returns false past the end.
__device__ inline bool box_point(const Box& b, int& k, int& j, int& i) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.size()) return false;
  int ni = b.ni(), nj = b.nj();
  i = b.i0 + (int)(t % ni);
  long r = t / ni;
  j = b.j0 + (int)(r % nj);
  k = b.k0 + (int)(r / nj);
  return true;
}

// Launch helper: one thread per (j, i) column of a box (column kernels).
#define PW_LAUNCH_COLS(kern, box, stream, ...)                                             \
  do {                                                                                     \
    long _n = (long)(box).nj() * (box).ni();                                               \
    if (_n > 0) {                                                                          \
      unsigned _g = (unsigned)((_n + PW_BLOCK - 1) / PW_BLOCK);                            \
      hipLaunchKernelGGL(kern, dim3(_g), dim3(PW_BLOCK), 0, (hipStream_t)(stream), (box), __VA_ARGS__); \
    }                                                                                      \
  } while (0)

__device__ inline bool box_column(const Box& b, int& j, int& i) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)b.nj() * b.ni()) return false;
  i = b.i0 + (int)(t % b.ni());
  j = b.j0 + (int)(t / b.ni());
  return true;
}

// Fortran SIGN(1., x) (the reference's fortran_sign): +1 for x >= 0, else -1.
__device__ inline float sgn1(float x) { return x >= 0.0f ? 1.0f : -1.0f; }

// Wicker-Skamarock upwind-biased face values (wrf_dynamics.py:1313-1329 with
// time_step > 0).  Argument order matches the reference's flux5_u / flux3_u.
__device__ inline float flux5(float qm3, float qm2, float qm1, float q0, float qp1, float qp2, float ua) {
  float f6 = (37.0f * (q0 + qm1) - 8.0f * (qp1 + qm2) + (qp2 + qm3)) / 60.0f;
  float d = (qp2 - qm3) - 5.0f * (qp1 - qm2) + 10.0f * (q0 - qm1);
  return f6 - (sgn1(ua) * d) / 60.0f;
}
__device__ inline float flux3(float qm2, float qm1, float q0, float qp1, float ua) {
  float f4 = (7.0f * (q0 + qm1) - (qp1 + qm2)) / 12.0f;
  float d = (qp1 - qm2) - 3.0f * (q0 - qm1);
  return f4 + (sgn1(ua) * d) / 12.0f;
}

struct Dims {
  int nz, ny, nx;
  __host__ __device__ long ix(int k, int j, int i) const { return ((long)k * ny + j) * nx + i; }
  __host__ __device__ long ix2(int j, int i) const { return (long)j * nx + i; }
  __host__ __device__ long plane() const { return (long)ny * nx; }
};

__host__ inline Dims dims_of(const pywrf_grid_t* g) { Dims d; d.nz = g->nz; d.ny = g->ny; d.nx = g->nx; return d; }

#define PW_RET() return (int)hipGetLastError()

__host__ __device__ static inline int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ static inline int imax(int a, int b) { return a > b ? a : b; }

// Sum of term(0..n-1) in the order torch's CPU reduction uses along a
// non-innermost dimension (the four-level cascade of its outer-sum kernel:
// blocks of 2^p terms summed in sequence, p = max(4, ceil(log2 n) / 4),
// partial sums carried up a level per block, the tail added last).  For n = 41
// this is (s[0..15] + s[16..31]) added to the sequential tail s[32..40].  The
// reference's column sums (e.g. divergence integrals) are computed by that
// kernel, and reproducing its order keeps ill-conditioned sums identical.
template <class Fn>
__device__ inline float torch_outer_sum(int n, Fn term) {
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  const int lp = (lg / 4) > 4 ? (lg / 4) : 4;
  const int step = 1 << lp, mask = step - 1;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int i = 0;
  while (i + step <= n) {
    for (int j = 0; j < step; ++j, ++i) acc[0] += term(i);
    for (int j = 1; j < 4; ++j) {
      acc[j] += acc[j - 1];
      acc[j - 1] = 0.0f;
      if ((i & (mask << (j * lp))) != 0) break;
    }
  }
  for (; i < n; ++i) acc[0] += term(i);
  for (int j = 1; j < 4; ++j) acc[0] += acc[j];
  return acc[0];
}
