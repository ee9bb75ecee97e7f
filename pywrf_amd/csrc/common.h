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
// MI355X has 8 XCDs, each with its own L2; consecutive workgroups of a launch
// are dealt to them round-robin.
#define PW_XCDS 8

// Block (bx, by) of a 2-D grid, renumbered in XCD order: the hardware hands
// consecutive workgroups to the PW_XCDS XCDs in turn, so neighbouring rows of
// a row-per-block grid would sit in different L2 caches.  Renumbering gives
// each XCD a contiguous range of the row-major block order (XCD x < rem takes
// q + 1 blocks, the others q, for n = q PW_XCDS + rem blocks): a bijection of
// the blocks, so every block is still run exactly once.
__device__ inline void xcd_block_2d(int& bx, int& by) {
  const int gx = (int)gridDim.x, n = gx * (int)gridDim.y;
  const int lin = (int)blockIdx.y * gx + (int)blockIdx.x;
  const int q = n / PW_XCDS, rem = n % PW_XCDS, x = lin % PW_XCDS, r = lin / PW_XCDS;
  const int ord = x < rem ? x * (q + 1) + r : rem * (q + 1) + (x - rem) * q + r;
  bx = ord % gx;
  by = ord / gx;
}

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

// Launch helpers.  A box launch is a 3-D grid: x covers i in 64-lane rows (one
// wavefront per row segment, so every load and store of a wave is one
// contiguous run of a row), y covers j four rows per workgroup, z covers k.
// Decoding needs no integer division.
#define PW_TX 64
#define PW_TY 4
#define PW_LAUNCH_BOX(kern, box, stream, ...)                                              \
  do {                                                                                     \
    if ((box).size() > 0) {                                                                \
      dim3 _g((unsigned)(((box).ni() + PW_TX - 1) / PW_TX), (unsigned)(((box).nj() + PW_TY - 1) / PW_TY), \
              (unsigned)(box).nk());                                                       \
      hipLaunchKernelGGL(kern, _g, dim3(PW_TX, PW_TY), 0, (hipStream_t)(stream), (box), __VA_ARGS__); \
    }                                                                                      \
  } while (0)

// Point of a PW_LAUNCH_BOX launch; false outside the box.
__device__ inline bool box_point(const Box& b, int& k, int& j, int& i) {
  i = b.i0 + (int)(blockIdx.x * PW_TX + threadIdx.x);
  j = b.j0 + (int)(blockIdx.y * PW_TY + threadIdx.y);
  k = b.k0 + (int)blockIdx.z;
  return i < b.i1 && j < b.j1;
}

// Launch helper: one thread per (j, i) column of a box (column kernels).
#define PW_LAUNCH_COLS(kern, box, stream, ...)                                             \
  do {                                                                                     \
    if ((box).nj() > 0 && (box).ni() > 0) {                                                \
      dim3 _g((unsigned)(((box).ni() + PW_TX - 1) / PW_TX), (unsigned)(((box).nj() + PW_TY - 1) / PW_TY)); \
      hipLaunchKernelGGL(kern, _g, dim3(PW_TX, PW_TY), 0, (hipStream_t)(stream), (box), __VA_ARGS__); \
    }                                                                                      \
  } while (0)

__device__ inline bool box_column(const Box& b, int& j, int& i) {
  i = b.i0 + (int)(blockIdx.x * PW_TX + threadIdx.x);
  j = b.j0 + (int)(blockIdx.y * PW_TY + threadIdx.y);
  return i < b.i1 && j < b.j1;
}

// Edge launches, for kernels that act only on a band of width w along the
// four sides of a box (boundary rows and columns): one thread per band point,
// each point visited exactly once (corners belong to the j-rows).  Per level
// the band holds 2*w*ni points of the first and last w rows, then 2*w*(nj-2w)
// points of the first and last w columns.  With w = 0 the launch is an
// ordinary box launch (used when the band would not fit in the box, or when
// the acting set of the kernel is not known to lie in the band).
#define PW_EDGE_BLOCK 256
__host__ __device__ inline long edge_count(const Box& b, int w) {
  return 2L * w * b.ni() + 2L * w * (b.nj() - 2 * w);
}

__device__ inline bool edge_point(const Box& b, int w, int& k, int& j, int& i) {
  if (w == 0) return box_point(b, k, j, i);
  int t = (int)(blockIdx.x * PW_EDGE_BLOCK + threadIdx.x);
  k = b.k0 + (int)blockIdx.z;
  const int ni = b.ni(), nj = b.nj(), rows = w * ni, h = nj - 2 * w;
  if (t < rows) { j = b.j0 + t / ni; i = b.i0 + t % ni; return true; }
  t -= rows;
  if (t < rows) { j = b.j1 - w + t / ni; i = b.i0 + t % ni; return true; }
  t -= rows;
  if (t < w * h) { j = b.j0 + w + t / w; i = b.i0 + t % w; return true; }
  t -= w * h;
  if (t < w * h) { j = b.j0 + w + t / w; i = b.i1 - w + t % w; return true; }
  return false;
}

template <class Kern, class... Args>
inline void pw_launch_edges(Kern kern, const Box& b, int w, pywrf_stream_t s, Args... args) {
  if (b.size() <= 0) return;
  if (w > 0 && 2 * w <= b.nj() && 2 * w <= b.ni()) {
    long n = edge_count(b, w);
    dim3 g((unsigned)((n + PW_EDGE_BLOCK - 1) / PW_EDGE_BLOCK), 1, (unsigned)b.nk());
    hipLaunchKernelGGL(kern, g, dim3(PW_EDGE_BLOCK), 0, (hipStream_t)s, b, w, args...);
  } else {
    dim3 g((unsigned)((b.ni() + PW_TX - 1) / PW_TX), (unsigned)((b.nj() + PW_TY - 1) / PW_TY), (unsigned)b.nk());
    hipLaunchKernelGGL(kern, g, dim3(PW_TX, PW_TY), 0, (hipStream_t)s, b, 0, args...);
  }
}

// Fortran SIGN(1., x) (the reference's fortran_sign): +1 for x >= 0, else -1.
__device__ inline float sgn1(float x) { return x >= 0.0f ? 1.0f : -1.0f; }

// Wicker-Skamarock upwind-biased face values (wrf_dynamics.py:1313-1329 with
// time_step > 0).  Argument order matches the reference's flux5_u / flux3_u.
//
// The divisions by 60 and 12 give the correctly rounded quotient x / d of an
// IEEE division, computed as a product by r = RN(1/d) and one fma correction
// (Markstein's correction step): q = RN(x r), e = x - q d (exact in an fma),
// RN(q + e r).  An exhaustive run over every float (the float_div_const check
// of the test suite) shows that for d = 60 and d = 12 this equals RN(x / d)
// bit for bit whenever |x| is finite and at least 2^-124, and for +-0 once the
// sign of x is copied onto it.  Inputs outside that range (subnormal x,
// infinities) and results that are subnormal or NaN take the true division,
// so the function equals x / d for every input.  It costs 4 VALU operations
// where the division sequence costs 11, one of them a quarter-rate rcp.
template <int D>
__device__ inline float div_const(float x) {
  constexpr float d = (float)D;
  constexpr float r = 1.0f / d;
  float q = x * r;
  float e = __builtin_fmaf(-q, d, x);
  float y = __builtin_copysignf(__builtin_fmaf(e, r, q), x);
  // fpclass bits: 1|2 NaN, 4|512 infinity, 16|128 subnormal
  if (__builtin_expect(__builtin_isfpclass(x, 4 | 16 | 128 | 512) | __builtin_isfpclass(y, 1 | 2 | 16 | 128), 0))
    y = x / d;
  return y;
}
__device__ inline float flux5(float qm3, float qm2, float qm1, float q0, float qp1, float qp2, float ua) {
  float f6 = div_const<60>(37.0f * (q0 + qm1) - 8.0f * (qp1 + qm2) + (qp2 + qm3));
  float d = (qp2 - qm3) - 5.0f * (qp1 - qm2) + 10.0f * (q0 - qm1);
  return f6 - div_const<60>(sgn1(ua) * d);
}
__device__ inline float flux3(float qm2, float qm1, float q0, float qp1, float ua) {
  float f4 = div_const<12>(7.0f * (q0 + qm1) - (qp1 + qm2));
  float d = (qp1 - qm2) - 3.0f * (q0 - qm1);
  return f4 + div_const<12>(sgn1(ua) * d);
}

// Offsets are 32-bit: every entry point rejects arrays of 2^30 or more points
// (pw_size_ok), so an element's byte offset fits in 32 bits.  AT / AT2 index
// through that unsigned byte offset added to the (wave-uniform) array base, the
// form global loads and stores take as a scalar base plus one 32-bit vector
// offset, with constant displacements folded into the instruction; unsigned
// offset arithmetic is exact modulo 2^32, so the compiler may reassociate it.
struct Dims {
  int nz, ny, nx;
  __host__ __device__ int ix(int k, int j, int i) const { return (k * ny + j) * nx + i; }
  __host__ __device__ int ix2(int j, int i) const { return j * nx + i; }
  __host__ __device__ int plane() const { return ny * nx; }
  // byte offsets of element (k, j, i) / (j, i)
  __device__ unsigned ob(int k, int j, int i) const {
    return (((unsigned)k * (unsigned)ny + (unsigned)j) * (unsigned)nx + (unsigned)i) * 4u;
  }
  __device__ unsigned ob2(int j, int i) const { return ((unsigned)j * (unsigned)nx + (unsigned)i) * 4u; }
};

template <class T>
__device__ __forceinline__ T& at_b(T* p, unsigned off) {
  return *(T*)((char*)p + off);
}
template <class T>
__device__ __forceinline__ const T& at_b(const T* p, unsigned off) {
  return *(const T*)((const char*)p + off);
}

// The whole array of a grid, as a box.
static inline Box whole(const pywrf_grid_t* g) { return mkbox(0, g->nz, 0, g->ny, 0, g->nx); }

__host__ inline bool pw_size_ok(const pywrf_grid_t* g) {
  return g && g->nz > 0 && g->ny > 0 && g->nx > 0 && (long)g->nz * g->ny * g->nx < (1L << 30);
}

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

// torch_outer_sum for a sum whose terms are +0 outside [lo, hi]: only the
// terms inside are evaluated, in their block positions.  Skipping +0 terms
// leaves every partial sum unchanged (the accumulators start at +0 and only
// non-negative or NaN terms are added), so the result is the same float.
template <class Fn>
__device__ inline float torch_outer_sum_range(int n, int lo, int hi, Fn term) {
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  const int lp = (lg / 4) > 4 ? (lg / 4) : 4;
  const int step = 1 << lp, mask = step - 1;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int i = 0;
  while (i + step <= n) {
    const int e = i + step;
    for (int m = imax(i, lo), me = imin(e, hi + 1); m < me; ++m) acc[0] += term(m);
    i = e;
    for (int j = 1; j < 4; ++j) {
      acc[j] += acc[j - 1];
      acc[j - 1] = 0.0f;
      if ((i & (mask << (j * lp))) != 0) break;
    }
  }
  for (int m = imax(i, lo), me = imin(n, hi + 1); m < me; ++m) acc[0] += term(m);
  for (int j = 1; j < 4; ++j) acc[0] += acc[j];
  return acc[0];
}
