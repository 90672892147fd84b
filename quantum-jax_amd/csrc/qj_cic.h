// Cloud-in-cell building blocks shared by the forward and adjoint particle
// kernels: FP64 indices/weights in the reference's operation order
// (quantum-stars.py:200-209), the corner order of :220-243, deposit and
// gather of one particle.  Translation units that include this header are
// compiled with -ffp-contract=off so every operation rounds separately.
#pragma once

#include "qj_common.h"
#include "qj_kernels.h"

namespace qj {

struct CicGeom {
  int nx, ny, nz;
  double dx, dy, dz, vol, Lx, Ly, Lz;
};

inline CicGeom make_geom(const Dims& d) {
  CicGeom g;
  g.nx = d.nx;
  g.ny = d.ny;
  g.nz = d.nz;
  g.dx = d.Lx / d.nx;
  g.dy = d.Ly / d.ny;
  g.dz = d.Lz / d.nz;
  g.vol = g.dx * g.dy * g.dz;  // quantum-stars.py:135
  g.Lx = d.Lx;
  g.Ly = d.Ly;
  g.Lz = d.Lz;
  return g;
}

// floored modulo of an integer-valued double (numpy/jnp.mod semantics).
// For a power-of-two n (every plan axis) it is the low bits of the two's
// complement integer; the general form needs a 64-bit division.
__host__ __device__ __forceinline__ int imod(double fi, int n) {
  long long ii = (long long)fi;
  if ((n & (n - 1)) == 0) return (int)(ii & (long long)(n - 1));
  long long r = ii % n;
  return (int)(r < 0 ? r + n : r);
}

// One axis of get_cic_indices_and_weights (quantum-stars.py:202-208).
// Contraction is switched off here as well, so that every translation unit
// that includes this header (the Poisson x pass of the fused star gather
// among them) computes the same indices as the particle kernels.
__device__ __forceinline__ void cic_axis(double x, double dx, int n, int& i0, int& i1, double& w0,
                                         double& w1, bool& ok) {
#pragma clang fp contract(off)
  double fi = floor((x - 0.5 * dx) / dx);
  double fip1 = fi + 1.0;
  w0 = ((fip1 + 0.5) * dx - x) / dx;
  w1 = (x - (fi + 0.5) * dx) / dx;
  ok = isfinite(fi) && fabs(fi) < 4.0e18;
  i0 = ok ? imod(fi, n) : 0;
  i1 = ok ? imod(fip1, n) : 0;
}

struct Cic {
  int ix[2], iy[2], iz[2];
  double wx[2], wy[2], wz[2];
  bool ok;
};

__device__ __forceinline__ Cic cic(const CicGeom& g, double x, double y, double z) {
  Cic c;
  bool a, b, e;
  cic_axis(x, g.dx, g.nx, c.ix[0], c.ix[1], c.wx[0], c.wx[1], a);
  cic_axis(y, g.dy, g.ny, c.iy[0], c.iy[1], c.wy[0], c.wy[1], b);
  cic_axis(z, g.dz, g.nz, c.iz[0], c.iz[1], c.wz[0], c.wz[1], e);
  c.ok = a && b && e;
  return c;
}

// Corner order of quantum-stars.py:220-243 and :264-287.
__device__ constexpr int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                          {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

template <class T>
__device__ __forceinline__ void deposit_one(const CicGeom& g, const Cic& c, double fac, T* rho) {
  if (!c.ok) return;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
    const double w = c.wx[sx] * c.wy[sy] * c.wz[sz] * fac;
    const int idx = (c.ix[sx] * g.ny + c.iy[sy]) * g.nz + c.iz[sz];
    unsafeAtomicAdd(&rho[idx], (T)w);
  }
}

// (a_x, a_y, a_z) of cell idx of the acceleration grid (12 or 24 bytes,
// 4- or 8-byte aligned: one dwordx3 load for float).
template <class T>
__device__ __forceinline__ void ld3(const T* __restrict__ A3, int idx, double o[3]) {
  const T* q = A3 + kAccComp * idx;
  o[0] = q[0];
  o[1] = q[1];
  o[2] = q[2];
}

// sum over the corners k of w_k a_k, in the corner order
__device__ __forceinline__ void interp_sum(const Cic& c, const double (&a)[8][3], double acc[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
    const double w = c.wx[sx] * c.wy[sy] * c.wz[sz];
    acc[0] += w * a[k][0];
    acc[1] += w * a[k][1];
    acc[2] += w * a[k][2];
  }
}

template <class T>
__device__ __forceinline__ void interp_one(const CicGeom& g, const Cic& c, const T* A3,
                                           double acc[3]) {
  if (!c.ok) {
    acc[0] = acc[1] = acc[2] = __builtin_nan("");
    return;
  }
  // all 8 corner loads are issued before the accumulation
  double a[8][3];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
    ld3(A3, (c.ix[sx] * g.ny + c.iy[sy]) * g.nz + c.iz[sz], a[k]);
  }
  interp_sum(c, a, acc);
}

// interp_one from the corner values a fused gather stored for the star:
// cvp[3 k + e] is component e at corner k, the grid value interp_one would
// load, so the result is the same to the bit.  96 (FP32) or 192 (FP64)
// contiguous bytes per star, 16-byte aligned.
template <class T>
__device__ __forceinline__ void interp_cv(const Cic& c, const T* __restrict__ cvp, double acc[3]) {
  if (!c.ok) {
    acc[0] = acc[1] = acc[2] = __builtin_nan("");
    return;
  }
  double a[8][3];
  if constexpr (sizeof(T) == 4) {
    const float4* q = reinterpret_cast<const float4*>(cvp);
    float f[24];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const float4 t = q[i];
      f[4 * i] = t.x;
      f[4 * i + 1] = t.y;
      f[4 * i + 2] = t.z;
      f[4 * i + 3] = t.w;
    }
#pragma unroll
    for (int i = 0; i < 24; ++i) a[i / 3][i % 3] = f[i];
  } else {
    const double2* q = reinterpret_cast<const double2*>(cvp);
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      const double2 t = q[i];
      a[(2 * i) / 3][(2 * i) % 3] = t.x;
      a[(2 * i + 1) / 3][(2 * i + 1) % 3] = t.y;
    }
  }
  interp_sum(c, a, acc);
}

// reals of corner values per star
constexpr int kCvReals = 24;

// floored modulo with numpy semantics (quantum-stars.py:312 jnp.mod(pos, L)).
// The remainder fmod(x, L) is exact; for 0 < L and -L < x < 2L it is x, or
// x - L when L <= x (exact by Sterbenz's lemma, L <= x <= 2L), so the
// library's iterative fmod is only called outside that range.
__host__ __device__ __forceinline__ double fmod_floor(double x, double L) {
  double r;
  if (L > 0.0 && x > -L && x < 2.0 * L)
    r = (x >= L) ? x - L : x;
  else
    r = fmod(x, L);
  if (r != 0.0) {
    if ((r < 0.0) != (L < 0.0)) r += L;
  } else {
    r = copysign(0.0, L);
  }
  return r;
}


}  // namespace qj
