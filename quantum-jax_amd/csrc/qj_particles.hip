// Cloud-in-cell particle kernels: FP64 indices/weights (bit-exact with the
// reference formula evaluated in IEEE FP64), mass deposit with hardware
// float atomics, 8-corner interpolation, and the fused star leapfrog.
//
// This translation unit is compiled with -ffp-contract=off: the index and
// weight sequence of quantum-stars.py:200-209 and the leapfrog updates of
// :304,311-312 must round exactly like separate IEEE operations.
#include "qj_common.h"
#include "qj_kernels.h"

namespace qj {

namespace {

struct CicGeom {
  int nx, ny, nz;
  double dx, dy, dz, vol, Lx, Ly, Lz;
};

CicGeom make_geom(const Dims& d) {
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

// floored modulo of an integer-valued double (numpy/jnp.mod semantics)
__device__ __forceinline__ int imod(double fi, int n) {
  long long ii = (long long)fi;
  long long r = ii % n;
  return (int)(r < 0 ? r + n : r);
}

// One axis of get_cic_indices_and_weights (quantum-stars.py:202-208).
__device__ __forceinline__ void cic_axis(double x, double dx, int n, int& i0, int& i1, double& w0,
                                         double& w1, bool& ok) {
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

// 4 reals of the interleaved grid cell idx (a_x, a_y, a_z, V).
template <class T>
__device__ __forceinline__ void ld4(const T* __restrict__ A4, int idx, double o[4]) {
  if constexpr (sizeof(T) == 4) {
    const float4 t = *reinterpret_cast<const float4*>(A4 + 4 * idx);
    o[0] = t.x;
    o[1] = t.y;
    o[2] = t.z;
    o[3] = t.w;
  } else {
    const double2 a = *reinterpret_cast<const double2*>(A4 + 4 * idx);
    const double2 b = *reinterpret_cast<const double2*>(A4 + 4 * idx + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
  }
}

template <class T>
__device__ __forceinline__ void interp_one(const CicGeom& g, const Cic& c, const T* A4,
                                           double acc[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
  if (!c.ok) {
    acc[0] = acc[1] = acc[2] = __builtin_nan("");
    return;
  }
  // all 8 corner loads are issued before the accumulation
  double a[8][4];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
    ld4(A4, (c.ix[sx] * g.ny + c.iy[sy]) * g.nz + c.iz[sz], a[k]);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
    const double w = c.wx[sx] * c.wy[sy] * c.wz[sz];
    acc[0] += w * a[k][0];
    acc[1] += w * a[k][1];
    acc[2] += w * a[k][2];
  }
}

// floored modulo with numpy semantics (quantum-stars.py:312 jnp.mod(pos, L))
__device__ __forceinline__ double fmod_floor(double x, double L) {
  double r = fmod(x, L);
  if (r != 0.0) {
    if ((r < 0.0) != (L < 0.0)) r += L;
  } else {
    r = copysign(0.0, L);
  }
  return r;
}

__global__ void k_cic_indices(CicGeom g, const double* __restrict__ pos, long long np,
                              int* __restrict__ i, int* __restrict__ ip1, double* __restrict__ w_i,
                              double* __restrict__ w_ip1) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
  i[3 * p] = c.ix[0];
  i[3 * p + 1] = c.iy[0];
  i[3 * p + 2] = c.iz[0];
  ip1[3 * p] = c.ix[1];
  ip1[3 * p + 1] = c.iy[1];
  ip1[3 * p + 2] = c.iz[1];
  w_i[3 * p] = c.wx[0];
  w_i[3 * p + 1] = c.wy[0];
  w_i[3 * p + 2] = c.wz[0];
  w_ip1[3 * p] = c.wx[1];
  w_ip1[3 * p + 1] = c.wy[1];
  w_ip1[3 * p + 2] = c.wz[1];
}

template <class T>
__global__ void k_deposit(CicGeom g, const double* __restrict__ pos,
                          const double* __restrict__ mass, double m_scalar, long long np,
                          T* __restrict__ rho) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
  const double fac = (mass ? mass[p] : m_scalar) / g.vol;  // quantum-stars.py:219
  deposit_one(g, c, fac, rho);
}

template <class T>
__global__ void k_interp(CicGeom g, const double* __restrict__ pos, long long np,
                         const T* __restrict__ A4, double* __restrict__ acc) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
  double a[3];
  interp_one(g, c, A4, a);
  acc[3 * p] = a[0];
  acc[3 * p + 1] = a[1];
  acc[3 * p + 2] = a[2];
}

// One star step up to the deposit: the closing half kick of the previous
// step (h_owed, same grid and positions), this step's opening half kick,
// the drift with floored periodic wrap, and the deposit at the new position.
template <class T>
__global__ void k_star_step(CicGeom g, double* __restrict__ pos, double* __restrict__ vel,
                            long long np, const T* __restrict__ A4, double h_owed, double h1,
                            double dt, const double* __restrict__ mass, double m_scalar,
                            T* __restrict__ rho) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
  double v[3] = {vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]};
  if (h_owed != 0.0 || h1 != 0.0) {
    Cic c = cic(g, x[0], x[1], x[2]);
    double a[3];
    interp_one(g, c, A4, a);
    if (h_owed != 0.0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = v[d] + a[d] * h_owed;  // quantum-stars.py:321
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = v[d] + a[d] * h1;  // quantum-stars.py:304
  }
  const double L[3] = {g.Lx, g.Ly, g.Lz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    x[d] = x[d] + v[d] * dt;        // quantum-stars.py:311
    x[d] = fmod_floor(x[d], L[d]);  // quantum-stars.py:312
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pos[3 * p + d] = x[d];
    vel[3 * p + d] = v[d];
  }
  Cic c = cic(g, x[0], x[1], x[2]);
  deposit_one(g, c, (mass ? mass[p] : m_scalar) / g.vol, rho);
}

// Closing half kick and dt_s = min |dx / vel| (quantum-stars.py:321,328).
template <class T>
__global__ void k_star_kick(CicGeom g, const double* __restrict__ pos, double* __restrict__ vel,
                            long long np, const T* __restrict__ A4, double h,
                            double* __restrict__ diag) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double dts = __builtin_inf();
  if (p < np) {
    Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    double a[3];
    interp_one(g, c, A4, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double v = vel[3 * p + d] + a[d] * h;
      vel[3 * p + d] = v;
      double r = fabs(g.dx / v);
      dts = (r < dts || r != r) ? r : dts;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double other = __shfl_xor(dts, o, 64);
    dts = (other < dts || other != other) ? other : dts;
  }
  if ((threadIdx.x & 63) == 0 && dts == dts) atomic_min_nonneg(&diag[kDiagDtS], dts);
}

inline int pblocks(long long np) { return (int)((np + 255) / 256); }

}  // namespace

template <class T>
hipError_t launch_cic_indices(const Dims& d, const double* pos, long long np, int* i, int* ip1,
                              double* w_i, double* w_ip1, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  k_cic_indices<<<pblocks(np), 256, 0, s>>>(make_geom(d), pos, np, i, ip1, w_i, w_ip1);
  return hipGetLastError();
}

template <class T>
hipError_t launch_deposit(const Dims& d, const double* pos, const double* mass, double m_scalar,
                          long long np, void* rho, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  k_deposit<T><<<pblocks(np), 256, 0, s>>>(make_geom(d), pos, mass, m_scalar, np, (T*)rho);
  return hipGetLastError();
}

template <class T>
hipError_t launch_interp(const Dims& d, const double* pos, long long np, const void* A4,
                         double* acc, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  k_interp<T><<<pblocks(np), 256, 0, s>>>(make_geom(d), pos, np, (const T*)A4, acc);
  return hipGetLastError();
}

template <class T>
hipError_t launch_star_step(const Dims& d, double* pos, double* vel, long long np, const void* A4,
                            double h_owed, double h1, double dt, const double* mass,
                            double m_scalar, void* rho, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  k_star_step<T><<<pblocks(np), 256, 0, s>>>(make_geom(d), pos, vel, np, (const T*)A4, h_owed,
                                             h1, dt, mass, m_scalar, (T*)rho);
  return hipGetLastError();
}

template <class T>
hipError_t launch_star_kick(const Dims& d, const double* pos, double* vel, long long np,
                            const void* A4, double h, double* diag, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  k_star_kick<T><<<pblocks(np), 256, 0, s>>>(make_geom(d), pos, vel, np, (const T*)A4, h, diag);
  return hipGetLastError();
}

#define QJ_PINST(T)                                                                              \
  template hipError_t launch_cic_indices<T>(const Dims&, const double*, long long, int*, int*,     \
                                            double*, double*, hipStream_t);                        \
  template hipError_t launch_deposit<T>(const Dims&, const double*, const double*, double,        \
                                        long long, void*, hipStream_t);                            \
  template hipError_t launch_interp<T>(const Dims&, const double*, long long, const void*,         \
                                       double*, hipStream_t);                                      \
  template hipError_t launch_star_step<T>(const Dims&, double*, double*, long long, const void*, \
                                          double, double, double, const double*, double, void*,   \
                                          hipStream_t);                                            \
  template hipError_t launch_star_kick<T>(const Dims&, const double*, double*, long long,          \
                                          const void*, double, double*, hipStream_t);

QJ_PINST(float)
QJ_PINST(double)

}  // namespace qj
