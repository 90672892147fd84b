// Cloud-in-cell particle kernels: FP64 indices/weights (bit-exact with the
// reference formula evaluated in IEEE FP64), mass deposit with hardware
// float atomics, 8-corner interpolation, and the fused star leapfrog.
//
// This translation unit is compiled with -ffp-contract=off: the index and
// weight sequence of quantum-stars.py:200-209 and the leapfrog updates of
// :304,311-312 must round exactly like separate IEEE operations.
#include "qj_cic.h"

namespace qj {

namespace {

struct k_cic_indices {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             long long np, int* __restrict__ i,
                                             int* __restrict__ ip1, double* __restrict__ w_i,
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
};

template <class T>
struct k_deposit {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             const double* __restrict__ mass, double m_scalar,
                                             long long np, T* __restrict__ rho) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    const double fac = (mass ? mass[p] : m_scalar) / g.vol;  // quantum-stars.py:219
    deposit_one(g, c, fac, rho);
  }
};

template <class T>
struct k_interp {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             long long np, const T* __restrict__ A4,
                                             double* __restrict__ acc) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    double a[3];
    interp_one(g, c, A4, a);
    acc[3 * p] = a[0];
    acc[3 * p + 1] = a[1];
    acc[3 * p + 2] = a[2];
  }
};

// One star step up to the deposit: the closing half kick of the previous
// step (h_owed, same grid and positions), this step's opening half kick,
// the drift with floored periodic wrap, and the deposit at the new position.
template <class T>
struct k_star_step {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, double* __restrict__ pos,
                                             double* __restrict__ vel, long long np,
                                             const T* __restrict__ A4, double h_owed, double h1,
                                             double dt, const double* __restrict__ mass,
                                             double m_scalar, T* __restrict__ rho) {
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
};

// Closing half kick and dt_s = min |dx / vel| (quantum-stars.py:321,328).
template <class T>
struct k_star_kick {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             double* __restrict__ vel, long long np,
                                             const T* __restrict__ A4, double h,
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
};

// d w_c / d x_e for the 8 corners: the per-axis weights are linear in the
// position with slopes -1/dx (w_i) and +1/dx (w_ip1) inside a cell.
__device__ __forceinline__ void corner_grad(const CicGeom& g, const Cic& c, int k, double gw[3]) {
  const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
  const double dx = (sx ? 1.0 : -1.0) / g.dx, dy = (sy ? 1.0 : -1.0) / g.dy,
               dz = (sz ? 1.0 : -1.0) / g.dz;
  gw[0] = dx * c.wy[sy] * c.wz[sz];
  gw[1] = c.wx[sx] * dy * c.wz[sz];
  gw[2] = c.wx[sx] * c.wy[sy] * dz;
}

// Adjoint of vel += h * interp(A4, pos) (and, first, of the drift
// pos += dt*vel when dt_xv != 0: vbar += dt_xv * xbar).  With
// acc_bar = h * vbar: abar[c] += w_c acc_bar (the transpose of the gather,
// into (a_x, a_y, a_z, -) cells) and xbar += sum_c grad w_c (a[c] . acc_bar).
template <class T>
struct k_interp_adj {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             long long np, const T* __restrict__ A4, double h,
                                             double* __restrict__ vbar, double dt_xv,
                                             double* __restrict__ xbar, T* __restrict__ abar4) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double vb[3], xb[3];
  #pragma unroll
    for (int d = 0; d < 3; ++d) {
      xb[d] = xbar[3 * p + d];
      vb[d] = vbar[3 * p + d];
      if (dt_xv != 0.0) vb[d] = vb[d] + dt_xv * xb[d];
      vbar[3 * p + d] = vb[d];
    }
    Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    if (!c.ok) return;
    const double ab[3] = {h * vb[0], h * vb[1], h * vb[2]};
  #pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
      const int idx = (c.ix[sx] * g.ny + c.iy[sy]) * g.nz + c.iz[sz];
      const double w = c.wx[sx] * c.wy[sy] * c.wz[sz];
  #pragma unroll
      for (int d = 0; d < 3; ++d) unsafeAtomicAdd(&abar4[4 * idx + d], (T)(w * ab[d]));
      double a[4];
      ld4(A4, idx, a);
      const double s = a[0] * ab[0] + a[1] * ab[1] + a[2] * ab[2];
      double gw[3];
      corner_grad(g, c, k, gw);
  #pragma unroll
      for (int e = 0; e < 3; ++e) xb[e] += gw[e] * s;
    }
  #pragma unroll
    for (int d = 0; d < 3; ++d) xbar[3 * p + d] = xb[d];
  }
};

// Adjoint of the deposit rho[c] += w_c m / vol: mbar += sum_c w_c rhobar[c] / vol,
// xbar += (m / vol) sum_c grad w_c rhobar[c].
template <class T>
struct k_deposit_adj {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             const double* __restrict__ mass, double m_scalar,
                                             long long np, const T* __restrict__ rhobar,
                                             double* __restrict__ mbar, double* __restrict__ xbar) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    if (!c.ok) return;
    const double fac = (mass ? mass[p] : m_scalar) / g.vol;
    double sm = 0.0, xb[3] = {0.0, 0.0, 0.0};
  #pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
      const double r = (double)rhobar[(c.ix[sx] * g.ny + c.iy[sy]) * g.nz + c.iz[sz]];
      sm += c.wx[sx] * c.wy[sy] * c.wz[sz] * r;
      double gw[3];
      corner_grad(g, c, k, gw);
  #pragma unroll
      for (int e = 0; e < 3; ++e) xb[e] += gw[e] * r;
    }
    mbar[p] += sm / g.vol;
  #pragma unroll
    for (int e = 0; e < 3; ++e) xbar[3 * p + e] += fac * xb[e];
  }
};

inline int pblocks(long long np) { return (int)((np + 255) / 256); }

}  // namespace

template <class T>
hipError_t launch_cic_indices(const Dims& d, const double* pos, long long np, int* i, int* ip1,
                              double* w_i, double* w_ip1, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_cic_indices>(pblocks(np), 256, s, make_geom(d), pos, np, i, ip1, w_i, w_ip1);
  return hipGetLastError();
}

template <class T>
hipError_t launch_deposit(const Dims& d, const double* pos, const double* mass, double m_scalar,
                          long long np, void* rho, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_deposit<T>>(pblocks(np), 256, s, make_geom(d), pos, mass, m_scalar, np, (T*)rho);
  return hipGetLastError();
}

template <class T>
hipError_t launch_interp(const Dims& d, const double* pos, long long np, const void* A4,
                         double* acc, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_interp<T>>(pblocks(np), 256, s, make_geom(d), pos, np, (const T*)A4, acc);
  return hipGetLastError();
}

template <class T>
hipError_t launch_star_step(const Dims& d, double* pos, double* vel, long long np, const void* A4,
                            double h_owed, double h1, double dt, const double* mass,
                            double m_scalar, void* rho, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_star_step<T>>(pblocks(np), 256, s, make_geom(d), pos, vel, np, (const T*)A4, h_owed,
                                             h1, dt, mass, m_scalar, (T*)rho);
  return hipGetLastError();
}

template <class T>
hipError_t launch_star_kick(const Dims& d, const double* pos, double* vel, long long np,
                            const void* A4, double h, double* diag, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_star_kick<T>>(pblocks(np), 256, s, make_geom(d), pos, vel, np, (const T*)A4, h, diag);
  return hipGetLastError();
}

template <class T>
hipError_t launch_interp_adj(const Dims& d, const double* pos, long long np, const void* A4,
                             double h, double* vbar, double dt_xv, double* xbar, void* abar4,
                             hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_interp_adj<T>>(pblocks(np), 256, s, make_geom(d), pos, np, (const T*)A4, h, vbar, dt_xv,
                                              xbar, (T*)abar4);
  return hipGetLastError();
}

template <class T>
hipError_t launch_deposit_adj(const Dims& d, const double* pos, const double* mass,
                              double m_scalar, long long np, const void* rhobar, double* mbar,
                              double* xbar, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_deposit_adj<T>>(pblocks(np), 256, s, make_geom(d), pos, mass, m_scalar, np,
                                               (const T*)rhobar, mbar, xbar);
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
                                          const void*, double, double*, hipStream_t);            \
  template hipError_t launch_interp_adj<T>(const Dims&, const double*, long long, const void*,     \
                                           double, double*, double, double*, void*, hipStream_t); \
  template hipError_t launch_deposit_adj<T>(const Dims&, const double*, const double*, double,     \
                                            long long, const void*, double*, double*, hipStream_t);

QJ_PINST(float)
QJ_PINST(double)

}  // namespace qj
