// Isothermal MUSCL-Hancock gas update of quantum-fluid.py:281-407 as one
// fused stencil kernel (periodic box, cubic cells of size dx).
//
// Per cell it evaluates, in the reference's operation order: centred
// gradients (:200-205), the ratio-form minmod limiter (:352-407), the
// primitive half-step predictor (:295-311), face extrapolation (:208-221),
// the Rusanov flux of the averaged state (:238-264) on the six faces and
// the conservative update (:224-235, :338-347).
//
// The face states of a neighbour need that neighbour's predictor, which
// needs its limited slopes: each thread evaluates the cell state of itself
// and its six face neighbours (a 2-cell stencil reach) from the cache.
#include "qj_common.h"
#include "qj_kernels.h"

namespace qj {

namespace {

template <class T>
struct HGeom {
  int nx, ny, nz;
  T dx, dt, cs, vol;
};

template <class T>
struct Cell {
  T p[4];      // predicted primitives rho', vx', vy', vz'
  T s[3][4];   // limited slopes [axis][var]
};

template <class T>
__device__ __forceinline__ int wrap(int i, int n) {
  return (i < 0) ? i + n : (i >= n ? i - n : i);
}

// quantum-fluid.py:352-407 for one variable along one axis
template <class T>
__device__ __forceinline__ T limit(T f, T fm, T fp, T g, T dx) {
  const T eps = (T)1.0e-12;
  const T denom = (g > (T)0) ? g + eps : ((g < (T)0) ? g - eps : eps);
  T num = (f - fm) / dx;
  T lim = fmax((T)0, fmin((T)1, num / denom));
  g = lim * g;
  num = -(f - fp) / dx;
  lim = fmax((T)0, fmin((T)1, num / denom));
  return lim * g;
}

template <class T>
__device__ __forceinline__ Cell<T> cell_state(const HGeom<T>& g, const T* const* f, int x, int y,
                                              int z) {
  const int n[3] = {g.nx, g.ny, g.nz};
  const int c[3] = {x, y, z};
  auto at = [&](int xx, int yy, int zz) { return (xx * g.ny + yy) * g.nz + zz; };
  const int ic = at(x, y, z);
  T v0[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v0[k] = f[k][ic];
  Cell<T> cs;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int cm[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]};
    cm[a] = wrap<T>(c[a] - 1, n[a]);
    cp[a] = wrap<T>(c[a] + 1, n[a]);
    const int im = at(cm[0], cm[1], cm[2]), ip = at(cp[0], cp[1], cp[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const T fm = f[k][im], fp = f[k][ip];
      const T grad = (fp - fm) / ((T)2 * g.dx);  // quantum-fluid.py:201-203
      cs.s[a][k] = limit<T>(v0[k], fm, fp, grad, g.dx);
    }
  }
  const T rho = v0[0], vx = v0[1], vy = v0[2], vz = v0[3];
  const T hdt = (T)0.5 * g.dt;
  const T c2 = g.cs * g.cs;
  // quantum-fluid.py:295-311
  cs.p[0] = rho - hdt * (vx * cs.s[0][0] + rho * cs.s[0][1] + vy * cs.s[1][0] + rho * cs.s[1][2] +
                         vz * cs.s[2][0] + rho * cs.s[2][3]);
  cs.p[1] = vx - hdt * (vx * cs.s[0][1] + vy * cs.s[1][1] + vz * cs.s[2][1] +
                        ((T)1 / rho) * cs.s[0][0] * g.cs * g.cs);
  cs.p[2] = vy - hdt * (vx * cs.s[0][2] + vy * cs.s[1][2] + vz * cs.s[2][2] +
                        ((T)1 / rho) * cs.s[1][0] * g.cs * g.cs);
  cs.p[3] = vz - hdt * (vx * cs.s[0][3] + vy * cs.s[1][3] + vz * cs.s[2][3] +
                        ((T)1 / rho) * cs.s[2][0] * g.cs * g.cs);
  (void)c2;
  return cs;
}

// Rusanov flux (quantum-fluid.py:238-264); u is the face-normal velocity.
template <class T>
__device__ __forceinline__ void flux(T rL, T uL, T vL, T wL, T rR, T uR, T vR, T wR, T cs, T F[4]) {
  const T rho_star = (T)0.5 * (rL + rR);
  const T mu = (T)0.5 * (rL * uL + rR * uR);
  const T mv = (T)0.5 * (rL * vL + rR * vR);
  const T mw = (T)0.5 * (rL * wL + rR * wR);
  const T P = rho_star * cs * cs;
  F[0] = mu;
  F[1] = mu * mu / rho_star + P;
  F[2] = mu * mv / rho_star;
  F[3] = mu * mw / rho_star;
  const T C = fmax(cs + fabs(uL), cs + fabs(uR));
  F[0] -= C * (T)0.5 * (rR - rL);
  F[1] -= C * (T)0.5 * (rR * uR - rL * uL);
  F[2] -= C * (T)0.5 * (rR * vR - rL * vL);
  F[3] -= C * (T)0.5 * (rR * wR - rL * wL);
}

// Flux through the +a face of the cell with state L, neighbour state R.
// Returns it in conserved-variable order (Mass, Momx, Momy, Momz).
template <class T>
__device__ __forceinline__ void face_flux(const Cell<T>& L, const Cell<T>& R, int a, T dx, T cs,
                                          T out[4]) {
  T fl[4], fr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    fl[k] = L.p[k] + (T)0.5 * L.s[a][k] * dx;  // quantum-fluid.py:209
    fr[k] = R.p[k] - (T)0.5 * R.s[a][k] * dx;  // quantum-fluid.py:210-211
  }
  T F[4];
  if (a == 0) {
    flux<T>(fl[0], fl[1], fl[2], fl[3], fr[0], fr[1], fr[2], fr[3], cs, F);
    out[0] = F[0]; out[1] = F[1]; out[2] = F[2]; out[3] = F[3];
  } else if (a == 1) {  // quantum-fluid.py:331-333
    flux<T>(fl[0], fl[2], fl[3], fl[1], fr[0], fr[2], fr[3], fr[1], cs, F);
    out[0] = F[0]; out[2] = F[1]; out[3] = F[2]; out[1] = F[3];
  } else {              // quantum-fluid.py:334-336
    flux<T>(fl[0], fl[3], fl[1], fl[2], fr[0], fr[3], fr[1], fr[2], cs, F);
    out[0] = F[0]; out[3] = F[1]; out[1] = F[2]; out[2] = F[3];
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_hydro(HGeom<T> g, const T* __restrict__ rho,
                                               const T* __restrict__ vx, const T* __restrict__ vy,
                                               const T* __restrict__ vz, T* __restrict__ rho_o,
                                               T* __restrict__ vx_o, T* __restrict__ vy_o,
                                               T* __restrict__ vz_o) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, x = blockIdx.z;
  if (z >= g.nz) return;
  const T* f[4] = {rho, vx, vy, vz};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int c[3] = {x, y, z};
  const Cell<T> me = cell_state(g, f, x, y, z);
  const int ic = (x * g.ny + y) * g.nz + z;
  // quantum-fluid.py:338 get_conserved with the un-predicted state
  const T r0 = rho[ic];
  T U[4] = {r0 * g.vol, r0 * vx[ic] * g.vol, r0 * vy[ic] * g.vol, r0 * vz[ic] * g.vol};
  const T fac = g.dt * g.dx * g.dx;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int cm[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]};
    cm[a] = wrap<T>(c[a] - 1, n[a]);
    cp[a] = wrap<T>(c[a] + 1, n[a]);
    const Cell<T> lo = cell_state(g, f, cm[0], cm[1], cm[2]);
    const Cell<T> hi = cell_state(g, f, cp[0], cp[1], cp[2]);
    T Fp[4], Fm[4];
    face_flux(me, hi, a, g.dx, g.cs, Fp);
    face_flux(lo, me, a, g.dx, g.cs, Fm);
    // quantum-fluid.py:226-233: F += -fac*flux; F += fac*roll(flux, 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      U[k] += -fac * Fp[k];
      U[k] += fac * Fm[k];
    }
  }
  // quantum-fluid.py:191-197
  rho_o[ic] = U[0] / g.vol;
  vx_o[ic] = U[1] / U[0];
  vy_o[ic] = U[2] / U[0];
  vz_o[ic] = U[3] / U[0];
}

template <class T>
__global__ __launch_bounds__(256) void k_gas_kick(long long n, T* __restrict__ vx,
                                                  T* __restrict__ vy, T* __restrict__ vz,
                                                  const T* __restrict__ A4, T h) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {  // quantum-fluid.py:274-276
    vx[i] += A4[4 * i] * h;
    vy[i] += A4[4 * i + 1] * h;
    vz[i] += A4[4 * i + 2] * h;
  }
}

}  // namespace

template <class T>
hipError_t launch_gas_kick(long long n, void* vx, void* vy, void* vz, const void* A4, double h,
                           hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_gas_kick<T><<<(int)blocks, 256, 0, s>>>(n, (T*)vx, (T*)vy, (T*)vz, (const T*)A4, (T)h);
  return hipGetLastError();
}

template <class T>
hipError_t launch_hydro(const Dims& d, const void* rho, const void* vx, const void* vy,
                        const void* vz, const void* A4, double kick, double dt, double cs,
                        void* rho_o, void* vx_o, void* vy_o, void* vz_o, hipStream_t s) {
  if (kick != 0.0) {
    // pending gravitational kick of the gas velocity (applied in place)
    hipError_t e = launch_gas_kick<T>((long long)d.nx * d.ny * d.nz, (void*)vx, (void*)vy,
                                      (void*)vz, A4, kick, s);
    if (e != hipSuccess) return e;
  }
  HGeom<T> g;
  g.nx = d.nx;
  g.ny = d.ny;
  g.nz = d.nz;
  const double dx = d.Lx / d.nx;
  g.dx = (T)dx;
  g.dt = (T)dt;
  g.cs = (T)cs;
  g.vol = (T)(dx * dx * dx);
  dim3 block(d.nz < 256 ? d.nz : 256);
  dim3 grid((d.nz + block.x - 1) / block.x, d.ny, d.nx);
  k_hydro<T><<<grid, block, 0, s>>>(g, (const T*)rho, (const T*)vx, (const T*)vy, (const T*)vz,
                                    (T*)rho_o, (T*)vx_o, (T*)vy_o, (T*)vz_o);
  return hipGetLastError();
}

template hipError_t launch_hydro<float>(const Dims&, const void*, const void*, const void*,
                                        const void*, const void*, double, double, double, void*,
                                        void*, void*, void*, hipStream_t);
template hipError_t launch_hydro<double>(const Dims&, const void*, const void*, const void*,
                                         const void*, const void*, double, double, double, void*,
                                         void*, void*, void*, hipStream_t);
template hipError_t launch_gas_kick<float>(long long, void*, void*, void*, const void*, double,
                                           hipStream_t);
template hipError_t launch_gas_kick<double>(long long, void*, void*, void*, const void*, double,
                                            hipStream_t);

}  // namespace qj
