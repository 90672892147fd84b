// Isothermal MUSCL-Hancock gas update of quantum-fluid.py:281-407 as one
// fused stencil kernel (periodic box, cubic cells of size dx).
//
// Per cell it evaluates, in the reference's operation order: centred
// gradients (:200-205), the ratio-form minmod limiter (:352-407), the
// primitive half-step predictor (:295-311), face extrapolation (:208-221),
// the Rusanov flux of the averaged state (:238-264) or, by option, the HLL
// flux on the six faces and the conservative update (:224-235, :338-347).
#include "../../include/qj_api.h"
#include "qj_common.h"
#include "qj_kernels.h"

namespace qj {

namespace {

template <class T>
struct HGeom {
  int nx, ny, nz;
  T dx, dt, cs, vol;
};

// Reciprocal: the hardware approximation for FP32 (within 1 ulp), an IEEE
// division for FP64.  Used where the reference divides several values by
// one denominator.
template <class T>
__device__ __forceinline__ T rcp(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
  else return (T)1 / x;
}

// clip(x, 0, 1) (one med3 in FP32)
template <class T>
__device__ __forceinline__ T clip01(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f);
  else return fmax((T)0, fmin((T)1, x));
}

// Centred gradient (quantum-fluid.py:201-203) and the ratio-form minmod
// limiter (:352-407) of one variable along one axis, in cell units: the
// limited slope times dx (the reference's g and eps = 1e-12 scaled by dx, so
// the 1/dx factors of its three ratios cancel).  The reference's denominator
// is g + eps (g > 0), g - eps (g < 0) or eps (g = 0); g + copysign(eps, g)
// differs only for g = -0, where the slope lim * g is -0 either way.
//
// The ratio form is minmod(a, b) of the one-sided differences a = f - fm,
// b = fp - f, up to the eps term: for a, b of one sign with |a| <= |b|,
// clip(a / g) g = a and clip(b / g) = 1 (g = (a + b) / 2); for opposite signs
// one of the ratios is negative and the slope is 0.  The eps term changes a
// slope by the relative amount eps / |g|, 1e-12 / |df/dx|, which is below
// FP32 resolution for every gradient above 1e-5 (and where it is not, the
// slope itself is below 1e-5 dx, under FP32 resolution of the face values).
// FP32 therefore takes minmod directly, med3(a, b, 0) (one instruction
// instead of a reciprocal and eleven others); FP64 keeps the ratio form.
template <class T>
__device__ __forceinline__ T limit(T f, T fm, T fp, T eps_dx) {
  if constexpr (sizeof(T) == 4) {
    (void)eps_dx;
    return __builtin_amdgcn_fmed3f(f - fm, fp - f, 0.0f);
  }
  T g = (T)0.5 * (fp - fm);
  const T denom = g + copysign(eps_dx, g);
  const T rd = rcp(denom);
  T lim = clip01((f - fm) * rd);
  g = lim * g;
  lim = clip01((fp - f) * rd);
  return lim * g;
}

// Rusanov flux (quantum-fluid.py:238-264); u is the face-normal velocity.
template <class T>
__device__ __forceinline__ void flux(T rL, T uL, T vL, T wL, T rR, T uR, T vR, T wR, T cs, T F[4]) {
  const T rho_star = (T)0.5 * (rL + rR);
  const T mu = (T)0.5 * (rL * uL + rR * uR);
  const T mv = (T)0.5 * (rL * vL + rR * vR);
  const T mw = (T)0.5 * (rL * wL + rR * wR);
  const T P = rho_star * cs * cs;
  const T ir = rcp(rho_star);
  F[0] = mu;
  F[1] = mu * mu * ir + P;
  F[2] = mu * mv * ir;
  F[3] = mu * mw * ir;
  // max(cs + |uL|, cs + |uR|) = cs + max(|uL|, |uR|) exactly (rounding is monotonic)
  const T C = cs + fmax(fabs(uL), fabs(uR));
  F[0] -= C * (T)0.5 * (rR - rL);
  F[1] -= C * (T)0.5 * (rR * uR - rL * uL);
  F[2] -= C * (T)0.5 * (rR * vR - rL * vL);
  F[3] -= C * (T)0.5 * (rR * wR - rL * wL);
}

// HLL flux (Harten, Lax and van Leer 1983) of the isothermal system with
// Davis's signal speeds S_L = min(uL, uR) - cs, S_R = max(uL, uR) + cs;
// U = (rho, rho u, rho v, rho w), F(U) = (rho u, rho u^2 + rho cs^2, rho u v,
// rho u w).  Not in the reference (its flux is the Rusanov form above).
template <class T>
__device__ __forceinline__ void flux_hll(T rL, T uL, T vL, T wL, T rR, T uR, T vR, T wR, T cs,
                                         T F[4]) {
  const T c2 = cs * cs;
  const T mL = rL * uL, mR = rR * uR;
  const T SL = fmin(uL, uR) - cs, SR = fmax(uL, uR) + cs;
  if (SL >= (T)0) {
    F[0] = mL;
    F[1] = mL * uL + rL * c2;
    F[2] = mL * vL;
    F[3] = mL * wL;
  } else if (SR <= (T)0) {
    F[0] = mR;
    F[1] = mR * uR + rR * c2;
    F[2] = mR * vR;
    F[3] = mR * wR;
  } else {
    // (SR F_L - SL F_R + SL SR (U_R - U_L)) / (SR - SL), grouped by state:
    // every component of F(U) - S U is (u - S) times a component of U (plus
    // rho cs^2 in the momentum flux), so F = aL U_L - aR U_R + c2 (SR rL - SL rR) / (SR - SL)
    // e_x with aL = SR (uL - SL) / (SR - SL), aR = SL (uR - SR) / (SR - SL).
    const T inv = rcp(SR - SL);
    const T aL = SR * (uL - SL) * inv, aR = SL * (uR - SR) * inv;
    F[0] = rL * aL - rR * aR;
    F[1] = mL * aL - mR * aR + c2 * (SR * rL - SL * rR) * inv;
    F[2] = (rL * vL) * aL - (rR * vR) * aR;
    F[3] = (rL * wL) * aL - (rR * wR) * aR;
  }
}

template <class T, int FLUX>
__device__ __forceinline__ void face_flux(T rL, T uL, T vL, T wL, T rR, T uR, T vR, T wR, T cs,
                                          T F[4]) {
  if constexpr (FLUX == QJ_FLUX_HLL) flux_hll<T>(rL, uL, vL, wL, rR, uR, vR, wR, cs, F);
  else flux<T>(rL, uL, vL, wL, rR, uR, vR, wR, cs, F);
}

// ---------------------------------------------------------------------------
// Streaming version: a workgroup owns a TY x TZ tile of (y, z) columns and
// marches along x over XC planes.  Raw planes (rho, v, with the owed gravity
// kick already applied) live in a 4-slot LDS ring with a 2-cell (y, z) halo;
// the predicted state and y/z slopes of the current plane's tile plus a
// 1-cell ring live in LDS (double-buffered), and each thread keeps its own
// column's predicted state and x slopes for planes x-1, x, x+1 in registers.
// Each cell state is evaluated once (plus the ring), instead of seven times.
// ---------------------------------------------------------------------------
// Max tile edges (y, z) per precision: FP32 runs 8 x 16 tiles (33 KB of LDS
// per 192-thread block, 0.58 ms at 256^3 with HLL; 8 x 8 ran 0.64 ms and
// 16 x 16, 57 KB and two blocks per CU, 0.84 ms); FP64 keeps 16 x 16 (its
// whole c128 gas step ran 2.31 ms with them and 2.43 ms with 8 x 8).
template <class T>
struct HTile {
  static constexpr int Y = sizeof(T) == 4 ? 8 : 16;    // tile edges
  static constexpr int Z = 16;
  static constexpr int RY = Y + 4, RZ = Z + 4;         // raw region (2-cell halo)
  static constexpr int SY = Y + 2, SZ = Z + 2;         // state region (1-cell ring)
  static constexpr int threads = (SY * SZ + 63) / 64 * 64;  // tile cells + ring, whole waves
};

// Element i of a field, addressed by a 32-bit byte offset from the array's
// base (every field is under 4 GiB): the loads and stores then take the
// scalar-base form, one shared offset for all the arrays of a cell, instead
// of 64-bit address arithmetic per array.
template <class T>
__device__ __forceinline__ const T& at(const T* __restrict__ p, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ T& at(T* __restrict__ p, unsigned i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (unsigned)sizeof(T));
}

// (a_x, a_y, a_z) of cell i of the acceleration grid
template <class T>
__device__ __forceinline__ void ld_acc(const T* __restrict__ A3, unsigned i, T a[3]) {
  const T* q = &at(A3, kAccComp * i);
  a[0] = q[0];
  a[1] = q[1];
  a[2] = q[2];
}

template <class T>
struct XState {
  T p[4];
  T sx[4];
};

template <class T, int FLUX>
struct k_hydro_stream {
  static constexpr int kThreads = HTile<T>::threads;
  // FP32: 29.8 KB of LDS fits five 192-thread workgroups per CU, which four
  // waves per SIMD (at most 128 registers) let run
  static constexpr int kMinWaves = sizeof(T) == 4 ? 4 : 0;
  static __device__ __forceinline__ void run(HGeom<T> g, int TY, int TZ, int XC,
                                             const T* __restrict__ rho, const T* __restrict__ vx,
                                             const T* __restrict__ vy, const T* __restrict__ vz,
                                             const T* __restrict__ A3, T kick,
                                             T* __restrict__ rho_o, T* __restrict__ vx_o,
                                             T* __restrict__ vy_o, T* __restrict__ vz_o,
                                             const double* __restrict__ dslot) {
    if (dslot) {  // step time and owed + opening kick from the device step slot
      g.dt = (T)dslot[kSlotDt];
      kick = (T)dslot[kSlotKick];
    }
    __shared__ T raw[4][4][HTile<T>::RY * HTile<T>::RZ];  // [slot][var][cell]
    __shared__ T st[12][HTile<T>::SY * HTile<T>::SZ];  // [p0..3, sy0..3, sz0..3][cell]
    // +y and +z face fluxes of a plane: [y, z][var][cell]
    __shared__ T fyz[2][4][HTile<T>::SY * HTile<T>::SZ];
    const int tid = threadIdx.x;
    const int nthr = TY * TZ;
    const int RZ = TZ + 4, RY = TY + 4, SZ = TZ + 2, SY = TY + 2;
    const int ntz = g.nz / TZ, nty = g.ny / TY;
    const int tile = blockIdx.x % (nty * ntz);
    const int xc = blockIdx.x / (nty * ntz);
    const int y0 = (tile / ntz) * TY, z0 = (tile % ntz) * TZ, x0 = xc * XC;
    const int ty = tid / TZ, tz = tid % TZ;
    const bool own = tid < nthr;
    // slopes are kept in cell units (times dx): hdx = dt / (2 dx) scales the
    // predictor, and the faces are p -+ slope / 2
    const T hdx = (T)0.5 * g.dt / g.dx, eps_dx = (T)1.0e-12 * g.dx;
    const T c2 = g.cs * g.cs;

    // unsigned cell indices (every grid has under 2^32 / 16 cells): a 32-bit
    // offset from the array's base, so the loads and stores take the
    // scalar-base form instead of 64-bit address arithmetic per array
    auto gidx = [&](int x, int y, int z) -> unsigned {
      return (((unsigned)(x & (g.nx - 1)) * (unsigned)g.ny + (unsigned)(y & (g.ny - 1))) *
              (unsigned)g.nz) + (unsigned)(z & (g.nz - 1));
    };
    // A plane of the raw region (tile + 2-cell halo, at most 2 cells per
    // thread) is fetched into registers one iteration before it is written to
    // its LDS ring slot, so the global loads overlap the states and fluxes.
    // The owed gravity kick is applied in put, not here: an add on the
    // loaded values would make the fetch wait for its loads, and the fetch
    // of the next plane is meant to stay in flight through the current one.
    struct Fetch {
      T q[2][4];
      T a[2][3];
    };
    auto fetch = [&](int x, Fetch& f) {
  #pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int c = tid + k * (int)blockDim.x;
        if (c >= RY * RZ) break;
        const int yy = c / RZ - 2, zz = c % RZ - 2;
        const unsigned i = gidx(x, y0 + yy, z0 + zz);
        f.q[k][0] = at(rho, i);
        f.q[k][1] = at(vx, i);
        f.q[k][2] = at(vy, i);
        f.q[k][3] = at(vz, i);
        if (kick != (T)0) ld_acc(A3, i, f.a[k]);
      }
    };
    auto put = [&](int x, const Fetch& f) {
      const int slot = x & 3;
  #pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int c = tid + k * (int)blockDim.x;
        if (c >= RY * RZ) break;
        const int o = (c / RZ) * RZ + (c % RZ);
        T q[4] = {f.q[k][0], f.q[k][1], f.q[k][2], f.q[k][3]};
        if (kick != (T)0) {  // owed gravity kick of the gas velocity, quantum-fluid.py:274-276
          q[1] += f.a[k][0] * kick;
          q[2] += f.a[k][1] * kick;
          q[3] += f.a[k][2] * kick;
        }
  #pragma unroll
        for (int v = 0; v < 4; ++v) raw[slot][v][o] = q[v];
      }
    };
    auto load_plane = [&](int x) {
      Fetch f;
      fetch(x, f);
      put(x, f);
    };
    // predicted state and limited slopes of region cell (yy, zz) at plane x
    auto cell = [&](int x, int yy, int zz, T p[4], T s[3][4]) {
      const int o = (yy + 2) * RZ + (zz + 2);
      const int sm = (x - 1) & 3, sc = x & 3, sp = (x + 1) & 3;
      T v0[4];
  #pragma unroll
      for (int k = 0; k < 4; ++k) v0[k] = raw[sc][k][o];
  #pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[0][k] = limit<T>(v0[k], raw[sm][k][o], raw[sp][k][o], eps_dx);
        s[1][k] = limit<T>(v0[k], raw[sc][k][o - RZ], raw[sc][k][o + RZ], eps_dx);
        s[2][k] = limit<T>(v0[k], raw[sc][k][o - 1], raw[sc][k][o + 1], eps_dx);
      }
      const T r = v0[0], u = v0[1], v = v0[2], w = v0[3];
      // quantum-fluid.py:295-311; the pressure terms (1/rho) rho_d cs cs as
      // (cs^2 / rho) rho_d: one product per component instead of three (the
      // FP32 rounding of that term moves by at most an ulp)
      const T irc2 = rcp(r) * c2;
      p[0] = r - hdx * (u * s[0][0] + r * s[0][1] + v * s[1][0] + r * s[1][2] + w * s[2][0] +
                        r * s[2][3]);
      p[1] = u - hdx * (u * s[0][1] + v * s[1][1] + w * s[2][1] + irc2 * s[0][0]);
      p[2] = v - hdx * (u * s[0][2] + v * s[1][2] + w * s[2][2] + irc2 * s[1][0]);
      p[3] = w - hdx * (u * s[0][3] + v * s[1][3] + w * s[2][3] + irc2 * s[2][0]);
    };
    // states of plane x for the tile and its ring into st; own column
    // (p, s_x) returned in xs
    // one cell per thread: the tile's cells on threads 0..nthr-1, the ring on
    // the next (2*TY + 2*TZ + 4) threads
    auto states = [&](int x, XState<T>& xs) {
      const int ring = SY * SZ - nthr;
      const int c = tid;
      if (c >= nthr + ring) return;
      int yy, zz;
      if (c < nthr) {
        yy = c / TZ;
        zz = c % TZ;
      } else {  // ring cell number r, walked around the tile
        const int r = c - nthr;
        if (r < SZ) { yy = -1; zz = r - 1; }
        else if (r < 2 * SZ) { yy = TY; zz = r - SZ - 1; }
        else { const int q = r - 2 * SZ; yy = q / 2; zz = (q & 1) ? TZ : -1; }
      }
      T p[4], s3[3][4];
      cell(x, yy, zz, p, s3);
      const int o = (yy + 1) * SZ + (zz + 1);
  #pragma unroll
      for (int k = 0; k < 4; ++k) {
        st[k][o] = p[k];
        st[4 + k][o] = s3[1][k];
        st[8 + k][o] = s3[2][k];
      }
      if (c < nthr) {
  #pragma unroll
        for (int k = 0; k < 4; ++k) {
          xs.p[k] = p[k];
          xs.sx[k] = s3[0][k];
        }
      }
    };

    // Flux through the +y (a = 1) or +z (a = 2) face of state cell o, from
    // the states in st (rotated flux arguments, quantum-fluid.py:331-336).
    auto face = [&](int a, int o, T F[4]) {
      const int step = (a == 1) ? SZ : 1;
      const int sl = 4 * a;  // s_y at 4..7, s_z at 8..11
      T fl[4], fr[4], f[4];
  #pragma unroll
      for (int k = 0; k < 4; ++k) {
        fl[k] = st[k][o] + (T)0.5 * st[sl + k][o];
        fr[k] = st[k][o + step] - (T)0.5 * st[sl + k][o + step];
      }
      if (a == 1) {
        face_flux<T, FLUX>(fl[0], fl[2], fl[3], fl[1], fr[0], fr[2], fr[3], fr[1], g.cs, f);
        F[0] = f[0]; F[2] = f[1]; F[3] = f[2]; F[1] = f[3];
      } else {
        face_flux<T, FLUX>(fl[0], fl[3], fl[1], fl[2], fr[0], fr[3], fr[1], fr[2], g.cs, f);
        F[0] = f[0]; F[3] = f[1]; F[1] = f[2]; F[2] = f[3];
      }
    };
    // Each y and z face of a plane evaluated once: a tile cell takes its +y
    // and +z faces (kept in Fy, Fz), the ring row below the tile its +y faces
    // (in Fy) and the ring column left of it its +z faces (in Fz), the
    // tile's -y and -z faces.  faces() only computes them; store_faces()
    // puts them into fyz after the plane's first barrier, which every update
    // of the plane before (the last readers of fyz) has passed, so a single
    // buffer serves (and the workgroup's LDS fits five workgroups per CU).
    // oy, oz: the cells of fyz they go to (-1: none).
    auto faces = [&](T Fy[4], T Fz[4], int& oy, int& oz) {
      const int c = tid;
      oy = oz = -1;
      if (c < nthr) {
        const int o = (ty + 1) * SZ + (tz + 1);
        face(1, o, Fy);
        face(2, o, Fz);
        oy = oz = o;
        return;
      }
      const int r = c - nthr;
      if (r < SZ) {  // ring row yy = -1: its +y face is the tile's -y face
        const int zz = r - 1;
        if (zz < 0 || zz >= TZ) return;
        face(1, zz + 1, Fy);
        oy = zz + 1;
        return;
      }
      const int q = r - 2 * SZ;
      if (q < 0 || (q & 1) || q / 2 >= TY) return;
      const int o = (q / 2 + 1) * SZ;  // ring column zz = -1 of row q / 2
      face(2, o, Fz);
      oz = o;
    };
    auto store_faces = [&](const T Fy[4], const T Fz[4], int oy, int oz) {
      if (oy >= 0) {
  #pragma unroll
        for (int k = 0; k < 4; ++k) fyz[0][k][oy] = Fy[k];
      }
      if (oz >= 0) {
  #pragma unroll
        for (int k = 0; k < 4; ++k) fyz[1][k][oz] = Fz[k];
      }
    };

    XState<T> xm, xc0, xp;
    // warm-up: raw planes x0-2 .. x0+1, states of x0-1 (own column) and x0
    for (int x = x0 - 2; x <= x0 + 1; ++x) load_plane(x);
    __syncthreads();
    if (own) {
      T p[4], s3[3][4];
      cell(x0 - 1, ty, tz, p, s3);
  #pragma unroll
      for (int k = 0; k < 4; ++k) {
        xm.p[k] = p[k];
        xm.sx[k] = s3[0][k];
      }
    }
    states(x0, xc0);
    const T fac = g.dt * g.dx * g.dx;
    // the -x face flux of plane x0; each later plane takes the +x flux of the
    // plane before (same thread, same column), so each x face is evaluated once
    T Fm[4];
    if (own) {
      T fl[4], fr[4];
  #pragma unroll
      for (int k = 0; k < 4; ++k) {
        fl[k] = xm.p[k] + (T)0.5 * xm.sx[k];
        fr[k] = xc0.p[k] - (T)0.5 * xc0.sx[k];
      }
      face_flux<T, FLUX>(fl[0], fl[1], fl[2], fl[3], fr[0], fr[1], fr[2], fr[3], g.cs, Fm);
    }
    Fetch pre;
    fetch(x0 + 2, pre);
    // the first put overwrites the slot of plane x0-2, read by the warm-up
    __syncthreads();
    const T ivol = (T)1 / g.vol;
    // One plane: cur holds the own column's state of plane x, nxt receives
    // that of plane x + 1; Fm is the -x face flux (the +x flux of the plane
    // before), Fp receives the +x flux.  The loop takes planes in pairs with
    // the roles swapped, so these registers rotate without copies.
    auto plane = [&](int x, const XState<T>& cur, XState<T>& nxt, const T* Fm, T* Fp) {
      // the y and z face fluxes of plane x from its states
      T Fy[4], Fz[4];
      int oy, oz;
      faces(Fy, Fz, oy, oz);
      // plane x+2 (fetched during the previous plane) into its slot, which
      // planes x-1 .. x+1 in use leave free; then fetch plane x+3
      put(x + 2, pre);
      if (x + 3 <= x0 + XC + 1) fetch(x + 3, pre);
      __syncthreads();
      store_faces(Fy, Fz, oy, oz);
      states(x + 1, nxt);
      __syncthreads();
      if (own) {
        const int ro = (ty + 2) * RZ + (tz + 2);
        const T r0 = raw[x & 3][0][ro];
        // quantum-fluid.py:338 get_conserved on the (kicked) state of plane x
        T U[4] = {r0 * g.vol, r0 * raw[x & 3][1][ro] * g.vol, r0 * raw[x & 3][2][ro] * g.vol,
                  r0 * raw[x & 3][3][ro] * g.vol};
        T fl[4], fr[4];
        // +x face from the own column's registers
  #pragma unroll
        for (int k = 0; k < 4; ++k) {
          fl[k] = cur.p[k] + (T)0.5 * cur.sx[k];
          fr[k] = nxt.p[k] - (T)0.5 * nxt.sx[k];
        }
        face_flux<T, FLUX>(fl[0], fl[1], fl[2], fl[3], fr[0], fr[1], fr[2], fr[3], g.cs, Fp);
  #pragma unroll
        for (int k = 0; k < 4; ++k) {
          U[k] += -fac * Fp[k];
          U[k] += fac * Fm[k];
        }
        // y and z faces: the +faces from registers, the -faces stored by the
        // neighbouring cell (or ring cell) of the same plane
  #pragma unroll
        for (int k = 0; k < 4; ++k) {
          U[k] += -fac * Fy[k];
          U[k] += fac * fyz[0][k][ty * SZ + (tz + 1)];
        }
  #pragma unroll
        for (int k = 0; k < 4; ++k) {
          U[k] += -fac * Fz[k];
          U[k] += fac * fyz[1][k][(ty + 1) * SZ + tz];
        }
        const unsigned i = gidx(x, y0 + ty, z0 + tz);
        const T iM = rcp(U[0]);
        at(rho_o, i) = U[0] * ivol;  // quantum-fluid.py:191-197 (M / vol)
        at(vx_o, i) = U[1] * iM;
        at(vy_o, i) = U[2] * iM;
        at(vz_o, i) = U[3] * iM;
      }
    };
    // XC is even (a power of two of at least 8, or 32)
    T Fp[4];
    for (int x = x0; x < x0 + XC; x += 2) {
      plane(x, xc0, xp, Fm, Fp);
      plane(x + 1, xp, xc0, Fp, Fm);
    }
  }
};

template <class T>
struct k_gas_kick {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(long long n, T* __restrict__ vx, T* __restrict__ vy,
                                             T* __restrict__ vz, const T* __restrict__ A3, T h) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {  // quantum-fluid.py:274-276
      vx[i] += A3[kAccComp * i] * h;
      vy[i] += A3[kAccComp * i + 1] * h;
      vz[i] += A3[kAccComp * i + 2] * h;
    }
  }
};

// Largest signal speed max(cs + |v|) over the cells (quantum-fluid-accretion.py:697-699,
// isothermal: c_eff = cs), with v the velocity after the owed kick
// v + h a (h = 0: the stored velocity).  NaN-propagating; each block stores
// its maximum in part[blockIdx.x] and k_cfl_finish reduces them (no
// contended atomics on one address).
template <class T>
struct k_gas_signal {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(long long n, const T* __restrict__ vx,
                                             const T* __restrict__ vy, const T* __restrict__ vz,
                                             const T* __restrict__ A3, T h,
                                             const double* __restrict__ hdev, double cs,
                                             double* __restrict__ part) {
    __shared__ double wmax[4];
    if (hdev) h = (T)*hdev;  // the kick owed on the device (step slot)
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    double m = 0.0;
    for (; i < n; i += stride) {
      const double c = A3 ? gas_signal<T>(vx[i], vy[i], vz[i], A3[kAccComp * i],
                                          A3[kAccComp * i + 1], A3[kAccComp * i + 2], h, cs)
                          : gas_signal<T>(vx[i], vy[i], vz[i], cs);
      m = nan_max(m, c);
    }
  #pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double other = __shfl_xor(m, o, 64);
      m = (other > m || other != other) ? other : m;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = wmax[0];
      for (int w = 1; w < 4; ++w) b = (wmax[w] > b || wmax[w] != wmax[w]) ? wmax[w] : b;
      part[blockIdx.x] = b;
    }
  }
};

// dt = min(dt_max, cfl dx / (Cmax + 1e-30)) (quantum-fluid-accretion.py:700-701);
// a NaN signal speed gives a NaN dt.  With advance, the step slot moves on:
// kick = owed + dt/2, owed = dt/2, dts[k] = dt.
// One block of 1024 threads, so that the partial maxima (2048 at 256^3) are
// read two per thread rather than in a chain of 32 loads per thread.
struct k_cfl_finish {
  static constexpr int kThreads = 1024;
  static __device__ __forceinline__ void run(const double* __restrict__ part, int nblk,
                                             double* out, double cfl, double dx, double dt_max,
                                             int advance, double* dts, int k) {
    __shared__ double wmax[kThreads / 64];
    double m = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kThreads) m = nan_max(m, part[i]);
  #pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = nan_max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
  #pragma unroll
      for (int w = 1; w < kThreads / 64; ++w) m = nan_max(m, wmax[w]);
      out[kSlotCmax] = m;
      const double c = m;
      const double dth = cfl * dx / (c + 1e-30);
      const double dt = (c != c) ? c : (dth < dt_max ? dth : dt_max);
      out[kSlotDt] = dt;
      if (advance) {
        out[kSlotKick] = out[kSlotOwed] + 0.5 * dt;
        out[kSlotOwed] = 0.5 * dt;
        dts[k] = dt;
      }
    }
  }
};

}  // namespace

template <class T>
hipError_t launch_gas_cfl(long long n, const void* vx, const void* vy, const void* vz,
                          const void* A3, double h, double cs, double dx, double cfl,
                          double dt_max, double* out, double* dts, int k, double* part,
                          int max_blocks, hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > max_blocks) blocks = max_blocks;
  // with dts, out is a whole step slot and the owed kick lives in it
  const double* hdev = (dts && A3) ? out + kSlotOwed : nullptr;
  launch<k_gas_signal<T>>((int)blocks, 256, s, n, (const T*)vx, (const T*)vy, (const T*)vz,
                          (const T*)A3, (T)h, hdev, cs, part);
  launch<k_cfl_finish>(1, k_cfl_finish::kThreads, s, (const double*)part, (int)blocks, out, cfl, dx, dt_max,
                       dts ? 1 : 0, dts, k);
  return hipGetLastError();
}

hipError_t launch_cfl_finish(const double* part, int nblk, double cfl, double dx, double dt_max,
                             double* slot, double* dts, int k, hipStream_t s) {
  launch<k_cfl_finish>(1, k_cfl_finish::kThreads, s, part, nblk, slot, cfl, dx, dt_max, 1, dts, k);
  return hipGetLastError();
}

template <class T>
hipError_t launch_gas_kick(long long n, void* vx, void* vy, void* vz, const void* A3, double h,
                           hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  launch<k_gas_kick<T>>((int)blocks, 256, s, n, (T*)vx, (T*)vy, (T*)vz, (const T*)A3, (T)h);
  return hipGetLastError();
}

template <class T>
hipError_t launch_hydro(const Dims& d, const void* rho, const void* vx, const void* vy,
                        const void* vz, const void* A3, double kick, double dt, double cs,
                        int flux_kind, void* rho_o, void* vx_o, void* vy_o, void* vz_o,
                        hipStream_t s, const double* dslot) {
  if ((kick != 0.0 || dslot) && A3 == nullptr) return hipErrorInvalidValue;
  if (flux_kind != QJ_FLUX_RUSANOV && flux_kind != QJ_FLUX_HLL) return hipErrorInvalidValue;
  HGeom<T> g;
  g.nx = d.nx;
  g.ny = d.ny;
  g.nz = d.nz;
  const double dx = d.Lx / d.nx;
  g.dx = (T)dx;
  g.dt = (T)dt;
  g.cs = (T)cs;
  g.vol = (T)(dx * dx * dx);
  const int TY = d.ny < HTile<T>::Y ? d.ny : HTile<T>::Y;
  const int TZ = d.nz < HTile<T>::Z ? d.nz : HTile<T>::Z;
  const int XC = d.nx < 32 ? d.nx : 32;
  const int nblk = (d.ny / TY) * (d.nz / TZ) * (d.nx / XC);
  const int nthr = ((TY + 2) * (TZ + 2) + 63) / 64 * 64;  // tile cells + ring, whole waves
  if ((TY + 4) * (TZ + 4) > 2 * nthr) return hipErrorInvalidConfiguration;  // 2 cells per thread
  if (nthr > k_hydro_stream<T, QJ_FLUX_HLL>::kThreads) return hipErrorInvalidConfiguration;
  if (flux_kind == QJ_FLUX_HLL)
    launch<k_hydro_stream<T, QJ_FLUX_HLL>>(nblk, nthr, s, g, TY, TZ, XC, (const T*)rho,
                                           (const T*)vx, (const T*)vy, (const T*)vz,
                                           (const T*)A3, (T)kick, (T*)rho_o, (T*)vx_o,
                                           (T*)vy_o, (T*)vz_o, dslot);
  else
    launch<k_hydro_stream<T, QJ_FLUX_RUSANOV>>(nblk, nthr, s, g, TY, TZ, XC, (const T*)rho,
                                               (const T*)vx, (const T*)vy, (const T*)vz,
                                               (const T*)A3, (T)kick, (T*)rho_o, (T*)vx_o,
                                               (T*)vy_o, (T*)vz_o, dslot);
  return hipGetLastError();
}

template hipError_t launch_hydro<float>(const Dims&, const void*, const void*, const void*,
                                        const void*, const void*, double, double, double, int,
                                        void*, void*, void*, void*, hipStream_t, const double*);
template hipError_t launch_hydro<double>(const Dims&, const void*, const void*, const void*,
                                         const void*, const void*, double, double, double, int,
                                         void*, void*, void*, void*, hipStream_t, const double*);
template hipError_t launch_gas_kick<float>(long long, void*, void*, void*, const void*, double,
                                           hipStream_t);
template hipError_t launch_gas_cfl<float>(long long, const void*, const void*, const void*,
                                          const void*, double, double, double, double, double,
                                          double*, double*, int, double*, int, hipStream_t);
template hipError_t launch_gas_cfl<double>(long long, const void*, const void*, const void*,
                                           const void*, double, double, double, double, double,
                                           double*, double*, int, double*, int, hipStream_t);
template hipError_t launch_gas_kick<double>(long long, void*, void*, void*, const void*, double,
                                            hipStream_t);

}  // namespace qj
