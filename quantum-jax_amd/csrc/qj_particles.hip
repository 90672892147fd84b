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
                                             long long np, const T* __restrict__ A3,
                                             double* __restrict__ acc) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    double a[3];
    interp_one(g, c, A3, a);
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
                                             const T* __restrict__ A3, double h_owed, double h1,
                                             double dt, const double* __restrict__ mass,
                                             double m_scalar, T* __restrict__ rho) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    double v[3] = {vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]};
    if (h_owed != 0.0 || h1 != 0.0) {
      Cic c = cic(g, x[0], x[1], x[2]);
      double a[3];
      interp_one(g, c, A3, a);
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

// Closing half kick and dt_s = min |dx / vel| (quantum-stars.py:321,328)
// over the finite values (a non-finite velocity is left out of the minimum;
// jnp.min would return NaN), one atomic per block.
template <class T>
struct k_star_kick {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             double* __restrict__ vel, long long np,
                                             const T* __restrict__ A3, double h,
                                             double* __restrict__ diag,
                                             const T* __restrict__ cv) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dts = __builtin_inf();
    if (p < np) {
      Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
      double a[3];
      if (cv)
        interp_cv(c, cv + kCvReals * p, a);
      else
        interp_one(g, c, A3, a);
  #pragma unroll
      for (int d = 0; d < 3; ++d) {
        double v = vel[3 * p + d] + a[d] * h;
        vel[3 * p + d] = v;
        double r = fabs(g.dx / v);
        dts = (r < dts) ? r : dts;  // NaN compares false: skipped
      }
    }
  #pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double other = __shfl_xor(dts, o, 64);
      dts = (other < dts) ? other : dts;
    }
    __shared__ double wmin[kThreads / 64];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dts;
    __syncthreads();
    if (threadIdx.x == 0) {
  #pragma unroll
      for (int w = 1; w < kThreads / 64; ++w) dts = (wmin[w] < dts) ? wmin[w] : dts;
      if (dts < __builtin_inf()) atomic_min_nonneg(&diag[kDiagDtS], dts);
    }
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

// ---------------------------------------------------------------------------
// Tile-binned deposit (owner computes).  The grid is cut into tiles of
// 2^lx x 2^ly x 2^lz cells.  Each particle is registered in every tile its
// CIC stencil touches (1 to 8, distinct); one block per tile then adds the
// corners that fall inside its tile in LDS and writes the whole tile with
// plain coalesced stores, zeros included.  No global float atomics, and
// every cell of rho is overwritten (no clearing needed beforehand).
//   count  : per-tile registration counter (zero between steps)
//   rank   : slot of each registration within its tile (8 per particle)
//   offset : exclusive scan of count, ntiles + 1 entries
//   sorted : registrations grouped by tile (particle indices)
// ---------------------------------------------------------------------------

// The distinct tiles touched by the stencil of c, in eight fixed slots
// k = 4a + 2b + e: slot k holds the tile of the stencil's a-th x, b-th y and
// e-th z tile and is used (bit k of the returned mask) when the stencil
// reaches a second tile along every axis whose index there is 1.  Slot 0 is
// always used.  Fixed slots keep ids[] in registers and let a thread issue
// all its registrations before it waits on any of them.
__device__ __forceinline__ unsigned stencil_tiles(const TileGeom& t, const Cic& c, int ids[8]) {
  const int ax[2] = {c.ix[0] >> t.lx, c.ix[1] >> t.lx};
  const int ay[2] = {c.iy[0] >> t.ly, c.iy[1] >> t.ly};
  const int az[2] = {c.iz[0] >> t.lz, c.iz[1] >> t.lz};
  const bool two[3] = {ax[1] != ax[0], ay[1] != ay[0], az[1] != az[0]};
  unsigned mask = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int a = k >> 2, b = (k >> 1) & 1, e = k & 1;
    ids[k] = (ax[a] * t.nty + ay[b]) * t.ntz + az[e];
    if ((!a || two[0]) && (!b || two[1]) && (!e || two[2])) mask |= 1u << k;
  }
  return mask;
}

// Column tiles in the one-row registration (StarBins::home): the tiles of
// the star's lower y row, one per z tile its z pair touches (slot 0: the z
// tile of iz0; slot 1: that of iz1 when it differs).
__device__ __forceinline__ unsigned home_tiles(const TileGeom& t, const Cic& c, int ids[2]) {
  const int ty = c.iy[0] >> t.ly;
  const int z0 = c.iz[0] >> t.lz, z1 = c.iz[1] >> t.lz;
  ids[0] = ty * t.ntz + z0;
  ids[1] = ty * t.ntz + z1;
  return z1 != z0 ? 3u : 1u;
}

// Absolute corner descriptor of the one-row registration: the star's two
// cell indices per axis, 9 bits each (x0, x1, y0, y1, z0, z1 from bit 0);
// a consumer tests each (y, z) corner pair against its own tile.
__device__ __forceinline__ unsigned long long home_desc(const Cic& c) {
  return (unsigned long long)c.ix[0] | ((unsigned long long)c.ix[1] << 9) |
         ((unsigned long long)c.iy[0] << 18) | ((unsigned long long)c.iy[1] << 27) |
         ((unsigned long long)c.iz[0] << 36) | ((unsigned long long)c.iz[1] << 45);
}

// Registers the star p in the tiles of its stencil: one counting atomic per
// used slot, all issued before the returned ranks are stored.
__device__ __forceinline__ void register_tiles(const TileGeom& tg, const Cic& c, long long p,
                                               int* __restrict__ count, int* __restrict__ rank) {
  int ids[8], r[8];
  const unsigned m = stencil_tiles(tg, c, ids);
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (m >> k & 1u) r[k] = atomicAdd(&count[ids[k] * kCountStride], 1);
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (m >> k & 1u) rank[8 * p + k] = r[k];
}

// Corner descriptor of a star for column tile t (all of x by 2^ly y rows by
// 2^lz z columns): bits 0-8 and 9-17 hold the star's two cell x indices,
// and the four 9-bit fields from bit 18 hold, for the (y, z) corner pairs
// (sy, sz) in the order 2 sy + sz, 1 + the column's index in the tile
// (dy 2^lz + dz), or 0 when the column is outside it.  The Poisson x pass
// reads these instead of the position.
__device__ __forceinline__ unsigned long long corner_desc(const CicGeom& g, const TileGeom& tg,
                                                         const Cic& c, int t) {
  (void)g;
  const int y0 = (t / tg.ntz) << tg.ly, z0 = (t % tg.ntz) << tg.lz;
  unsigned long long d = (unsigned long long)c.ix[0] | ((unsigned long long)c.ix[1] << 9);
#pragma unroll
  for (int yz = 0; yz < 4; ++yz) {
    const int dy = c.iy[yz >> 1] - y0, dz = c.iz[yz & 1] - z0;
    const bool in = (unsigned)dy < (1u << tg.ly) && (unsigned)dz < (1u << tg.lz);
    const unsigned f = in ? (unsigned)((dy << tg.lz) | dz) + 1u : 0u;
    d |= (unsigned long long)f << (18 + 9 * yz);
  }
  return d;
}

// Adjoint of vel += h * interp(A3, pos) (and, first, of the drift
// pos += dt*vel when dt_xv != 0: vbar += dt_xv * xbar).  With
// acc_bar = h * vbar: abar[c] += w_c acc_bar (the transpose of the gather,
// into (a_x, a_y, a_z, -) cells) and xbar += sum_c grad w_c (a[c] . acc_bar).
// BIN: instead of adding h vbar into abar3 with atomics, register the star
// in the tiles of its stencil (the tile-binned deposit adds it afterwards).
template <class T, bool BIN = false>
struct k_interp_adj {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, const double* __restrict__ pos,
                                             long long np, const T* __restrict__ A3, double h,
                                             double* __restrict__ vbar, double dt_xv,
                                             double* __restrict__ xbar, T* __restrict__ abar3,
                                             TileGeom tg, int* __restrict__ count,
                                             int* __restrict__ rank) {
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
    if constexpr (BIN) {  // count == nullptr: registrations taped by the forward
      if (count) register_tiles(tg, c, p, count, rank);
    }
  #pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
      const int idx = (c.ix[sx] * g.ny + c.iy[sy]) * g.nz + c.iz[sz];
      const double w = c.wx[sx] * c.wy[sy] * c.wz[sz];
      if constexpr (!BIN) {
  #pragma unroll
        for (int d = 0; d < 3; ++d) unsafeAtomicAdd(&abar3[kAccComp * idx + d], (T)(w * ab[d]));
      }
      double a[3];
      ld3(A3, idx, a);
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


// Star step with registration counting instead of the deposit.  MOVE = false
// only counts the current positions (the deposit of the initial state).
template <class T, bool MOVE>
struct k_star_step_bin {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, TileGeom tg, double* __restrict__ pos,
                                             double* __restrict__ vel, long long np,
                                             const T* __restrict__ A3, double h_owed, double h1,
                                             double dt, int* __restrict__ count,
                                             int* __restrict__ rank, const T* __restrict__ cv,
                                             int home) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    if constexpr (MOVE) {
      double v[3] = {vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]};
      if (h_owed != 0.0 || h1 != 0.0) {
        Cic c = cic(g, x[0], x[1], x[2]);
        double a[3];
        if (cv)
          interp_cv(c, cv + kCvReals * p, a);
        else
          interp_one(g, c, A3, a);
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
    }
    Cic c = cic(g, x[0], x[1], x[2]);
    if (!c.ok) return;  // non-finite positions deposit nothing
    if (home) {
      int ids[2], r[2];
      const unsigned m = home_tiles(tg, c, ids);
  #pragma unroll
      for (int k = 0; k < 2; ++k)
        if (m >> k & 1u) r[k] = atomicAdd(&count[ids[k] * kCountStride], 1);
  #pragma unroll
      for (int k = 0; k < 2; ++k)
        if (m >> k & 1u) rank[8 * p + k] = r[k];
    } else {
      register_tiles(tg, c, p, count, rank);
    }
  }
};

// Exclusive scan of the tile counts (one block of 1024 threads); resets count.
struct k_bin_scan {
  static constexpr int kThreads = 1024;
  static __device__ __forceinline__ void run(int* __restrict__ count, int* __restrict__ offset,
                                             int ntiles) {
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int per = (ntiles + 1023) / 1024;
    const int b = tid * per, e = min(b + per, ntiles);
    // Up to kHeld counts per thread (ntiles <= 8192) are loaded once, all in
    // flight together, and held in registers for the second pass.
    constexpr int kHeld = 8;
    const bool held_all = per <= kHeld;
    int held[kHeld];
    int local = 0;
    if (held_all) {
#pragma unroll
      for (int r = 0; r < kHeld; ++r) held[r] = (b + r < e) ? count[(b + r) * kCountStride] : 0;
#pragma unroll
      for (int r = 0; r < kHeld; ++r) local += held[r];
    } else {
      for (int i = b; i < e; ++i) local += count[i * kCountStride];
    }
    // inclusive scan of the per-thread sums within the wave, then across waves
    int incl = local;
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; ++w) base += wsum[w];
    int run = base + incl - local;
    if (held_all) {
#pragma unroll
      for (int r = 0; r < kHeld; ++r) {
        if (b + r < e) {
          offset[b + r] = run;
          run += held[r];
          count[(b + r) * kCountStride] = 0;
        }
      }
    } else {
      for (int i = b; i < e; ++i) {
        const int c = count[i * kCountStride];
        offset[i] = run;
        run += c;
        count[i * kCountStride] = 0;
      }
    }
    if (tid == 1023) offset[ntiles] = base + incl;
  }
};

// Registrations into their tile's slots (the ranks of the counting pass).
// With desc (column tiles of the fused star gather), each registration also
// gets the star's corner descriptor relative to its tile (corner_desc); with
// wts, the star's eight corner weights m w / vol as the per-tile deposit
// computes them (the same products in the same order), so the deposit fused
// into the source pass reads 8 reals per registration instead of computing
// the CIC weights of the star again.
template <class T>
struct k_bin_scatter {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, TileGeom tg,
                                             const double* __restrict__ pos, long long np,
                                             const int* __restrict__ rank,
                                             const int* __restrict__ offset,
                                             int* __restrict__ sorted,
                                             unsigned long long* __restrict__ desc,
                                             T* __restrict__ wts, const double* __restrict__ mass,
                                             double m_scalar, int home) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
    if (!c.ok) return;
    if (home) {
      int ids[2], o[2];
      const unsigned m = home_tiles(tg, c, ids);
      const unsigned long long d = home_desc(c);
#pragma unroll
      for (int k = 0; k < 2; ++k)
        if (m >> k & 1u) o[k] = offset[ids[k]] + rank[8 * p + k];
#pragma unroll
      for (int k = 0; k < 2; ++k)
        if (m >> k & 1u) {
          sorted[o[k]] = (int)p;
          desc[o[k]] = d;
        }
    } else {
      int ids[8], o[8];
      const unsigned m = stencil_tiles(tg, c, ids);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (m >> k & 1u) o[k] = offset[ids[k]] + rank[8 * p + k];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (m >> k & 1u) sorted[o[k]] = (int)p;
      if (desc) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (m >> k & 1u) desc[o[k]] = corner_desc(g, tg, c, ids[k]);
      }
    }
    if (wts) {
      const double fac = (mass ? mass[p] : m_scalar) / g.vol;
      T w[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
        w[4 * sy + 2 * sz + sx] = (T)(c.wx[sx] * c.wy[sy] * c.wz[sz] * fac);
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) wts[8 * p + e] = w[e];
    }
  }
};

// One block per tile: LDS accumulation of the corners inside the tile (the
// corner values of deposit_one), then the whole tile is stored.
// VOL: LDS cells per tile (the tile volume rounded up to a supported size),
// so small tiles take small LDS and more blocks fit per CU.
template <class T, int VOL = kTileMaxCells>
struct k_tile_deposit {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, TileGeom tg,
                                             const double* __restrict__ pos,
                                             const double* __restrict__ mass, double m_scalar,
                                             const int* __restrict__ offset,
                                             const int* __restrict__ sorted, T* __restrict__ rho) {
    __shared__ T acc[VOL];
    const int t = blockIdx.x;
    const int tz = t % tg.ntz, ty = (t / tg.ntz) % tg.nty, tx = t / (tg.ntz * tg.nty);
    const int vol = 1 << (tg.lx + tg.ly + tg.lz);
    for (int l = threadIdx.x; l < vol; l += blockDim.x) acc[l] = (T)0;
    __syncthreads();
    const int b = offset[t], e = offset[t + 1];
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
      const int p = sorted[i];
      Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
      const double fac = (mass ? mass[p] : m_scalar) / g.vol;
  #pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
        const int ix = c.ix[sx], iy = c.iy[sy], iz = c.iz[sz];
        if ((ix >> tg.lx) != tx || (iy >> tg.ly) != ty || (iz >> tg.lz) != tz) continue;
        const double w = c.wx[sx] * c.wy[sy] * c.wz[sz] * fac;
        const int l = (((ix & ((1 << tg.lx) - 1)) << tg.ly | (iy & ((1 << tg.ly) - 1))) << tg.lz) |
                      (iz & ((1 << tg.lz) - 1));
        atomicAdd(&acc[l], (T)w);
      }
    }
    __syncthreads();
    const int x0 = tx << tg.lx, y0 = ty << tg.ly, z0 = tz << tg.lz;
    const int mz = (1 << tg.lz) - 1, my = (1 << tg.ly) - 1;
    for (int l = threadIdx.x; l < vol; l += blockDim.x) {
      const int lz = l & mz, ly = (l >> tg.lz) & my, lx = l >> (tg.lz + tg.ly);
      rho[((long long)(x0 + lx) * g.ny + (y0 + ly)) * g.nz + (z0 + lz)] = acc[l];
    }
  }
};

// Deterministic form of k_tile_deposit: the corner values are added as
// 64-bit fixed-point integers, whose sum does not depend on the order of
// the additions, so every run gives the same bits.  The tile's scale 2^sh
// is a power of two chosen from an upper bound on any cell's total, B =
// (stars registered in the tile) max|m| / vol < 2^ex, so that |sum| 2^sh <
// 2^62; each corner value is rounded to 2^-sh (about 2^-62 B, far below
// FP64 resolution for tiles of up to thousands of stars) and the exact sum
// is rounded once to T.  A non-finite mass in the tile makes the tile NaN.
template <class T, int VOL = kTileMaxCells>
struct k_tile_deposit_det {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, TileGeom tg,
                                             const double* __restrict__ pos,
                                             const double* __restrict__ mass, double m_scalar,
                                             const int* __restrict__ offset,
                                             const int* __restrict__ sorted, T* __restrict__ rho) {
    __shared__ unsigned long long acc[VOL];
    __shared__ double wmax[kThreads / 64];
    const int t = blockIdx.x;
    const int tz = t % tg.ntz, ty = (t / tg.ntz) % tg.nty, tx = t / (tg.ntz * tg.nty);
    const int vol = 1 << (tg.lx + tg.ly + tg.lz);
    for (int l = threadIdx.x; l < vol; l += blockDim.x) acc[l] = 0ull;
    const int b = offset[t], e = offset[t + 1];
    // max |m| over the tile's stars (NaN-propagating; order-independent)
    double mx = mass ? 0.0 : fabs(m_scalar);
    if (mass) {
      for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
        const double a = fabs(mass[sorted[i]]);
        mx = (a > mx || a != a) ? a : mx;
      }
    }
  #pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(mx, o, 64);
      mx = (y > mx || y != y) ? y : mx;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
  #pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) mx = (wmax[w] > mx || wmax[w] != wmax[w]) ? wmax[w] : mx;
    const double bound = (double)(e - b) * mx / g.vol;
    const bool bad = !isfinite(bound);
    int ex = 0;
    if (bound > 0.0 && !bad) (void)frexp(bound, &ex);  // bound < 2^ex
    const int sh = 62 - ex;
    if (bound > 0.0 && !bad) {
      for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
        const int p = sorted[i];
        Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        const double fac = (mass ? mass[p] : m_scalar) / g.vol;
  #pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
          const int ix = c.ix[sx], iy = c.iy[sy], iz = c.iz[sz];
          if ((ix >> tg.lx) != tx || (iy >> tg.ly) != ty || (iz >> tg.lz) != tz) continue;
          const double w = c.wx[sx] * c.wy[sy] * c.wz[sz] * fac;
          const int l = (((ix & ((1 << tg.lx) - 1)) << tg.ly | (iy & ((1 << tg.ly) - 1))) << tg.lz) |
                        (iz & ((1 << tg.lz) - 1));
          // two's-complement addition: signed values add exactly
          atomicAdd(&acc[l], (unsigned long long)(long long)rint(ldexp(w, sh)));
        }
      }
    }
    __syncthreads();
    const int x0 = tx << tg.lx, y0 = ty << tg.ly, z0 = tz << tg.lz;
    const int mz = (1 << tg.lz) - 1, my = (1 << tg.ly) - 1;
    for (int l = threadIdx.x; l < vol; l += blockDim.x) {
      const int lz = l & mz, ly = (l >> tg.lz) & my, lx = l >> (tg.lz + tg.ly);
      const T v = bad ? (T)NAN : (T)ldexp((double)(long long)acc[l], -sh);
      rho[((long long)(x0 + lx) * g.ny + (y0 + ly)) * g.nz + (z0 + lz)] = v;
    }
  }
};

// Vector form of the per-tile deposit for the adjoint of the star kicks:
// abar[c] = sum over stars of w_c h vbar (the corner values of the atomic
// interpolation adjoint), written as whole (a_x, a_y, a_z) tiles.
template <class T, int VOL = kTileMaxCells>
struct k_tile_deposit_vec {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, TileGeom tg,
                                             const double* __restrict__ pos,
                                             const double* __restrict__ vbar, double h,
                                             const int* __restrict__ offset,
                                             const int* __restrict__ sorted,
                                             T* __restrict__ abar3) {
    __shared__ T acc[3][VOL];
    const int t = blockIdx.x;
    const int tz = t % tg.ntz, ty = (t / tg.ntz) % tg.nty, tx = t / (tg.ntz * tg.nty);
    const int vol = 1 << (tg.lx + tg.ly + tg.lz);
    for (int l = threadIdx.x; l < vol; l += blockDim.x) acc[0][l] = acc[1][l] = acc[2][l] = (T)0;
    __syncthreads();
    const int b = offset[t], e = offset[t + 1];
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
      const int p = sorted[i];
      Cic c = cic(g, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
      const double ab[3] = {h * vbar[3 * p], h * vbar[3 * p + 1], h * vbar[3 * p + 2]};
  #pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
        const int ix = c.ix[sx], iy = c.iy[sy], iz = c.iz[sz];
        if ((ix >> tg.lx) != tx || (iy >> tg.ly) != ty || (iz >> tg.lz) != tz) continue;
        const double w = c.wx[sx] * c.wy[sy] * c.wz[sz];
        const int l = (((ix & ((1 << tg.lx) - 1)) << tg.ly | (iy & ((1 << tg.ly) - 1))) << tg.lz) |
                      (iz & ((1 << tg.lz) - 1));
  #pragma unroll
        for (int d = 0; d < 3; ++d) atomicAdd(&acc[d][l], (T)(w * ab[d]));
      }
    }
    __syncthreads();
    const int x0 = tx << tg.lx, y0 = ty << tg.ly, z0 = tz << tg.lz;
    const int mz = (1 << tg.lz) - 1, my = (1 << tg.ly) - 1;
    for (int l = threadIdx.x; l < vol; l += blockDim.x) {
      const int lz = l & mz, ly = (l >> tg.lz) & my, lx = l >> (tg.lz + tg.ly);
      const long long idx = ((long long)(x0 + lx) * g.ny + (y0 + ly)) * g.nz + (z0 + lz);
      T* o = abar3 + kAccComp * idx;
      o[0] = acc[0][l];
      o[1] = acc[1][l];
      o[2] = acc[2][l];
    }
  }
};

// Accretion of the field onto black hole 0 (quantum-bh.py:306-339): the cell
// that holds it (floored, per-axis cell size dx as in the reference), the
// rate dM/dt = 32 pi^2 (G m)^3 (m/hbar) |psi|^2 / (c^3 v (1 - exp(-xi))) with
// v = sqrt(sigma^2 + |vel|^2) and xi = 2 pi G m (m/hbar) / v, the amplitude
// at that cell lowered by sqrt(dM/dt dt / vol) at fixed phase, and dM/dt dt
// added to the mass array (the reference adds the scalar to every entry).
// One thread, FP64.
template <class T>
struct k_bh_accretion {
  static constexpr int kThreads = 64;
  static __device__ __forceinline__ void run(CicGeom g, cpx<T>* __restrict__ psi,
                                             const double* __restrict__ pos,
                                             const double* __restrict__ vel,
                                             double* __restrict__ m_bh, long long n_bh, double G,
                                             double c, double sigma, double mph, double dt) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int ii[3];
    const int n[3] = {g.nx, g.ny, g.nz};
  #pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double fi = floor((pos[d] - 0.5 * g.dx) / g.dx);
      if (!(isfinite(fi) && fabs(fi) < 4.0e18)) return;
      ii[d] = imod(fi, n[d]);
    }
    const long long idx = ((long long)ii[0] * g.ny + ii[1]) * g.nz + ii[2];
    const cpx<T> p = psi[idx];
    const double amp = sqrt((double)p.r * p.r + (double)p.i * p.i);
    const double th = atan2((double)p.i, (double)p.r);
    const double v = sqrt(sigma * sigma + vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    const double Gm = G * m_bh[0];
    const double xi = 2.0 * M_PI * Gm * mph / v;
    const double dMdt = 32.0 * M_PI * M_PI * (Gm * Gm * Gm) * mph * (amp * amp) /
                        (c * c * c * v * (1.0 - exp(-xi)));
    const double na = amp - sqrt(dMdt * dt / g.vol);
    psi[idx] = {(T)(na * cos(th)), (T)(na * sin(th))};
    const double dm = dMdt * dt;
    for (long long b = 0; b < n_bh; ++b) m_bh[b] = m_bh[b] + dm;
  }
};

// Ambient Bondi sink of quantum-fluid-accretion.py:440-510 (fixed-grid
// sink radius r_acc = r_acc_mult dx, ambient density and gas velocity
// averaged over a thin shell outside it, Bondi/BHL target rate, removal
// weighted by a Gaussian kernel times the density inside r_acc, capped by
// the gas available there).  One block over the box of cells whose centres
// can lie within r_hi of the hole; every cell outside the box has a zero
// mask in the reference as well.  FP64 sums in a fixed order.
// bh = (M_eff, x, y, z, vx, vy, vz); out = (dMdt_target, rho_inf, r_B,
// dM_cap, r_acc, M_avail, dM_uncl, shell cell count).
struct BondiPrm {
  double cs, G, dt, r_acc_mult, lam, fmax, ann_lo, ann_hi;
  int box;  // cells per axis of the box (<= n)
};

template <class T>
struct k_bondi_sink {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(CicGeom g, BondiPrm q, T* __restrict__ rho,
                                             const T* __restrict__ vx, const T* __restrict__ vy,
                                             const T* __restrict__ vz,
                                             const double* __restrict__ bh,
                                             double* __restrict__ out) {
    // the reference's cubic box: dx and L of the x axis for every axis
    const double dx = g.dx, L = g.Lx;
    const int n = g.nx;
    const double M = bh[0], xb[3] = {bh[1], bh[2], bh[3]};
    const double r_acc = q.r_acc_mult * dx;
    const double rB = q.G * M / (q.cs * q.cs + 1e-30);
    const double rK = fmin(0.5 * r_acc, fmax(0.5 * dx, rB));
    const double r_lo = fmax(q.ann_lo * r_acc, r_acc + 0.25 * dx);
    const double r_hi = fmax(q.ann_hi * r_acc, r_lo + 0.5 * dx);
    const int W = q.box, cells = W * W * W;
    int lo[3];
  #pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double c = floor(xb[d] / dx - 0.5);
      lo[d] = (isfinite(c) ? (int)c : 0) - (W - 1) / 2;
    }
    // the cell (c) of the box: wrapped indices and distance to the hole
    auto at = [&](int c, long long& idx, double& R) {
      int ii[3] = {c / (W * W), (c / W) % W, c % W};
      double r2 = 0.0;
  #pragma unroll
      for (int d = 0; d < 3; ++d) {
        int i = (lo[d] + ii[d]) % n;
        if (i < 0) i += n;
        ii[d] = i;
        const double dd = fabs(((double)i + 0.5) * dx - xb[d]);  // :384-386
        const double m = fmin(dd, L - dd);
        r2 += m * m;
      }
      idx = ((long long)ii[0] * g.ny + ii[1]) * g.nz + ii[2];
      R = sqrt(r2);
    };
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // shell n, rho, vx, vy, vz; sink rho; W rho
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
      long long idx;
      double R;
      at(c, idx, R);
      const double r = (double)rho[idx];
      if (R >= r_lo && R <= r_hi) {
        acc[0] += 1.0;
        acc[1] += r;
        if (vx) {
          acc[2] += (double)vx[idx];
          acc[3] += (double)vy[idx];
          acc[4] += (double)vz[idx];
        }
      }
      if (R <= r_acc) {
        const double z = R / (rK + 1e-30);
        acc[5] += r;
        acc[6] += exp(-0.5 * z * z) * r;
      }
    }
    __shared__ double red[7][kThreads / 64];
    __shared__ double tot[7];
  #pragma unroll
    for (int k = 0; k < 7; ++k) {
      double v = wave_sum(acc[k]);
      if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
      double v = 0.0;
      for (int w = 0; w < kThreads / 64; ++w) v += red[threadIdx.x][w];
      tot[threadIdx.x] = v;
    }
    __syncthreads();
    const double ns = tot[0];
    const double rho_inf = tot[1] / (ns + 1e-30);
    double denom = q.cs * q.cs * q.cs;
    if (vx) {
      const double dvx = bh[4] - tot[2] / (ns + 1e-30), dvy = bh[5] - tot[3] / (ns + 1e-30),
                   dvz = bh[6] - tot[4] / (ns + 1e-30);
      const double c2v2 = q.cs * q.cs + (dvx * dvx + dvy * dvy + dvz * dvz);
      denom = c2v2 * sqrt(c2v2);  // (cs^2 + v_rel^2)^1.5
    }
    const double GM = q.G * M;
    const double dMdt = 4.0 * M_PI * q.lam * (GM * GM) * rho_inf / (denom + 1e-30);
    const double dM_uncl = dMdt * q.dt;
    const double M_avail = tot[5] * (dx * dx * dx);
    const double dM_cap = fmin(dM_uncl, q.fmax < 0.0 ? M_avail : q.fmax * M_avail);
    const double sumWRho = tot[6] + 1e-30;
    const double take = dM_cap / (dx * dx * dx) / sumWRho;
    __syncthreads();
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
      long long idx;
      double R;
      at(c, idx, R);
      if (R <= r_acc) {
        const double z = R / (rK + 1e-30);
        const double r = (double)rho[idx];
        rho[idx] = (T)fmax(0.0, r - take * (exp(-0.5 * z * z) * r));
      }
    }
    if (threadIdx.x == 0) {
      out[0] = dMdt;
      out[1] = rho_inf;
      out[2] = rB;
      out[3] = dM_cap;
      out[4] = r_acc;
      out[5] = M_avail;
      out[6] = dM_uncl;
      out[7] = ns;
    }
  }
};

inline int pblocks(long long np) { return (int)((np + 255) / 256); }

// Runs stmt with VOL_ = the smallest supported LDS volume holding a tile.
#define QJ_TILE_VOL(tg, ...)                                                 \
  do {                                                                       \
    const int vol_ = 1 << ((tg).lx + (tg).ly + (tg).lz);                     \
    if (vol_ <= 512) { constexpr int VOL_ = 512; __VA_ARGS__; }              \
    else if (vol_ <= 1024) { constexpr int VOL_ = 1024; __VA_ARGS__; }       \
    else if (vol_ <= 2048) { constexpr int VOL_ = 2048; __VA_ARGS__; }       \
    else { constexpr int VOL_ = kTileMaxCells; __VA_ARGS__; }                \
  } while (0)

// QJ_TILE_VOL with the column tiles' larger volume as well (deposit kernels
// whose LDS holds one real or one 64-bit word per cell).
#define QJ_TILE_VOL_COL(tg, ...)                                             \
  do {                                                                       \
    const int vcol_ = 1 << ((tg).lx + (tg).ly + (tg).lz);                    \
    if (vcol_ > kTileMaxCells) {                                             \
      constexpr int VOL_ = kTileColMaxCells;                                 \
      __VA_ARGS__;                                                           \
    } else {                                                                 \
      QJ_TILE_VOL(tg, __VA_ARGS__);                                          \
    }                                                                        \
  } while (0)

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
hipError_t launch_interp(const Dims& d, const double* pos, long long np, const void* A3,
                         double* acc, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_interp<T>>(pblocks(np), 256, s, make_geom(d), pos, np, (const T*)A3, acc);
  return hipGetLastError();
}

template <class T>
hipError_t launch_star_step(const Dims& d, double* pos, double* vel, long long np, const void* A3,
                            double h_owed, double h1, double dt, const double* mass,
                            double m_scalar, void* rho, hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_star_step<T>>(pblocks(np), 256, s, make_geom(d), pos, vel, np, (const T*)A3, h_owed,
                                             h1, dt, mass, m_scalar, (T*)rho);
  return hipGetLastError();
}

template <class T>
hipError_t launch_star_kick(const Dims& d, const double* pos, double* vel, long long np,
                            const void* A3, double h, double* diag, hipStream_t s,
                            const void* cv) {
  if (np <= 0) return hipSuccess;
  launch<k_star_kick<T>>(pblocks(np), 256, s, make_geom(d), pos, vel, np, (const T*)A3, h, diag,
                         (const T*)cv);
  return hipGetLastError();
}

template <class T>
hipError_t launch_interp_adj(const Dims& d, const double* pos, long long np, const void* A3,
                             double h, double* vbar, double dt_xv, double* xbar, void* abar3,
                             hipStream_t s) {
  if (np <= 0) return hipSuccess;
  launch<k_interp_adj<T, false>>(pblocks(np), 256, s, make_geom(d), pos, np, (const T*)A3, h,
                                 vbar, dt_xv, xbar, (T*)abar3, TileGeom{}, (int*)nullptr,
                                 (int*)nullptr);
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

TileGeom make_tiles(const Dims& d) {
  auto lg = [](int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
  };
  TileGeom t;
  t.lx = lg(d.nx < kTileX ? d.nx : kTileX);
  t.ly = lg(d.ny < kTileY ? d.ny : kTileY);
  t.lz = lg(d.nz < kTileZ ? d.nz : kTileZ);
  // small grids: shorter z extents (down to 8) so that the per-tile deposit
  // still has about 2048 blocks (8 per CU)
  while (t.lz > 3 && ((d.nx >> t.lx) * (d.ny >> t.ly) * (d.nz >> t.lz)) < 2048) --t.lz;
  t.ntx = d.nx >> t.lx;
  t.nty = d.ny >> t.ly;
  t.ntz = d.nz >> t.lz;
  t.ntiles = t.ntx * t.nty * t.ntz;
  return t;
}

TileGeom make_column_tiles(const Dims& d, int cols, int rows) {
  auto lg = [](int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
  };
  // cols columns as R y rows by Z z columns, Z = min(cols / rows, nz); with
  // two rows a star's y pair lies in one tile for every other y, 1.5 tiles
  // per star on average where one-row tiles make 2
  if (rows < 1) rows = 1;
  int Z = cols / rows < d.nz ? cols / rows : d.nz;
  if (Z < 1) Z = 1;
  int R = cols / Z;
  if (R > d.ny) R = d.ny;
  TileGeom t;
  t.lx = lg(d.nx);
  t.ly = lg(R);
  t.lz = lg(Z);
  t.ntx = 1;
  t.nty = d.ny >> t.ly;
  t.ntz = d.nz >> t.lz;
  t.ntiles = t.nty * t.ntz;
  return t;
}

namespace {
// count and offset entries for either tile kind (column tiles have at most
// as many tiles as the smallest Poisson x workgroup covers columns: 16)
int max_tiles(const Dims& d) {
  const int a = make_tiles(d).ntiles, b = make_column_tiles(d, 16, 1).ntiles;
  return a > b ? a : b;
}
size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

size_t star_scratch_bytes(const Dims& d, long long np) {
  const int nt = max_tiles(d);
  return al256((size_t)nt * kCountStride * 4) + al256((size_t)(nt + 1) * 4) +
         2 * al256((size_t)np * 8 * 4);
}

size_t star_scratch_total_bytes(const Dims& d, long long np, int real_bytes) {
  return star_scratch_bytes(d, np) + al256((size_t)np * kCvReals * real_bytes) +
         al256((size_t)np * 8 * sizeof(unsigned long long)) + al256((size_t)np * 8 * real_bytes);
}

StarBins carve_star_scratch(const Dims& d, long long np, void* scratch, int real_bytes,
                            int cols, int rows) {
  const int nt = max_tiles(d);
  StarBins sb;
  char* c = (char*)scratch;
  const size_t cb = al256((size_t)nt * kCountStride * 4);
  sb.count = (int*)c;
  sb.offset = (int*)(c + cb);
  sb.sorted = (int*)(c + cb + al256((size_t)(nt + 1) * 4));
  sb.rank = (int*)((char*)sb.sorted + al256((size_t)np * 8 * 4));
  sb.cv = real_bytes ? (void*)(c + star_scratch_bytes(d, np)) : nullptr;
  sb.desc = (real_bytes && cols)
                ? (unsigned long long*)((char*)sb.cv + al256((size_t)np * kCvReals * real_bytes))
                : nullptr;
  sb.wts = sb.desc ? (void*)((char*)sb.desc + al256((size_t)np * 8 * sizeof(unsigned long long)))
                   : nullptr;
  sb.tg = cols ? make_column_tiles(d, cols, rows) : make_tiles(d);
  sb.ntiles = sb.tg.ntiles;
  sb.home = 0;
  return sb;
}

template <class T>
hipError_t launch_star_step_binned(const Dims& d, double* pos, double* vel, long long np,
                                   const void* A3, double h_owed, double h1, double dt, int move,
                                   const StarBins& sb, hipStream_t s, const void* cv) {
  if (np <= 0) return hipSuccess;
  const TileGeom tg = sb.tg;
  if (sb.home && (tg.ntx != 1 || !sb.desc)) return hipErrorInvalidValue;  // column tiles only
  if (move)
    launch<k_star_step_bin<T, true>>(pblocks(np), 256, s, make_geom(d), tg, pos, vel, np,
                                     (const T*)A3, h_owed, h1, dt, sb.count, sb.rank,
                                     (const T*)cv, sb.home);
  else
    launch<k_star_step_bin<T, false>>(pblocks(np), 256, s, make_geom(d), tg, pos, vel, np,
                                      (const T*)A3, h_owed, h1, dt, sb.count, sb.rank,
                                      (const T*)nullptr, sb.home);
  return hipGetLastError();
}

template <class T>
hipError_t launch_tile_deposit(const Dims& d, const double* pos, const double* mass,
                               double m_scalar, long long np, const StarBins& sb, void* rho,
                               hipStream_t s, int det) {
  const TileGeom tg = sb.tg;
  // the fused deposit reads the corner weights of the registrations; the
  // per-tile deposit kernels read lists of every tile a star touches, not
  // the one-row lists
  if (!rho && (!sb.desc || !sb.wts || tg.ntx != 1)) return hipErrorInvalidValue;
  if (sb.home && rho) return hipErrorInvalidValue;
  launch<k_bin_scan>(1, 1024, s, sb.count, sb.offset, tg.ntiles);
  if (np > 0)
    launch<k_bin_scatter<T>>(pblocks(np), 256, s, make_geom(d), tg, pos, np, (const int*)sb.rank,
                             (const int*)sb.offset, sb.sorted, sb.desc,
                             rho ? (T*)nullptr : (T*)sb.wts, mass, m_scalar, sb.home);
  if (!rho) return hipGetLastError();
  if (det)
    QJ_TILE_VOL_COL(tg, launch<k_tile_deposit_det<T, VOL_>>(tg.ntiles, 256, s, make_geom(d), tg,
                                                            pos, mass, m_scalar,
                                                            (const int*)sb.offset,
                                                            (const int*)sb.sorted, (T*)rho));
  else
    QJ_TILE_VOL_COL(tg, launch<k_tile_deposit<T, VOL_>>(tg.ntiles, 256, s, make_geom(d), tg, pos,
                                                        mass, m_scalar, (const int*)sb.offset,
                                                        (const int*)sb.sorted, (T*)rho));
  return hipGetLastError();
}

template <class T>
hipError_t launch_bh_accretion(const Dims& d, void* psi, const double* pos, const double* vel,
                               double* m_bh, long long n_bh, double G, double c, double sigma,
                               double mph, double dt, hipStream_t s) {
  launch<k_bh_accretion<T>>(1, 64, s, make_geom(d), (cpx<T>*)psi, pos, vel, m_bh, n_bh, G, c,
                            sigma, mph, dt);
  return hipGetLastError();
}

template <class T>
hipError_t launch_bondi_sink(const Dims& d, void* rho, const void* vx, const void* vy,
                             const void* vz, const double* bh, double cs, double G, double dt,
                             double r_acc_mult, double lam, double fmax_, double ann_lo,
                             double ann_hi, double* out, hipStream_t s) {
  const double dx = d.Lx / d.nx, r_acc = r_acc_mult * dx;
  const double r_lo = fmax(ann_lo * r_acc, r_acc + 0.25 * dx);
  const double r_hi = fmax(ann_hi * r_acc, r_lo + 0.5 * dx);
  // cells whose centres can lie within r_hi of the hole, per axis
  int box = 2 * (int)ceil(r_hi / dx) + 3;
  if (box > d.nx) box = d.nx;
  BondiPrm q{cs, G, dt, r_acc_mult, lam, fmax_, ann_lo, ann_hi, box};
  launch<k_bondi_sink<T>>(1, 256, s, make_geom(d), q, (T*)rho, (const T*)vx, (const T*)vy,
                          (const T*)vz, bh, out);
  return hipGetLastError();
}

template <class T>
hipError_t launch_interp_adj_binned(const Dims& d, const double* pos, long long np,
                                    const void* A3, double h, double* vbar, double dt_xv,
                                    double* xbar, void* abar3, const StarBins& sb, bool taped,
                                    hipStream_t s) {
  const TileGeom tg = sb.tg;
  if (tg.lx + tg.ly + tg.lz > 12) return hipErrorInvalidValue;  // the deposit's own tiles only
  if (np > 0)
    launch<k_interp_adj<T, true>>(pblocks(np), 256, s, make_geom(d), pos, np, (const T*)A3, h,
                                  vbar, dt_xv, xbar, (T*)nullptr, tg,
                                  taped ? (int*)nullptr : sb.count, sb.rank);
  if (!taped) {  // else offset and sorted hold the forward's registrations of pos
    launch<k_bin_scan>(1, 1024, s, sb.count, sb.offset, tg.ntiles);
    if (np > 0)
      launch<k_bin_scatter<T>>(pblocks(np), 256, s, make_geom(d), tg, pos, np,
                               (const int*)sb.rank, (const int*)sb.offset, sb.sorted,
                               (unsigned long long*)nullptr, (T*)nullptr,
                               (const double*)nullptr, 0.0, 0);
  }
  QJ_TILE_VOL(tg, launch<k_tile_deposit_vec<T, VOL_>>(tg.ntiles, 256, s, make_geom(d), tg, pos,
                                                      (const double*)vbar, h,
                                                      (const int*)sb.offset,
                                                      (const int*)sb.sorted, (T*)abar3));
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
                                          const void*, double, double*, hipStream_t, const void*);\
  template hipError_t launch_interp_adj<T>(const Dims&, const double*, long long, const void*,     \
                                           double, double*, double, double*, void*, hipStream_t); \
  template hipError_t launch_deposit_adj<T>(const Dims&, const double*, const double*, double,     \
                                            long long, const void*, double*, double*, hipStream_t); \
  template hipError_t launch_star_step_binned<T>(const Dims&, double*, double*, long long,          \
                                                 const void*, double, double, double, int,         \
                                                 const StarBins&, hipStream_t, const void*);       \
  template hipError_t launch_tile_deposit<T>(const Dims&, const double*, const double*, double,    \
                                             long long, const StarBins&, void*, hipStream_t, int);\
  template hipError_t launch_interp_adj_binned<T>(const Dims&, const double*, long long,           \
                                                  const void*, double, double*, double, double*,   \
                                                  void*, const StarBins&, bool, hipStream_t);      \
  template hipError_t launch_bondi_sink<T>(const Dims&, void*, const void*, const void*,          \
                                           const void*, const double*, double, double, double,    \
                                           double, double, double, double, double, double*,       \
                                           hipStream_t);                                          \
  template hipError_t launch_bh_accretion<T>(const Dims&, void*, const double*, const double*,    \
                                             double*, long long, double, double, double, double,  \
                                             double, hipStream_t);

QJ_PINST(float)
QJ_PINST(double)

}  // namespace qj
