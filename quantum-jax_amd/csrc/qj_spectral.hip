// Spectral pass kernels of the KDK step: kinetic drift sweeps, the Poisson
// source pass, the half-spectrum Poisson passes with fused spectral
// gradients, and the pointwise potential kick.
//
// Field layout in HBM: C-ordered (nx, ny, nz), z contiguous, complex values
// interleaved.  The Poisson solve works on a half spectrum along x,
// H[kx][y][z] with kx = 0..nx/2 (real input => Hermitian along x).
//
// Sweep plan of one spectral operator:
//   drift  : IFFT_a * diag(D_a) * FFT_a for a = z, y, x.  exp(-i dt |k|^2/(2 mph))
//            factorises over the axes, so each axis is one read+write sweep.
//   Poisson: x (rho -> half spectrum, fused with the x drift sweep) ->
//            y (FFT) -> z (FFT, -4 pi G/k^2, IFFT; dV/dz spectrum) ->
//            y (IFFT; dV/dy spectrum) -> x (Hermitian-completed IFFT to
//            V, a_x, a_y, a_z and max |a|).
#include "qj_cic.h"
#include "qj_fft.h"
#include "qj_kernels.h"

namespace qj {

namespace {

__host__ __device__ inline int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_twiddles(cpx<T>* tw, int n) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double s, c;
  sincospi(2.0 * (double)p / (double)n, &s, &c);
  tw[p] = {(T)c, (T)(-s)};
}

// D[k] = exp(-i * c * kphys^2) / n, kphys = 2 pi/L * freq(k); c = dt/(2 mph).
// quantum-stars.py:308 (one axis of the separable kinetic phase).  With
// dslot, c is a factor and dt comes from the device step slot.
template <class T>
__device__ __forceinline__ void k_drift_table_body(cpx<T>* tab, int n, double L, double c,
                                                   const double* dslot, int k) {
  if (k >= n) return;
  if (dslot) c *= dslot[kSlotDt];
  double kp = 2.0 * M_PI / L * (double)freq_index(k, n);
  double th = reduce_2pi(c * kp * kp);
  double s, co;
  sincos(th, &s, &co);
  double inv = 1.0 / (double)n;
  tab[k] = {(T)(co * inv), (T)(-s * inv)};
}

template <class T>
__global__ void k_drift_table(cpx<T>* tab, int n, double L, double c, const double* dslot) {
  k_drift_table_body<T>(tab, n, L, c, dslot, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// The three axes' tables in one launch (blockIdx.y = axis), for steps that
// take dt from the device step slot and so rebuild them every step.
template <class T>
struct DriftTables3 {
  cpx<T>* tab[3];
  int n[3];
  double L[3];
};

template <class T>
__global__ void k_drift_tables3(DriftTables3<T> t, double c, const double* dslot) {
  const int a = blockIdx.y;
  k_drift_table_body<T>(t.tab[a], t.n[a], t.L[a], c, dslot,
                        (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// ---------------------------------------------------------------------------
// pointwise kick: psi <- exp(-i coef (V + Vb)) psi  (quantum-stars.py:301)
// ---------------------------------------------------------------------------
template <class T>
struct k_kick {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ psi, const T* __restrict__ V,
                                             const T* __restrict__ Vb, const T* __restrict__ Vx,
                                             double coef, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
      double v = (double)V[i];
      if (Vb) v += (double)Vb[i];
      if (Vx) v += (double)Vx[i];
      psi[i] = cmul(psi[i], expmi<T>(coef * v));
    }
  }
};

// ---------------------------------------------------------------------------
// drift sweep along the contiguous axis, optionally preceded by the kick
// ---------------------------------------------------------------------------
// EXTRA = false drops the optional terms (b rho and the external potential in
// the kick, the conjugate tables of the drift adjoint) that the stepper's
// plain and stars steps never use: 233 -> 201 registers, 77 -> 67 us at
// 256^3 c64 in isolation.
template <class T, int N, bool KICK, bool KADJ = false, bool EXTRA = true>
struct k_drift_c {
  static constexpr int kThreads = 256;
  // A block walks pencil groups blockIdx.x, blockIdx.x + gridDim.x, ...; the
  // next group's psi (and V) rows are loaded while the current group is
  // transformed, so each wave keeps memory traffic in flight through its FFTs.
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ psi, PassGeom g,
                                             const cpx<T>* __restrict__ D,
                                             const cpx<T>* __restrict__ tw, const T* __restrict__ V,
                                             const T* __restrict__ Vb,
                                             const T* __restrict__ Vx, double kick_coef,
                                             int conj_d, const double* __restrict__ dslot,
                                             const cpx<T>* __restrict__ pref,
                                             T* __restrict__ Vbar, cpx<T>* __restrict__ pout) {
    using S = PencilShape<N>;
    __shared__ cpx<T> lds[S::kLds];
    if (KICK && dslot) kick_coef *= dslot[kSlotKick];  // kick time from the device step slot
    Lane<N, true> ln(threadIdx.x);
    const int se = (int)g.s_elem;
    const int ngroups = (g.ncols + S::C - 1) / S::C;
    cpx<T> nv[S::E];
    T nth[KICK ? S::E : 1];
    auto fetch = [&](int grp) {
      const int q = grp * S::C + ln.col;
      const bool act = q < g.ncols;
      const int base = act ? (int)pencil_base(g, q) : 0;
  #pragma unroll
      for (int m = 0; m < S::E; ++m) {
        const int idx = base + (ln.j + S::TT * m) * se;
        nv[m] = act ? psi[idx] : cpx<T>{0, 0};
        if constexpr (KICK) nth[m] = act ? V[idx] : (T)0;
      }
    };
    int grp = blockIdx.x;
    if (grp < ngroups) fetch(grp);
    cpx<T>* my = lds + ln.col * S::E * S::P;
    for (; grp < ngroups; grp += gridDim.x) {
      const int q = grp * S::C + ln.col;
      const bool act = q < g.ncols;
      const int base = act ? (int)pencil_base(g, q) : 0;
      cpx<T> v[S::E];
  #pragma unroll
      for (int m = 0; m < S::E; ++m) {
        v[m] = nv[m];
        if constexpr (KICK) {
          if (act) {
            const int idx = base + (ln.j + S::TT * m) * se;
            double th = (double)nth[m];
            if constexpr (EXTRA) {
              if (Vb) th += (double)Vb[idx];
              if (Vx) th += (double)Vx[idx];
            }
            v[m] = cmul(v[m], expmi<T>(kick_coef * th));
          }
        }
      }
      if (grp + (int)gridDim.x < ngroups) fetch(grp + gridDim.x);
      pencil_fft<N, -1, T>(v, ln.j, my, tw);
  #pragma unroll
      for (int m = 0; m < S::E; ++m) {
        const cpx<T> d0 = D[ln.j + S::TT * m];
        v[m] = cmul(v[m], (EXTRA && conj_d) ? conjg(d0) : d0);
      }
      pencil_fft<N, +1, T>(v, ln.j, my, tw);
      if constexpr (KADJ) {
        // adjoint of the next kick, psi_bar <- exp(+i c V) psi_bar, with
        // Vbar = c Im(conj(psi_bar) psi_ref) (the formula of k_kick_adj)
        if (act) {
  #pragma unroll
          for (int m = 0; m < S::E; ++m) {
            const int idx = base + (ln.j + S::TT * m) * se;
            const double vt = (double)V[idx] + (Vx ? (double)Vx[idx] : 0.0);
            const cpx<T> b = cmulc(v[m], expmi<T>(kick_coef * vt));
            v[m] = b;
            const cpx<T> r = pref[idx];
            Vbar[idx] = (T)(kick_coef * ((double)b.r * r.i - (double)b.i * r.r));
          }
        }
      }
      if (act) {
        // out of place when pout is given (the rows were read from psi)
        cpx<T>* dst = pout ? pout : psi;
  #pragma unroll
        for (int m = 0; m < S::E; ++m) dst[base + (ln.j + S::TT * m) * se] = v[m];
      }
    }
  }
};

// Common set-up of a strided sweep: VW neighbouring pencils per lane,
// pencil slot fastest across lanes.
#define QJ_STRIDED_SETUP(T, N)                                           \
  using S = PencilShape<N>;                                              \
  constexpr int VW = StridedVW<T, N>::value;                             \
  constexpr int E = S::E;                                                \
  constexpr int TT = S::TT;                                              \
  __shared__ cpx<T> lds[S::kLdsStrided * VW];                            \
  const int col = threadIdx.x % S::C, j = threadIdx.x / S::C;            \
  const int q0 = (blockIdx.x * S::C + col) * VW;                         \
  const bool act = q0 < g.ncols;                                         \
  const int base = act ? (int)pencil_base(g, q0) : 0;                    \
  const int se = (int)g.s_elem;                                          \
  (void)E;

template <int VW, int E, class T>
__device__ __forceinline__ void load_rows(cpx<T> (&v)[VW][E], const cpx<T>* __restrict__ p,
                                          int base, int se, int j, int TT, bool act) {
#pragma unroll
  for (int m = 0; m < E; ++m) {
    cpx<T> t[VW];
    if (act) {
      ldc<VW>(p, base + (j + TT * m) * se, t);
    } else {
#pragma unroll
      for (int w = 0; w < VW; ++w) t[w] = {0, 0};
    }
#pragma unroll
    for (int w = 0; w < VW; ++w) v[w][m] = t[w];
  }
}

template <int VW, int E, class T>
__device__ __forceinline__ void store_rows(const cpx<T> (&v)[VW][E], cpx<T>* __restrict__ p,
                                           int base, int se, int j, int TT, bool act) {
  if (!act) return;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    cpx<T> t[VW];
#pragma unroll
    for (int w = 0; w < VW; ++w) t[w] = v[w][m];
    stc<VW>(p, base + (j + TT * m) * se, t);
  }
}

// ---------------------------------------------------------------------------
// In-place strided sweep of one pencil group per block: op(v, j, col, base,
// act) transforms the group held in registers.  The kernels that use it
// exchange through one LDS buffer that a lane's pencils take turns with (the
// transform's phased form), so their LDS (34.8 KB at 256 points) and
// registers allow four waves per SIMD; walking several groups per block
// with the next group's rows prefetched, at two waves per SIMD, was slower
// (drift_y 56 us against 50 at 256^3 c64).
// ---------------------------------------------------------------------------
template <class T, int N, class Op>
__device__ __forceinline__ void strided_sweep(cpx<T>* __restrict__ a, const PassGeom& g, Op&& op) {
  using S = PencilShape<N>;
  constexpr int VW = StridedVW<T, N>::value;
  constexpr int E = S::E;
  constexpr int TT = S::TT;
  const int col = threadIdx.x % S::C, j = threadIdx.x / S::C;
  const int se = (int)g.s_elem;
  const int q0 = (blockIdx.x * S::C + col) * VW;
  const bool act = q0 < g.ncols;
  const int base = act ? (int)pencil_base(g, q0) : 0;
  cpx<T> v[VW][E];
  load_rows<VW, E>(v, a, base, se, j, TT, act);
  op(v, j, col, base, act);
  store_rows<VW, E>(v, a, base, se, j, TT, act);
}

// drift sweep along a strided axis (x or y); with RHO the density adjoint
// psi_bar += 2 rhobar psi_ref (the formula of k_rho_adj) is applied to the
// rows as they arrive
template <class T, int N, bool RHO = false>
struct k_drift_s {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ psi, PassGeom g,
                                             const cpx<T>* __restrict__ D,
                                             const cpx<T>* __restrict__ tw, int conj_d,
                                             const T* __restrict__ rb,
                                             const cpx<T>* __restrict__ pref) {
    using S = PencilShape<N>;
    constexpr int VW = StridedVW<T, N>::value;
    __shared__ cpx<T> lds[S::kLdsStrided];
    const int se = (int)g.s_elem;
    strided_sweep<T, N>(psi, g, [&](cpx<T>(&v)[VW][S::E], int j, int col, int base, bool act) {
      if constexpr (RHO) {
        if (act) {
  #pragma unroll
          for (int m = 0; m < S::E; ++m) {
            const int idx = base + (j + S::TT * m) * se;
  #pragma unroll
            for (int w = 0; w < VW; ++w) {
              const T f = (T)2 * rb[idx + w];
              const cpx<T> r = pref[idx + w];
              v[w][m].r += f * r.r;
              v[w][m].i += f * r.i;
            }
          }
        }
      }
      pencil_fft_strided<N, -1, T, VW, true>(v, j, col, lds, tw);
  #pragma unroll
      for (int m = 0; m < S::E; ++m) {
        const cpx<T> d0 = D[j + S::TT * m];
        const cpx<T> d = conj_d ? conjg(d0) : d0;
  #pragma unroll
        for (int w = 0; w < VW; ++w) v[w][m] = cmul(v[w][m], d);
      }
      pencil_fft_strided<N, +1, T, VW, true>(v, j, col, lds, tw);
    });
  }
};

// ---------------------------------------------------------------------------
// FP64 source mass: each block of the source pass stores its partial sum; the
// Poisson z pass (the next kernel in stream order, which needs the k = 0
// bin) adds the partials with a compensated sum in its first block.  The
// kernel boundary makes the partials visible, so no grid-wide fence is
// needed (an agent-scope fence per block writes back the XCD's L2).
// ---------------------------------------------------------------------------
__device__ void block_partial(double local, bool bad, double* partials, double* diag) {
  __shared__ double red[4];
  __shared__ int flag[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double s = wave_sum(local);
  int anybad = __any(bad);
  if (lane == 0) {
    red[wv] = s;
    flag[wv] = anybad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
    if (flag[0] | flag[1] | flag[2] | flag[3])
      atomicExch(reinterpret_cast<unsigned long long*>(&diag[kDiagNaN]),
                 (unsigned long long)__double_as_longlong(1.0));
  }
}

__device__ __forceinline__ void two_sum_acc(double& s, double& c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
  s = t;
}

// Compensated (Neumaier per thread, error-free merges across lanes and
// waves) sum of x[0..n) by one 256-thread block, in a fixed order; the
// result is returned to every thread.
__device__ double block_compensated_sum(const double* __restrict__ x, int n) {
  __shared__ double ws[4], wc[4];
  __shared__ double total;
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) two_sum_acc(s, c, x[i]);
  for (int off = 32; off > 0; off >>= 1) {
    const double so = __shfl_down(s, off), co = __shfl_down(c, off);
    two_sum_acc(s, c, so);
    c += co;
  }
  if ((threadIdx.x & 63) == 0) {
    ws[threadIdx.x >> 6] = s;
    wc[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = ws[0], Cc = wc[0];
    for (int w = 1; w < 4; ++w) {
      two_sum_acc(S, Cc, ws[w]);
      Cc += wc[w];
    }
    total = S + Cc;
  }
  __syncthreads();
  return total;
}

// ---------------------------------------------------------------------------
// Poisson source pass along x (optionally fused with the x drift sweep):
//   rho = |psi|^2 + src; FP64 mass; forward x-FFT of (rho - rho_bar);
//   store kx = 0..nx/2 into H.
// DEP (stars steps with the fused deposit): the workgroup's columns are a
// column tile of the star registrations (the layout of the Poisson x
// gather); the pass adds the CIC corner weights of the tile's stars
// (quantum-stars.py:226-247, m w / vol, which the scatter wrote per star)
// at the corners inside the tile into LDS, adds the tile to |psi|^2 and
// writes it to rho_s, so the star density is not written by one kernel and
// read back by the next, and the deposit needs no launch.
// ---------------------------------------------------------------------------
// The registration lists a column tile's workgroup serves: its own, and with
// the one-row registration (each star registered only in the tiles of its
// lower y row) the list of the tile row below, whose stars' upper y corners
// lie in this tile.  Entry v of the two is slot slot(v).
struct TileLists {
  int a0, n0, a1, n;
  __device__ __forceinline__ int slot(int v) const { return v < n0 ? a0 + v : a1 + (v - n0); }
};
__device__ __forceinline__ TileLists tile_lists(const int* __restrict__ offset, int home,
                                                int lnz, int lz) {
  TileLists L;
  L.a0 = offset[blockIdx.x];
  L.n0 = offset[blockIdx.x + 1] - L.a0;
  L.a1 = 0;
  L.n = L.n0;
  if (home) {
    const int nbz = 1 << (lnz - lz);
    const int nty = (int)gridDim.x / nbz;  // one workgroup per tile
    if (nty > 1) {
      const int tb = (((int)blockIdx.x / nbz + nty - 1) % nty) * nbz + (int)blockIdx.x % nbz;
      L.a1 = offset[tb];
      L.n += offset[tb + 1] - L.a1;
    }
  }
  return L;
}

// Column + 1 of the star's (y, z) corner pair yz = 2 sy + sz in this
// workgroup's tile, or 0 outside it: from a tile-relative descriptor
// (corner_desc) or, home, an absolute one (home_desc) tested against the
// tile at rows y0.. (R of them) and columns z0.. (2^lz).
__device__ __forceinline__ int corner_column(unsigned long long d, int yz, int home, int y0,
                                             int z0, int R, int lz) {
  if (!home) return (int)((d >> (18 + 9 * yz)) & 511u);
  const int dy = (int)((d >> (18 + 9 * (yz >> 1))) & 511u) - y0;
  const int dz = (int)((d >> (36 + 9 * (yz & 1))) & 511u) - z0;
  return ((unsigned)dy < (unsigned)R && (unsigned)dz < (1u << lz)) ? ((dy << lz) | dz) + 1 : 0;
}

template <class T>
struct DepositArgs {
  const int* offset;
  const int* sorted;
  const unsigned long long* desc;  // corner descriptors of the registrations
  const T* wts;                    // 8 corner weights per star
  T* rho;                          // rho_s, the tile's star density (or null)
  int lz;                          // log2 of the tile's z columns
  int lnz;                         // log2 nz
  int home;                        // one-row registration (TileLists)
};

// the eight corner weights of star i
template <class T>
__device__ __forceinline__ void load_wts8(const T* __restrict__ w, long long i, T (&o)[8]) {
  if constexpr (sizeof(T) == 4) {
    const float4* q = reinterpret_cast<const float4*>(w + 8 * i);
    const float4 a = q[0], b = q[1];
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
  } else {
    const double2* q = reinterpret_cast<const double2*>(w + 8 * i);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double2 a = q[e];
      o[2 * e] = a.x;
      o[2 * e + 1] = a.y;
    }
  }
}

template <class T, int N, bool DRIFT, bool DEP = false>
struct k_source_x {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ psi, T* __restrict__ src,
                                             int clear_src, T* __restrict__ brho, double b_coeff,
                                             double rho_bar, PassGeom g,
                                             const cpx<T>* __restrict__ D,
                                             const cpx<T>* __restrict__ tw, cpx<T>* __restrict__ H,
                                             double* __restrict__ partials,
                                             double* __restrict__ diag,
                                             cpx<T>* __restrict__ pout, DepositArgs<T> dp) {
    // one exchange buffer that the lane's pencils take turns with (half the
    // LDS of QJ_STRIDED_SETUP): the pass holds 162 registers, so the LDS
    // decides between two and three waves per SIMD
    using S = PencilShape<N>;
    constexpr int VW = StridedVW<T, N>::value;
    constexpr int E = S::E;
    constexpr int TT = S::TT;
    constexpr int CV = S::C * VW;  // columns of the workgroup
    __shared__ cpx<T> lds[S::kLdsStrided];
    const int col = threadIdx.x % S::C, j = threadIdx.x / S::C;
    int q0 = (blockIdx.x * S::C + col) * VW;
    bool act = q0 < g.ncols;
    // the fused deposit: the thread's first registration of the tile is
    // loaded before psi and its star's weights after psi's rows, so both
    // latencies hide behind the drift transforms
    TileLists dl{};
    int dp0 = 0;
    unsigned long long dd = 0;
    T dw[8];
    if constexpr (DEP) {
      // column tile t = blockIdx.x at (t / (nz >> lz), t % (nz >> lz)) in
      // tile units, its CV columns as rows of 2^lz z columns (a lane's VW
      // columns adjacent in z)
      const int nbz = 1 << (dp.lnz - dp.lz);
      const int c0 = col * VW;
      const int y = (blockIdx.x / nbz) * (CV >> dp.lz) + (c0 >> dp.lz);
      const int z = ((blockIdx.x % nbz) << dp.lz) + (c0 & ((1 << dp.lz) - 1));
      act = (y << dp.lnz) < g.ncols;
      q0 = (y << dp.lnz) + z;
      dl = tile_lists(dp.offset, dp.home, dp.lnz, dp.lz);
      // loaded by every thread (the lists' first slot when the thread has no
      // entry, slot 0 when they are empty): no branch around the loads.  An
      // empty list's slot 0 may be stale, so its star is taken to be star 0.
      const int v0 = threadIdx.x;
      const int ic = v0 < dl.n ? dl.slot(v0) : (dl.n > 0 ? dl.slot(0) : 0);
      const int p0 = dp.sorted[ic];
      dp0 = dl.n > 0 ? p0 : 0;
      dd = dp.desc[ic];
    }
    const int base = act ? (int)pencil_base(g, q0) : 0;
    const int se = (int)g.s_elem;
    cpx<T> v[VW][E];
    T rho[VW][E];
  #pragma unroll
    for (int m = 0; m < E; ++m)
  #pragma unroll
      for (int w = 0; w < VW; ++w) rho[w][m] = (T)0;
    // the star density rows are loaded with psi's, so their latency is hidden
    // behind the drift transforms (LDS, not registers, limits the blocks per
    // CU): 120 -> 88 us at 256^3 c64
    T sv[VW][E];
    if (!DEP && src && act) {
  #pragma unroll
      for (int m = 0; m < E; ++m) {
        T t[VW];
        ldr<VW>(src, base + (j + TT * m) * se, t);
  #pragma unroll
        for (int w = 0; w < VW; ++w) sv[w][m] = t[w];
      }
    }
    if (psi) {
      load_rows<VW, E>(v, psi, base, se, j, TT, act);
      if constexpr (DEP) load_wts8(dp.wts, dp0, dw);
      if constexpr (DRIFT) {
        pencil_fft_strided<N, -1, T, VW, true>(v, j, col, lds, tw);
  #pragma unroll
        for (int m = 0; m < E; ++m) {
          const cpx<T> d = D[j + TT * m];
  #pragma unroll
          for (int w = 0; w < VW; ++w) v[w][m] = cmul(v[w][m], d);
        }
        pencil_fft_strided<N, +1, T, VW, true>(v, j, col, lds, tw);
        // out of place when pout is given (the rows were read from psi)
        store_rows<VW, E>(v, pout ? pout : psi, base, se, j, TT, act);
      }
  #pragma unroll
      for (int m = 0; m < E; ++m)
  #pragma unroll
        for (int w = 0; w < VW; ++w)
          rho[w][m] = fmad(v[w][m].r, v[w][m].r, v[w][m].i * v[w][m].i);
    }
    if constexpr (DEP) {
      // The exchange buffer is free between the drift and the source
      // transform: the tile's density is summed there (cell (x, column c) at
      // x CV + c, float atomics as in k_tile_deposit), added to the thread's
      // cells and written to rho_s.  Short pencils (one thread per pencil)
      // have no exchange buffer and take one of their own.
      constexpr bool kOwn = N * CV * sizeof(T) > sizeof(lds);
      __shared__ T own_tile[kOwn ? N * CV : 1];
      T* acc = kOwn ? own_tile : reinterpret_cast<T*>(lds);
      __syncthreads();  // the drift's last exchange reads are done
      for (int l = threadIdx.x; l < N * CV; l += kThreads) acc[l] = (T)0;
      __syncthreads();
      const int v0 = threadIdx.x;
      const int nbzt = 1 << (dp.lnz - dp.lz);
      const int Rt = CV >> dp.lz;
      const int y0t = ((int)blockIdx.x / nbzt) * Rt, z0t = ((int)blockIdx.x % nbzt) << dp.lz;
      for (int v = v0; v < dl.n; v += kThreads) {
        unsigned long long d = dd;
        T w[8];
  #pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = dw[e];
        if (v != v0) {
          const int i = dl.slot(v);
          d = dp.desc[i];
          load_wts8(dp.wts, dp.sorted[i], w);
        }
        // the descriptor: x of the two x corners (9 bits each), then per
        // (y, z) corner pair yz = 2 sy + sz its tile column + 1, or 0 outside
        const int x0 = (int)(d & 511u) * CV, x1 = (int)((d >> 9) & 511u) * CV;
  #pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
          const int f = corner_column(d, yz, dp.home, y0t, z0t, Rt, dp.lz);
          if (f != 0) {
            atomicAdd(&acc[x0 + f - 1], w[2 * yz]);
            atomicAdd(&acc[x1 + f - 1], w[2 * yz + 1]);
          }
        }
      }
      __syncthreads();
      if (act) {
  #pragma unroll
        for (int m = 0; m < E; ++m) {
          const int xr = j + TT * m;
          T t[VW];
  #pragma unroll
          for (int w = 0; w < VW; ++w) {
            t[w] = acc[xr * CV + col * VW + w];
            rho[w][m] += t[w];
          }
          if (dp.rho) str<VW>(dp.rho, base + xr * se, t);
        }
      }
      // the source transform's first exchange waits at a barrier before it
      // writes the buffer
    }
    double local = 0.0;
    bool bad = false;
  #pragma unroll
    for (int m = 0; m < E; ++m) {
      const int idx = base + (j + TT * m) * se;
      if (act) {
        if (!DEP && src) {
  #pragma unroll
          for (int w = 0; w < VW; ++w) rho[w][m] += sv[w][m];
          if (clear_src) {
            T z[VW];
  #pragma unroll
            for (int w = 0; w < VW; ++w) z[w] = (T)0;
            str<VW>(src, idx, z);
          }
        }
        if (brho) {
          T t[VW];
  #pragma unroll
          for (int w = 0; w < VW; ++w) t[w] = (T)(b_coeff * (double)rho[w][m]);
          str<VW>(brho, idx, t);
        }
  #pragma unroll
        for (int w = 0; w < VW; ++w) {
          local += (double)rho[w][m];
          bad |= !isfinite(rho[w][m]);
        }
      }
  #pragma unroll
      for (int w = 0; w < VW; ++w) v[w][m] = {act ? (T)((double)rho[w][m] - rho_bar) : (T)0, (T)0};
    }
    if constexpr (VW == 2) {
      // the lane's two real pencils as one complex transform Z of
      // z = r0 + i r1, split with the mirrored bin (one LDS exchange):
      // R0[k] = (Z[k] + conj Z[N-k]) / 2, R1[k] = (Z[k] - conj Z[N-k]) / 2i
      cpx<T> z[1][E];
  #pragma unroll
      for (int m = 0; m < E; ++m) z[0][m] = {v[0][m].r, v[1][m].r};
      pencil_fft_strided<N, -1, T, 1>(z, j, col, lds, tw);
      __syncthreads();  // the transform's last exchange is read out
  #pragma unroll
      for (int m = 0; m < E; ++m) lds[(j + TT * m) * S::C + col] = z[0][m];
      __syncthreads();
      if (act) {
  #pragma unroll
        for (int m = 0; m < E; ++m) {
          const int k = j + TT * m;
          if (k <= N / 2) {
            const cpx<T> a = z[0][m], b = lds[((N - k) & (N - 1)) * S::C + col];
            // b is read unconjugated: conj(b) enters through the signs
            cpx<T> t[2];
            t[0] = {(T)0.5 * (a.r + b.r), (T)0.5 * (a.i - b.i)};
            t[1] = {(T)0.5 * (a.i + b.i), (T)0.5 * (b.r - a.r)};
            stc<2>(H, base + k * se, t);
          }
        }
      }
    } else {
      pencil_fft_strided<N, -1, T, VW, true>(v, j, col, lds, tw);
      if (act) {
  #pragma unroll
        for (int m = 0; m < E; ++m) {
          const int k = j + TT * m;
          if (k <= N / 2) {
            cpx<T> t[VW];
  #pragma unroll
            for (int w = 0; w < VW; ++w) t[w] = v[w][m];
            stc<VW>(H, base + k * se, t);
          }
        }
      }
    }
    block_partial(local, bad, partials, diag);
  }
};

// ---------------------------------------------------------------------------
// generic in-place complex FFT sweep along a strided axis
// ---------------------------------------------------------------------------
template <class T, int N, int DIR>
struct k_fft_s {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ a, PassGeom g,
                                             const cpx<T>* __restrict__ tw) {
    using S = PencilShape<N>;
    constexpr int VW = StridedVW<T, N>::value;
    __shared__ cpx<T> lds[S::kLdsStrided];
    strided_sweep<T, N>(a, g, [&](cpx<T>(&v)[VW][S::E], int j, int col, int, bool) {
      pencil_fft_strided<N, DIR, T, VW, true>(v, j, col, lds, tw);
    });
  }
};

// ---------------------------------------------------------------------------
// Poisson z pass on the half spectrum: FFT_z, V_hat = -4piG rho_hat/k^2
// (quantum-stars.py:191, k = 0 divided by 1 with the FP64 source sum),
// IFFT_z of V_hat and, with GRAD, of i kz V_hat (Nyquist derivative 0).
// lp_kc2 > 0: V_hat is multiplied by the low-pass factor
// exp(-(k^2 / lp_kc2)^4) of quantum-fluid-accretion.py:143-145 (lp_kc2 =
// k_cut^2 + 1e-30; the k = 0 bin keeps factor 1).
// ---------------------------------------------------------------------------
template <class T, int N, bool GRAD, bool LP = false>
struct k_poisson_z {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ H, cpx<T>* __restrict__ Uz,
                                             PassGeom g, const cpx<T>* __restrict__ tw, int ny_log2,
                                             int nx, int ny, double kfx, double kfy, double kfz,
                                             double four_pi_G, double rho_bar, double ntot,
                                             int dc_mode, const double* __restrict__ partials,
                                             int nparts, double* __restrict__ diag,
                                             double lp_kc2) {
    using S = PencilShape<N>;
    __shared__ cpx<T> lds[S::kLds];
    // block 0 holds the kx = ky = 0 pencil: FP64 source mass for the k = 0 bin
    double mass = 0.0;
    if (blockIdx.x == 0) {
      mass = block_compensated_sum(partials, nparts);
      if (threadIdx.x == 0) diag[kDiagMass] = mass;
    }
    Lane<N, true> ln(threadIdx.x);
    const int q = blockIdx.x * S::C + ln.col;
    const bool act = q < g.ncols;
    const int base = act ? (int)pencil_base(g, q) : 0;
    cpx<T>* my = lds + ln.col * S::E * S::P;
    cpx<T> v[S::E];
  #pragma unroll
    for (int m = 0; m < S::E; ++m) v[m] = act ? H[base + ln.j + S::TT * m] : cpx<T>{0, 0};
    pencil_fft<N, -1, T>(v, ln.j, my, tw);
    const int x = q >> ny_log2, y = q & (ny - 1);
    const double kx = kfx * (double)freq_index(x, nx);
    const double ky = kfy * (double)freq_index(y, ny);
    const double kxy2 = kx * kx + ky * ky;
    const double inv_ntot = 1.0 / ntot;
    cpx<T> w[GRAD ? S::E : 1];
  #pragma unroll
    for (int m = 0; m < S::E; ++m) {
      const int kzi = ln.j + S::TT * m;
      const double kz = kfz * (double)freq_index(kzi, N);
      const double k2 = kxy2 + kz * kz;
      double gk = -four_pi_G * inv_ntot / (k2 == 0.0 ? 1.0 : k2);
      if constexpr (LP) {  // a compile-time form: the exp costs the plain pass 80 VGPRs
        double r = k2 / lp_kc2;
        r *= r;
        gk *= exp(-(r * r));
      }
      const T Gk = (T)gk;
      v[m] = scale(v[m], Gk);
      if (x == 0 && y == 0 && kzi == 0) {
        double dc = (dc_mode == 0) ? -four_pi_G * (mass - ntot * rho_bar) * inv_ntot : 0.0;
        v[m] = {(T)dc, (T)0};
      }
      if constexpr (GRAD) {
        const T kzt = (kzi == N / 2) ? (T)0 : (T)kz;
        w[m] = {-kzt * v[m].i, kzt * v[m].r};
      }
    }
    pencil_fft<N, +1, T>(v, ln.j, my, tw);
    if (act) {
  #pragma unroll
      for (int m = 0; m < S::E; ++m) H[base + ln.j + S::TT * m] = v[m];
    }
    if constexpr (GRAD) {
      pencil_fft<N, +1, T>(w, ln.j, my, tw);
      if (act) {
  #pragma unroll
        for (int m = 0; m < S::E; ++m) Uz[base + ln.j + S::TT * m] = w[m];
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Poisson y pass (inverse): U -> IFFT_y U (in place); with GRAD also
// IFFT_y(i ky U) -> Uy and IFFT_y Uz (in place).
// ---------------------------------------------------------------------------
template <class T, int N, bool GRAD, bool WIDE = true>
struct k_poisson_y {
  // the wide form (one pencil per lane) for the gradient pass, which holds
  // two spectra per pencil, in 256-thread workgroups (16 pencils): 63.8 ->
  // 59.4 us at 256^3 against 512 threads (the Poisson x pass keeps 512)
  using Cfg = StridedCfg<T, N, GRAD && WIDE, 256>;
  static constexpr int kThreads = Cfg::TH;
  // four 256-thread blocks per CU need four waves per SIMD (128 VGPRs)
  static constexpr int kMinWaves = Cfg::kWide ? 4 : 0;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ H, cpx<T>* __restrict__ Uz,
                                             cpx<T>* __restrict__ Uy, PassGeom g,
                                             const cpx<T>* __restrict__ tw, double kfy) {
    using S = PencilShape<N, Cfg::TH>;
    constexpr int VW = Cfg::VW;
    constexpr int E = S::E;
    constexpr int TT = S::TT;
    constexpr int TH = Cfg::TH;
    __shared__ cpx<T> lds[S::kLdsStrided * VW];
    const int col = threadIdx.x % S::C, j = threadIdx.x / S::C;
    const int q0 = (blockIdx.x * S::C + col) * VW;
    const bool act = q0 < g.ncols;
    const int base = act ? (int)pencil_base(g, q0) : 0;
    const int se = (int)g.s_elem;
    cpx<T> v[VW][E];
    load_rows<VW, E>(v, H, base, se, j, TT, act);
    if constexpr (GRAD) {
      cpx<T> w[VW][E];
  #pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = j + TT * m;
        const T ky = (k == N / 2) ? (T)0 : (T)(kfy * (double)freq_index(k, N));
  #pragma unroll
        for (int x = 0; x < VW; ++x) w[x][m] = {-ky * v[x][m].i, ky * v[x][m].r};
      }
      pencil_fft_strided<N, +1, T, VW, false, TH>(v, j, col, lds, tw);
      store_rows<VW, E>(v, H, base, se, j, TT, act);
      pencil_fft_strided<N, +1, T, VW, false, TH>(w, j, col, lds, tw);
      store_rows<VW, E>(w, Uy, base, se, j, TT, act);
      load_rows<VW, E>(v, Uz, base, se, j, TT, act);
      pencil_fft_strided<N, +1, T, VW, false, TH>(v, j, col, lds, tw);
      store_rows<VW, E>(v, Uz, base, se, j, TT, act);
    } else {
      pencil_fft_strided<N, +1, T, VW, false, TH>(v, j, col, lds, tw);
      store_rows<VW, E>(v, H, base, se, j, TT, act);
    }
  }
};

// Rows of the full x spectrum from the half spectrum (Hermitian along x).
template <int N, int VW, int E, class T>
__device__ __forceinline__ void load_herm_rows(cpx<T> (&v)[VW][E], const cpx<T>* __restrict__ a,
                                               int base, int se, int j, int TT, bool act) {
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = j + TT * m;
    cpx<T> t[VW];
    if (act) {
      ldc<VW>(a, base + (k <= N / 2 ? k : N - k) * se, t);
      if (k > N / 2) {
#pragma unroll
        for (int w = 0; w < VW; ++w) t[w].i = -t[w].i;
      }
    } else {
#pragma unroll
      for (int w = 0; w < VW; ++w) t[w] = {0, 0};
    }
#pragma unroll
    for (int w = 0; w < VW; ++w) v[w][m] = t[w];
  }
}

// Wide form (sixteen lanes per 256-point pencil, one pencil per lane): the
// lanes of a wave hold rows j of the stride-16 decomposition in Hermitian
// partner pairs, j and 16 - j (rows 0 and 8 are their own partners), each
// over the workgroup's columns.  Element m of row j is bin k = j + 16 m,
// whose mirror N - k is element 15 - m of row 16 - j (row 0: element 16 - m
// of itself), so the rows load only the bins k <= N/2 and take the others
// from the partner row.  512 threads, 32 columns: a wave holds one pair,
// row w in lanes 0-31 and its partner in lanes 32-63 (wave 0: rows 0 and 8);
// 256 threads, 16 columns: a wave holds two pairs (a, 16 - a) with
// a = 2 w + lane / 32, partner rows 16 lanes apart (a = 0: rows 0 and 8).
template <class S>
__device__ __forceinline__ void paired_row(int& j, int& col) {
  static_assert(S::TT == 16 && ((S::kThreads == 512 && S::C == 32) ||
                                (S::kThreads == 256 && S::C == 16)),
                "wide 256-point layout only");
  const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  if constexpr (S::C == 32) {
    col = lane & 31;
    j = (lane < 32) ? w : (w == 0 ? 8 : 16 - w);
  } else {
    col = lane & 15;
    const int a = 2 * w + (lane >> 5);
    j = (lane & 16) ? (a == 0 ? 8 : 16 - a) : a;
  }
}

template <int N, int E, int C>
__device__ __forceinline__ void load_herm_paired(cpx<float> (&v)[1][E],
                                                 const cpx<float>* __restrict__ a, int base,
                                                 int se, int j, bool act) {
  static_assert(N == 256 && E == 16, "the paired rows need 256-point pencils of 16 lanes");
  constexpr int PD = C;  // lane distance to the partner row (32 or 16)
  // own bins: m = 0..7 (k = j + 16 m <= 127), and k = 128 for row 0
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    cpx<float> t = {0.f, 0.f};
    if (act) t = a[base + (j + 16 * m) * se];
    v[0][m] = t;
  }
  {
    cpx<float> t = {0.f, 0.f};
    if (act && j == 0) t = a[base + (N / 2) * se];
    v[0][8] = t;
  }
  // mirrored bins: the conjugate of the partner's (or the row's own) bin
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const cpx<float> own = v[0][e];
    const cpx<float> p = {__shfl_xor(own.r, PD, 64), __shfl_xor(own.i, PD, 64)};
    if (j == 8)
      v[0][15 - e] = {own.r, -own.i};
    else if (j != 0)
      v[0][15 - e] = {p.r, -p.i};
    else if (e >= 1)
      v[0][16 - e] = {own.r, -own.i};
  }
}

// ---------------------------------------------------------------------------
// Poisson x pass (inverse, Hermitian-completed): V = IFFT_x(UU); with GRAD
// a_x + i a_y = -IFFT_x(i kx UU + i Uy), a_z = -IFFT_x(Uzz) (V and a_z from
// one transform, V + i a_z = IFFT_x(UU - i Uzz)), written as the
// cell-interleaved grid A3 = (a_x, a_y, a_z), and max |a| folded into
// the running maximum diag[kDiagAMax] with one atomic per block.
// quantum-stars.py:192 and :256-260.
// SIG (gas steps of the CFL stepper): the pass also loads the gas velocity
// of its cells and writes the block maximum of the CFL signal speed
// cs + |v + h a| (quantum-fluid-accretion.py:696-699) to sg.part[block], so
// the next step's time step needs no pass over v and the grid of its own.
// ---------------------------------------------------------------------------
template <class T>
struct SigArgs {
  const T* vx;
  const T* vy;
  const T* vz;
  const double* owed;  // h, the kick owed after this step
  double cs;
  double* part;
};

// CVG (stars steps): instead of writing A3, the pass hands each star
// registered in its column tile the accelerations at its CIC corners (the
// StarGather contract): a star needs 8 of the grid's cells, 1 in 64 at one
// star per 64 cells, and the star step then reads 96 contiguous bytes per
// star instead of gathering 128-byte lines of A3 at random.
template <class T>
struct GatherArgs {
  const int* offset;
  const int* sorted;
  const unsigned long long* desc;
  T* cv;
  int lz;    // log2 of the z columns of a column tile (its y rows: C VW / 2^lz)
  int lnz;   // log2 nz
  int home;  // one-row registration: own and lower-row lists, absolute descriptors
};


template <class T, int N, bool GRAD, bool WIDE = true, bool SIG = false, bool CVG = false>
struct k_poisson_x {
  using Cfg = StridedCfg<T, N, WIDE>;
  static constexpr int kThreads = Cfg::TH;
  // two 512-thread workgroups per CU (LDS) need four waves per SIMD
  static constexpr int kMinWaves = (Cfg::kWide && GRAD) ? 4 : 0;
  static_assert(!(SIG && CVG) && (GRAD || !CVG), "the gather serves the stars gradient pass");
  static __device__ __forceinline__ void run(const cpx<T>* __restrict__ H,
                                             const cpx<T>* __restrict__ Uy,
                                             const cpx<T>* __restrict__ Uzz, PassGeom g,
                                             const cpx<T>* __restrict__ tw, double kfx,
                                             T* __restrict__ V, T* __restrict__ A3,
                                             const T* __restrict__ addend,
                                             double* __restrict__ diag, SigArgs<T> sg,
                                             GatherArgs<T> ga) {
    using S = PencilShape<N, Cfg::TH>;
    constexpr int VW = Cfg::VW;
    constexpr int E = S::E;
    constexpr int TT = S::TT;
    __shared__ cpx<T> lds[S::kLdsStrided * VW];
    int col = threadIdx.x % S::C, j = threadIdx.x / S::C;
    if constexpr (Cfg::kWide) paired_row<S>(j, col);
    int q0 = (blockIdx.x * S::C + col) * VW;
    bool act = q0 < g.ncols;
    if constexpr (CVG) {
      // the workgroup's columns form its column tile: C VW columns as rows
      // of 2^lz z columns (a lane's VW columns adjacent in z), tile t at
      // (t / (nz >> lz), t % (nz >> lz)) in tile units
      const int nbz = 1 << (ga.lnz - ga.lz);
      const int c0 = col * VW;
      const int y = (blockIdx.x / nbz) * ((S::C * VW) >> ga.lz) + (c0 >> ga.lz);
      const int z = ((blockIdx.x % nbz) << ga.lz) + (c0 & ((1 << ga.lz) - 1));
      act = (y << ga.lnz) < g.ncols;
      q0 = (y << ga.lnz) + z;
    }
    const int base = act ? (int)pencil_base(g, q0) : 0;
    const int se = (int)g.s_elem;
    auto load_herm = [&](cpx<T>(&x)[VW][E], const cpx<T>* __restrict__ a) {
      if constexpr (Cfg::kWide)
        load_herm_paired<N, E, S::C>(x, a, base, se, j, act);
      else
        load_herm_rows<N, VW, E>(x, a, base, se, j, TT, act);
    };
    // the fused star gather: the thread's first registration of the tile is
    // loaded before the spectra, so its latency hides behind the transforms
    TileLists gl{};
    int gp0 = 0;
    unsigned long long gd0 = 0;
    if constexpr (CVG) {
      gl = tile_lists(ga.offset, ga.home, ga.lnz, ga.lz);
      if ((int)threadIdx.x < gl.n) {
        const int i0 = gl.slot(threadIdx.x);
        gp0 = ga.sorted[i0];
        gd0 = ga.desc[i0];
      }
    }
    cpx<T> v[VW][E];
    load_herm(v, H);
    if constexpr (!GRAD) {
      pencil_fft_strided<N, +1, T, VW, false, Cfg::TH>(v, j, col, lds, tw);
      if (act) {
  #pragma unroll
        for (int m = 0; m < E; ++m) {
          const int idx = base + (j + TT * m) * se;
          T t[VW];
          if (addend) ldr<VW>(addend, idx, t);
  #pragma unroll
          for (int w = 0; w < VW; ++w) t[w] = addend ? v[w][m].r + t[w] : v[w][m].r;
          str<VW>(V, idx, t);
        }
      }
    } else {
      // Two x transforms for four real outputs, all three spectra loaded up
      // front: u = i kx UU + i Uy transforms to -(a_x + i a_y), and
      // UU - i Uzz to V + i a_z (V and a_z are real: their spectra are
      // Hermitian along x).
      cpx<T> u[VW][E];
      load_herm(u, Uy);
      {
        cpx<T> z[VW][E];
        load_herm(z, Uzz);
  #pragma unroll
        for (int m = 0; m < E; ++m) {
          const int k = j + TT * m;
          const T kx = (k == N / 2) ? (T)0 : (T)(kfx * (double)freq_index(k, N));
  #pragma unroll
          for (int w = 0; w < VW; ++w) {
            u[w][m] = {-kx * v[w][m].i - u[w][m].i, kx * v[w][m].r + u[w][m].r};
            v[w][m] = {v[w][m].r + z[w][m].i, v[w][m].i - z[w][m].r};
          }
        }
      }
      pencil_fft_strided<N, +1, T, VW, false, Cfg::TH>(u, j, col, lds, tw);
      pencil_fft_strided<N, +1, T, VW, false, Cfg::TH>(v, j, col, lds, tw);
      // Gas velocity of the pencil's cells, loaded after the transforms (live
      // through them it would cost the pass its second workgroup per CU) in
      // groups of four rows.  The signal is reduced as the largest squared
      // speed |v + h a|^2: cs + sqrt(.) is monotonic, so cs + sqrt(max) is
      // the max of gas_signal to the bit, with one square root per thread.
      constexpr int MG = SIG ? 4 : 1;
      T amax = (T)0;
      double smax2 = 0.0;
      const T h = SIG ? (T)*sg.owed : (T)0;
      if (act) {
  #pragma unroll
        for (int m0 = 0; m0 < E; m0 += MG) {
          T gv[SIG ? 3 : 1][MG][VW];
          if constexpr (SIG) {
  #pragma unroll
            for (int mm = 0; mm < MG; ++mm) {
              const int idx = base + (j + TT * (m0 + mm)) * se;
              ldr<VW>(sg.vx, idx, gv[0][mm]);
              ldr<VW>(sg.vy, idx, gv[1][mm]);
              ldr<VW>(sg.vz, idx, gv[2][mm]);
            }
          }
  #pragma unroll
          for (int mm = 0; mm < MG; ++mm) {
            const int m = m0 + mm;
            const int idx = base + (j + TT * m) * se;
            T t[VW];
  #pragma unroll
            for (int w = 0; w < VW; ++w) t[w] = v[w][m].r;
            str<VW>(V, idx, t);
  #pragma unroll
            for (int w = 0; w < VW; ++w) {
              const T a0 = -u[w][m].r, a1 = -u[w][m].i, a2 = v[w][m].i;
              if constexpr (!CVG) {
                T* o = A3 + kAccComp * (idx + w);
                o[0] = a0;
                o[1] = a1;
                o[2] = a2;
              }
              const T f0 = fabs(a0), f1 = fabs(a1), f2 = fabs(a2);
              amax = (f0 > amax || f0 != f0) ? f0 : amax;
              amax = (f1 > amax || f1 != f1) ? f1 : amax;
              amax = (f2 > amax || f2 != f2) ? f2 : amax;
              if constexpr (SIG)
                smax2 = nan_max(smax2, gas_signal_sq<T>(gv[0][mm][w], gv[1][mm][w],
                                                        gv[2][mm][w], a0, a1, a2, h));
            }
          }
        }
      }
      double smax = SIG ? (double)sg.cs + sqrt(smax2) : 0.0;
      // block max (NaN-propagating), one atomic per block: device-scope
      // atomics on one address serialise, so one per wave would stretch the
      // kernel's tail
      double am = (double)amax;
  #pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        double other = __shfl_xor(am, o, 64);
        am = (other > am || other != other) ? other : am;
      }
      __shared__ double wmax[kThreads / 64];
      if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = am;
      __syncthreads();
      if (threadIdx.x == 0) {
  #pragma unroll
        for (int w = 1; w < kThreads / 64; ++w) am = (wmax[w] > am || wmax[w] != wmax[w]) ? wmax[w] : am;
        atomic_max_nonneg(&diag[kDiagAMax], am);
      }
      if constexpr (CVG) {
        // The transform exchange buffer is free now: the workgroup's
        // accelerations are staged through it half a pencil (x < N/2, then
        // x >= N/2) at a time, cell (x, column c) at 3 (x' C V + c), and each
        // thread serves whole registrations of the tile.
        constexpr int CV = S::C * VW;  // columns of the workgroup
        constexpr int HX = N / 2;
        // short pencils (one thread per pencil) have no exchange buffer
        constexpr bool kOwn = HX * CV * 3 > 2 * S::kLdsStrided * VW;
        __shared__ T own_buf[kOwn ? HX * CV * 3 : 1];
        T* la = kOwn ? own_buf : reinterpret_cast<T*>(lds);
        const int v0 = threadIdx.x;
        const int p0 = gp0;
        const unsigned long long d0 = gd0;
        // the tile (rows y0.. of R, columns z0.. of 2^lz) for the absolute
        // descriptors of the one-row registration
        const int nbzt = 1 << (ga.lnz - ga.lz);
        const int Rt = CV >> ga.lz;
        const int y0t = ((int)blockIdx.x / nbzt) * Rt, z0t = ((int)blockIdx.x % nbzt) << ga.lz;
  #pragma unroll
        for (int h = 0; h < 2; ++h) {
          __syncthreads();  // the last readers of the buffer are done
          if (act) {
  #pragma unroll
            for (int mm = 0; mm < E / 2; ++mm) {
              const int m = h * (E / 2) + mm;
              const int xl = j + TT * mm;  // x - h HX
  #pragma unroll
              for (int w = 0; w < VW; ++w) {
                T* o = la + 3 * (xl * CV + col * VW + w);
                o[0] = -u[w][m].r;
                o[1] = -u[w][m].i;
                o[2] = v[w][m].i;
              }
            }
          }
          __syncthreads();
          for (int v = v0; v < gl.n; v += kThreads) {
            const int i = gl.slot(v);
            const int p = (v == v0) ? p0 : ga.sorted[i];
            const unsigned long long d = (v == v0) ? d0 : ga.desc[i];
            T* out = ga.cv + kCvReals * (long long)p;
  #pragma unroll
            for (int k = 0; k < 8; ++k) {
              const int sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
              const int xl = (int)((d >> (9 * sx)) & 511u) - h * HX;
              const int f = corner_column(d, 2 * sy + sz, ga.home, y0t, z0t, Rt, ga.lz);
              if ((unsigned)xl < (unsigned)HX && f != 0) {
                const T* q = la + 3 * (xl * CV + f - 1);
                out[3 * k] = q[0];
                out[3 * k + 1] = q[1];
                out[3 * k + 2] = q[2];
              }
            }
          }
        }
      }
      if constexpr (SIG) {
  #pragma unroll
        for (int o = 32; o > 0; o >>= 1) smax = nan_max(smax, __shfl_xor(smax, o, 64));
        __shared__ double wsig[kThreads / 64];
        if ((threadIdx.x & 63) == 0) wsig[threadIdx.x >> 6] = smax;
        __syncthreads();
        if (threadIdx.x == 0) {
  #pragma unroll
          for (int w = 1; w < kThreads / 64; ++w) smax = nan_max(smax, wsig[w]);
          sg.part[blockIdx.x] = smax;
        }
      }
    }
  }
};

// ===========================================================================
// adjoint passes
// ===========================================================================

// psi_bar <- exp(+i coef V) psi_bar; Vbar = coef * Im(conj(psi_bar) psi_ref).
// (Adjoint of psi_out = exp(-i coef V) psi_in: the phase factor is unitary
// and dL/dV = coef Im(conj(psi_bar_out) psi_out) = coef Im(conj(psi_bar_in) psi_in).)
template <class T>
struct k_kick_adj {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ pb, const cpx<T>* __restrict__ pr,
                                             const T* __restrict__ V, const T* __restrict__ Vx,
                                             double coef, T* __restrict__ Vbar, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
      // the phase includes the static external potential, which has no adjoint
      const double vt = (double)V[i] + (Vx ? (double)Vx[i] : 0.0);
      const cpx<T> e = expmi<T>(coef * vt);             // exp(-i coef V)
      const cpx<T> b = cmulc(pb[i], e);                 // * exp(+i coef V)
      pb[i] = b;
      const cpx<T> r = pr[i];
      Vbar[i] = (T)(coef * ((double)b.r * r.i - (double)b.i * r.r));
    }
  }
};

// psi_bar += 2 rhobar psi_ref  (adjoint of rho = |psi|^2)
template <class T>
struct k_rho_adj {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ pb, const cpx<T>* __restrict__ pr,
                                             const T* __restrict__ rb, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
      const T f = (T)2 * rb[i];
      const cpx<T> r = pr[i];
      cpx<T> b = pb[i];
      b.r += f * r.r;
      b.i += f * r.i;
      pb[i] = b;
    }
  }
};

// Adjoint Poisson source along x: half spectra of
//   T1 = FFT_x(Vbar) + i kx FFT_x(abar_x),  T2 = FFT_x(abar_y),  T3 = FFT_x(abar_z).
template <class T, int N>
struct k_adj_src_x {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const T* __restrict__ Vbar,
                                             const T* __restrict__ abar3, PassGeom g,
                                             const cpx<T>* __restrict__ tw, double kfx,
                                             cpx<T>* __restrict__ T1, cpx<T>* __restrict__ T2,
                                             cpx<T>* __restrict__ T3) {
    QJ_STRIDED_SETUP(T, N)
    cpx<T> v[VW][E], u[VW][E];
    auto load_real = [&](cpx<T>(&o)[VW][E], int comp) {
  #pragma unroll
      for (int m = 0; m < E; ++m) {
        const int idx = base + (j + TT * m) * se;
  #pragma unroll
        for (int w = 0; w < VW; ++w) {
          T x = (T)0;
          if (act) x = comp < 0 ? Vbar[idx + w] : abar3[kAccComp * (idx + w) + comp];
          o[w][m] = {x, (T)0};
        }
      }
    };
    auto store_half = [&](const cpx<T>(&o)[VW][E], cpx<T>* dst) {
      if (!act) return;
  #pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = j + TT * m;
        if (k <= N / 2) {
          cpx<T> t[VW];
  #pragma unroll
          for (int w = 0; w < VW; ++w) t[w] = o[w][m];
          stc<VW>(dst, base + k * se, t);
        }
      }
    };
    load_real(v, -1);
    pencil_fft_strided<N, -1, T, VW>(v, j, col, lds, tw);
    load_real(u, 0);
    pencil_fft_strided<N, -1, T, VW>(u, j, col, lds, tw);
  #pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = j + TT * m;
      const T kx = (k == N / 2) ? (T)0 : (T)(kfx * (double)freq_index(k, N));
  #pragma unroll
      for (int w = 0; w < VW; ++w) v[w][m] = {v[w][m].r - kx * u[w][m].i, v[w][m].i + kx * u[w][m].r};
    }
    store_half(v, T1);
    load_real(v, 1);
    pencil_fft_strided<N, -1, T, VW>(v, j, col, lds, tw);
    store_half(v, T2);
    load_real(v, 2);
    pencil_fft_strided<N, -1, T, VW>(v, j, col, lds, tw);
    store_half(v, T3);
  }
};

// Adjoint Poisson y pass (forward FFT): T1 <- FFT_y(T1) + i ky FFT_y(T2),
// T3 <- FFT_y(T3).
template <class T, int N>
struct k_adj_y {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ T1, const cpx<T>* __restrict__ T2,
                                             cpx<T>* __restrict__ T3, PassGeom g,
                                             const cpx<T>* __restrict__ tw, double kfy) {
    QJ_STRIDED_SETUP(T, N)
    cpx<T> v[VW][E], u[VW][E];
    load_rows<VW, E>(v, T1, base, se, j, TT, act);
    pencil_fft_strided<N, -1, T, VW>(v, j, col, lds, tw);
    load_rows<VW, E>(u, T2, base, se, j, TT, act);
    pencil_fft_strided<N, -1, T, VW>(u, j, col, lds, tw);
  #pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = j + TT * m;
      const T ky = (k == N / 2) ? (T)0 : (T)(kfy * (double)freq_index(k, N));
  #pragma unroll
      for (int w = 0; w < VW; ++w) v[w][m] = {v[w][m].r - ky * u[w][m].i, v[w][m].i + ky * u[w][m].r};
    }
    store_rows<VW, E>(v, T1, base, se, j, TT, act);
    load_rows<VW, E>(v, T3, base, se, j, TT, act);
    pencil_fft_strided<N, -1, T, VW>(v, j, col, lds, tw);
    store_rows<VW, E>(v, T3, base, se, j, TT, act);
  }
};

// Adjoint Poisson z pass: S = FFT_z(T1) + i kz FFT_z(T3); S *= G(k)/ntot;
// T1 <- IFFT_z(S).
template <class T, int N>
struct k_adj_z {
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(cpx<T>* __restrict__ T1, const cpx<T>* __restrict__ T3,
                                             PassGeom g, const cpx<T>* __restrict__ tw, int ny_log2,
                                             int nx, int ny, double kfx, double kfy, double kfz,
                                             double four_pi_G, double ntot, int dc_mode) {
    using S = PencilShape<N>;
    __shared__ cpx<T> lds[S::kLds];
    Lane<N, true> ln(threadIdx.x);
    const int q = blockIdx.x * S::C + ln.col;
    const bool act = q < g.ncols;
    const int base = act ? (int)pencil_base(g, q) : 0;
    cpx<T>* my = lds + ln.col * S::E * S::P;
    cpx<T> v[S::E], u[S::E];
  #pragma unroll
    for (int m = 0; m < S::E; ++m) v[m] = act ? T1[base + ln.j + S::TT * m] : cpx<T>{0, 0};
    pencil_fft<N, -1, T>(v, ln.j, my, tw);
  #pragma unroll
    for (int m = 0; m < S::E; ++m) u[m] = act ? T3[base + ln.j + S::TT * m] : cpx<T>{0, 0};
    pencil_fft<N, -1, T>(u, ln.j, my, tw);
    const int x = q >> ny_log2, y = q & (ny - 1);
    const double kx = kfx * (double)freq_index(x, nx);
    const double ky = kfy * (double)freq_index(y, ny);
    const double kxy2 = kx * kx + ky * ky;
  #pragma unroll
    for (int m = 0; m < S::E; ++m) {
      const int kzi = ln.j + S::TT * m;
      const double kz = kfz * (double)freq_index(kzi, N);
      const T kzt = (kzi == N / 2) ? (T)0 : (T)kz;
      cpx<T> sv = {v[m].r - kzt * u[m].i, v[m].i + kzt * u[m].r};
      const double k2 = kxy2 + kz * kz;
      double Gk = -four_pi_G / (k2 == 0.0 ? 1.0 : k2);
      if (k2 == 0.0 && dc_mode != 0) Gk = 0.0;
      v[m] = scale(sv, (T)(Gk / ntot));
    }
    pencil_fft<N, +1, T>(v, ln.j, my, tw);
    if (act) {
  #pragma unroll
      for (int m = 0; m < S::E; ++m) T1[base + ln.j + S::TT * m] = v[m];
    }
  }
};

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
template <int N>
inline int nblocks(int ncols) {
  return (ncols + PencilShape<N>::C - 1) / PencilShape<N>::C;
}
// Blocks of a group-walking contiguous sweep: two per CU (256 CUs), the
// blocks the kernel's registers (two waves per SIMD) keep resident at once,
// each taking several pencil groups.  1024 blocks ran as two rounds of 512
// and took 88 us at 256^3 c64 instead of 83.
template <int N>
inline int persist_blocks(int ncols, int cap = 0) {
  const int n = nblocks<N>(ncols);
  const int c = cap > 0 ? cap : 512;
  return n < c ? n : c;
}
template <class T, int N>
inline int nblocks_s(int ncols) {
  const int per = PencilShape<N>::C * StridedVW<T, N>::value;
  return (ncols + per - 1) / per;
}

#define QJ_DISPATCH_N(n, ...)                      \
  switch (n) {                                     \
    case 8: { constexpr int N_ = 8; __VA_ARGS__; } break;     \
    case 16: { constexpr int N_ = 16; __VA_ARGS__; } break;   \
    case 32: { constexpr int N_ = 32; __VA_ARGS__; } break;   \
    case 64: { constexpr int N_ = 64; __VA_ARGS__; } break;   \
    case 128: { constexpr int N_ = 128; __VA_ARGS__; } break; \
    case 256: { constexpr int N_ = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int N_ = 512; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;          \
  }

// pencil geometries over a full (nxh, ny, nz) field
inline PassGeom geom_x(const Dims& d) {  // pencils along x, q = y*nz + z
  return PassGeom{0, 1, (long long)d.ny * d.nz, ilog2(d.ny * d.nz), d.ny * d.nz};
}
inline PassGeom geom_y(const Dims& d, int nxh) {  // pencils along y, q = x*nz + z
  return PassGeom{(long long)d.ny * d.nz, 1, d.nz, ilog2(d.nz), nxh * d.nz};
}
inline PassGeom geom_z(const Dims& d, int nxh) {  // pencils along z, q = x*ny + y
  return PassGeom{0, d.nz, 1, 30, nxh * d.ny};
}

}  // namespace

template <class T>
hipError_t launch_twiddles(void* tw, int n, hipStream_t s) {
  k_twiddles<T><<<(n + 255) / 256, 256, 0, s>>>((cpx<T>*)tw, n);
  return hipGetLastError();
}

template <class T>
hipError_t launch_drift_table(void* tab, int n, double L, double c, hipStream_t s,
                              const double* dslot) {
  k_drift_table<T><<<(n + 255) / 256, 256, 0, s>>>((cpx<T>*)tab, n, L, c, dslot);
  return hipGetLastError();
}

template <class T>
hipError_t launch_drift_tables3(void* const tab[3], const int n[3], const double L[3], double c,
                                hipStream_t s, const double* dslot) {
  DriftTables3<T> t;
  int nmax = 0;
  for (int a = 0; a < 3; ++a) {
    t.tab[a] = (cpx<T>*)tab[a];
    t.n[a] = n[a];
    t.L[a] = L[a];
    nmax = n[a] > nmax ? n[a] : nmax;
  }
  k_drift_tables3<T><<<dim3((nmax + 255) / 256, 3), 256, 0, s>>>(t, c, dslot);
  return hipGetLastError();
}

template <class T>
hipError_t launch_kick(void* psi, const void* V, const void* Vb, const void* Vx, double coef,
                       long long n, hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  launch<k_kick<T>>((int)blocks, 256, s, (cpx<T>*)psi, (const T*)V, (const T*)Vb, (const T*)Vx,
                    coef, n);
  return hipGetLastError();
}

template <class T>
hipError_t launch_drift_axis(const Dims& d, const SpecTables<T>& tb, int axis, void* psi,
                             const void* V, const void* Vb, const void* Vx, double kick_coef,
                             int kick, hipStream_t s, int conj_d, const double* dslot,
                             void* psi_out) {
  cpx<T>* p = (cpx<T>*)psi;
  const T* v = (const T*)V;
  const T* vb = (const T*)Vb;
  const T* vx = (const T*)Vx;
  if (axis == 2) {
    PassGeom g = geom_z(d, d.nx);
    const cpx<T>* D = (const cpx<T>*)tb.drift_z;
    const cpx<T>* tw = (const cpx<T>*)tb.tw_z;
    const bool extra = vb || vx || conj_d;
    QJ_DISPATCH_N(d.nz, {
      const int nb = persist_blocks<N_>(g.ncols, tb.zblocks);
      if (kick && extra)
        launch<k_drift_c<T, N_, true>>(nb, 256, s, p, g, D, tw, v, vb, vx, kick_coef, conj_d,
                                       dslot, (const cpx<T>*)nullptr, (T*)nullptr,
                                       (cpx<T>*)psi_out);
      else if (kick)
        launch<k_drift_c<T, N_, true, false, false>>(nb, 256, s, p, g, D, tw, v, vb, vx,
                                                     kick_coef, conj_d, dslot,
                                                     (const cpx<T>*)nullptr, (T*)nullptr,
                                                     (cpx<T>*)psi_out);
      else if (extra)
        launch<k_drift_c<T, N_, false>>(nb, 256, s, p, g, D, tw, v, vb, vx, kick_coef, conj_d,
                                        dslot, (const cpx<T>*)nullptr, (T*)nullptr,
                                        (cpx<T>*)psi_out);
      else
        launch<k_drift_c<T, N_, false, false, false>>(nb, 256, s, p, g, D, tw, v, vb, vx,
                                                      kick_coef, conj_d, dslot,
                                                      (const cpx<T>*)nullptr, (T*)nullptr,
                                                      (cpx<T>*)psi_out);
    });
  } else {
    // the kick is fused into the z sweep only, and only the z sweep runs out of place
    if (kick || psi_out) return hipErrorInvalidValue;
    const bool y = axis == 1;
    PassGeom g = y ? geom_y(d, d.nx) : geom_x(d);
    const cpx<T>* D = (const cpx<T>*)(y ? tb.drift_y : tb.drift_x);
    const cpx<T>* tw = (const cpx<T>*)(y ? tb.tw_y : tb.tw_x);
    QJ_DISPATCH_N(y ? d.ny : d.nx, {
      launch<k_drift_s<T, N_>>(nblocks_s<T, N_>(g.ncols), 256, s, p, g, D, tw, conj_d,
                               (const T*)nullptr, (const cpx<T>*)nullptr);
    });
  }
  return hipGetLastError();
}

template <class T>
int source_x_blocks_t(const Dims& d) {
  const int ncols = d.ny * d.nz;
  switch (d.nx) {
    case 8: return nblocks_s<T, 8>(ncols);
    case 16: return nblocks_s<T, 16>(ncols);
    case 32: return nblocks_s<T, 32>(ncols);
    case 64: return nblocks_s<T, 64>(ncols);
    case 128: return nblocks_s<T, 128>(ncols);
    case 256: return nblocks_s<T, 256>(ncols);
    case 512: return nblocks_s<T, 512>(ncols);
    default: return 1;
  }
}

template <class T>
int poisson_x_blocks_t(const Dims& d) {
  const int ncols = d.ny * d.nz;
  int nb = 1;
  auto take = [&](int a, int b) { nb = a > b ? a : b; };
  switch (d.nx) {
    case 8: take(StridedCfg<T, 8, true>::blocks(ncols), StridedCfg<T, 8, false>::blocks(ncols)); break;
    case 16: take(StridedCfg<T, 16, true>::blocks(ncols), StridedCfg<T, 16, false>::blocks(ncols)); break;
    case 32: take(StridedCfg<T, 32, true>::blocks(ncols), StridedCfg<T, 32, false>::blocks(ncols)); break;
    case 64: take(StridedCfg<T, 64, true>::blocks(ncols), StridedCfg<T, 64, false>::blocks(ncols)); break;
    case 128: take(StridedCfg<T, 128, true>::blocks(ncols), StridedCfg<T, 128, false>::blocks(ncols)); break;
    case 256: take(StridedCfg<T, 256, true>::blocks(ncols), StridedCfg<T, 256, false>::blocks(ncols)); break;
    case 512: take(StridedCfg<T, 512, true>::blocks(ncols), StridedCfg<T, 512, false>::blocks(ncols)); break;
    default: break;
  }
  return nb;
}

#if !defined(QJ_PART) || QJ_PART == 0
int poisson_x_blocks(const Dims& d) {
  const int a = poisson_x_blocks_t<float>(d), b = poisson_x_blocks_t<double>(d);
  return a > b ? a : b;
}

namespace {
template <class T>
int gather_columns_t(const Dims& d, bool wide) {
  int cv = 0;
  auto take = [&](int c, int vw) { cv = c * vw; };
  switch (d.nx) {
#define QJ_GC(n)                                                                         \
  case n:                                                                                \
    if (wide) take(StridedCfg<T, n, true>::C, StridedCfg<T, n, true>::VW);               \
    else take(StridedCfg<T, n, false>::C, StridedCfg<T, n, false>::VW);                  \
    break;
    QJ_GC(8) QJ_GC(16) QJ_GC(32) QJ_GC(64) QJ_GC(128) QJ_GC(256) QJ_GC(512)
#undef QJ_GC
    default: return 0;
  }
  const long long ncols = (long long)d.ny * d.nz;
  const long long vol = (long long)d.nx * (cv < ncols ? cv : ncols);
  return vol <= kTileColMaxCells ? cv : 0;
}
}  // namespace

int poisson_x_gather_columns(const Dims& d, int c64, int wide) {
  return c64 ? gather_columns_t<float>(d, wide != 0) : gather_columns_t<double>(d, wide != 0);
}

namespace {
template <class T>
int source_x_columns_t(const Dims& d) {
  switch (d.nx) {
#define QJ_SC(n) \
  case n: return PencilShape<n>::C * StridedVW<T, n>::value;
    QJ_SC(8) QJ_SC(16) QJ_SC(32) QJ_SC(64) QJ_SC(128) QJ_SC(256) QJ_SC(512)
#undef QJ_SC
    default: return 0;
  }
}
}  // namespace

int source_x_columns(const Dims& d, int c64) {
  return c64 ? source_x_columns_t<float>(d) : source_x_columns_t<double>(d);
}

int source_x_blocks(const Dims& d) {
  // partials are sized for the larger of the two precisions' grids
  const int a = source_x_blocks_t<float>(d), b = source_x_blocks_t<double>(d);
  return a > b ? a : b;
}
#endif

template <class T>
hipError_t launch_source_x(const Dims& d, const SpecTables<T>& tb, int drift, void* psi,
                           void* src, int clear_src, void* brho, double b_coeff, double rho_bar,
                           void* H, double* partials, double* diag, hipStream_t s,
                           void* psi_out, const StarDeposit* dep) {
  if (psi_out && !drift) return hipErrorInvalidValue;  // only the drift writes psi
  PassGeom g = geom_x(d);
  const cpx<T>* D = (const cpx<T>*)tb.drift_x;
  const cpx<T>* tw = (const cpx<T>*)tb.tw_x;
  DepositArgs<T> da{};
  if (dep) {
    // the tiles must be the workgroups' columns: cols columns as rows of
    // 2^lz z columns, one tile per workgroup
    const int cols = source_x_columns(d, sizeof(T) == 4);
    const long long ncols = (long long)d.ny * d.nz;
    int lnz = 0;
    while ((1 << lnz) < d.nz) ++lnz;
    if (src || !psi || !dep->offset || !dep->sorted || !dep->desc || !dep->wts ||
        (long long)d.nx * (cols < ncols ? cols : ncols) > kTileColMaxCells || dep->lz < 0 ||
        dep->lz > lnz ||
        (1 << dep->lz) > cols || dep->ntiles != source_x_blocks_t<T>(d) ||
        ((1 << (dep->ly + dep->lz)) != cols && dep->ntiles != 1))
      return hipErrorInvalidValue;
    da = DepositArgs<T>{dep->offset, dep->sorted, dep->desc, (const T*)dep->wts, (T*)dep->rho,
                        dep->lz, lnz, dep->home};
  }
  QJ_DISPATCH_N(d.nx, {
    if (drift && dep)
      launch<k_source_x<T, N_, true, true>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (cpx<T>*)psi, (T*)src, clear_src, (T*)brho, b_coeff, rho_bar, g, D, tw, (cpx<T>*)H,
          partials, diag, (cpx<T>*)psi_out, da);
    else if (dep)
      launch<k_source_x<T, N_, false, true>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (cpx<T>*)psi, (T*)src, clear_src, (T*)brho, b_coeff, rho_bar, g, D, tw, (cpx<T>*)H,
          partials, diag, (cpx<T>*)psi_out, da);
    else if (drift)
      launch<k_source_x<T, N_, true>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (cpx<T>*)psi, (T*)src, clear_src, (T*)brho, b_coeff, rho_bar, g, D, tw, (cpx<T>*)H,
          partials, diag, (cpx<T>*)psi_out, da);
    else
      launch<k_source_x<T, N_, false>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (cpx<T>*)psi, (T*)src, clear_src, (T*)brho, b_coeff, rho_bar, g, D, tw, (cpx<T>*)H,
          partials, diag, (cpx<T>*)psi_out, da);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_fft_y_half(const Dims& d, const SpecTables<T>& tb, void* H, hipStream_t s) {
  PassGeom g = geom_y(d, d.nx / 2 + 1);
  const cpx<T>* tw = (const cpx<T>*)tb.tw_y;
  QJ_DISPATCH_N(d.ny, {
    launch<k_fft_s<T, N_, -1>>(nblocks_s<T, N_>(g.ncols), 256, s, (cpx<T>*)H, g, tw);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_poisson_z(const Dims& d, const SpecTables<T>& tb, int grad, void* H, void* Uz,
                            double four_pi_G, double rho_bar, int dc_mode,
                            const double* partials, double* diag, hipStream_t s, double lp_kc2) {
  const int nparts = source_x_blocks_t<T>(d);
  PassGeom g = geom_z(d, d.nx / 2 + 1);
  const cpx<T>* tw = (const cpx<T>*)tb.tw_z;
  const double kfx = 2.0 * M_PI / d.Lx, kfy = 2.0 * M_PI / d.Ly, kfz = 2.0 * M_PI / d.Lz;
  const double ntot = (double)d.nx * d.ny * d.nz;
  const int nyl = ilog2(d.ny);
  QJ_DISPATCH_N(d.nz, {
    if (grad && lp_kc2 > 0.0)
      launch<k_poisson_z<T, N_, true, true>>(nblocks<N_>(g.ncols), 256, s,
          (cpx<T>*)H, (cpx<T>*)Uz, g, tw, nyl, d.nx, d.ny, kfx, kfy, kfz, four_pi_G, rho_bar,
          ntot, dc_mode, partials, nparts, diag, lp_kc2);
    else if (grad)
      launch<k_poisson_z<T, N_, true>>(nblocks<N_>(g.ncols), 256, s,
          (cpx<T>*)H, (cpx<T>*)Uz, g, tw, nyl, d.nx, d.ny, kfx, kfy, kfz, four_pi_G, rho_bar,
          ntot, dc_mode, partials, nparts, diag, lp_kc2);
    else
      launch<k_poisson_z<T, N_, false>>(nblocks<N_>(g.ncols), 256, s,
          (cpx<T>*)H, (cpx<T>*)Uz, g, tw, nyl, d.nx, d.ny, kfx, kfy, kfz, four_pi_G, rho_bar,
          ntot, dc_mode, partials, nparts, diag, lp_kc2);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_poisson_y(const Dims& d, const SpecTables<T>& tb, int grad, void* H, void* Uz,
                            void* Uy, hipStream_t s) {
  PassGeom g = geom_y(d, d.nx / 2 + 1);
  const cpx<T>* tw = (const cpx<T>*)tb.tw_y;
  const double kfy = 2.0 * M_PI / d.Ly;
  QJ_DISPATCH_N(d.ny, {
    if (grad && !tb.wide)
      launch<k_poisson_y<T, N_, true, false>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (cpx<T>*)H, (cpx<T>*)Uz, (cpx<T>*)Uy, g, tw, kfy);
    else if (grad)
      launch<k_poisson_y<T, N_, true>>(k_poisson_y<T, N_, true>::Cfg::blocks(g.ncols),
          k_poisson_y<T, N_, true>::Cfg::TH, s, (cpx<T>*)H, (cpx<T>*)Uz, (cpx<T>*)Uy, g, tw, kfy);
    else
      launch<k_poisson_y<T, N_, false>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (cpx<T>*)H, (cpx<T>*)Uz, (cpx<T>*)Uy, g, tw, kfy);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_poisson_x(const Dims& d, const SpecTables<T>& tb, int grad, const void* H,
                            const void* Uy, const void* Uzz, void* V, void* A3,
                            const void* addend, double* diag, hipStream_t s,
                            GasSignalFold* sig, const StarGather* gather) {
  PassGeom g = geom_x(d);
  const cpx<T>* tw = (const cpx<T>*)tb.tw_x;
  const double kfx = 2.0 * M_PI / d.Lx;
  if (sig && (!grad || !sig->v[0] || !sig->v[1] || !sig->v[2] || !sig->owed || !sig->part))
    return hipErrorInvalidValue;
  if (gather && (!grad || sig || !gather->offset || !gather->sorted || !gather->desc ||
                 !gather->cv ||
                 poisson_x_gather_columns(d, sizeof(T) == 4, tb.wide) == 0))
    return hipErrorInvalidValue;
  SigArgs<T> sa{};
  if (sig) sa = SigArgs<T>{(const T*)sig->v[0], (const T*)sig->v[1], (const T*)sig->v[2],
                           sig->owed, sig->cs, sig->part};
  GatherArgs<T> ga{};
  if (gather) {
    int lnz = 0;
    while ((1 << lnz) < d.nz) ++lnz;
    if (gather->lz < 0 || gather->lz > lnz) return hipErrorInvalidValue;
    ga = GatherArgs<T>{gather->offset, gather->sorted, gather->desc, (T*)gather->cv, gather->lz,
                       lnz, gather->home};
  }
  QJ_DISPATCH_N(d.nx, {
    using Cfg = StridedCfg<T, N_, true>;
    using Nar = StridedCfg<T, N_, false>;
    const int nb = Cfg::blocks(g.ncols);
    const bool narrow = grad && !tb.wide;
    if (sig) {
      sig->nblk = narrow ? Nar::blocks(g.ncols) : nb;
      if (sig->nblk > sig->cap) return hipErrorInvalidValue;
    }
    if (narrow && sig)
      launch<k_poisson_x<T, N_, true, false, true>>(Nar::blocks(g.ncols), Nar::TH, s,
          (const cpx<T>*)H, (const cpx<T>*)Uy, (const cpx<T>*)Uzz, g, tw, kfx, (T*)V, (T*)A3,
          (const T*)addend, diag, sa, ga);
    else if (narrow && gather)
      launch<k_poisson_x<T, N_, true, false, false, true>>(Nar::blocks(g.ncols), Nar::TH, s,
          (const cpx<T>*)H, (const cpx<T>*)Uy, (const cpx<T>*)Uzz, g, tw, kfx, (T*)V, (T*)A3,
          (const T*)addend, diag, sa, ga);
    else if (narrow)
      launch<k_poisson_x<T, N_, true, false>>(Nar::blocks(g.ncols), Nar::TH, s,
          (const cpx<T>*)H, (const cpx<T>*)Uy, (const cpx<T>*)Uzz, g, tw, kfx, (T*)V, (T*)A3,
          (const T*)addend, diag, sa, ga);
    else if (grad && sig)
      launch<k_poisson_x<T, N_, true, true, true>>(nb, Cfg::TH, s,
          (const cpx<T>*)H, (const cpx<T>*)Uy, (const cpx<T>*)Uzz, g, tw, kfx, (T*)V, (T*)A3,
          (const T*)addend, diag, sa, ga);
    else if (grad && gather)
      launch<k_poisson_x<T, N_, true, true, false, true>>(nb, Cfg::TH, s,
          (const cpx<T>*)H, (const cpx<T>*)Uy, (const cpx<T>*)Uzz, g, tw, kfx, (T*)V, (T*)A3,
          (const T*)addend, diag, sa, ga);
    else if (grad)
      launch<k_poisson_x<T, N_, true>>(nb, Cfg::TH, s,
          (const cpx<T>*)H, (const cpx<T>*)Uy, (const cpx<T>*)Uzz, g, tw, kfx, (T*)V, (T*)A3,
          (const T*)addend, diag, sa, ga);
    else
      launch<k_poisson_x<T, N_, false>>(nb, Cfg::TH, s,
          (const cpx<T>*)H, (const cpx<T>*)Uy, (const cpx<T>*)Uzz, g, tw, kfx, (T*)V, (T*)A3,
          (const T*)addend, diag, sa, ga);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_drift_adj_axis(const Dims& d, const SpecTables<T>& tb, int axis, void* psi_bar,
                                 const void* rhobar, const void* rho_ref, const void* V,
                                 const void* Vx, double coef, const void* kick_ref, void* Vbar,
                                 hipStream_t s) {
  cpx<T>* p = (cpx<T>*)psi_bar;
  if (axis == 2) {
    if (!Vbar) return launch_drift_axis<T>(d, tb, 2, psi_bar, nullptr, nullptr, nullptr, 0.0, 0, s, 1);
    PassGeom g = geom_z(d, d.nx);
    const cpx<T>* D = (const cpx<T>*)tb.drift_z;
    const cpx<T>* tw = (const cpx<T>*)tb.tw_z;
    QJ_DISPATCH_N(d.nz, {
      launch<k_drift_c<T, N_, false, true>>(persist_blocks<N_>(g.ncols, tb.zblocks), 256, s, p, g, D, tw,
                                            (const T*)V, (const T*)nullptr, (const T*)Vx, coef, 1,
                                            (const double*)nullptr, (const cpx<T>*)kick_ref,
                                            (T*)Vbar, (cpx<T>*)nullptr);
    });
    return hipGetLastError();
  }
  if (!rhobar) return launch_drift_axis<T>(d, tb, axis, psi_bar, nullptr, nullptr, nullptr, 0.0, 0, s, 1);
  const bool y = axis == 1;
  PassGeom g = y ? geom_y(d, d.nx) : geom_x(d);
  const cpx<T>* D = (const cpx<T>*)(y ? tb.drift_y : tb.drift_x);
  const cpx<T>* tw = (const cpx<T>*)(y ? tb.tw_y : tb.tw_x);
  QJ_DISPATCH_N(y ? d.ny : d.nx, {
    launch<k_drift_s<T, N_, true>>(nblocks_s<T, N_>(g.ncols), 256, s, p, g, D, tw, 1,
                                   (const T*)rhobar, (const cpx<T>*)rho_ref);
  });
  return hipGetLastError();
}

template <class T>
hipError_t launch_kick_adj(void* psi_bar, const void* psi_ref, const void* V, const void* Vx,
                           double coef, void* Vbar, long long n, hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  launch<k_kick_adj<T>>((int)blocks, 256, s, (cpx<T>*)psi_bar, (const cpx<T>*)psi_ref,
                        (const T*)V, (const T*)Vx, coef, (T*)Vbar, n);
  return hipGetLastError();
}

template <class T>
hipError_t launch_rho_adj(void* psi_bar, const void* psi_ref, const void* rhobar, long long n,
                          hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  launch<k_rho_adj<T>>((int)blocks, 256, s, (cpx<T>*)psi_bar, (const cpx<T>*)psi_ref,
                                           (const T*)rhobar, n);
  return hipGetLastError();
}

template <class T>
hipError_t launch_adj_poisson(const Dims& d, const SpecTables<T>& tb, const void* Vbar,
                              const void* abar3, double four_pi_G, int dc_mode, void* H, void* U2,
                              void* U3, void* rhobar, hipStream_t s) {
  const double kfx = 2.0 * M_PI / d.Lx, kfy = 2.0 * M_PI / d.Ly, kfz = 2.0 * M_PI / d.Lz;
  const int nxh = d.nx / 2 + 1;
  {
    PassGeom g = geom_x(d);
    const cpx<T>* tw = (const cpx<T>*)tb.tw_x;
    QJ_DISPATCH_N(d.nx, {
      launch<k_adj_src_x<T, N_>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (const T*)Vbar, (const T*)abar3, g, tw, kfx, (cpx<T>*)H, (cpx<T>*)U2, (cpx<T>*)U3);
    });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  {
    PassGeom g = geom_y(d, nxh);
    const cpx<T>* tw = (const cpx<T>*)tb.tw_y;
    QJ_DISPATCH_N(d.ny, {
      launch<k_adj_y<T, N_>>(nblocks_s<T, N_>(g.ncols), 256, s,
          (cpx<T>*)H, (const cpx<T>*)U2, (cpx<T>*)U3, g, tw, kfy);
    });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  {
    PassGeom g = geom_z(d, nxh);
    const cpx<T>* tw = (const cpx<T>*)tb.tw_z;
    const double ntot = (double)d.nx * d.ny * d.nz;
    QJ_DISPATCH_N(d.nz, {
      launch<k_adj_z<T, N_>>(nblocks<N_>(g.ncols), 256, s,
          (cpx<T>*)H, (const cpx<T>*)U3, g, tw, ilog2(d.ny), d.nx, d.ny, kfx, kfy, kfz,
          four_pi_G, ntot, dc_mode);
    });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipError_t e = launch_poisson_y<T>(d, tb, 0, H, nullptr, nullptr, s);
  if (e != hipSuccess) return e;
  return launch_poisson_x<T>(d, tb, 0, H, nullptr, nullptr, rhobar, nullptr, nullptr, nullptr, s);
}

#define QJ_INST(T)                                                                              \
  template hipError_t launch_twiddles<T>(void*, int, hipStream_t);                              \
  template hipError_t launch_drift_table<T>(void*, int, double, double, hipStream_t,             \
                                            const double*);           \
  template hipError_t launch_drift_tables3<T>(void* const[3], const int[3], const double[3],     \
                                              double, hipStream_t, const double*);              \
  template hipError_t launch_kick<T>(void*, const void*, const void*, const void*, double,       \
                                     long long, hipStream_t);                                   \
  template hipError_t launch_drift_axis<T>(const Dims&, const SpecTables<T>&, int, void*,        \
                                           const void*, const void*, const void*, double, int,  \
                                           hipStream_t, int, const double*, void*);             \
  template hipError_t launch_source_x<T>(const Dims&, const SpecTables<T>&, int, void*, void*,  \
                                         int, void*, double, double, void*, double*, double*,   \
                                         hipStream_t, void*, const StarDeposit*);               \
  template hipError_t launch_fft_y_half<T>(const Dims&, const SpecTables<T>&, void*,            \
                                           hipStream_t);                                        \
  template hipError_t launch_poisson_z<T>(const Dims&, const SpecTables<T>&, int, void*, void*, \
                                          double, double, int, const double*, double*,          \
                                          hipStream_t, double);                                 \
  template hipError_t launch_poisson_y<T>(const Dims&, const SpecTables<T>&, int, void*, void*, \
                                          void*, hipStream_t);                                  \
  template hipError_t launch_poisson_x<T>(const Dims&, const SpecTables<T>&, int, const void*,  \
                                          const void*, const void*, void*, void*, const void*,  \
                                          double*, hipStream_t, GasSignalFold*,                 \
                                          const StarGather*);

#define QJ_INST_ADJ(T)                                                                          \
  template hipError_t launch_kick_adj<T>(void*, const void*, const void*, const void*, double,   \
                                         void*,                                                 \
                                         long long, hipStream_t);                               \
  template hipError_t launch_rho_adj<T>(void*, const void*, const void*, long long, hipStream_t); \
  template hipError_t launch_drift_adj_axis<T>(const Dims&, const SpecTables<T>&, int, void*,   \
                                               const void*, const void*, const void*,           \
                                               const void*, double, const void*, void*,         \
                                               hipStream_t);                                    \
  template hipError_t launch_adj_poisson<T>(const Dims&, const SpecTables<T>&, const void*,     \
                                            const void*, double, int, void*, void*, void*,      \
                                            void*, hipStream_t);

// The build compiles this file once per QJ_PART (0..3) so that the four
// instantiation sets build in parallel; without QJ_PART all are emitted.
#if !defined(QJ_PART) || QJ_PART == 0
QJ_INST(float)
#endif
#if !defined(QJ_PART) || QJ_PART == 1
QJ_INST(double)
#endif
#if !defined(QJ_PART) || QJ_PART == 2
QJ_INST_ADJ(float)
#endif
#if !defined(QJ_PART) || QJ_PART == 3
QJ_INST_ADJ(double)
#endif

}  // namespace qj
