// Register/LDS pencil FFT for gfx950.
//
// An N-point pencil (N a power of two, 8..512) is spread over TT = N/E
// threads; thread j holds the E elements x[j + TT*m], m = 0..E-1.  The
// transform is a two-stage Stockham decomposition N = E * TT:
//   stage 1: an E-point DFT in registers over m (stride TT), twiddle
//            W_N^(j*k1);
//   exchange through LDS (one write, one read, block barrier);
//   stage 2: E/TT independent TT-point DFTs in registers.
// Output element k1 + E*k2 lands in thread (k1 % TT), register
// (k1 / TT) + (E/TT)*k2, i.e. in exactly the input distribution
// "thread j holds k = j + TT*m".  A forward transform, a diagonal k-space
// multiply and an inverse transform therefore chain in registers with no
// extra data movement.
#pragma once

#include "qj_common.h"

namespace qj {

// Elements per thread for an N-point pencil, in workgroups of TH threads.
template <int N, int TH = 256>
struct PencilShape {
  static_assert((N & (N - 1)) == 0 && N >= 8 && N <= 512, "pencil length must be 2^k in [8,512]");
  static constexpr int E = (N <= 16) ? N : (N <= 256 ? 16 : 32);
  static constexpr int TT = N / E;        // threads per pencil
  static constexpr int kThreads = TH;     // workgroup size
  static constexpr int C = kThreads / TT; // pencils per workgroup
  static constexpr int P = TT + 1;        // padded LDS row (k1-major)
  static constexpr int kLds = (TT > 1) ? C * E * P : 1;  // complex elements
  // Column-fastest layout for the strided lane mapping (neighbouring lanes
  // hold neighbouring pencils): element (k1, j) of pencil col lives at
  // k1*KS + j*C + col.  The 16-element pad per k1 block puts the pencil
  // groups a wave reads at once into disjoint LDS banks.
  static constexpr int KS = TT * C + 16;
  static constexpr int kLdsStrided = (TT > 1) ? E * KS : 1;
};

// In-register E-point DFT, natural order in and out (radix-2/4 four-step
// recursion with compile-time twiddles).
template <int E, int DIR, class T>
__device__ __forceinline__ void dft(cpx<T>* v) {
  if constexpr (E == 1) {
    return;
  } else if constexpr (E == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (E == 4) {
    cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
    cpx<T> t2 = v[1] + v[3];
    cpx<T> t3 = twc<1, 4, DIR>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
  } else {
    constexpr int A = (E >= 16) ? 4 : 2;
    constexpr int B = E / A;
    cpx<T> y[B][A];
    // step 1: A-point DFTs over x[B*a + b], then twiddle W_E^(b*ka)
    sfor<0, B>([&](auto bc) {
      constexpr int b = decltype(bc)::value;
      cpx<T> t[A];
      sfor<0, A>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        t[a] = v[B * a + b];
      });
      dft<A, DIR, T>(t);
      sfor<0, A>([&](auto kc) {
        constexpr int ka = decltype(kc)::value;
        y[b][ka] = twc<b * ka, E, DIR>(t[ka]);
      });
    });
    // step 2: B-point DFTs over b -> X[ka + A*kb]
    sfor<0, A>([&](auto kc) {
      constexpr int ka = decltype(kc)::value;
      cpx<T> t[B];
      sfor<0, B>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        t[b] = y[b][ka];
      });
      dft<B, DIR, T>(t);
      sfor<0, B>([&](auto kb) {
        constexpr int k = decltype(kb)::value;
        v[ka + A * k] = t[k];
      });
    });
  }
}

// N-point pencil transform over the TT threads that hold one pencil.
//   v    : the thread's E elements (x[j + TT*m])
//   j    : thread rank inside the pencil, 0..TT-1
//   lds  : CONTIG: this pencil's private slice (E*P complex elements);
//          strided: the workgroup's shared exchange buffer, and col is the
//          pencil's slot (column-fastest layout, see PencilShape::KS)
//   tw   : W_N^p = exp(-2*pi*i*p/N), p = 0..N-1 (conjugated for DIR = +1)
// Every thread of the workgroup must call this (it contains barriers).
template <int N, int DIR, class T, bool CONTIG = true>
__device__ __forceinline__ void pencil_fft(cpx<T>* v, int j, cpx<T>* lds,
                                           const cpx<T>* __restrict__ tw, int col = 0) {
  using S = PencilShape<N>;
  constexpr int E = S::E, TT = S::TT, P = S::P;
  dft<E, DIR, T>(v);
  if constexpr (TT > 1) {
#pragma unroll
    for (int k1 = 1; k1 < E; ++k1) {
      cpx<T> w = tw[(j * k1) & (N - 1)];
      v[k1] = (DIR < 0) ? cmul(v[k1], w) : cmulc(v[k1], w);
    }
    __syncthreads();  // previous readers of this LDS buffer are done
    if constexpr (CONTIG) {
#pragma unroll
      for (int k1 = 0; k1 < E; ++k1) lds[k1 * P + j] = v[k1];
    } else {
#pragma unroll
      for (int k1 = 0; k1 < E; ++k1) lds[k1 * S::KS + j * S::C + col] = v[k1];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E / TT; ++q) {
      cpx<T> u[TT];
      const int k1 = j + TT * q;
#pragma unroll
      for (int b = 0; b < TT; ++b)
        u[b] = CONTIG ? lds[k1 * P + b] : lds[k1 * S::KS + b * S::C + col];
      dft<TT, DIR, T>(u);
#pragma unroll
      for (int kb = 0; kb < TT; ++kb) v[q + (E / TT) * kb] = u[kb];
    }
  }
}

// Several pencils per lane in the strided layout: the VW pencils of a lane
// are neighbours in memory (q0, q0+1), so one row of them is one 16-byte
// access for c64, and the pencils share the stage twiddles and the two
// block barriers of the exchange.  All VW pencils use the column-fastest
// LDS layout, one kLdsStrided block each.
template <int N, int DIR, class T, int VW, bool PHASED = false, int TH = 256>
__device__ __forceinline__ void pencil_fft_strided(cpx<T> (*v)[PencilShape<N>::E], int j, int col,
                                                   cpx<T>* lds, const cpx<T>* __restrict__ tw) {
  using S = PencilShape<N, TH>;
  constexpr int E = S::E, TT = S::TT, C = S::C, KS = S::KS, L = S::kLdsStrided;
#pragma unroll
  for (int w = 0; w < VW; ++w) dft<E, DIR, T>(v[w]);
  if constexpr (TT > 1) {
#pragma unroll
    for (int k1 = 1; k1 < E; ++k1) {
      const cpx<T> t = tw[(j * k1) & (N - 1)];
#pragma unroll
      for (int w = 0; w < VW; ++w) v[w][k1] = (DIR < 0) ? cmul(v[w][k1], t) : cmulc(v[w][k1], t);
    }
    if constexpr (PHASED) {
      // one exchange buffer, the VW pencils take turns (half the LDS)
#pragma unroll
      for (int w = 0; w < VW; ++w) {
        __syncthreads();
#pragma unroll
        for (int k1 = 0; k1 < E; ++k1) lds[k1 * KS + j * C + col] = v[w][k1];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < E / TT; ++q) {
          cpx<T> u[TT];
          const int k1 = j + TT * q;
#pragma unroll
          for (int b = 0; b < TT; ++b) u[b] = lds[k1 * KS + b * C + col];
          dft<TT, DIR, T>(u);
#pragma unroll
          for (int kb = 0; kb < TT; ++kb) v[w][q + (E / TT) * kb] = u[kb];
        }
      }
      return;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < VW; ++w)
#pragma unroll
      for (int k1 = 0; k1 < E; ++k1) lds[w * L + k1 * KS + j * C + col] = v[w][k1];
    __syncthreads();
#pragma unroll
    for (int w = 0; w < VW; ++w) {
#pragma unroll
      for (int q = 0; q < E / TT; ++q) {
        cpx<T> u[TT];
        const int k1 = j + TT * q;
#pragma unroll
        for (int b = 0; b < TT; ++b) u[b] = lds[w * L + k1 * KS + b * C + col];
        dft<TT, DIR, T>(u);
#pragma unroll
        for (int kb = 0; kb < TT; ++kb) v[w][q + (E / TT) * kb] = u[kb];
      }
    }
  }
}

// Pencils per lane of the strided sweeps: two for c64 at 256 points (one
// 16-byte access per row, shared twiddles and barriers), one otherwise.  At
// 128 points and below a single box has too few pencil groups to fill the
// CUs with two per lane (256 blocks at 128^3); one per lane doubles the
// blocks and cut the 128^3 stars step by 9%.
template <class T, int N>
struct StridedVW {
  static constexpr int value = (sizeof(T) == 4 && N == 256) ? 2 : 1;
};

// Workgroup size and pencils per lane of a strided sweep.  The wide form
// (c64 at 256 points) gives each lane one pencil in 512-thread workgroups:
// the same 32-column row pieces as two pencils per lane in 256 threads, with
// half the registers per lane.  A sweep that holds several spectra per
// pencil then runs four waves per SIMD instead of two (the Poisson x pass:
// 225 -> 123 VGPRs, 136 -> 121 us at 256^3 in isolation).
// WTH: workgroup size of the wide form, 512 threads (32 pencils, two rows
// of a pencil per wave) or 256 (16 pencils, four rows per wave).
template <class T, int N, bool WIDE, int WTH = 512>
struct StridedCfg {
  static constexpr bool kWide = WIDE && sizeof(T) == 4 && N == 256;
  static constexpr int TH = kWide ? WTH : 256;
  static constexpr int VW = kWide ? 1 : StridedVW<T, N>::value;
  static constexpr int C = PencilShape<N, TH>::C;
  static int blocks(int ncols) { return (ncols + C * VW - 1) / (C * VW); }
};

// VW neighbouring complex (or real) values starting at element i.
template <int VW, class T>
__device__ __forceinline__ void ldc(const cpx<T>* __restrict__ p, int i, cpx<T>* o) {
  if constexpr (VW == 2 && sizeof(T) == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p + i);
    o[0] = {t.x, t.y};
    o[1] = {t.z, t.w};
  } else {
#pragma unroll
    for (int w = 0; w < VW; ++w) o[w] = p[i + w];
  }
}
template <int VW, class T>
__device__ __forceinline__ void stc(cpx<T>* __restrict__ p, int i, const cpx<T>* o) {
  if constexpr (VW == 2 && sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p + i) = make_float4(o[0].r, o[0].i, o[1].r, o[1].i);
  } else {
#pragma unroll
    for (int w = 0; w < VW; ++w) p[i + w] = o[w];
  }
}
template <int VW, class T>
__device__ __forceinline__ void ldr(const T* __restrict__ p, int i, T* o) {
  if constexpr (VW == 2 && sizeof(T) == 4) {
    const float2 t = *reinterpret_cast<const float2*>(p + i);
    o[0] = t.x;
    o[1] = t.y;
  } else {
#pragma unroll
    for (int w = 0; w < VW; ++w) o[w] = p[i + w];
  }
}
template <int VW, class T>
__device__ __forceinline__ void str(T* __restrict__ p, int i, const T* o) {
  if constexpr (VW == 2 && sizeof(T) == 4) {
    *reinterpret_cast<float2*>(p + i) = make_float2(o[0], o[1]);
  } else {
#pragma unroll
    for (int w = 0; w < VW; ++w) p[i + w] = o[w];
  }
}

// Signed frequency index of bin k of an n-point transform, Nyquist negative
// (quantum-stars.py:142-148: ifftshift(arange(-n/2, n/2))).
__device__ __forceinline__ int freq_index(int k, int n) { return (k < n / 2) ? k : k - n; }

// Pass geometry: pencil q starts at
//   base(q) = (q >> inner_log2) * s_outer + (q & (2^inner_log2 - 1)) * s_inner
// and its elements are s_elem apart.
struct PassGeom {
  long long s_outer;
  long long s_inner;
  long long s_elem;
  int inner_log2;
  int ncols;
};

__device__ __forceinline__ long long pencil_base(const PassGeom& g, int q) {
  return (long long)(q >> g.inner_log2) * g.s_outer +
         (long long)(q & ((1 << g.inner_log2) - 1)) * g.s_inner;
}

// Thread -> (pencil slot, rank) mapping inside a 256-thread workgroup.
// STRIDED: pencil slot fastest across lanes (neighbouring pencils are
// neighbouring addresses: x and y passes).  CONTIGUOUS: rank fastest
// (a pencil's elements are neighbouring addresses: z passes).
template <int N, bool CONTIG>
struct Lane {
  int col, j;
  __device__ __forceinline__ Lane(int tid) {
    using S = PencilShape<N>;
    if constexpr (CONTIG) {
      j = tid % S::TT;
      col = tid / S::TT;
    } else {
      col = tid % S::C;
      j = tid / S::C;
    }
  }
};

}  // namespace qj
