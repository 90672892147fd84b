// Launch layer: every compute kernel of the library is a body type
//   struct K { static constexpr int kThreads = ...;
//              static __device__ void run(args...); };
// launched through launch<K>(grid, block, stream, args...).  Two
// __global__ wrappers instantiate each body:
//   k_one<K>  - one launch, arguments by value (the ordinary path);
//   k_many<K> - one launch over a batch of independent problems of the same
//               shape: blockIdx.z selects the problem, whose argument pack is
//               read from device memory (packs[blockIdx.z]).
// A thread-local Recorder, when installed, captures launches (body, grid,
// block, argument bytes) instead of issuing them.  The ensemble runner
// records one KDK step per member this way, zips members with identical
// launch sequences into k_many launches and captures those in a hipGraph;
// the composition code (kdk_steps) stays the single description of a step.
#pragma once

#include <hip/hip_runtime.h>
#include <string.h>

#include <type_traits>
#include <utility>
#include <vector>

namespace qj {

// Aggregate argument pack with a layout that host and device agree on.
template <class... A>
struct Pack;
template <>
struct Pack<> {};
template <class H, class... R>
struct Pack<H, R...> {
  H h;
  Pack<R...> r;
};

inline Pack<> make_pack() { return {}; }
template <class H, class... R>
Pack<H, R...> make_pack(H h, R... r) {
  return Pack<H, R...>{h, make_pack(r...)};
}

template <int I, class H, class... R>
__device__ __forceinline__ const auto& pack_get(const Pack<H, R...>& p) {
  if constexpr (I == 0)
    return p.h;
  else
    return pack_get<I - 1>(p.r);
}

template <class K, class... A, int... I>
__device__ __forceinline__ void pack_apply(const Pack<A...>& p, std::integer_sequence<int, I...>) {
  K::run(pack_get<I>(p)...);
}

template <class K, class... A>
__global__ __launch_bounds__(K::kThreads) void k_one(A... a) {
  K::run(a...);
}

template <class K, class... A>
__global__ __launch_bounds__(K::kThreads) void k_many(const Pack<A...>* __restrict__ packs) {
  const Pack<A...> p = packs[blockIdx.z];
  pack_apply<K>(p, std::make_integer_sequence<int, sizeof...(A)>{});
}

// A body may ask for a minimum number of waves per SIMD (K::kMinWaves): the
// compiler then keeps its registers within 512 / kMinWaves.  Such bodies get
// wrappers of their own, so the others keep their default schedules.
template <class K, class = void>
struct MinWaves {
  static constexpr int value = 0;
};
template <class K>
struct MinWaves<K, std::void_t<decltype(K::kMinWaves)>> {
  static constexpr int value = K::kMinWaves;
};

template <class K, class... A>
__global__ __launch_bounds__(K::kThreads, MinWaves<K>::value) void k_one_w(A... a) {
  K::run(a...);
}

template <class K, class... A>
__global__ __launch_bounds__(K::kThreads, MinWaves<K>::value) void k_many_w(
    const Pack<A...>* __restrict__ packs) {
  const Pack<A...> p = packs[blockIdx.z];
  pack_apply<K>(p, std::make_integer_sequence<int, sizeof...(A)>{});
}

// One captured launch: the batched wrapper, its geometry and the bytes of
// its argument pack (pack_align is the pack's alignment).
struct LaunchRec {
  const void* fn;
  dim3 grid;
  dim3 block;
  size_t pack_align;
  std::vector<unsigned char> pack;
};

struct Recorder {
  std::vector<LaunchRec> recs;
};

inline thread_local Recorder* g_recorder = nullptr;

// Issues (or records) one launch; errors surface through hipGetLastError.
template <class K, class... A>
void launch(dim3 grid, dim3 block, hipStream_t s, A... a) {
  if (grid.x == 0) return;
  if (g_recorder) {
    const Pack<A...> p = make_pack(a...);
    LaunchRec r;
    if constexpr (MinWaves<K>::value > 0)
      r.fn = reinterpret_cast<const void*>(&k_many_w<K, A...>);
    else
      r.fn = reinterpret_cast<const void*>(&k_many<K, A...>);
    r.grid = grid;
    r.block = block;
    r.pack_align = alignof(Pack<A...>);
    r.pack.resize(sizeof(p));
    memcpy(r.pack.data(), &p, sizeof(p));
    g_recorder->recs.push_back(std::move(r));
    return;
  }
  if constexpr (MinWaves<K>::value > 0)
    k_one_w<K, A...><<<grid, block, 0, s>>>(a...);
  else
    k_one<K, A...><<<grid, block, 0, s>>>(a...);
}

// Installs a recorder for the lifetime of the object.
struct RecordScope {
  Recorder* prev;
  explicit RecordScope(Recorder* r) : prev(g_recorder) { g_recorder = r; }
  ~RecordScope() { g_recorder = prev; }
};

}  // namespace qj
