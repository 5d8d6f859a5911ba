// Per-step fluid neighbour list of the fused step, built by the setup pass
// and walked by every later neighbour pass of the step.
//
// The 27-cell stencil of a particle is 9 x-runs (rows dy, dz = -1..1, cells
// x-1..x+1). Slots are in cell order, x fastest, so the three runs of one dz
// ("slab") of every particle of a block of kBlock consecutive slots lie in
// one slot interval, the block's span of that slab: from the first run start
// to the last run end over the block's particles (about three rows of
// particles). A list entry is the neighbour's offset from the span start of
// its slab (16 bits), so the list costs 2 bytes per pair; the global slot is
// span base + entry, and a pass that stages a span in LDS reads the
// neighbour at LDS index = entry.
//
// Layout: entry e of particle i at idx[e * n + i] (a wavefront reads or writes
// 64 consecutive entries), entries ordered by slab; info[i] holds the entry
// count of each slab (8 bits each) and kListOverflow when the particle has
// more than kNbrCap neighbours or a span of its block exceeds 16 bits (its
// passes then walk the grid with the exact predicate instead).
#pragma once

#include "scene.hpp"

namespace prtcl_amd {

constexpr uint32_t kListOverflow = 1u << 31;
// LDS bytes of one staged field array of one span (the real4 fields of 1536
// particles in FP32, 768 in FP64); a block whose span is larger reads the
// neighbours' fields from global memory instead
constexpr uint32_t kSlabBytes = 24u * 1024u;

template <class T>
constexpr uint32_t slab_cap() {
  return kSlabBytes / static_cast<uint32_t>(sizeof(vec4_t<T>));
}

struct BlockRanges {
  uint32_t base[3];  // first slot of the span of slab s (dz = s - 1)
  uint32_t len[3];   // its length (0: no particle of the block has that layer)
  uint32_t spare[2];
};

struct NbrList {
  uint16_t *idx;
  uint32_t *info;
  BlockRanges *blk;  // one per block of kBlock slots
};

// The span table of one block in registers (selected by slab without
// indexing, so it never lives in scratch memory).
struct Spans {
  uint32_t b0, b1, b2, l0, l1, l2;
  __device__ __forceinline__ static Spans of(const BlockRanges &R) {
    return Spans{R.base[0], R.base[1], R.base[2], R.len[0], R.len[1], R.len[2]};
  }
  __device__ __forceinline__ uint32_t base(int s) const { return s == 0 ? b0 : (s == 1 ? b1 : b2); }
  __device__ __forceinline__ uint32_t len(int s) const { return s == 0 ? l0 : (s == 1 ? l1 : l2); }
  __device__ __forceinline__ uint32_t max_len() const {
    return l0 > l1 ? (l0 > l2 ? l0 : l2) : (l1 > l2 ? l1 : l2);
  }
};

// Cooperative copy of the span of slab s into LDS: dst[u] = src[base + u]
// (all threads of the block call it).
template <class V>
__device__ __forceinline__ void stage_span(const Spans &R, int s, const V *__restrict__ src,
                                           V *dst) {
  const V *__restrict__ p = src + R.base(s);
  const uint32_t len = R.len(s);
  for (uint32_t u = threadIdx.x; u < len; u += kBlock)
    dst[u] = p[u];
}

// Visits the global slots of the list entries of particle i (not
// overflowed) in list order: fn(j). R is the span table of i's block.
template <class Fn>
__device__ __forceinline__ void list_visit(const NbrList &L, const Spans &R, size_t n, int i,
                                           uint32_t info, Fn &&fn) {
  const uint16_t *__restrict__ idx = L.idx + i;
  const uint32_t c0 = info & 0xFFu;
  const uint32_t c01 = c0 + ((info >> 8) & 0xFFu);
  const uint32_t c = c01 + ((info >> 16) & 0xFFu);
  const uint32_t d1 = R.b1 - R.b0, d2 = R.b2 - R.b1;
  for (uint32_t e = 0; e < c; ++e)
    fn(R.b0 + idx[e * n] + (e >= c0 ? d1 : 0u) + (e >= c01 ? d2 : 0u));
}

// The same walk in groups of kUnroll entries, with the memory traffic of a
// group issued before its arithmetic: the indices first, then every gather
// (v[k] = load(j_k)), then use(v[k]) in order. The slot is
// b0 + entry + [e >= c0] (b1 - b0) + [e >= c01] (b2 - b1) (two selects
// against register values).
template <int kUnroll, class Load, class Use>
__device__ __forceinline__ void list_gather(const NbrList &L, const Spans &R, size_t n, int i,
                                            uint32_t info, Load &&load, Use &&use) {
  const uint16_t *__restrict__ idx = L.idx + i;
  const uint32_t c0 = info & 0xFFu;
  const uint32_t c01 = c0 + ((info >> 8) & 0xFFu);
  const uint32_t c = c01 + ((info >> 16) & 0xFFu);
  const uint32_t d1 = R.b1 - R.b0, d2 = R.b2 - R.b1;
  auto slot = [&](uint32_t e, uint32_t u) {
    return R.b0 + u + (e >= c0 ? d1 : 0u) + (e >= c01 ? d2 : 0u);
  };
  using V = decltype(load(0u));
  uint32_t e = 0;
  for (; e + kUnroll <= c; e += kUnroll) {
    uint32_t u[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k)
      u[k] = idx[(e + k) * n];
    V v[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; ++k)
      v[k] = load(slot(e + k, u[k]));
    // keep every gather of the group ahead of the arithmetic (the scheduler
    // would otherwise consume each gather before issuing the next one)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < kUnroll; ++k)
      use(v[k]);
  }
  for (; e < c; ++e)
    use(load(slot(e, idx[e * n])));
}

}  // namespace prtcl_amd
