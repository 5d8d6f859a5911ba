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
// Layout: chunks of 8 entries; entry e of particle i at
// idx[((e / 8) * n + i) * 8 + e % 8], so one 16-byte load per lane fetches 8
// entries and a wavefront's load is 1 KB contiguous. Entries are ordered by
// slab; info[i] holds the entry
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

// the same with the value of global slot g given by at(g)
template <class V, class At>
__device__ __forceinline__ void stage_span_fn(const Spans &R, int s, At &&at, V *dst) {
  const size_t base = R.base(s);
  const uint32_t len = R.len(s);
  for (uint32_t u = threadIdx.x; u < len; u += kBlock)
    dst[u] = at(base + u);
}

constexpr uint32_t kChunk = 8;  // list entries per 16-byte chunk

// the 8 entries of chunk g of particle i
__device__ __forceinline__ uint4 list_chunk(const NbrList &L, size_t n, int i, uint32_t g) {
  return reinterpret_cast<const uint4 *>(L.idx)[static_cast<size_t>(g) * n +
                                                static_cast<size_t>(i)];
}

// entry k (0..7) of a chunk
__device__ __forceinline__ uint32_t chunk_entry(const uint4 &q, int k) {
  const uint32_t w = k < 2 ? q.x : (k < 4 ? q.y : (k < 6 ? q.z : q.w));
  return (k & 1) ? (w >> 16) : (w & 0xFFFFu);
}

// entry e of particle i (single load)
__device__ __forceinline__ uint32_t list_entry(const NbrList &L, size_t n, int i, uint32_t e) {
  return L.idx[((static_cast<size_t>(e) / kChunk) * n + static_cast<size_t>(i)) * kChunk +
               e % kChunk];
}

// entries of a (not overflowed) list: the sum of its three slab counts
__device__ __forceinline__ uint32_t list_count(uint32_t info) {
  return (info & 0xFFu) + ((info >> 8) & 0xFFu) + ((info >> 16) & 0xFFu);
}

// Visits the list entries of particle i (not overflowed) in list order:
// fn(e, j) with e the entry's index and j its global slot. R is the span
// table of i's block. The slot is b0 + entry + [e >= c0] (b1 - b0) +
// [e >= c01] (b2 - b1) (two selects against register values).
template <class Fn>
__device__ __forceinline__ void list_visit_indexed(const NbrList &L, const Spans &R, size_t n,
                                                   int i, uint32_t info, Fn &&fn) {
  const uint32_t c0 = info & 0xFFu;
  const uint32_t c01 = c0 + ((info >> 8) & 0xFFu);
  const uint32_t c = c01 + ((info >> 16) & 0xFFu);
  const uint32_t d1 = R.b1 - R.b0, d2 = R.b2 - R.b1;
  for (uint32_t g = 0; g * kChunk < c; ++g) {
    const uint4 q = list_chunk(L, n, i, g);
#pragma unroll
    for (int k = 0; k < static_cast<int>(kChunk); ++k) {
      const uint32_t e = g * kChunk + k;
      if (e < c)
        fn(e, R.b0 + chunk_entry(q, k) + (e >= c0 ? d1 : 0u) + (e >= c01 ? d2 : 0u));
    }
  }
}

// the same with the slot only: fn(j)
template <class Fn>
__device__ __forceinline__ void list_visit(const NbrList &L, const Spans &R, size_t n, int i,
                                           uint32_t info, Fn &&fn) {
  list_visit_indexed(L, R, n, i, info, [&](uint32_t, uint32_t j) { fn(j); });
}

// The same walk one chunk (8 entries) at a time with the memory traffic of
// a chunk issued before its arithmetic: the chunk, then the gathers
// (v[k] = load(j_k)) of kGroup entries, then use(v[k]) in order; the next
// chunk is fetched while the current one is in flight. kGroup < 8 keeps fewer
// gathered values live (more waves per SIMD) for wide records.
template <int kGroup = 8, class Load, class Use>
__device__ __forceinline__ void list_gather(const NbrList &L, const Spans &R, size_t n, int i,
                                            uint32_t info, Load &&load, Use &&use) {
  static_assert(kChunk % kGroup == 0, "group must divide the chunk");
  const uint32_t c0 = info & 0xFFu;
  const uint32_t c01 = c0 + ((info >> 8) & 0xFFu);
  const uint32_t c = c01 + ((info >> 16) & 0xFFu);
  const uint32_t d1 = R.b1 - R.b0, d2 = R.b2 - R.b1;
  using V = decltype(load(0u));
  uint4 q = c > 0 ? list_chunk(L, n, i, 0) : make_uint4(0u, 0u, 0u, 0u);
  for (uint32_t g = 0; g * kChunk < c; ++g) {
    const uint4 cur = q;
    if ((g + 1) * kChunk < c)
      q = list_chunk(L, n, i, g + 1);
#pragma unroll
    for (int h = 0; h < static_cast<int>(kChunk); h += kGroup) {
      V v[kGroup];
#pragma unroll
      for (int k = 0; k < kGroup; ++k) {
        const uint32_t e = g * kChunk + h + k;
        if (e < c)
          v[k] = load(R.b0 + chunk_entry(cur, h + k) + (e >= c0 ? d1 : 0u) +
                      (e >= c01 ? d2 : 0u));
      }
      // keep every gather of the group ahead of the arithmetic (the
      // scheduler would otherwise consume each gather before issuing the
      // next one)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < kGroup; ++k)
        if (g * kChunk + h + k < c)
          use(v[k]);
    }
  }
}

}  // namespace prtcl_amd
