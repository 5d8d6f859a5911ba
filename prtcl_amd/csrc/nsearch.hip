// Uniform-grid neighbour search for MI355X: cell keys, a counting sort with a
// wavefront prefix scan, a deterministic stable fix-up and the gathers that
// reorder particle data into cell order.
//
// Semantics follow the reference grid (src/prtcl/util/grouped_uniform_grid.hpp):
// cell edge = radius, cell index floor(double(x) / radius) per axis
// (:621-632), neighbours = particles of the 27 cells around the particle's
// cell whose FP32 squared distance is < double(radius)^2 (:407-422). The
// reference orders cells along a Morton curve with a comparison sort; here the
// cells live in a dense box (x fastest) and are sorted by counting, which
// yields the same cells and the same neighbour sets in a different order.
// Particles outside the box are clamped to its faces; clamping is monotone
// per axis, so two particles within one cell of each other stay within one
// clamped cell and no neighbour is lost.
#include "scene.hpp"

namespace prtcl_amd {

namespace {

constexpr int kScanBS = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBS * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x % kWave;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t u = __shfl_up(v, o, kWave);
    if (lane >= o)
      v += u;
  }
  return v;
}

// exclusive block scan over kScanBS threads; returns the exclusive prefix of
// this thread and the block total in *total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wsum[kScanBS / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1)
    wsum[w] = incl;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanBS / kWave; ++k) {
    uint32_t s = wsum[k];
    if (k < w)
      base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + incl - v;
}

__device__ __forceinline__ void k_scan_reduce_body(const uint32_t *count, uint32_t n,
                                                         uint32_t *tile_sums, unsigned bid) {
  const size_t base = static_cast<size_t>(bid) * kScanTile;
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    size_t i = base + static_cast<size_t>(k) * kScanBS + threadIdx.x;
    if (i < n)
      s += count[i];
  }
  s = block_sum<kScanBS>(s);
  if (threadIdx.x == 0)
    tile_sums[bid] = s;
}

__global__ __launch_bounds__(kScanBS) void k_scan_reduce(const uint32_t *count, uint32_t n,
                                                         uint32_t *tile_sums) {
  k_scan_reduce_body(count, n, tile_sums, blockIdx.x);
}

// one block: exclusive scan of the tile sums in place, total into start[n]
__global__ __launch_bounds__(kScanBS) void k_scan_top(uint32_t *tile_sums, uint32_t ntiles,
                                                      uint32_t *start, uint32_t n) {
  uint32_t carry = 0;
  for (uint32_t b = 0; b < ntiles; b += kScanBS) {
    uint32_t i = b + threadIdx.x;
    uint32_t v = i < ntiles ? tile_sums[i] : 0u;
    uint32_t tot;
    uint32_t ex = block_excl_scan(v, &tot);
    if (i < ntiles)
      tile_sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0)
    start[n] = carry;
}

// exclusive scan of tile bid after the prefix of the tiles before it;
// clears the counts and returns the tile's total
__device__ __forceinline__ uint32_t k_scan_down_body(uint32_t *count, uint32_t n,
                                                     uint32_t prefix, uint32_t *start,
                                                     unsigned bid) {
  const size_t base = static_cast<size_t>(bid) * kScanTile +
                      static_cast<size_t>(threadIdx.x) * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t s = 0;
  const bool full = base + kScanItems <= n;  // 16 items in range: 4 x 16-byte accesses
  if (full) {
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      const uint4 q = reinterpret_cast<const uint4 *>(count + base)[k];
      vals[4 * k] = q.x;
      vals[4 * k + 1] = q.y;
      vals[4 * k + 2] = q.z;
      vals[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      size_t i = base + k;
      vals[k] = i < n ? count[i] : 0u;
    }
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    s += vals[k];
  uint32_t tot;
  uint32_t ex = block_excl_scan(s, &tot) + prefix;
  if (full) {
#pragma unroll
    for (int k = 0; k < kScanItems / 4; ++k) {
      uint4 q;
      q.x = ex;
      ex += vals[4 * k];
      q.y = ex;
      ex += vals[4 * k + 1];
      q.z = ex;
      ex += vals[4 * k + 2];
      q.w = ex;
      ex += vals[4 * k + 3];
      reinterpret_cast<uint4 *>(start + base)[k] = q;
      reinterpret_cast<uint4 *>(count + base)[k] = make_uint4(0u, 0u, 0u, 0u);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      size_t i = base + k;
      if (i < n) {
        start[i] = ex;
        count[i] = 0u;
      }
      ex += vals[k];
    }
  }
  return tot;
}

// one launch for the scan of every tile: each block sums the (unscanned)
// totals of the tiles before it itself, and the last block writes the total
// count into start[n]
__global__ __launch_bounds__(kScanBS) void k_scan_down(uint32_t *count, uint32_t n,
                                                       const uint32_t *tile_sums,
                                                       uint32_t *start) {
  const unsigned bid = blockIdx.x;
  uint32_t p = strided_sum<kScanBS>(tile_sums, bid);
  uint32_t all;
  block_excl_scan(p, &all);  // the block total, in every thread
  p = all;
  const uint32_t tot = k_scan_down_body(count, n, p, start, bid);
  if (bid + 1 == gridDim.x && threadIdx.x == 0)
    start[n] = p + tot;
}

template <class T>
__device__ __forceinline__ void k_cell_keys_body(GridDims g, const vec4_t<T> *x, int n,
                                                      uint32_t *cell, uint32_t *count,
                                                      uint32_t *rank, unsigned bid) {
  const int i = bid * kBlock + threadIdx.x;
  uint32_t lin = 0xFFFFFFFFu;
  if (i < n) {
    const vec4_t<T> p = x[i];
    const double comp[3] = {static_cast<double>(p.x), static_cast<double>(p.y),
                            static_cast<double>(p.z)};
    const int org[3] = {g.ox, g.oy, g.oz};
    const int dim[3] = {g.nx, g.ny, g.nz};
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      // floor(double(x) / r) as grouped_uniform_grid.hpp:626-628, relative to
      // the box origin, clamped to the box (fmax maps NaN to 0)
      double f = floor(comp[a] / g.radius) - static_cast<double>(org[a]);
      f = fmin(fmax(f, 0.0), static_cast<double>(dim[a] - 1));
      c[a] = static_cast<uint32_t>(f);
    }
    cell[i] = c[0] | (c[1] << 11) | (c[2] << 22);
    lin = c[0] + static_cast<uint32_t>(g.nx) * (c[1] + static_cast<uint32_t>(g.ny) * c[2]);
  }
  // One atomic per distinct cell of the wavefront (particles arrive in the
  // previous step's cell order, so a wavefront covers only a few cells): the
  // lanes of a cell are grouped by ballots first, then the first lane of
  // every group adds the group's size in one atomic instruction for all
  // groups, and the lanes of a group take consecutive ranks after the base.
  const int lane = static_cast<int>(__lane_id());
  unsigned long long pending = __ballot(i < n);
  int leader_of = lane;
  uint32_t in_group = 0, group_size = 0;
  while (pending) {
    const int leader = __ffsll(static_cast<long long>(pending)) - 1;
    const uint32_t key = __shfl(lin, leader, kWave);
    const unsigned long long same = __ballot(lin == key) & pending;
    if ((same >> lane) & 1ull) {
      leader_of = leader;
      in_group = static_cast<uint32_t>(__popcll(same & ((1ull << lane) - 1ull)));
    }
    if (lane == leader)
      group_size = static_cast<uint32_t>(__popcll(same));
    pending &= ~same;
  }
  uint32_t base = 0;
  if (i < n && leader_of == lane)
    base = atomicAdd(&count[lin], group_size);
  const uint32_t r = __shfl(base, leader_of, kWave) + in_group;
  if (i < n)
    rank[i] = r;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cell_keys(GridDims g, const vec4_t<T> *x, int n,
                                                      uint32_t *cell, uint32_t *count,
                                                      uint32_t *rank) {
  k_cell_keys_body<T>(g, x, n, cell, count, rank, blockIdx.x);
}

__device__ __forceinline__ void k_scatter_body(GridDims g, const uint32_t *cell,
                                                    const uint32_t *rank,
                                                    const uint32_t *start, int n,
                                                    uint32_t *perm, unsigned bid) {
  const int i = bid * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const uint32_t c = cell[i];
  const uint32_t lin = (c & 0x7FFu) + static_cast<uint32_t>(g.nx) *
                                          (((c >> 11) & 0x7FFu) +
                                           static_cast<uint32_t>(g.ny) * (c >> 22));
  perm[start[lin] + rank[i]] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(kBlock) void k_scatter(GridDims g, const uint32_t *cell,
                                                    const uint32_t *rank,
                                                    const uint32_t *start, int n,
                                                    uint32_t *perm) {
  k_scatter_body(g, cell, rank, start, n, perm, blockIdx.x);
}

// Stable placement: the atomic ranks of k_cell_keys are scheduling
// dependent, so every particle recomputes its rank inside its cell as the
// number of cell members with a smaller previous slot. The result is a pure
// function of the previous order (a stable counting sort).
__device__ __forceinline__ void k_stable_rank_body(GridDims g, const uint32_t *cell,
                                                        const uint32_t *start,
                                                        const uint32_t *perm, int n,
                                                        uint32_t *perm_out, unsigned bid) {
  const int s = bid * kBlock + threadIdx.x;
  if (s >= n)
    return;
  const uint32_t o = perm[s];
  const uint32_t c = cell[o];
  const uint32_t lin = (c & 0x7FFu) + static_cast<uint32_t>(g.nx) *
                                          (((c >> 11) & 0x7FFu) +
                                           static_cast<uint32_t>(g.ny) * (c >> 22));
  const uint32_t b = start[lin], e = start[lin + 1];
  uint32_t r = 0;
  for (uint32_t k = b; k < e; ++k)
    r += perm[k] < o ? 1u : 0u;
  perm_out[b + r] = o;
}

__global__ __launch_bounds__(kBlock) void k_stable_rank(GridDims g, const uint32_t *cell,
                                                        const uint32_t *start,
                                                        const uint32_t *perm, int n,
                                                        uint32_t *perm_out) {
  k_stable_rank_body(g, cell, start, perm, n, perm_out, blockIdx.x);
}

// k_stable_rank and the reorder of the carried state in one pass: the thread
// of unstable slot s finds the stable slot p of its particle o = perm[s] as
// k_stable_rank does and moves x, v, id and cell of o to p (the same
// permutation as gathering through k_stable_rank's result, without writing it)
template <class V>
__device__ __forceinline__ void k_place_state_body(GridDims g, const uint32_t *__restrict__ cell,
                                                   const uint32_t *__restrict__ start,
                                                   const uint32_t *__restrict__ perm, int n,
                                                   V *__restrict__ xd, const V *__restrict__ xs,
                                                   V *__restrict__ vd, const V *__restrict__ vs,
                                                   uint32_t *__restrict__ id_d,
                                                   const uint32_t *__restrict__ id_s,
                                                   uint32_t *__restrict__ c_d, unsigned bid) {
  const int s = bid * kBlock + threadIdx.x;
  if (s >= n)
    return;
  const uint32_t o = perm[s];
  const uint32_t c = cell[o];
  const uint32_t lin = (c & 0x7FFu) + static_cast<uint32_t>(g.nx) *
                                          (((c >> 11) & 0x7FFu) +
                                           static_cast<uint32_t>(g.ny) * (c >> 22));
  const uint32_t b = start[lin], e = start[lin + 1];
  uint32_t r = 0;
  for (uint32_t k = b; k < e; ++k)
    r += perm[k] < o ? 1u : 0u;
  const uint32_t p = b + r;
  xd[p] = xs[o];
  vd[p] = vs[o];
  id_d[p] = id_s[o];
  c_d[p] = c;
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_place_state(GridDims g, const uint32_t *cell,
                                                        const uint32_t *start,
                                                        const uint32_t *perm, int n, V *xd,
                                                        const V *xs, V *vd, const V *vs,
                                                        uint32_t *id_d, const uint32_t *id_s,
                                                        uint32_t *c_d) {
  k_place_state_body<V>(g, cell, start, perm, n, xd, xs, vd, vs, id_d, id_s, c_d, blockIdx.x);
}

template <int W>
struct Words {
  uint32_t w[W];
};

template <class E>
__global__ __launch_bounds__(kBlock) void k_gather(E *__restrict__ dst,
                                                   const E *__restrict__ src,
                                                   const uint32_t *__restrict__ perm, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    dst[i] = src[perm[i]];
}

// the state carried between steps, reordered in one pass
template <class V>
__device__ __forceinline__ void k_gather_state_body(V *__restrict__ xd, const V *__restrict__ xs,
                                                         V *__restrict__ vd, const V *__restrict__ vs,
                                                         uint32_t *__restrict__ id_d,
                                                         const uint32_t *__restrict__ id_s,
                                                         uint32_t *__restrict__ c_d,
                                                         const uint32_t *__restrict__ c_s,
                                                         const uint32_t *__restrict__ perm, int n, unsigned bid) {
  const int i = bid * kBlock + threadIdx.x;
  if (i >= n)
    return;
  const uint32_t o = perm[i];
  xd[i] = xs[o];
  vd[i] = vs[o];
  id_d[i] = id_s[o];
  c_d[i] = c_s[o];
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_gather_state(V *__restrict__ xd, const V *__restrict__ xs,
                                                         V *__restrict__ vd, const V *__restrict__ vs,
                                                         uint32_t *__restrict__ id_d,
                                                         const uint32_t *__restrict__ id_s,
                                                         uint32_t *__restrict__ c_d,
                                                         const uint32_t *__restrict__ c_s,
                                                         const uint32_t *__restrict__ perm, int n) {
  k_gather_state_body<V>(xd, xs, vd, vs, id_d, id_s, c_d, c_s, perm, n, blockIdx.x);
}

// ---- batched wrappers: one launch over the blocks of every scene of a batch
template <class T>
__global__ __launch_bounds__(kBlock) void k_cell_keys_b(const BatchEntry<T> *tab,
                                                        const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_cell_keys_body<T>(E.S.grid, E.S.xm, E.S.nf, E.S.fcell, E.count, E.rank,
                      u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kScanBS) void k_scan_reduce_b(const BatchEntry<T> *tab,
                                                           const uint32_t *tmap) {
  const uint32_t u = tmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_scan_reduce_body(E.count, E.S.grid.ncells, E.scan_tmp, u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kScanBS) void k_scan_top_b(const BatchEntry<T> *tab) {
  const BatchEntry<T> &E = tab[blockIdx.x];
  uint32_t carry = 0;
  for (uint32_t b = 0; b < E.ntiles; b += kScanBS) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t v = i < E.ntiles ? E.scan_tmp[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan(v, &tot);
    if (i < E.ntiles)
      E.scan_tmp[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0)
    E.start[E.S.grid.ncells] = carry;
}

template <class T>
__global__ __launch_bounds__(kScanBS) void k_scan_down_b(const BatchEntry<T> *tab,
                                                         const uint32_t *tmap) {
  const uint32_t u = tmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_scan_down_body(E.count, E.S.grid.ncells, E.scan_tmp[u & kBatchLocalMask], E.start,
                   u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scatter_b(const BatchEntry<T> *tab,
                                                      const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_scatter_body(E.S.grid, E.S.fcell, E.rank, E.start, E.S.nf, E.perm, u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_stable_rank_b(const BatchEntry<T> *tab,
                                                          const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_stable_rank_body(E.S.grid, E.S.fcell, E.start, E.perm, E.S.nf, E.rank,
                     u & kBatchLocalMask);
}

// cell_dst must not be the perm buffer: perm is read by every thread
template <class T>
__global__ __launch_bounds__(kBlock) void k_place_state_b(const BatchEntry<T> *tab,
                                                          const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_place_state_body<vec4_t<T>>(E.S.grid, E.S.fcell, E.start, E.perm, E.S.nf, E.x_dst, E.S.xm,
                                E.v_dst, E.S.v, E.id_dst,
                                reinterpret_cast<const uint32_t *>(E.S.id), E.cell_dst,
                                u & kBatchLocalMask);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_gather_state_b(const BatchEntry<T> *tab,
                                                           const uint32_t *pmap) {
  const uint32_t u = pmap[blockIdx.x];
  const BatchEntry<T> &E = tab[u >> kBatchLocalBits];
  k_gather_state_body<vec4_t<T>>(E.x_dst, E.S.xm, E.v_dst, E.S.v, E.id_dst,
                                 reinterpret_cast<const uint32_t *>(E.S.id), E.cell_dst,
                                 E.S.fcell, E.rank, E.S.nf, u & kBatchLocalMask);
}

}  // namespace

size_t scan_tmp_words(uint32_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void scan_counts(hipStream_t st, uint32_t *count, uint32_t *start, uint32_t n,
                 uint32_t *tmp) {
  const uint32_t ntiles = (n + kScanTile - 1) / kScanTile;
  if (ntiles == 0) {
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBS), 0, st, tmp, 0u, start, 0u);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kScanBS), 0, st, count, n, tmp);
  hipLaunchKernelGGL(k_scan_down, dim3(ntiles), dim3(kScanBS), 0, st, count, n, tmp, start);
}

template <class T>
const uint32_t *GridOps<T>::sort_group(hipStream_t st, const GridDims &g, const vec4_t<T> *x, int n,
                            uint32_t *cell, uint32_t *count, uint32_t *start,
                            uint32_t *rank, uint32_t *perm, uint32_t *scan_tmp) {
  if (n > 0)
    hipLaunchKernelGGL(k_cell_keys<T>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, g, x, n,
                       cell, count, rank);
  scan_counts(st, count, start, g.ncells, scan_tmp);
  if (n > 0) {
    // scatter by the atomic ranks into `perm`, then the stable placement into
    // `rank` (the ranks are consumed by the scatter): the result is in `rank`
    hipLaunchKernelGGL(k_scatter, dim3(grid_blocks(n)), dim3(kBlock), 0, st, g, cell, rank,
                       start, n, perm);
    hipLaunchKernelGGL(k_stable_rank, dim3(grid_blocks(n)), dim3(kBlock), 0, st, g, cell,
                       start, perm, n, rank);
  }
  return rank;
}

template <class T>
void GridOps<T>::sort_state(hipStream_t st, const GridDims &g, const vec4_t<T> *x, int n,
                            uint32_t *cell, uint32_t *count, uint32_t *start, uint32_t *rank,
                            uint32_t *perm, uint32_t *scan_tmp, vec4_t<T> *xd, vec4_t<T> *vd,
                            const vec4_t<T> *vs, uint32_t *id_d, const uint32_t *id_s) {
  if (n > 0)
    hipLaunchKernelGGL(k_cell_keys<T>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, g, x, n,
                       cell, count, rank);
  scan_counts(st, count, start, g.ncells, scan_tmp);
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(grid_blocks(n)), dim3(kBlock), 0, st, g, cell, rank,
                       start, n, perm);
    // the atomic ranks are consumed: the cells go to `rank`
    hipLaunchKernelGGL(k_place_state<vec4_t<T>>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, g,
                       cell, start, perm, n, xd, x, vd, vs, id_d, id_s, rank);
  }
}

template <class T>
void GridOps<T>::gather4(hipStream_t st, void *dst, const void *src, const uint32_t *perm,
                         int n) {
  gather_elems(st, dst, src, perm, n, 4);
}

template <class T>
void GridOps<T>::gather_elems(hipStream_t st, void *dst, const void *src,
                              const uint32_t *perm, int n, int elem_bytes) {
  if (n <= 0)
    return;
  dim3 gr(grid_blocks(n)), bl(kBlock);
  switch (elem_bytes) {
    case 4:
      hipLaunchKernelGGL(k_gather<uint32_t>, gr, bl, 0, st, static_cast<uint32_t *>(dst),
                         static_cast<const uint32_t *>(src), perm, n);
      break;
    case 8:
      hipLaunchKernelGGL(k_gather<uint2>, gr, bl, 0, st, static_cast<uint2 *>(dst),
                         static_cast<const uint2 *>(src), perm, n);
      break;
    case 16:
      hipLaunchKernelGGL(k_gather<uint4>, gr, bl, 0, st, static_cast<uint4 *>(dst),
                         static_cast<const uint4 *>(src), perm, n);
      break;
    case 32:
      hipLaunchKernelGGL(k_gather<Words<8>>, gr, bl, 0, st, static_cast<Words<8> *>(dst),
                         static_cast<const Words<8> *>(src), perm, n);
      break;
    case 36:
      hipLaunchKernelGGL(k_gather<Words<9>>, gr, bl, 0, st, static_cast<Words<9> *>(dst),
                         static_cast<const Words<9> *>(src), perm, n);
      break;
    case 72:
      hipLaunchKernelGGL(k_gather<Words<18>>, gr, bl, 0, st, static_cast<Words<18> *>(dst),
                         static_cast<const Words<18> *>(src), perm, n);
      break;
    default:
      throw HipError{hipErrorInvalidValue, "gather_elems: element size"};
  }
}

template <class T>
void GridOps<T>::gather_state(hipStream_t st, vec4_t<T> *xd, const vec4_t<T> *xs,
                              vec4_t<T> *vd, const vec4_t<T> *vs, uint32_t *id_d,
                              const uint32_t *id_s, uint32_t *c_d, const uint32_t *c_s,
                              const uint32_t *perm, int n) {
  if (n > 0)
    hipLaunchKernelGGL(k_gather_state<vec4_t<T>>, dim3(grid_blocks(n)), dim3(kBlock), 0, st,
                       xd, xs, vd, vs, id_d, id_s, c_d, c_s, perm, n);
}

template <class T>
void GridOps<T>::sort_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                            unsigned npblocks, const uint32_t *tmap, unsigned ntiles,
                            unsigned nscenes) {
  if (npblocks > 0)
    hipLaunchKernelGGL(k_cell_keys_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  if (ntiles > 0)
    hipLaunchKernelGGL(k_scan_reduce_b<T>, dim3(ntiles), dim3(kScanBS), 0, st, tab, tmap);
  if (nscenes > 0)
    hipLaunchKernelGGL(k_scan_top_b<T>, dim3(nscenes), dim3(kScanBS), 0, st, tab);
  if (ntiles > 0)
    hipLaunchKernelGGL(k_scan_down_b<T>, dim3(ntiles), dim3(kScanBS), 0, st, tab, tmap);
  if (npblocks > 0) {
    hipLaunchKernelGGL(k_scatter_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
    hipLaunchKernelGGL(k_place_state_b<T>, dim3(npblocks), dim3(kBlock), 0, st, tab, pmap);
  }
}

unsigned scan_tiles(uint32_t ncells) { return (ncells + kScanTile - 1) / kScanTile; }

template struct GridOps<float>;
template struct GridOps<double>;

}  // namespace prtcl_amd
