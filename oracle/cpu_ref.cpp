// ORACLE — test infrastructure only.
//
// A CPU/OpenMP restatement of prtcl's SPH step hot path, written from the
// reference's algorithm description (file:line citations below point into the
// siretty/prtcl sources). It is the checker for the HIP product and the CPU
// baseline that bench.py times; the product never links or calls it.
//
// What is restated, and from where:
//   * grouped uniform grid + Morton order + neighbour query:
//       src/prtcl/util/grouped_uniform_grid.hpp:126-422,621-632,
//       src/prtcl/util/morton_order.hpp:39-82,
//       src/prtcl/util/neighborhood.cpp:100-167
//   * permutation of all varying fields into grid order:
//       src/prtcl/util/neighborhood.cpp:106-154, grouped_uniform_grid.hpp:394-399,
//       src/prtcl/data/varying_manager.hpp:86-102 (gather: new[i] = old[perm[i]])
//   * cubic spline kernel: src/prtcl/math/kernel/cubic_spline_kernel.hpp:43-102,
//       src/prtcl/math/kernel/kernel_facade.hpp:204-229,
//       src/prtcl/math/constants.hpp:75-78 (gradient cut-off sqrt(FLT_EPSILON))
//   * schemes (generated OpenMP C++): density.hpp:160-215, gravity.hpp:133,
//       viscosity.hpp:171-240, symplectic_euler.hpp:172-270,
//       aiast12.hpp:125-175, iisph.hpp:188-516, wkbb18.hpp:194-425
//   * CG: src/prtcl/solver/cg_openmp.hpp:32-236
//   * step order: share/scripts/prtcl-lua.example.lua:363-462
//
// Arithmetic follows the reference expression trees in FP32 without FMA
// contraction (the reference is built for baseline x86-64 without
// -ffast-math). Eigen's fixed-size-3 reductions (dot, squaredNorm) are
// restated as x0 + (x1 + x2); that association order is an assumption about
// Eigen 3.3.7 that nothing in /root/reference pins ("parity unpinned" at the
// ulp level). OpenMP reductions are restated as sums over fixed chunks of
// 1024 particles combined in chunk order, so results do not depend on the
// thread count (the reference's association order is unspecified).

#include "cpu_ref.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <parallel/algorithm>
#include <string>
#include <utility>
#include <vector>

namespace {

using real = float;

// ---------------------------------------------------------------------------
// small tensor helpers (restating the Eigen operations used on the path,
// src/prtcl/math/impl/eigen.hpp:124-175, 258-264, 292-314)
// ---------------------------------------------------------------------------

inline real dot3(const real *a, const real *b) {
  return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
inline real sqnorm3(const real *a) { return dot3(a, a); }
inline real norm3(const real *a) { return std::sqrt(sqnorm3(a)); }
inline void sub3(const real *a, const real *b, real *out) {
  out[0] = a[0] - b[0];
  out[1] = a[1] - b[1];
  out[2] = a[2] - b[2];
}

// Planted mistakes for the invariant tests (tests/test_oracle_invariants.py).
// Built with -DORACLE_MUTANT=k, the hook numbered k returns its planted value
// (a wrong factor, sign or index) instead of the reference's, so the tests can
// show that each invariant catches such a mistake. The default build
// (ORACLE_MUTANT 0) is the restatement itself: every hook returns its first
// argument and folds away at compile time.
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
template <class V>
constexpr V mutant(int k, V reference, V planted) {
  return ORACLE_MUTANT == k ? planted : reference;
}
const real kZero3[3] = {0, 0, 0};

// reciprocal_or_zero (eigen.hpp:292-300)
inline real reciprocal_or_zero(real x, real eps) {
  if (std::abs(x) > eps)
    return 1 / x;
  return 0;
}
// unit_step_l (eigen.hpp:306-314): 1 if eps < x
inline real unit_step_l(real eps, real x) { return (eps < x) ? real{1} : real{0}; }

// 3x3 inverse by the adjugate / determinant (cofactor) formula; column-major
// storage (element (r,c) at r + 3c) like Eigen's default storage order.
inline real cof3(const real *m, int i, int j) {
  auto at = [m](int r, int c) { return m[r + 3 * c]; };
  int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return at(i1, j1) * at(i2, j2) - at(i1, j2) * at(i2, j1);
}
inline void invert3(const real *m, real *out) {
  real c00 = cof3(m, 0, 0), c10 = cof3(m, 1, 0), c20 = cof3(m, 2, 0);
  real det = c00 * m[0] + (c10 * m[1] + c20 * m[2]);
  real invdet = real{1} / det;
  // inverse(r, c) = cofactor(c, r) / det
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      out[r + 3 * c] = cof3(m, c, r) * invdet;
}

// Moore-Penrose pseudo-inverse of a 3x3 matrix (column-major), invert_pm of
// eigen.hpp:266-276. Eigen takes it from a complete orthogonal decomposition
// (not available here; its rank counts pivots above 3 eps of the largest);
// this restatement uses a one-sided Jacobi SVD (Hestenes) in double: rotate
// column pairs of B = A (accumulating V) until they are orthogonal, then
// sigma_k = |b_k| and pinv(A) = sum over sigma_k > 3 eps sigma_max of
// v_k b_k^T / sigma_k^2. Pinned against numpy.linalg.pinv in the tests.
inline void pinv3(const real *m, real *out) {
  double b[9], v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int k = 0; k < 9; ++k)
    b[k] = static_cast<double>(m[k]);
  for (int sweep = 0; sweep < 12; ++sweep) {
    bool rotated = false;
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      double al = 0, be = 0, ga = 0;
      for (int r = 0; r < 3; ++r) {
        al += b[r + 3 * p] * b[r + 3 * p];
        be += b[r + 3 * q] * b[r + 3 * q];
        ga += b[r + 3 * p] * b[r + 3 * q];
      }
      if (!(std::abs(ga) > 1e-17 * std::sqrt(al * be)))
        continue;
      rotated = true;
      const double zeta = (be - al) / (2.0 * ga);
      const double t = std::copysign(1.0, zeta) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
      const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
      for (int r = 0; r < 3; ++r) {
        const double bp = b[r + 3 * p], bq = b[r + 3 * q];
        b[r + 3 * p] = c * bp - sn * bq;
        b[r + 3 * q] = sn * bp + c * bq;
        const double vp = v[r + 3 * p], vq = v[r + 3 * q];
        v[r + 3 * p] = c * vp - sn * vq;
        v[r + 3 * q] = sn * vp + c * vq;
      }
    }
    if (!rotated)
      break;
  }
  double s2[3], smax2 = 0;
  for (int k = 0; k < 3; ++k) {
    s2[k] = b[3 * k] * b[3 * k] + b[3 * k + 1] * b[3 * k + 1] + b[3 * k + 2] * b[3 * k + 2];
    smax2 = std::max(smax2, s2[k]);
  }
  const double thr = 3.0 * static_cast<double>(std::numeric_limits<real>::epsilon());
  double inv[3];
  for (int k = 0; k < 3; ++k)
    inv[k] = (s2[k] > thr * thr * smax2 && s2[k] > 0.0) ? 1.0 / s2[k] : 0.0;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      out[r + 3 * c] = static_cast<real>(v[r] * b[c] * inv[0] + v[r + 3] * b[c + 3] * inv[1] +
                                         v[r + 6] * b[c + 6] * inv[2]);
}

// L g for a 3x3 L (column-major) with Eigen's fixed-size row order
// l0 g0 + (l1 g1 + l2 g2)
inline void matvec3(const real *L, const real *g, real *out) {
  for (int r = 0; r < 3; ++r)
    out[r] = L[r] * g[0] + (L[r + 3] * g[1] + L[r + 6] * g[2]);
}

// ---------------------------------------------------------------------------
// cubic spline kernel (Monaghan 1992), as cubic_spline_kernel.hpp:43-102 and
// kernel_facade.hpp:204-229
// ---------------------------------------------------------------------------

inline real constpow(real base, int e) {  // util/constpow.hpp:9-18
  real r = 1;
  for (; e > 0; --e)
    r *= base;
  return r;
}
inline real kernel_alpha() {
  return static_cast<real>(
      1.L / 4.L *
      0.318309886183790671537767526745028724068919291480912897495334688117793595L);
}
inline real evalq(real q) {
  real result = 0;
  if (q < 1)
    result -= 4 * constpow(1 - q, 3);
  if (q < 2)
    result += constpow(2 - q, 3);
  return kernel_alpha() * result;
}
inline real evaldq(real q) {
  real result = 0;
  if (q < 1)
    result += 12 * constpow(1 - q, 2);
  if (q < 2)
    result -= 3 * constpow(2 - q, 2);
  return kernel_alpha() * result;
}
inline real kernel_eps() { return std::sqrt(std::numeric_limits<real>::epsilon()); }

inline real kernel_w(const real *v, real h) {
  real r = norm3(v);
  return evalq(std::abs(r) / h) / constpow(h, 3);
}
inline void kernel_grad(const real *v, real h, real *out) {
  real r = norm3(v);
  if (r < kernel_eps()) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  real dr = -std::copysign(evaldq(std::abs(r) / h) / constpow(h, 4), r);
  real f = dr / r;
  out[0] = f * v[0];
  out[1] = f * v[1];
  out[2] = f * v[2];
}

// ---------------------------------------------------------------------------
// Morton order (morton_order.hpp:39-82)
// ---------------------------------------------------------------------------

using GridIndex = std::array<int32_t, 3>;

inline uint32_t unsign(int32_t index) {
  constexpr int32_t smax = std::numeric_limits<int32_t>::max();
  if (index < 0)
    return static_cast<uint32_t>(index + smax);
  return static_cast<uint32_t>(index) + static_cast<uint32_t>(smax);
}
inline bool less_msb(uint32_t lhs, uint32_t rhs) {
  return (lhs < rhs) && (lhs < (lhs ^ rhs));
}
inline int morton_compare(const GridIndex &lhs, const GridIndex &rhs) {
  size_t msd = 0;
  for (size_t dim = 1; dim < 3; ++dim)
    if (less_msb(unsign(lhs[msd]) ^ unsign(rhs[msd]),
                 unsign(lhs[dim]) ^ unsign(rhs[dim])))
      msd = dim;
  uint32_t l = unsign(lhs[msd]), r = unsign(rhs[msd]);
  return (l < r) ? -1 : (l > r) ? 1 : 0;
}
inline bool morton_less(const GridIndex &a, const GridIndex &b) {
  return morton_compare(a, b) < 0;
}

// the 13 "sparse" adjacent cell offsets (grouped_uniform_grid.hpp:51-68); the
// other 13 are their mirrors at index 26 - 1 - k (:331-333)
constexpr int32_t kSparseOffsets[13][3] = {
    {1, 0, 0},  {0, 1, 0},  {0, 0, 1},   {1, 1, 0},  {1, 0, 1},
    {0, 1, 1},  {1, 1, 1},  {0, -1, 1},  {1, 0, -1}, {1, 1, -1},
    {1, -1, -1}, {1, -1, 0}, {1, -1, 1}};

// ---------------------------------------------------------------------------
// groups
// ---------------------------------------------------------------------------

constexpr int kBoundary = 0;  // group "b"
constexpr int kFluid = 1;     // group "f"

struct Field {
  std::string name;
  int ncomp;
  std::vector<real> data;
};

struct GroupData {
  size_t n = 0;
  std::vector<Field> fields;  // all varying real fields
  std::vector<int64_t> id;    // carried original index

  Field *find(const std::string &name) {
    for (auto &f : fields)
      if (f.name == name)
        return &f;
    return nullptr;
  }
  real *ptr(const std::string &name) { return find(name)->data.data(); }
  void add(const std::string &name, int ncomp) {
    fields.push_back(Field{name, ncomp, std::vector<real>(n * ncomp, real{0})});
  }
};

// ---------------------------------------------------------------------------
// grouped uniform grid (grouped_uniform_grid.hpp:77-638) for two groups
// ---------------------------------------------------------------------------

struct RawIndex {
  int32_t group;
  int32_t index;
};

class Grid {
 public:
  double radius = 1;
  std::vector<std::vector<int32_t>> raw_to_sorted;
  std::vector<RawIndex> sorted_to_raw;
  std::vector<int32_t> sorted_to_cell;
  std::vector<std::pair<int32_t, int32_t>> cell_range;
  std::vector<GridIndex> cell_grid;
  std::vector<std::array<int32_t, 26>> cell_adj;

  // positions per group (float3 AoS), counts per group
  std::array<const real *, 2> x{nullptr, nullptr};
  std::array<size_t, 2> n{0, 0};

  GridIndex x_to_gi(const real *p) const {  // :621-632
    GridIndex gi;
    for (int a = 0; a < 3; ++a)
      gi[a] = static_cast<int32_t>(std::floor(static_cast<double>(p[a]) / radius));
    return gi;
  }
  GridIndex gi_of(const RawIndex &r) const {
    return x_to_gi(x[r.group] + 3 * static_cast<size_t>(r.index));
  }

  void update() {
    // update_sorted_to_raw (:139-213)
    bool reset = false;
    raw_to_sorted.resize(2);
    size_t total = 0;
    for (int g = 0; g < 2; ++g) {
      size_t old_count = raw_to_sorted[g].size();
      raw_to_sorted[g].resize(n[g]);
      if (old_count != n[g])
        reset = true;
      total += n[g];
    }
    if (reset) {
      sorted_to_raw.resize(total);
      size_t s = 0;
      for (int g = 0; g < 2; ++g)
        for (size_t r = 0; r < n[g]; ++r)
          sorted_to_raw[s++] = RawIndex{g, static_cast<int32_t>(r)};
    }
    // The reference sorts with an unstable parallel sort whose comparator
    // recomputes both grid indices; a stable sort with the same comparator
    // gives one of the admissible orders, deterministically.
    __gnu_parallel::stable_sort(
        sorted_to_raw.begin(), sorted_to_raw.end(),
        [this](const RawIndex &a, const RawIndex &b) {
          return morton_less(gi_of(a), gi_of(b));
        });
#pragma omp parallel for
    for (size_t s = 0; s < sorted_to_raw.size(); ++s)
      raw_to_sorted[sorted_to_raw[s].group][sorted_to_raw[s].index] =
          static_cast<int32_t>(s);

    // update_sorted_to_cell (:220-257)
    sorted_to_cell.assign(sorted_to_raw.size(), 0);
    cell_range.clear();
    cell_grid.clear();
    cell_adj.clear();
    if (sorted_to_raw.empty())
      return;
    size_t used = 1;
#pragma omp parallel for reduction(+ : used)
    for (size_t s = 1; s < sorted_to_raw.size(); ++s) {
      if (gi_of(sorted_to_raw[s]) == gi_of(sorted_to_raw[s - 1])) {
        sorted_to_cell[s] = 0;
      } else {
        sorted_to_cell[s] = 1;
        used += 1;
      }
    }
    cell_range.assign(used, {-1, -1});
    cell_grid.assign(used, GridIndex{0, 0, 0});
    std::array<int32_t, 26> invalid;
    invalid.fill(-1);
    cell_adj.assign(used, invalid);

    // update_cell_to_grid_and_sorted_range (:264-306), serial running sum
    cell_range[0].first = 0;
    cell_grid[0] = gi_of(sorted_to_raw[0]);
    for (size_t s = 1; s < sorted_to_raw.size(); ++s) {
      int32_t value = sorted_to_cell[s];
      int32_t c = sorted_to_cell[s - 1] + value;
      sorted_to_cell[s] = c;
      if (value == 1) {
        cell_range[c - 1].second = static_cast<int32_t>(s);
        cell_range[c].first = static_cast<int32_t>(s);
        cell_grid[c] = gi_of(sorted_to_raw[s]);
      }
    }
    cell_range[used - 1].second = static_cast<int32_t>(sorted_to_raw.size());

    // update_cell_to_adjacent_cells (:312-337)
#pragma omp parallel for
    for (size_t c = 0; c < used; ++c) {
      for (int k = 0; k < 13; ++k) {
        GridIndex gi = cell_grid[c];
        for (int a = 0; a < 3; ++a)
          gi[a] += kSparseOffsets[k][a];
        int32_t j = find_cell(gi);
        if (j >= 0) {
          cell_adj[c][k] = j;
          cell_adj[j][26 - 1 - k] = static_cast<int32_t>(c);
        }
      }
    }
  }

  int32_t find_cell(const GridIndex &gi) const {  // :347-359
    auto it = std::lower_bound(cell_grid.begin(), cell_grid.end(), gi, morton_less);
    if (it == cell_grid.end() || *it != gi)
      return -1;
    return static_cast<int32_t>(it - cell_grid.begin());
  }

  // neighbors(group, index, data, fn) (:407-422, :447-492)
  template <class Fn>
  void neighbors(int g, size_t i, Fn fn) const {
    const double r2 = (1.0 * radius) * radius;
    const real *xi = x[g] + 3 * i;
    auto visit = [&](int32_t c) {
      if (c < 0)
        return;
      for (int32_t s = cell_range[c].first; s < cell_range[c].second; ++s) {
        const RawIndex &r = sorted_to_raw[s];
        real d[3];
        sub3(xi, x[r.group] + 3 * static_cast<size_t>(r.index), d);
        real dist = sqnorm3(d);
        if (static_cast<double>(dist) < r2)
          fn(r.group, static_cast<size_t>(r.index));
      }
    };
    int32_t c = sorted_to_cell[raw_to_sorted[g][i]];
    visit(c);
    for (int k = 0; k < 26; ++k)
      visit(cell_adj[c][k]);
  }

  // compute_group_permutations (:394-399)
  void group_permutations(std::array<std::vector<size_t>, 2> &perm) const {
    std::array<size_t, 2> it{0, 0};
    for (int g = 0; g < 2; ++g)
      perm[g].assign(n[g], 0);
    for (const auto &r : sorted_to_raw)
      perm[r.group][it[r.group]++] = static_cast<size_t>(r.index);
  }
};

// deterministic chunked sum (see the file header)
template <class Acc, class Fn>
Acc chunked_sum(size_t n, Fn fn) {
  const size_t chunk = 1024;
  size_t nchunks = (n + chunk - 1) / chunk;
  std::vector<Acc> part(nchunks, Acc{0});
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < nchunks; ++c) {
    Acc acc = 0;
    size_t e = std::min(n, (c + 1) * chunk);
    for (size_t i = c * chunk; i < e; ++i)
      acc += static_cast<Acc>(fn(i));
    part[c] = acc;
  }
  Acc total = 0;
  for (size_t c = 0; c < nchunks; ++c)
    total += part[c];
  return total;
}

// ---------------------------------------------------------------------------
// scene
// ---------------------------------------------------------------------------

struct Scene {
  oracle_params prm{};
  double clock = 0;  // VirtualClock<double> seconds (virtual_clock.hpp:8-35)
  real current_time = 0;
  real aprde = 0;
  int32_t nprde = 0;
  real max_speed = 0;
  int32_t wkbb18_iterations = 0;
  // pt16 fluid uniforms (pt16.hpp:160-175)
  real pt16_maximum_error = static_cast<real>(1e-3);
  int32_t pt16_maximum_iterations = 500;
  int32_t pt16_iterations = 0;
  real strain_rate_viscosity = 0;
  int steps_since_permute = 0;
  std::array<GroupData, 2> grp;
  Grid grid;
  bool grid_valid = false;

  // CG buffers (cg_openmp.hpp:239); function-local static in the reference
  std::vector<real> cg_x, cg_y, cg_r, cg_p, cg_q, cg_b;

  Scene(size_t nf, size_t nb) {
    grp[kFluid].n = nf;
    grp[kBoundary].n = nb;
    for (const char *nm : {"position", "velocity", "acceleration", "iisph_helper_c",
                           "wkbb18_velocity_delta"})
      grp[kFluid].add(nm, 3);
    for (const char *nm : {"density", "pressure", "mass", "iisph_source_term",
                           "iisph_diagonal_element", "iisph_right_hand_side",
                           "time_of_birth"})
      grp[kFluid].add(nm, 1);
    grp[kFluid].add("wkbb18_system_diagonal", 9);
    grp[kFluid].add("gradient_correction", 9);
    for (const char *nm : {"pt16_normalizer", "pt16_density", "pt16_same_density",
                           "pt16_vorticity_diffusion_diagonal",
                           "pt16_velocity_reconstruction_diagonal"})
      grp[kFluid].add(nm, 1);
    for (const char *nm : {"vorticity", "pt16_vorticity_diffusion_rhs",
                           "pt16_velocity_reconstruction_rhs"})
      grp[kFluid].add(nm, 3);
    grp[kFluid].add("target_velocity_gradient", 9);
    grp[kBoundary].add("position", 3);
    grp[kBoundary].add("volume", 1);
    for (int g = 0; g < 2; ++g) {
      grp[g].id.resize(grp[g].n);
      std::iota(grp[g].id.begin(), grp[g].id.end(), int64_t{0});
    }
    prm.radius = 1;
    prm.iisph_min_iterations = 2;
    prm.iisph_max_iterations = 50;
    prm.iisph_tolerance = 1e-3f;
    prm.permute_every = 8;
  }

  void bind_grid() {
    grid.radius = prm.radius;
    for (int g = 0; g < 2; ++g) {
      grid.x[g] = grp[g].ptr("position");
      grid.n[g] = grp[g].n;
    }
  }

  // Neighborhood::CopyNeighbors into per-thread per-group vectors
  void copy_neighbors(int g, size_t i, std::array<std::vector<size_t>, 2> &nb) const {
    nb[0].clear();
    nb[1].clear();
    grid.neighbors(g, i, [&nb](int ng, size_t ni) { nb[ng].push_back(ni); });
  }

  template <class Fn>
  void foreach_fluid(Fn fn) {
    const size_t n = grp[kFluid].n;
#pragma omp parallel
    {
      std::array<std::vector<size_t>, 2> nb;
      nb[0].reserve(100);
      nb[1].reserve(100);
#pragma omp for
      for (size_t i = 0; i < n; ++i) {
        copy_neighbors(kFluid, i, nb);
        fn(i, nb);
      }
    }
  }

  // ---- aiast12.compute_volume (aiast12.hpp:125-175)
  void compute_volume() {
    auto &B = grp[kBoundary];
    real *x = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale;
    const size_t n = B.n;
#pragma omp parallel
    {
      std::array<std::vector<size_t>, 2> nb;
#pragma omp for
      for (size_t i = 0; i < n; ++i) {
        copy_neighbors(kBoundary, i, nb);
        V[i] = 0;
        for (size_t j : nb[kBoundary]) {
          if (mutant<bool>(9, false, j == i))
            continue;
          real d[3];
          sub3(x + 3 * i, x + 3 * j, d);
          V[i] += kernel_w(d, h);
        }
        V[i] = real{1} / V[i];
      }
    }
  }

  // ---- density.compute_density (density.hpp:160-215)
  void compute_density() {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    real *x = F.ptr("position"), *rho = F.ptr("density"), *m = F.ptr("mass");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density;
    foreach_fluid([&](size_t i, auto &nb) {
      rho[i] = 0;
      for (size_t j : nb[kFluid]) {
        real d[3];
        sub3(x + 3 * i, x + 3 * j, d);
        rho[i] += m[j] * kernel_w(d, h);
      }
      for (size_t j : nb[kBoundary]) {
        real d[3];
        sub3(x + 3 * i, bx + 3 * j, d);
        rho[i] += (V[j] * rho0) * kernel_w(d, h);
      }
    });
  }

  // ---- gravity.initialize_acceleration (gravity.hpp:133)
  void initialize_acceleration() {
    real *a = grp[kFluid].ptr("acceleration");
    const size_t n = grp[kFluid].n;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k)
        a[3 * i + k] = prm.gravity[k];
  }

  // ---- gravity.accumulate_acceleration (gravity.hpp:157-186): a += g
  void accumulate_acceleration() {
    real *a = grp[kFluid].ptr("acceleration");
    const size_t n = grp[kFluid].n;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k)
        a[3 * i + k] += prm.gravity[k];
  }

  // ---- gravity.initialize_axial_acceleration (gravity.hpp:189-215;
  //      gravity.prtcl:26-33): the acceleration of magnitude |g| towards the
  //      axis through gravity_center along g, zero on the axis
  void initialize_axial_acceleration() {
    real *x = grp[kFluid].ptr("position"), *a = grp[kFluid].ptr("acceleration");
    const size_t n = grp[kFluid].n;
    const real *g = prm.gravity, *c = prm.gravity_center;
    const real eps = static_cast<real>(1e-9);
    auto rcp0 = [&](real v) { return std::abs(v) > eps ? static_cast<real>(1) / v : real(0); };
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
      real xc[3], gn[3], xo[3];
      sub3(x + 3 * i, c, xc);
      const real ng = norm3(g);
      const real rg = rcp0(ng);
      for (int k = 0; k < 3; ++k)
        gn[k] = g[k] * rg;
      const real d = dot3(xc, gn);
      for (int k = 0; k < 3; ++k)
        xo[k] = xc[k] - d * gn[k];
      const real ro = rcp0(norm3(xo));
      for (int k = 0; k < 3; ++k)
        a[3 * i + k] = ((-ng) * xo[k]) * ro;
    }
  }

  // ---- viscosity.accumulate_acceleration (viscosity.hpp:171-240)
  void viscosity_accumulate() {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    real *x = F.ptr("position"), *v = F.ptr("velocity"), *a = F.ptr("acceleration");
    real *rho = F.ptr("density"), *m = F.ptr("mass");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density;
    const real mu = prm.fluid_viscosity, mub = prm.boundary_viscosity;
    foreach_fluid([&](size_t i, auto &nb) {
      for (size_t j : nb[kFluid]) {
        real d[3], dv[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        sub3(v + 3 * i, mutant<const real *>(6, v + 3 * j, kZero3), dv);
        kernel_grad(d, h, gw);
        real s = (((static_cast<real>(10) * (mu / rho[i])) *
                   (m[j] / rho[mutant<size_t>(5, j, i)])) *
                  dot3(dv, d)) /
                 (sqnorm3(d) + ((static_cast<real>(0.01) * h) * h));
        for (int k = 0; k < 3; ++k)
          a[3 * i + k] += s * gw[k];
      }
      for (size_t j : nb[kBoundary]) {
        real d[3], gw[3];
        sub3(x + 3 * i, bx + 3 * j, d);
        kernel_grad(d, h, gw);
        real vi[3] = {v[3 * i] - 0, v[3 * i + 1] - 0, v[3 * i + 2] - 0};
        real s = (((static_cast<real>(10) * (mub / rho[i])) * ((rho0 * V[j]) / rho[i])) *
                  dot3(vi, d)) /
                 (sqnorm3(d) + ((static_cast<real>(0.01) * h) * h));
        for (int k = 0; k < 3; ++k)
          a[3 * i + k] += s * gw[k];
      }
    });
  }

  // ---- symplectic_euler.integrate_velocity_with_hard_fade (:172-200)
  void integrate_velocity_with_hard_fade() {
    auto &F = grp[kFluid];
    real *v = F.ptr("velocity"), *a = F.ptr("acceleration"), *tb = F.ptr("time_of_birth");
    const real dt = prm.time_step, t = current_time, fade = prm.fade_duration;
    const size_t n = F.n;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
      real H = unit_step_l(static_cast<real>(0), (t - tb[i]) - fade);
      for (int k = 0; k < 3; ++k)
        v[3 * i + k] += (dt * a[3 * i + k]) * H;
    }
  }

  // ---- symplectic_euler.integrate_position (:236-270); running max
  void integrate_position() {
    auto &F = grp[kFluid];
    real *x = F.ptr("position"), *v = F.ptr("velocity");
    const real dt = prm.time_step;
    const size_t n = F.n;
    real ms = max_speed;
#pragma omp parallel for schedule(static) reduction(max : ms)
    for (size_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k)
        x[3 * i + k] += dt * v[3 * i + k];
      ms = std::max(ms, norm3(v + 3 * i));
    }
    max_speed = ms;
  }

  // ---- iisph.setup (iisph.hpp:188-355)
  void iisph_setup() {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    real *x = F.ptr("position"), *v = F.ptr("velocity"), *rho = F.ptr("density");
    real *p = F.ptr("pressure"), *m = F.ptr("mass"), *c = F.ptr("iisph_helper_c");
    real *s = F.ptr("iisph_source_term"), *AA = F.ptr("iisph_diagonal_element");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density, dt = prm.time_step;

    // pass 1: helper c (:197-249)
    foreach_fluid([&](size_t i, auto &nb) {
      real *ci = c + 3 * i;
      ci[0] = ci[1] = ci[2] = 0;
      for (size_t j : nb[kFluid]) {
        real d[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        kernel_grad(d, h, gw);
        real f = m[j] / (rho[i] * rho[i]);
        for (int k = 0; k < 3; ++k)
          ci[k] -= f * gw[k];
      }
      for (size_t j : nb[kBoundary]) {
        real d[3], gw[3];
        sub3(x + 3 * i, bx + 3 * j, d);
        kernel_grad(d, h, gw);
        real f = (mutant<real>(1, static_cast<real>(0.7) * static_cast<real>(2),
                               static_cast<real>(0.7)) *
                  rho0) *
                 (V[j] / (rho[i] * rho[i]));
        for (int k = 0; k < 3; ++k)
          ci[k] -= f * gw[k];
      }
    });

    // pass 2: p, s, AA and nprde (:252-354)
    foreach_fluid([&](size_t i, auto &nb) {
      p[i] = 0;
      s[i] = 0;
      AA[i] = 0;
      for (size_t j : nb[kFluid]) {
        real d[3], dv[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        sub3(v + 3 * i, mutant<const real *>(8, v + 3 * j, kZero3), dv);
        kernel_grad(d, h, gw);
        s[i] += m[j] * dot3(dv, gw);
        AA[i] += m[j] * dot3(c + 3 * i, gw);
        AA[i] -= mutant<real>(2, 1, -1) * ((m[j] * (m[i] / (rho[i] * rho[i]))) * sqnorm3(gw));
      }
      for (size_t j : nb[kBoundary]) {
        real d[3], gw[3];
        sub3(x + 3 * i, bx + 3 * j, d);
        kernel_grad(d, h, gw);
        real vi[3] = {v[3 * i] - 0, v[3 * i + 1] - 0, v[3 * i + 2] - 0};
        s[i] += (rho0 * V[j]) * dot3(vi, gw);
        AA[i] += (rho0 * V[j]) * dot3(c + 3 * i, gw);
      }
      s[i] = (rho0 - rho[i]) - ((dt * (rho0 / rho[i])) * s[i]);
      AA[i] *= (dt * dt) * (rho0 / rho[i]);
    });
    nprde = static_cast<int32_t>(F.n);
  }

  // ---- iisph.iteration_pressure_acceleration (iisph.hpp:358-423)
  void iisph_pressure_acceleration() {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    real *x = F.ptr("position"), *a = F.ptr("acceleration"), *rho = F.ptr("density");
    real *p = F.ptr("pressure"), *m = F.ptr("mass");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density;
    foreach_fluid([&](size_t i, auto &nb) {
      real *ai = a + 3 * i;
      ai[0] = ai[1] = ai[2] = 0;
      for (size_t j : nb[kFluid]) {
        real d[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        kernel_grad(d, h, gw);
        const size_t jr = mutant<size_t>(7, j, i);
        real f = m[j] * ((p[i] / (rho[i] * rho[i])) + (p[j] / (rho[jr] * rho[jr])));
        for (int k = 0; k < 3; ++k)
          ai[k] -= f * gw[k];
      }
      for (size_t j : nb[kBoundary]) {
        real d[3], gw[3];
        sub3(x + 3 * i, bx + 3 * j, d);
        kernel_grad(d, h, gw);
        real f = ((static_cast<real>(0.7) * rho0) * V[j]) *
                 ((mutant<real>(3, 2, 1) * p[i]) / (rho[i] * rho[i]));
        for (int k = 0; k < 3; ++k)
          ai[k] -= f * gw[k];
      }
    });
  }

  // ---- iisph.iteration_pressure (iisph.hpp:426-516)
  void iisph_pressure() {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    real *x = F.ptr("position"), *a = F.ptr("acceleration"), *rho = F.ptr("density");
    real *p = F.ptr("pressure"), *m = F.ptr("mass"), *s = F.ptr("iisph_source_term");
    real *AA = F.ptr("iisph_diagonal_element"), *Ap = F.ptr("iisph_right_hand_side");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density, dt = prm.time_step;
    const real omega = prm.iisph_relaxation;
    foreach_fluid([&](size_t i, auto &nb) {
      Ap[i] = 0;
      for (size_t j : nb[kFluid]) {
        real d[3], da[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        sub3(a + 3 * i, a + 3 * j, da);
        kernel_grad(d, h, gw);
        Ap[i] += m[j] * dot3(da, gw);
      }
      for (size_t j : nb[kBoundary]) {
        real d[3], gw[3];
        sub3(x + 3 * i, bx + 3 * j, d);
        kernel_grad(d, h, gw);
        real ai[3] = {a[3 * i] - 0, a[3 * i + 1] - 0, a[3 * i + 2] - 0};
        Ap[i] += mutant<real>(4, 1, -1) * ((rho0 * V[j]) * dot3(ai, gw));
      }
      Ap[i] *= (dt * dt) * (rho0 / rho[i]);
      p[i] = std::max(static_cast<real>(0),
                      p[i] + ((omega * (s[i] - Ap[i])) *
                              reciprocal_or_zero(AA[i], static_cast<real>(1e-9)))) *
             unit_step_l(static_cast<real>(1e-9), std::abs(AA[i]));
    });
    // reduce aprde += ((Ap - s) / rho0) * [1e-9 < p]   (:505-507)
    auto term = [&](size_t i) {
      return ((Ap[i] - s[i]) / rho0) * unit_step_l(static_cast<real>(1e-9), p[i]);
    };
    if (prm.fp64_reductions)
      aprde = static_cast<real>(static_cast<double>(aprde) +
                                chunked_sum<double>(F.n, term));
    else
      aprde = aprde + chunked_sum<real>(F.n, term);
  }

  // ---- correction.identity_gradient_correction (correction.hpp:133-158)
  void correction_identity() {
    auto &F = grp[kFluid];
    real *L = F.ptr("gradient_correction");
    for (size_t i = 0; i < F.n; ++i)
      for (int k = 0; k < 9; ++k)
        L[9 * i + k] = (k % 4 == 0) ? real{1} : real{0};
  }

  // ---- correction.compute_gradient_correction (correction.hpp:161-217)
  void correction_compute() {
    auto &F = grp[kFluid];
    real *x = F.ptr("position"), *rho = F.ptr("density"), *m = F.ptr("mass");
    real *L = F.ptr("gradient_correction");
    const real h = prm.smoothing_scale;
    foreach_fluid([&](size_t i, auto &nb) {
      real M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (size_t j : nb[kFluid]) {  // every fluid group (:196)
        real d[3], e[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        sub3(x + 3 * j, x + 3 * i, e);
        kernel_grad(d, h, gw);
        const real k = m[j] / rho[i];
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row)
            M[row + 3 * col] += k * (gw[row] * e[col]);
      }
      pinv3(M, L + 9 * i);
    });
  }

  // ---- pt16 (pt16.hpp:213-760; pt16.prtcl)
  // setup: same-group and smoothed densities, the velocity gradient, its
  // spin / expansion / shear decomposition, vorticity and the target
  // velocity gradient without the spin
  void pt16_setup() {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    real *x = F.ptr("position"), *v = F.ptr("velocity"), *m = F.ptr("mass");
    real *rho = F.ptr("density"), *same = F.ptr("pt16_same_density");
    real *prho = F.ptr("pt16_density"), *om = F.ptr("vorticity");
    real *tvg = F.ptr("target_velocity_gradient");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density, xi = strain_rate_viscosity;
    foreach_fluid([&](size_t i, auto &nb) {
      real d[3];
      same[i] = 0;
      for (size_t j : nb[kFluid]) {  // own group (:246-255)
        sub3(x + 3 * i, x + 3 * j, d);
        same[i] += m[j] * kernel_w(d, h);
      }
      prho[i] = 0;
      for (size_t j : nb[kFluid]) {  // every fluid group (:259-269)
        sub3(x + 3 * i, x + 3 * j, d);
        prho[i] += (m[j] / rho0) * kernel_w(d, h);
      }
      for (size_t j : nb[kBoundary]) {
        sub3(x + 3 * i, bx + 3 * j, d);
        prho[i] += V[j] * kernel_w(d, h);
      }
      prho[i] *= rho0;
    });
    foreach_fluid([&](size_t i, auto &nb) {
      real vg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (size_t j : nb[kFluid]) {  // own group (:305-316)
        real d[3], dv[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        sub3(v + 3 * j, v + 3 * i, dv);
        kernel_grad(d, h, gw);
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r)
            vg[r + 3 * c] += m[j] * (dv[r] * gw[c]);
      }
      for (int k = 0; k < 9; ++k)
        vg[k] /= rho[i];
      const real div = vg[0] + (vg[4] + vg[8]);
      real R[9], Vm[9], S[9];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
          R[r + 3 * c] = (vg[r + 3 * c] - vg[c + 3 * r]) / static_cast<real>(2);
          Vm[r + 3 * c] = (div / static_cast<real>(3)) * ((r == c) ? real{1} : real{0});
          S[r + 3 * c] = ((vg[r + 3 * c] + vg[c + 3 * r]) / static_cast<real>(2)) - Vm[r + 3 * c];
        }
      // vector_from_cross_product_matrix: (R(2,1), R(0,2), R(1,0)) (eigen.hpp:341-346)
      om[3 * i + 0] = static_cast<real>(2) * R[2 + 3 * 1];
      om[3 * i + 1] = static_cast<real>(2) * R[0 + 3 * 2];
      om[3 * i + 2] = static_cast<real>(2) * R[1 + 3 * 0];
      // unit_step_r(0, a, b) = [0 <= a or 0 <= b] (eigen.hpp:320-328)
      const real step = (real{0} <= rho[i] - rho0 || real{0} <= -div) ? real{1} : real{0};
      for (int k = 0; k < 9; ++k)
        tvg[9 * i + k] = Vm[k] * step + xi * S[k];
    });
  }

  // tvg += cross_product_matrix_from_vector(0.5 w) (eigen.hpp:330-339)
  void pt16_spin() {
    auto &F = grp[kFluid];
    real *om = F.ptr("vorticity"), *tvg = F.ptr("target_velocity_gradient");
    for (size_t i = 0; i < F.n; ++i) {
      const real a0 = static_cast<real>(0.5) * om[3 * i], a1 = static_cast<real>(0.5) * om[3 * i + 1],
                 a2 = static_cast<real>(0.5) * om[3 * i + 2];
      const real C[9] = {0, a2, -a1, -a2, 0, a0, a1, -a0, 0};  // column-major
      for (int k = 0; k < 9; ++k)
        tvg[9 * i + k] += C[k];
    }
  }

  // CGOpenMP<T> (scalar) solve of cg_openmp.hpp:32-215 over the fluid:
  // system(out, w), precond(out, w), b, x = guess on entry; returns the loop
  // counter
  template <class Sys, class Pre>
  size_t scalar_cg(size_t n, Sys &&system, Pre &&precond, const std::vector<real> &b,
                   std::vector<real> &x, real tol, size_t max_k) {
    std::vector<real> r(n), y(n), p(n), q(n);
    auto dot = [&](const std::vector<real> &l, const std::vector<real> &rr) -> real {
      auto term = [&](size_t i) { return l[i] * rr[i]; };
      if (prm.fp64_reductions)
        return static_cast<real>(chunked_sum<double>(n, term));
      return chunked_sum<real>(n, term);
    };
    system(r.data(), x.data());
    for (size_t k = 0; k < n; ++k)
      r[k] -= b[k];
    const real b_nsq = dot(b, b);
    const real thr = std::max(tol * tol * b_nsq, std::numeric_limits<real>::min());
    precond(y.data(), r.data());
    for (size_t k = 0; k < n; ++k)
      p[k] = -y[k];
    real r_nsq = dot(r, r);
    real next_r_dot_y = dot(r, y);
    auto ok = [&](size_t k, real cur) {
      if (k < 1)
        return true;
      if (k > max_k)
        return false;
      return !(cur < thr);
    };
    size_t iter;
    for (iter = 0; ok(iter, r_nsq); ++iter) {
      const real prev_r_dot_y = next_r_dot_y;
      system(q.data(), p.data());
      const real p_dot_q = dot(p, q);
      if (std::abs(p_dot_q) < tol)
        break;
      const real alpha = prev_r_dot_y / p_dot_q;
      for (size_t k = 0; k < n; ++k) {
        x[k] += alpha * p[k];
        r[k] += alpha * q[k];
      }
      r_nsq = dot(r, r);
      precond(y.data(), r.data());
      next_r_dot_y = dot(r, y);
      if (std::abs(prev_r_dot_y) < tol)
        break;
      const real beta = next_r_dot_y / prev_r_dot_y;
      for (size_t k = 0; k < n; ++k)
        p[k] = -y[k] + beta * p[k];
    }
    return iter;
  }

  // the two pt16 solves; which 0: vorticity diffusion (diagonal coefficient
  // pt16_same_density, result into vorticity), 1: velocity reconstruction
  // (pt16_density, result into velocity); preconditioner
  // 1 / (pt16_same_density - m W(0)) for both
  void pt16_solve(int which) {
    auto &F = grp[kFluid];
    const size_t n = F.n;
    real *x = F.ptr("position"), *v = F.ptr("velocity"), *m = F.ptr("mass");
    real *same = F.ptr("pt16_same_density"), *prho = F.ptr("pt16_density");
    real *om = F.ptr("vorticity"), *tvg = F.ptr("target_velocity_gradient");
    real *orhs = F.ptr("pt16_vorticity_diffusion_rhs");
    real *vrhs = F.ptr("pt16_velocity_reconstruction_rhs"), *tb = F.ptr("time_of_birth");
    const real h = prm.smoothing_scale, xi = strain_rate_viscosity;
    const real zero3[3] = {0, 0, 0};
    const real w0 = kernel_w(zero3, h);
    if (which == 0) {
      foreach_fluid([&](size_t i, auto &nb) {  // :395-414
        real acc[3] = {0, 0, 0};
        for (size_t j : nb[kFluid]) {
          real d[3];
          sub3(x + 3 * i, x + 3 * j, d);
          const real w = kernel_w(d, h);
          for (int k = 0; k < 3; ++k)
            acc[k] += (m[j] * (om[3 * i + k] - om[3 * j + k])) * w;
        }
        for (int k = 0; k < 3; ++k)
          orhs[3 * i + k] = acc[k] * xi;
      });
    } else {
      foreach_fluid([&](size_t i, auto &nb) {  // :588-606
        real acc[3] = {0, 0, 0};
        for (size_t j : nb[kFluid]) {
          real d[3], M[9];
          sub3(x + 3 * i, x + 3 * j, d);
          for (int k = 0; k < 9; ++k)
            M[k] = tvg[9 * i + k] + tvg[9 * j + k];
          const real mh = m[j] * static_cast<real>(0.5);
          const real w = kernel_w(d, h);
          for (int r = 0; r < 3; ++r)
            acc[r] += (mh * (M[r] * d[0] + (M[r + 3] * d[1] + M[r + 6] * d[2]))) * w;
        }
        for (int k = 0; k < 3; ++k)
          vrhs[3 * i + k] = acc[k];
      });
    }
    real *field = which == 0 ? om : v;
    const real *rhs = which == 0 ? orhs : vrhs;
    const real *diag = which == 0 ? same : prho;
    const real t = current_time, fade = prm.fade_duration;
    for (int dim = 0; dim < 3; ++dim) {
      std::vector<real> b(n), xs(n);
      for (size_t i = 0; i < n; ++i) {
        b[i] = rhs[3 * i + dim];
        xs[i] = field[3 * i + dim];
      }
      auto system = [&](real *out, const real *w) {
        foreach_fluid([&](size_t i, auto &nb) {
          real acc = diag[i] * w[i];
          for (size_t j : nb[kFluid]) {
            real d[3];
            sub3(x + 3 * i, x + 3 * j, d);
            acc -= (m[j] * kernel_w(d, h)) * w[j];
          }
          out[i] = acc;
        });
      };
      auto precond = [&](real *out, const real *w) {
        for (size_t i = 0; i < n; ++i)
          out[i] = w[i] / (same[i] - m[i] * w0);
      };
      const size_t it = scalar_cg(n, system, precond, b, xs, pt16_maximum_error,
                                  static_cast<size_t>(std::max(pt16_maximum_iterations, 0)));
      for (size_t i = 0; i < n; ++i) {  // apply (:491-502, 700-717)
        const real age = t - tb[i];
        field[3 * i + dim] = xs[i] * unit_step_l(real{0}, age - fade) +
                             field[3 * i + dim] * ((real{0} <= fade - age) ? real{1} : real{0});
      }
      pt16_iterations = static_cast<int32_t>(it);
    }
    if (which == 0)
      pt16_spin();
  }

  // ---- wkbb18.compute_diagonal (wkbb18.hpp:194-270); gc: wkbb18_gc, the
  // gradient is L_i gradW (wkbb18_gc.hpp:236-262)
  void wkbb18_compute_diagonal(bool gc = false) {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    real *x = F.ptr("position"), *rho = F.ptr("density"), *m = F.ptr("mass");
    real *sp = F.ptr("wkbb18_system_diagonal");
    real *Lc = F.ptr("gradient_correction");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density, dt = prm.time_step;
    const real mu = prm.fluid_viscosity, mub = prm.boundary_viscosity;
    foreach_fluid([&](size_t i, auto &nb) {
      real D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      auto grad = [&](const real *d, real *gw) {
        kernel_grad(d, h, gw);
        if (gc) {
          const real g0[3] = {gw[0], gw[1], gw[2]};
          matvec3(Lc + 9 * i, g0, gw);
        }
      };
      for (size_t j : nb[kFluid]) {  // own group only (:230)
        real d[3], gw[3];
        sub3(x + 3 * i, x + 3 * j, d);
        grad(d, gw);
        real f = ((static_cast<real>(10) * (mu / rho[i])) * (m[j] / rho[j])) /
                 (sqnorm3(d) + ((static_cast<real>(0.01) * h) * h));
        const real *dc = mutant<const real *>(11, d, x + 3 * i);
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row)
            D[row + 3 * col] += f * (gw[row] * dc[col]);
      }
      for (size_t j : nb[kBoundary]) {
        real d[3], gw[3];
        sub3(x + 3 * i, bx + 3 * j, d);
        grad(d, gw);
        real f = ((static_cast<real>(10) * (mub / rho[i])) * ((rho0 * V[j]) / rho[i])) /
                 (sqnorm3(d) + ((static_cast<real>(0.01) * h) * h));
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row)
            D[row + 3 * col] += f * (gw[row] * d[col]);
      }
      real M[9];
      for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row)
          M[row + 3 * col] = ((row == col) ? real{1} : real{0}) - dt * D[row + 3 * col];
      invert3(M, sp + 9 * i);
    });
  }

  // ---- wkbb18.accumulate_acceleration (wkbb18.hpp:273-425) with the CG of
  // cg_openmp.hpp:32-236
  // gc: wkbb18_gc, every term is (s L_i) gradW (wkbb18_gc.hpp:362-390)
  void wkbb18_accumulate_acceleration(bool gc = false) {
    auto &F = grp[kFluid];
    auto &B = grp[kBoundary];
    const size_t n = F.n;
    real *Lc = F.ptr("gradient_correction");
    // l += (s L_i) gradW, or s gradW without the correction
    auto term = [&](size_t i, real s, const real *gw, real *l) {
      if (gc) {
        real sl[9], t3[3];
        for (int k = 0; k < 9; ++k)
          sl[k] = s * Lc[9 * i + k];
        matvec3(sl, gw, t3);
        for (int k = 0; k < 3; ++k)
          l[k] += t3[k];
      } else {
        for (int k = 0; k < 3; ++k)
          l[k] += s * gw[k];
      }
    };
    if (n == 0) {
      wkbb18_iterations = 0;
      return;
    }
    real *x = F.ptr("position"), *v = F.ptr("velocity"), *a = F.ptr("acceleration");
    real *rho = F.ptr("density"), *m = F.ptr("mass"), *vd = F.ptr("wkbb18_velocity_delta");
    real *sp = F.ptr("wkbb18_system_diagonal"), *tb = F.ptr("time_of_birth");
    real *bx = B.ptr("position"), *V = B.ptr("volume");
    const real h = prm.smoothing_scale, rho0 = prm.rest_density, dt = prm.time_step;
    const real mu = prm.fluid_viscosity, mub = prm.boundary_viscosity;
    const real tol = prm.wkbb18_maximum_error;
    const size_t max_k = static_cast<size_t>(prm.wkbb18_maximum_iterations);
    const real t = current_time, fade = prm.fade_duration;

    for (auto *vec : {&cg_x, &cg_y, &cg_r, &cg_p, &cg_q, &cg_b})
      vec->assign(3 * n, real{0});

    // product_s: out_i = w_i - dt * (sum_f ... + sum_b ...)   (:342-390)
    auto system = [&](real *out, const real *w) {
      foreach_fluid([&](size_t i, auto &nb) {
        real l[3] = {0, 0, 0};
        for (size_t j : nb[kFluid]) {
          real d[3], dw[3], gw[3];
          sub3(x + 3 * i, x + 3 * j, d);
          sub3(w + 3 * i, mutant<const real *>(10, w + 3 * j, kZero3), dw);
          kernel_grad(d, h, gw);
          real f = (((static_cast<real>(10) * (mu / rho[i])) * (m[j] / rho[j])) *
                    dot3(dw, d)) /
                   (sqnorm3(d) + ((static_cast<real>(0.01) * h) * h));
          term(i, f, gw, l);
        }
        for (size_t j : nb[kBoundary]) {
          real d[3], gw[3];
          sub3(x + 3 * i, bx + 3 * j, d);
          kernel_grad(d, h, gw);
          real f = (((static_cast<real>(10) * (mub / rho[i])) * ((rho0 * V[j]) / rho[i])) *
                    dot3(w + 3 * i, d)) /
                   (sqnorm3(d) + ((static_cast<real>(0.01) * h) * h));
          term(i, f, gw, l);
        }
        for (int k = 0; k < 3; ++k)
          out[3 * i + k] = w[3 * i + k] - dt * l[k];
      });
    };
    // product_p: out_i = sp_i * w_i  (column-major 3x3 times vector)
    auto precond = [&](real *out, const real *w) {
#pragma omp parallel for
      for (size_t i = 0; i < n; ++i) {
        const real *M = sp + 9 * i;
        const real *wi = w + 3 * i;
        for (int r = 0; r < 3; ++r)  // row r of M dotted with w (x0 + (x1 + x2))
          out[3 * i + r] = M[r] * wi[0] + (M[r + 3] * wi[1] + M[r + 6] * wi[2]);
      }
    };
    auto dot = [&](const std::vector<real> &l, const std::vector<real> &r) -> real {
      auto term = [&](size_t i) {
        const real *a3 = l.data() + 3 * i, *b3 = r.data() + 3 * i;
        return a3[0] * b3[0] + (a3[1] * b3[1] + a3[2] * b3[2]);  // sum(cmul)
      };
      if (prm.fp64_reductions)
        return static_cast<real>(chunked_sum<double>(n, term));
      return chunked_sum<real>(n, term);
    };

    // x = guess (= v); b = rhs (= v - dt * a_b with a_b = 0); r = A x - b
    std::copy(v, v + 3 * n, cg_x.begin());
    {
      // setup_r (:296-334): the boundary term uses l_v_f_b = 0, so its dot is 0
      std::vector<real> lab(3 * n, real{0});
      foreach_fluid([&](size_t i, auto &nb) {
        real l[3] = {0, 0, 0};
        const real zero3[3] = {0, 0, 0};
        for (size_t j : nb[kBoundary]) {
          real d[3], gw[3];
          sub3(x + 3 * i, bx + 3 * j, d);
          kernel_grad(d, h, gw);
          real f = (((static_cast<real>(10) * (mub / rho[i])) * ((rho0 * V[j]) / rho[i])) *
                    dot3(zero3, d)) /
                   (sqnorm3(d) + ((static_cast<real>(0.01) * h) * h));
          term(i, f, gw, l);
        }
        for (int k = 0; k < 3; ++k)
          cg_b[3 * i + k] = v[3 * i + k] - dt * l[k];
      });
    }
    system(cg_r.data(), cg_x.data());
    for (size_t k = 0; k < 3 * n; ++k)
      cg_r[k] -= cg_b[k];

    const real b_nsq = dot(cg_b, cg_b);
    const real thr = std::max(tol * tol * b_nsq, std::numeric_limits<real>::min());

    precond(cg_y.data(), cg_r.data());
    for (size_t k = 0; k < 3 * n; ++k)
      cg_p[k] = -cg_y[k];
    real r_nsq = dot(cg_r, cg_r);
    real next_r_dot_y = dot(cg_r, cg_y);

    const size_t min_k = 1;
    auto ok = [&](size_t k, real cur) {
      if (k < min_k)
        return true;
      if (k > max_k)
        return false;
      if (cur < thr)
        return false;
      return true;
    };
    size_t iter;
    for (iter = 0; ok(iter, r_nsq); ++iter) {
      const real prev_r_dot_y = next_r_dot_y;
      system(cg_q.data(), cg_p.data());
      const real p_dot_q = dot(cg_p, cg_q);
      if (std::abs(p_dot_q) < tol)
        break;
      const real alpha = prev_r_dot_y / p_dot_q;
#pragma omp parallel for schedule(static)
      for (size_t k = 0; k < 3 * n; ++k) {
        cg_x[k] += alpha * cg_p[k];
        cg_r[k] += alpha * cg_q[k];
      }
      r_nsq = dot(cg_r, cg_r);
      precond(cg_y.data(), cg_r.data());
      next_r_dot_y = dot(cg_r, cg_y);
      if (std::abs(prev_r_dot_y) < tol)
        break;
      const real beta = next_r_dot_y / prev_r_dot_y;
#pragma omp parallel for schedule(static)
      for (size_t k = 0; k < 3 * n; ++k)
        cg_p[k] = -cg_y[k] + beta * cg_p[k];
    }
    // apply (:401-413)
#pragma omp parallel for
    for (size_t i = 0; i < n; ++i) {
      real H = unit_step_l(static_cast<real>(0), (t - tb[i]) - fade);
      for (int k = 0; k < 3; ++k) {
        vd[3 * i + k] = (cg_x[3 * i + k] - v[3 * i + k]) * H;
        a[3 * i + k] = vd[3 * i + k] / dt;
      }
    }
    wkbb18_iterations = static_cast<int32_t>(iter);
  }

  // ---- Neighborhood::Update / Permute
  void nsearch_update() {
    bind_grid();
    grid.update();
    grid_valid = true;
  }
  void nsearch_permute() {
    std::array<std::vector<size_t>, 2> perm;
    grid.group_permutations(perm);
    for (int g = 0; g < 2; ++g) {
      auto &G = grp[g];
      const auto &pm = perm[g];
      for (auto &f : G.fields) {
        std::vector<real> old = f.data;
        const int nc = f.ncomp;
#pragma omp parallel for
        for (size_t i = 0; i < G.n; ++i)
          for (int k = 0; k < nc; ++k)
            f.data[nc * i + k] = old[nc * pm[i] + k];
      }
      std::vector<int64_t> old = G.id;
      for (size_t i = 0; i < G.n; ++i)
        G.id[i] = old[pm[i]];
    }
    nsearch_update();
  }

  int run(const std::string &proc) {
    if (proc == "aiast12.compute_volume")
      compute_volume();
    else if (proc == "density.compute_density")
      compute_density();
    else if (proc == "gravity.initialize_acceleration")
      initialize_acceleration();
    else if (proc == "gravity.initialize_axial_acceleration")
      initialize_axial_acceleration();
    else if (proc == "gravity.accumulate_acceleration")
      accumulate_acceleration();
    else if (proc == "viscosity.accumulate_acceleration")
      viscosity_accumulate();
    else if (proc == "symplectic_euler.integrate_velocity_with_hard_fade")
      integrate_velocity_with_hard_fade();
    else if (proc == "symplectic_euler.integrate_position")
      integrate_position();
    else if (proc == "iisph.setup")
      iisph_setup();
    else if (proc == "iisph.iteration_pressure_acceleration")
      iisph_pressure_acceleration();
    else if (proc == "iisph.iteration_pressure")
      iisph_pressure();
    else if (proc == "wkbb18.compute_diagonal")
      wkbb18_compute_diagonal();
    else if (proc == "wkbb18.accumulate_acceleration")
      wkbb18_accumulate_acceleration();
    else if (proc == "correction.identity_gradient_correction")
      correction_identity();
    else if (proc == "correction.compute_gradient_correction")
      correction_compute();
    else if (proc == "wkbb18_gc.compute_diagonal")
      wkbb18_compute_diagonal(true);
    else if (proc == "wkbb18_gc.accumulate_acceleration")
      wkbb18_accumulate_acceleration(true);
    else if (proc == "pt16.setup")
      pt16_setup();
    else if (proc == "pt16.solve_vorticity_diffusion")
      pt16_solve(0);
    else if (proc == "pt16.vorticity_preservation")
      pt16_spin();
    else if (proc == "pt16.solve_velocity_reconstruction")
      pt16_solve(1);
    else
      return -1;
    return 0;
  }

  // example.lua:363-462
  void step(oracle_step_info *info) {
    steps_since_permute += 1;
    current_time = static_cast<real>(clock);
    if (!grid_valid) {
      nsearch_update();
    } else {
      nsearch_update();
      if (prm.permute_every > 0 && steps_since_permute >= prm.permute_every) {
        nsearch_permute();
        steps_since_permute = 0;
      }
    }
    compute_density();
    if (prm.implicit_viscosity == 2)  // prtcl-lua.bunny.lua:279-281
      correction_compute();
    if (prm.axial_gravity)  // example.lua:390-391
      initialize_axial_acceleration();
    else
      initialize_acceleration();
    if (prm.explicit_viscosity)
      viscosity_accumulate();
    integrate_velocity_with_hard_fade();
    iisph_setup();
    int iters = 0;
    real rel = 0;
    if (nprde >= 1) {
      for (int it = 1; it <= prm.iisph_max_iterations; ++it) {
        iters = it;
        aprde = 0;
        iisph_pressure_acceleration();
        iisph_pressure();
        rel = static_cast<real>(static_cast<double>(aprde) / static_cast<double>(nprde));
        if (it >= prm.iisph_min_iterations && rel < prm.iisph_tolerance)
          break;
      }
    }
    integrate_velocity_with_hard_fade();
    if (prm.implicit_viscosity == 1 || prm.implicit_viscosity == 2) {  // wkbb18 / _gc
      wkbb18_compute_diagonal(prm.implicit_viscosity == 2);
      wkbb18_accumulate_acceleration(prm.implicit_viscosity == 2);
      integrate_velocity_with_hard_fade();
    } else if (prm.implicit_viscosity == 3) {  // pt16 (example.lua:435-445)
      pt16_setup();
      pt16_maximum_error = prm.pt16_vorticity_maximum_error;
      pt16_maximum_iterations = prm.pt16_vorticity_maximum_iterations;
      pt16_solve(0);
      pt16_maximum_error = prm.pt16_velocity_maximum_error;
      pt16_maximum_iterations = prm.pt16_velocity_maximum_iterations;
      pt16_solve(1);
    }
    integrate_position();
    clock += static_cast<double>(prm.time_step);
    if (info) {
      info->iisph_iterations = iters;
      info->iisph_relative_error = rel;
      info->wkbb18_iterations = prm.implicit_viscosity == 3 ? pt16_iterations
                                : prm.implicit_viscosity ? wkbb18_iterations : 0;
      info->maximum_speed = max_speed;
    }
  }
};

int group_of(const char *g) {
  if (!g)
    return -1;
  std::string s(g);
  if (s == "fluid" || s == "f")
    return kFluid;
  if (s == "boundary" || s == "b")
    return kBoundary;
  return -1;
}

template <class Query>
int64_t build_csr(Scene *sc, int g, int ng, int64_t *offsets, int64_t *indices,
                  Query query) {
  const size_t n = sc->grp[g].n;
  std::vector<std::vector<int64_t>> rows(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (size_t i = 0; i < n; ++i) {
    query(i, rows[i]);
    std::sort(rows[i].begin(), rows[i].end());
  }
  int64_t total = 0;
  for (size_t i = 0; i < n; ++i) {
    if (offsets)
      offsets[i] = total;
    total += static_cast<int64_t>(rows[i].size());
  }
  if (offsets)
    offsets[n] = total;
  if (indices)
    for (size_t i = 0; i < n; ++i)
      std::copy(rows[i].begin(), rows[i].end(), indices + offsets[i]);
  (void)ng;
  return total;
}

}  // namespace

extern "C" {

void *oracle_scene_create(int64_t n_fluid, int64_t n_boundary) {
  return new Scene(static_cast<size_t>(n_fluid), static_cast<size_t>(n_boundary));
}
void oracle_scene_destroy(void *scene) { delete static_cast<Scene *>(scene); }
void oracle_set_params(void *scene, const oracle_params *p) {
  auto *sc = static_cast<Scene *>(scene);
  sc->prm = *p;
  sc->grid.radius = p->radius;
  sc->strain_rate_viscosity = p->strain_rate_viscosity;
  sc->pt16_maximum_error = p->pt16_vorticity_maximum_error;
  sc->pt16_maximum_iterations = p->pt16_vorticity_maximum_iterations;
}
void oracle_set_time(void *scene, double t) { static_cast<Scene *>(scene)->clock = t; }
double oracle_get_time(void *scene) { return static_cast<Scene *>(scene)->clock; }

int oracle_set_field(void *scene, const char *group, const char *name, const void *data,
                     int64_t count) {
  auto *sc = static_cast<Scene *>(scene);
  int g = group_of(group);
  if (g < 0)
    return -1;
  auto &G = sc->grp[g];
  if (std::string(name) == "id") {
    if (count != static_cast<int64_t>(G.n))
      return -2;
    std::memcpy(G.id.data(), data, sizeof(int64_t) * G.n);
    return 0;
  }
  Field *f = G.find(name);
  if (!f)
    return -3;
  if (count != static_cast<int64_t>(f->data.size()))
    return -2;
  std::memcpy(f->data.data(), data, sizeof(real) * f->data.size());
  if (std::string(name) == "position")
    sc->grid_valid = false;
  return 0;
}

int oracle_get_field(void *scene, const char *group, const char *name, void *out,
                     int64_t count) {
  auto *sc = static_cast<Scene *>(scene);
  int g = group_of(group);
  if (g < 0)
    return -1;
  auto &G = sc->grp[g];
  if (std::string(name) == "id") {
    if (count != static_cast<int64_t>(G.n))
      return -2;
    std::memcpy(out, G.id.data(), sizeof(int64_t) * G.n);
    return 0;
  }
  Field *f = G.find(name);
  if (!f)
    return -3;
  if (count != static_cast<int64_t>(f->data.size()))
    return -2;
  std::memcpy(out, f->data.data(), sizeof(real) * f->data.size());
  return 0;
}

double oracle_get_global(void *scene, const char *name) {
  auto *sc = static_cast<Scene *>(scene);
  std::string s(name);
  if (s == "iisph_aprde")
    return sc->aprde;
  if (s == "iisph_nprde")
    return sc->nprde;
  if (s == "maximum_speed")
    return sc->max_speed;
  if (s == "wkbb18_iterations")
    return sc->wkbb18_iterations;
  if (s == "current_time")
    return sc->current_time;
  if (s == "pt16_maximum_error")
    return sc->pt16_maximum_error;
  if (s == "pt16_maximum_iterations")
    return sc->pt16_maximum_iterations;
  if (s == "pt16_iterations")
    return sc->pt16_iterations;
  if (s == "strain_rate_viscosity")
    return sc->strain_rate_viscosity;
  return std::nan("");
}

void oracle_set_global(void *scene, const char *name, double value) {
  auto *sc = static_cast<Scene *>(scene);
  std::string s(name);
  if (s == "iisph_aprde")
    sc->aprde = static_cast<real>(value);
  else if (s == "iisph_nprde")
    sc->nprde = static_cast<int32_t>(value);
  else if (s == "maximum_speed")
    sc->max_speed = static_cast<real>(value);
  else if (s == "current_time")
    sc->current_time = static_cast<real>(value);
  else if (s == "pt16_maximum_error")
    sc->pt16_maximum_error = static_cast<real>(value);
  else if (s == "pt16_maximum_iterations")
    sc->pt16_maximum_iterations = static_cast<int32_t>(value);
  else if (s == "strain_rate_viscosity")
    sc->strain_rate_viscosity = static_cast<real>(value);
}

void oracle_nsearch_update(void *scene) { static_cast<Scene *>(scene)->nsearch_update(); }
void oracle_nsearch_permute(void *scene) { static_cast<Scene *>(scene)->nsearch_permute(); }

int oracle_run(void *scene, const char *procedure) {
  return static_cast<Scene *>(scene)->run(procedure);
}

void oracle_step(void *scene, oracle_step_info *info) {
  static_cast<Scene *>(scene)->step(info);
}

int64_t oracle_neighbors_csr(void *scene, int g, int ng, int64_t *offsets,
                             int64_t *indices) {
  auto *sc = static_cast<Scene *>(scene);
  if (!sc->grid_valid)
    sc->nsearch_update();
  return build_csr(sc, g, ng, offsets, indices, [&](size_t i, std::vector<int64_t> &row) {
    sc->grid.neighbors(g, i, [&](int gg, size_t j) {
      if (gg == ng)
        row.push_back(static_cast<int64_t>(j));
    });
  });
}

int64_t oracle_brute_neighbors_csr(void *scene, int g, int ng, int64_t *offsets,
                                   int64_t *indices) {
  auto *sc = static_cast<Scene *>(scene);
  const real *xg = sc->grp[g].ptr("position");
  const real *xn = sc->grp[ng].ptr("position");
  const size_t nn = sc->grp[ng].n;
  const double r = sc->prm.radius;
  const double r2 = (1.0 * r) * r;
  return build_csr(sc, g, ng, offsets, indices, [&](size_t i, std::vector<int64_t> &row) {
    for (size_t j = 0; j < nn; ++j) {
      real d[3];
      sub3(xg + 3 * i, xn + 3 * j, d);
      if (static_cast<double>(sqnorm3(d)) < r2)
        row.push_back(static_cast<int64_t>(j));
    }
  });
}

int oracle_morton_compare(const int32_t *lhs, const int32_t *rhs) {
  GridIndex a{lhs[0], lhs[1], lhs[2]}, b{rhs[0], rhs[1], rhs[2]};
  return morton_compare(a, b);
}

float oracle_kernel_w(const float *v, float h) { return kernel_w(v, h); }
void oracle_kernel_grad(const float *v, float h, float *out) { kernel_grad(v, h, out); }
void oracle_invert3(const float *m, float *out) { invert3(m, out); }
void oracle_pinv3(const float *m, float *out) { pinv3(m, out); }

int64_t oracle_sorted_order(void *scene, int32_t *group, int64_t *index) {
  auto *sc = static_cast<Scene *>(scene);
  const auto &s2r = sc->grid.sorted_to_raw;
  if (group && index)
    for (size_t s = 0; s < s2r.size(); ++s) {
      group[s] = s2r[s].group;
      index[s] = s2r[s].index;
    }
  return static_cast<int64_t>(s2r.size());
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
