// Device-resident scene (see scene_impl.hpp): HBM arena, grid build,
// host-index <-> slot transfers, procedure dispatch and the IISPH loop with a
// device-side convergence decision.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>

#include "scene_impl.hpp"
#include "sph.hpp"
#include "fused.hpp"
#include "wkbb18.hpp"
#include "pt16.hpp"

namespace prtcl_amd {

void *CtxResources::ensure_staging(size_t bytes) {
  if (bytes > staging_bytes) {
    if (staging)
      PRTCL_HIP_CHECK(hipFree(staging));
    staging = nullptr;
    staging_bytes = 0;
    PRTCL_HIP_CHECK(hipMalloc(&staging, bytes));
    staging_bytes = bytes;
  }
  return staging;
}

void CtxResources::ensure_batch(size_t bytes) {
  if (!batch_ev)
    PRTCL_HIP_CHECK(hipEventCreateWithFlags(&batch_ev, hipEventDisableTiming));
  if (bytes <= batch_bytes)
    return;
  // kernels of earlier batches may still read the old tables, and the
  // captured batch iterations point at them
  PRTCL_HIP_CHECK(hipStreamSynchronize(stream));
  for (GraphCache &g : batch_graphs)
    g.clear();
  if (batch_host)
    PRTCL_HIP_CHECK(hipHostFree(batch_host));
  if (batch_dev)
    PRTCL_HIP_CHECK(hipFree(batch_dev));
  batch_host = batch_dev = nullptr;
  batch_bytes = 0;
  const size_t b = std::max<size_t>(bytes, 64 << 10);
  PRTCL_HIP_CHECK(hipHostMalloc(&batch_host, b, hipHostMallocDefault));
  PRTCL_HIP_CHECK(hipMalloc(&batch_dev, b));
  batch_bytes = b;
}

void CtxResources::release() {
  for (GraphCache &g : batch_graphs)
    g.clear();
  if (batch_host)
    (void)hipHostFree(batch_host);
  if (batch_dev)
    (void)hipFree(batch_dev);
  if (batch_ev)
    (void)hipEventDestroy(batch_ev);
  batch_host = batch_dev = nullptr;
  batch_ev = nullptr;
  batch_bytes = 0;
}

void GraphCache::clear() {
  for (Entry &e : entries_) {
    if (e.exec)
      (void)hipGraphExecDestroy(e.exec);
    if (e.graph)
      (void)hipGraphDestroy(e.graph);
  }
  entries_.clear();
}

void GraphCache::launch(hipStream_t st, const std::string &key,
                        const std::function<void()> &enqueue) {
  Entry *hit = nullptr;
  for (Entry &e : entries_)
    if (e.key == key)
      hit = &e;
  if (!hit) {
    if (entries_.size() >= kMaxEntries) {
      // evict the least recently used graph once nothing is in flight
      PRTCL_HIP_CHECK(hipStreamSynchronize(st));
      auto lru = std::min_element(entries_.begin(), entries_.end(),
                                  [](const Entry &a, const Entry &b) { return a.used < b.used; });
      (void)hipGraphExecDestroy(lru->exec);
      (void)hipGraphDestroy(lru->graph);
      entries_.erase(lru);
    }
    Entry e;
    e.key = key;
    PRTCL_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g)
        (void)hipGraphDestroy(g);
      throw;
    }
    PRTCL_HIP_CHECK(hipStreamEndCapture(st, &e.graph));
    hipError_t err = hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0);
    if (err != hipSuccess) {
      (void)hipGraphDestroy(e.graph);
      throw HipError{err, "hipGraphInstantiate"};
    }
    entries_.push_back(e);
    hit = &entries_.back();
  }
  hit->used = ++clock_;
  PRTCL_HIP_CHECK(hipGraphLaunch(hit->exec, st));
}

// tags of the captured launch sequences (or-ed with the element size)
constexpr uint32_t kGraphIteration = 0x100u;       // one scene's IISPH iteration
constexpr uint32_t kGraphBatchIteration = 0x200u;  // a batch's IISPH iteration

const char *stage_name(int s) {
  static const char *names[kStageCount] = {
      "nsearch_sort",      "density",       "gravity_integrate_v", "viscosity",
      "iisph_setup",       "iisph_pressure_acceleration", "iisph_pressure",
      "integrate_position", "wkbb18_diagonal", "pcg_spmv", "pcg_vector",
      "gradient_correction", "wkbb18_diagonal_setup", "pt16_setup"};
  return (s >= 0 && s < kStageCount) ? names[s] : "?";
}

hipEvent_t Profiler::get() {
  if (pool.empty()) {
    hipEvent_t e;
    PRTCL_HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = pool.back();
  pool.pop_back();
  return e;
}

int Profiler::begin(hipStream_t st, int stage, int tag) {
  if (!on)
    return -1;
  Rec r{stage, tag, get(), get()};
  PRTCL_HIP_CHECK(hipEventRecord(r.a, st));
  recs.push_back(r);
  return static_cast<int>(recs.size()) - 1;
}

void Profiler::end(hipStream_t st, int rec) {
  if (!on || rec < 0)
    return;
  PRTCL_HIP_CHECK(hipEventRecord(recs[rec].b, st));
}

void Profiler::resolve(int valid_iters) {
  for (auto &r : recs) {
    PRTCL_HIP_CHECK(hipEventSynchronize(r.b));
    if (r.tag < 0 || r.tag < valid_iters) {
      float t = 0;
      PRTCL_HIP_CHECK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.stage] += t;
      launches[r.stage] += 1;
    }
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  recs.clear();
}

void Profiler::reset() {
  resolve();
  for (int k = 0; k < kStageCount; ++k) {
    ms[k] = 0;
    launches[k] = 0;
  }
}

namespace {

// ---- transfer kernels: host order (index id[s]) <-> slot order ------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_upload(T *dst, int stride, int off,
                                                   const T *src, int ncomp,
                                                   const int32_t *id, int n) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n)
    return;
  const long h = id ? id[s] : s;
  for (int k = 0; k < ncomp; ++k)
    dst[static_cast<long>(s) * stride + off + k] = src[h * ncomp + k];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_download(T *dst, int ncomp, const T *src,
                                                     int stride, int off, const int32_t *id,
                                                     int n) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n)
    return;
  const long h = id ? id[s] : s;
  for (int k = 0; k < ncomp; ++k)
    dst[h * ncomp + k] = src[static_cast<long>(s) * stride + off + k];
}

__global__ void k_iota(int32_t *id, int n) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s < n)
    id[s] = s;
}

// a permute: the ids move to keep (the permutation the host may ask for)
// and the slots become the new ids
__global__ void k_take_ids(int32_t *id, int32_t *keep, int n) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s < n) {
    keep[s] = id[s];
    id[s] = s;
  }
}

// control-block operations (kernels, not memsets, so they can live in graphs)
enum CtlOp { kCtlInit = 0, kCtlBeginSolve = 1, kCtlSetNprde = 2, kCtlSetAprde = 3,
             kCtlSetMaxSpeed = 4, kCtlSetNprdeValue = 5, kCtlBeginSolveAfterFirst = 6 };

__global__ void k_ctl(Ctl *c, int op, double dv, int i0, int i1, int i2, int i3,
                      unsigned long long *mirror) {
  switch (op) {
    case kCtlInit:  // the arena is zero-filled at allocation
      c->host_mirror = mirror;
      break;
    case kCtlBeginSolve:  // i3 >= 0: nprde first, in the same launch
    case kCtlBeginSolveAfterFirst:
      if (i3 >= 0)
        c->nprde = i3;
      begin_solve(c, dv, i0, i1, static_cast<unsigned>(i2), op == kCtlBeginSolveAfterFirst ? 1 : 0);
      break;
    case kCtlSetNprde:
    case kCtlSetNprdeValue:
      c->nprde = i0;
      break;
    case kCtlSetAprde:
      c->aprde = dv;
      break;
    case kCtlSetMaxSpeed:
      c->max_speed_bits_f = __float_as_uint(static_cast<float>(dv));
      c->max_speed_bits_d = static_cast<unsigned long long>(__double_as_longlong(dv));
      break;
  }
}

// the convergence step of IISPH iteration 1 cannot end a solve of at least 2
// iterations with a cap of at least 2 (example.lua:418-419)
inline bool first_iteration_cannot_stop(int min_it, int max_it) {
  return min_it >= 2 && max_it >= 2;
}

// nprde and the start of the IISPH solve of every scene of a batch (one
// thread per scene)
template <class T>
__global__ void k_begin_solve_b(const BatchEntry<T> *tab, unsigned nscenes) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nscenes)
    return;
  const BatchEntry<T> &E = tab[k];
  E.S.ctl->nprde = E.nprde;
  begin_solve(E.S.ctl, E.tol, E.min_it, E.max_it, E.seq, E.first_preset);
}

struct Arena {
  size_t off = 0;
  template <class P>
  void take(P *&ptr, size_t count) {
    off = (off + 255) & ~size_t(255);
    ptr = reinterpret_cast<P *>(static_cast<uintptr_t>(off));
    off += sizeof(P) * count;
  }
  template <class P>
  static void rebase(P *&ptr, char *base) {
    ptr = reinterpret_cast<P *>(base + reinterpret_cast<uintptr_t>(ptr));
  }
};

template <class T>
T pred_threshold(double radius) {
  const double R = (1.0 * radius) * radius;  // constpow(radius, 2)
  if constexpr (sizeof(T) == 4) {
    float f = static_cast<float>(R);
    if (static_cast<double>(f) < R)
      f = std::nextafterf(f, std::numeric_limits<float>::infinity());
    return f;  // d < R  <=>  d < f for every float d
  } else {
    return R;
  }
}

template <class T>
T kernel_alpha_t() {
  return static_cast<T>(
      1.L / 4.L *
      0.318309886183790671537767526745028724068919291480912897495334688117793595L);
}

template <class T>
class SceneImpl final : public SceneBase {
 public:
  using V4 = vec4_t<T>;

  SceneImpl(const prtcl_hip_scene_desc &d, CtxResources *r) : res_(r) {
    desc = d;
    nf_ = static_cast<int>(d.n_fluid);
    nb_ = static_cast<int>(d.n_boundary);
    const T h = static_cast<T>(d.smoothing_scale);
    radius_ = d.radius > 0 ? d.radius : 2.0 * static_cast<double>(h);
    pt16_xi_ = d.pt16_strain_rate_viscosity;
    pt16_tol_ = d.pt16_vorticity_maximum_error;
    pt16_max_k_ = d.pt16_vorticity_maximum_iterations;
    allocate();
  }

  ~SceneImpl() override {
    if (mirror_host_)
      (void)hipHostFree(mirror_host_);
    if (perm_host_)
      (void)hipHostFree(perm_host_);
    if (arena_)
      (void)hipFree(arena_);
    if (grid_mem_)
      (void)hipFree(grid_mem_);
  }

  static uint64_t particle_bytes(const prtcl_hip_scene_desc &d) {
    SceneImpl<T> *none = nullptr;
    (void)none;
    Arena A;
    Ptrs P;
    layout(A, P, static_cast<int>(d.n_fluid), static_cast<int>(d.n_boundary),
           d.implicit_viscosity != 0, d.implicit_viscosity == 3);
    return A.off;
  }

  uint64_t grid_bytes_for_domain() const override { return grid_bytes_; }

  // ---------------------------------------------------------------- transfer
  struct Loc {
    T *base;
    int stride, off, ncomp;
  };
  Loc loc(int group, int field) {
    T *tp = nullptr;
    if (group == PRTCL_HIP_FLUID) {
      switch (field) {
        case PRTCL_HIP_F_POSITION: return {reinterpret_cast<T *>(P_.xm), 4, 0, 3};
        case PRTCL_HIP_F_VELOCITY: return {reinterpret_cast<T *>(P_.v), 4, 0, 3};
        case PRTCL_HIP_F_ACCELERATION: return {reinterpret_cast<T *>(P_.a), 4, 0, 3};
        case PRTCL_HIP_F_DENSITY: return {P_.rho, 1, 0, 1};
        case PRTCL_HIP_F_PRESSURE: return {P_.p, 1, 0, 1};
        case PRTCL_HIP_F_MASS: return {reinterpret_cast<T *>(P_.xm), 4, 3, 1};
        case PRTCL_HIP_F_IISPH_HELPER_C: return {reinterpret_cast<T *>(P_.c), 4, 0, 3};
        case PRTCL_HIP_F_IISPH_SOURCE_TERM: return {P_.s, 1, 0, 1};
        case PRTCL_HIP_F_IISPH_DIAGONAL_ELEMENT: return {P_.AA, 1, 0, 1};
        case PRTCL_HIP_F_IISPH_RIGHT_HAND_SIDE: return {P_.Ap, 1, 0, 1};
        case PRTCL_HIP_F_TIME_OF_BIRTH: return {reinterpret_cast<T *>(P_.v), 4, 3, 1};
        case PRTCL_HIP_F_WKBB18_VELOCITY_DELTA: return {reinterpret_cast<T *>(P_.vd), 4, 0, 3};
        case PRTCL_HIP_F_WKBB18_SYSTEM_DIAGONAL: return {P_.sp, 9, 0, 9};
        case PRTCL_HIP_F_GRADIENT_CORRECTION: return {P_.gc, 9, 0, 9};
        default: break;
      }
      if (desc.implicit_viscosity == 3) {
        switch (field) {
          case PRTCL_HIP_F_PT16_NORMALIZER: return {P_.pone, 1, 0, 1};
          case PRTCL_HIP_F_PT16_DENSITY: return {P_.prho, 1, 0, 1};
          case PRTCL_HIP_F_PT16_SAME_DENSITY: return {P_.psrho, 1, 0, 1};
          case PRTCL_HIP_F_VORTICITY: return {reinterpret_cast<T *>(P_.om), 4, 0, 3};
          case PRTCL_HIP_F_TARGET_VELOCITY_GRADIENT: return {P_.tvg, 9, 0, 9};
          case PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_DIAGONAL: return {P_.odiag, 1, 0, 1};
          case PRTCL_HIP_F_PT16_VORTICITY_DIFFUSION_RHS:
            return {reinterpret_cast<T *>(P_.orhs), 4, 0, 3};
          case PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_DIAGONAL: return {P_.vdiag, 1, 0, 1};
          case PRTCL_HIP_F_PT16_VELOCITY_RECONSTRUCTION_RHS:
            return {reinterpret_cast<T *>(P_.vrhs), 4, 0, 3};
          default: break;
        }
      }
      switch (field) {
        default: break;
      }
    } else if (group == PRTCL_HIP_BOUNDARY) {
      switch (field) {
        case PRTCL_HIP_F_POSITION: return {reinterpret_cast<T *>(P_.bxV), 4, 0, 3};
        case PRTCL_HIP_F_VOLUME: return {reinterpret_cast<T *>(P_.bxV), 4, 3, 1};
        default: break;
      }
    }
    return {tp, 0, 0, 0};
  }

  void upload(int group, int field, const void *host, size_t count) override {
    ensure_resident();
    Loc L = loc(group, field);
    const int n = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    if (!L.base)
      throw std::invalid_argument("field does not exist for this group");
    if (count != static_cast<size_t>(n) * L.ncomp)
      throw std::invalid_argument("count does not match the group size");
    if (n == 0)
      return;
    hipStream_t st = res_->stream;
    void *stg = res_->ensure_staging(sizeof(T) * count);
    PRTCL_HIP_CHECK(hipMemcpyAsync(stg, host, sizeof(T) * count, hipMemcpyHostToDevice, st));
    const int32_t *ids = group == PRTCL_HIP_FLUID ? P_.id : P_.bid;
    hipLaunchKernelGGL(k_upload<T>, dim3(grid_blocks(n)), dim3(kBlock), 0, st, L.base, L.stride,
                       L.off, static_cast<const T *>(stg), L.ncomp, ids, n);
    PRTCL_HIP_CHECK(hipGetLastError());
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
    if (field == PRTCL_HIP_F_POSITION) {
      // host-side bounding box for the grid domain
      const T *hp = static_cast<const T *>(host);
      double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
          double x = static_cast<double>(hp[3 * i + k]);
          if (std::isfinite(x)) {
            lo[k] = std::min(lo[k], x);
            hi[k] = std::max(hi[k], x);
          }
        }
      auto &B = group == PRTCL_HIP_FLUID ? bb_f_ : bb_b_;
      for (int k = 0; k < 3; ++k) {
        B.lo[k] = lo[k];
        B.hi[k] = hi[k];
      }
      B.valid = lo[0] <= hi[0];
      if (group == PRTCL_HIP_BOUNDARY)
        bgrid_dirty_ = true;
      else
        fgrid_dirty_ = true;
    }
  }

  void download(int group, int field, void *host, size_t count) override {
    ensure_resident();
    Loc L = loc(group, field);
    const int n = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    if (!L.base)
      throw std::invalid_argument("field does not exist for this group");
    if (count != static_cast<size_t>(n) * L.ncomp)
      throw std::invalid_argument("count does not match the group size");
    if (n == 0)
      return;
    hipStream_t st = res_->stream;
    void *stg = res_->ensure_staging(sizeof(T) * count);
    const int32_t *ids = group == PRTCL_HIP_FLUID ? P_.id : P_.bid;
    hipLaunchKernelGGL(k_download<T>, dim3(grid_blocks(n)), dim3(kBlock), 0, st,
                       static_cast<T *>(stg), L.ncomp, L.base, L.stride, L.off, ids, n);
    PRTCL_HIP_CHECK(hipGetLastError());
    PRTCL_HIP_CHECK(hipMemcpyAsync(host, stg, sizeof(T) * count, hipMemcpyDeviceToHost, st));
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  }

  int field_components(int group, int field) override {
    ensure_resident();
    if (group != PRTCL_HIP_FLUID && group != PRTCL_HIP_BOUNDARY)
      return 0;
    // the WKBB18 arrays exist in every scene (scratch of other passes), but
    // only the viscous schemes bind them as fields
    const int iv = desc.implicit_viscosity;
    if ((field == PRTCL_HIP_F_WKBB18_VELOCITY_DELTA ||
         field == PRTCL_HIP_F_WKBB18_SYSTEM_DIAGONAL) && iv != 1 && iv != 2)
      return 0;
    if (field == PRTCL_HIP_F_GRADIENT_CORRECTION && iv != 2)
      return 0;
    const Loc L = loc(group, field);
    return L.base ? L.ncomp : 0;
  }

  void download_async(int group, int field, void *pinned, size_t count) override {
    ensure_resident();
    Loc L = loc(group, field);
    const int n = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    if (!L.base)
      throw std::invalid_argument("field does not exist for this group");
    if (count != static_cast<size_t>(n) * L.ncomp)
      throw std::invalid_argument("count does not match the group size");
    if (n == 0)
      return;
    hipStream_t st = res_->stream;
    // the staging buffer is reused in stream order: this kernel runs after
    // the previous copy out of it has finished
    if (res_->staging_bytes < sizeof(T) * count)
      throw std::logic_error("staging buffer smaller than the field");
    const int32_t *ids = group == PRTCL_HIP_FLUID ? P_.id : P_.bid;
    hipLaunchKernelGGL(k_download<T>, dim3(grid_blocks(n)), dim3(kBlock), 0, st,
                       static_cast<T *>(res_->staging), L.ncomp, L.base, L.stride, L.off, ids, n);
    PRTCL_HIP_CHECK(hipGetLastError());
    PRTCL_HIP_CHECK(hipMemcpyAsync(pinned, res_->staging, sizeof(T) * count,
                                   hipMemcpyDeviceToHost, st));
  }

  void gather_device(int group, int field, void *dev, size_t count) override {
    ensure_resident();
    Loc L = loc(group, field);
    const int n = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    if (!L.base)
      throw std::invalid_argument("field does not exist for this group");
    if (count != static_cast<size_t>(n) * L.ncomp)
      throw std::invalid_argument("count does not match the group size");
    if (n == 0)
      return;
    const int32_t *ids = group == PRTCL_HIP_FLUID ? P_.id : P_.bid;
    hipLaunchKernelGGL(k_download<T>, dim3(grid_blocks(n)), dim3(kBlock), 0, res_->stream,
                       static_cast<T *>(dev), L.ncomp, L.base, L.stride, L.off, ids, n);
    PRTCL_HIP_CHECK(hipGetLastError());
  }

  void slots(int group, int32_t *slot_of_host) override {
    ensure_resident();
    const int n = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    std::vector<int32_t> id(n);
    if (n > 0) {
      PRTCL_HIP_CHECK(hipMemcpyAsync(id.data(), group == PRTCL_HIP_FLUID ? P_.id : P_.bid,
                                     4 * static_cast<size_t>(n), hipMemcpyDeviceToHost,
                                     res_->stream));
      PRTCL_HIP_CHECK(hipStreamSynchronize(res_->stream));
    }
    for (int s = 0; s < n; ++s)
      slot_of_host[id[s]] = s;
  }

  void set_slots(int group, const int32_t *slot_of_host) override {
    ensure_resident();
    const int n = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    std::vector<int32_t> id(n, -1);
    for (int h = 0; h < n; ++h) {
      const int32_t s = slot_of_host[h];
      if (s < 0 || s >= n || id[s] >= 0)
        throw std::invalid_argument("slots are not a permutation of the particles");
      id[s] = h;
    }
    if (n > 0) {
      PRTCL_HIP_CHECK(hipMemcpyAsync(group == PRTCL_HIP_FLUID ? P_.id : P_.bid, id.data(),
                                     4 * static_cast<size_t>(n), hipMemcpyHostToDevice,
                                     res_->stream));
      PRTCL_HIP_CHECK(hipStreamSynchronize(res_->stream));
    }
    if (group == PRTCL_HIP_FLUID)
      fgrid_dirty_ = true;
    else
      bgrid_dirty_ = true;
  }

  double exact_clock() const override { return clock_; }
  void set_exact_clock(double t) override { clock_ = t; }
  int steps_since_permute() const override { return steps_since_permute_; }
  void set_steps_since_permute(int k) override { steps_since_permute_ = k; }

  // ------------------------------------------------------------------ globals
  Ctl read_ctl() {
    Ctl c;
    PRTCL_HIP_CHECK(hipMemcpyAsync(&c, P_.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, res_->stream));
    PRTCL_HIP_CHECK(hipStreamSynchronize(res_->stream));
    return c;
  }
  void ctl_op(int op, double dv = 0, int i0 = 0, int i1 = 0, int i2 = 0, int i3 = -1) {
    hipLaunchKernelGGL(k_ctl, dim3(1), dim3(1), 0, res_->stream, P_.ctl, op, dv, i0, i1, i2,
                       i3, mirror_dev_);
  }

  double get_global(int g) override {
    ensure_resident();
    if (g == PRTCL_HIP_G_CURRENT_TIME)
      return static_cast<double>(static_cast<T>(clock_));
    Ctl c = read_ctl();
    switch (g) {
      case PRTCL_HIP_G_IISPH_APRDE: return c.aprde;
      case PRTCL_HIP_G_IISPH_NPRDE: return c.nprde;
      case PRTCL_HIP_G_MAXIMUM_SPEED: {
        if constexpr (sizeof(T) == 4) {
          float f;
          std::memcpy(&f, &c.max_speed_bits_f, 4);
          return f;
        } else {
          double d;
          std::memcpy(&d, &c.max_speed_bits_d, 8);
          return d;
        }
      }
      case PRTCL_HIP_G_WKBB18_ITERATIONS: return c.cg_iter;
      case PRTCL_HIP_G_IISPH_ITERATIONS: return c.iters;
      case PRTCL_HIP_G_IISPH_REL_ERROR: return c.rel;
      case PRTCL_HIP_G_PT16_MAXIMUM_ERROR: return static_cast<double>(static_cast<T>(pt16_tol_));
      case PRTCL_HIP_G_PT16_MAXIMUM_ITERATIONS: return pt16_max_k_;
      case PRTCL_HIP_G_PT16_ITERATIONS: return pt16_iters_;
      case PRTCL_HIP_G_STRAIN_RATE_VISCOSITY: return static_cast<double>(static_cast<T>(pt16_xi_));
      default: throw std::out_of_range("global does not exist");
    }
  }
  void set_global(int g, double v) override {
    ensure_resident();
    switch (g) {
      case PRTCL_HIP_G_CURRENT_TIME: clock_ = v; break;
      case PRTCL_HIP_G_IISPH_APRDE: ctl_op(kCtlSetAprde, v); break;
      case PRTCL_HIP_G_IISPH_NPRDE: ctl_op(kCtlSetNprdeValue, 0, static_cast<int>(v)); break;
      case PRTCL_HIP_G_MAXIMUM_SPEED: ctl_op(kCtlSetMaxSpeed, v); break;
      case PRTCL_HIP_G_PT16_MAXIMUM_ERROR: pt16_tol_ = v; break;
      case PRTCL_HIP_G_PT16_MAXIMUM_ITERATIONS: pt16_max_k_ = static_cast<int>(v); break;
      case PRTCL_HIP_G_PT16_ITERATIONS: pt16_iters_ = static_cast<int>(v); break;
      case PRTCL_HIP_G_STRAIN_RATE_VISCOSITY: pt16_xi_ = v; break;
      default: throw std::out_of_range("global cannot be set");
    }
    PRTCL_HIP_CHECK(hipGetLastError());
  }

  // ------------------------------------------------------------------- grid
  // dense cell box around the uploaded particles (one cell of margin)
  GridDims plan_grid() const {
    double lo[3], hi[3];
    bool any = false;
    for (int k = 0; k < 3; ++k) {
      lo[k] = HUGE_VAL;
      hi[k] = -HUGE_VAL;
    }
    for (auto *B : {&bb_f_, &bb_b_})
      if (B->valid) {
        any = true;
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::min(lo[k], B->lo[k]);
          hi[k] = std::max(hi[k], B->hi[k]);
        }
      }
    if (!any)
      for (int k = 0; k < 3; ++k)
        lo[k] = hi[k] = 0.0;
    const int maxd[3] = {kMaxAxisCellsXY, kMaxAxisCellsXY, kMaxAxisCellsZ};
    int org[3], dim[3];
    for (int k = 0; k < 3; ++k) {
      const double a = std::floor(lo[k] / radius_) - 1.0;
      const double b = std::floor(hi[k] / radius_) + 1.0;
      org[k] = static_cast<int>(std::max(a, -1.0e9));
      dim[k] = static_cast<int>(std::min(std::max(b - a + 1.0, 1.0), static_cast<double>(maxd[k])));
    }
    GridDims g{};
    g.nx = dim[0];
    g.ny = dim[1];
    g.nz = dim[2];
    g.ox = org[0];
    g.oy = org[1];
    g.oz = org[2];
    g.ncells = static_cast<uint32_t>(dim[0]) * dim[1] * dim[2];
    g.radius = radius_;
    return g;
  }

  template <class GP>
  static uint64_t grid_layout(Arena &A, GP &G, size_t nc) {
    A.take(G.fstart, nc + 1);
    A.take(G.bstart, nc + 1);
    A.take(G.count, nc);
    A.take(G.scan_tmp, scan_tmp_words(static_cast<uint32_t>(nc)));
    return A.off;
  }

  void ensure_grid() {
    if (grid_mem_)
      return;
    gd_ = plan_grid();
    const size_t nc = gd_.ncells;
    Arena A;
    grid_bytes_ = grid_layout(A, G_, nc);
    if (budget_check && !budget_check(grid_bytes_))
      throw std::length_error("grid exceeds the HBM budget");
    PRTCL_HIP_CHECK(hipMalloc(&grid_mem_, grid_bytes_));
    char *base = static_cast<char *>(grid_mem_);
    Arena::rebase(G_.fstart, base);
    Arena::rebase(G_.bstart, base);
    Arena::rebase(G_.count, base);
    Arena::rebase(G_.scan_tmp, base);
    bytes += grid_bytes_;
    PRTCL_HIP_CHECK(hipMemsetAsync(G_.count, 0, sizeof(uint32_t) * nc, res_->stream));
    bgrid_dirty_ = true;
  }

  template <class E>
  void gather_swap(E *&arr, E *&tmp, const uint32_t *perm, int n, int elem_bytes) {
    GridOps<T>::gather_elems(res_->stream, tmp, arr, perm, n, elem_bytes);
    std::swap(arr, tmp);
  }

  void build_boundary_grid() {
    hipStream_t st = res_->stream;
    const uint32_t *pm = GridOps<T>::sort_group(st, gd_, P_.bxV, nb_, P_.bcell, G_.count,
                                                G_.bstart, P_.brank, P_.bperm, G_.scan_tmp);
    if (nb_ > 0) {
      gather_swap(P_.bxV, P_.btmpV4, pm, nb_, sizeof(V4));
      gather_swap(P_.bid, P_.btmpI, pm, nb_, 4);
      gather_swap(P_.bcell, P_.bperm, pm, nb_, 4);  // bperm was scratch of the sort
    }
    bgrid_dirty_ = false;
  }

  // sort the fluid into cell order; all=true also reorders every derived
  // field (the drop-in Neighborhood::Update), otherwise only the state that
  // carries over between steps (x, m, v, t_birth, id)
  void build_fluid_grid(bool all) {
    hipStream_t st = res_->stream;
    if (!all) {
      // the step's sort: only the carried state moves, placed by the
      // stable-rank pass itself (x -> xq, v -> tmpV4, id -> tmpI, cell ->
      // frank, then swap)
      GridOps<T>::sort_state(st, gd_, P_.xm, nf_, P_.fcell, G_.count, G_.fstart, P_.frank,
                             P_.fperm, G_.scan_tmp, P_.xq, P_.tmpV4, P_.v,
                             reinterpret_cast<uint32_t *>(P_.tmpI),
                             reinterpret_cast<const uint32_t *>(P_.id));
      if (nf_ > 0) {
        std::swap(P_.xm, P_.xq);
        std::swap(P_.v, P_.tmpV4);
        std::swap(P_.id, P_.tmpI);
        std::swap(P_.fcell, P_.frank);
      }
      fgrid_dirty_ = false;
      grid_built_ = true;
      return;
    }
    const uint32_t *pm = GridOps<T>::sort_group(st, gd_, P_.xm, nf_, P_.fcell, G_.count,
                                                G_.fstart, P_.frank, P_.fperm, G_.scan_tmp);
    if (nf_ > 0) {
      // x -> xq (rebuilt from x by the step before it is read), v -> tmpV4,
      // id -> tmpI, cell -> fperm (scratch of the sort), then swap
      GridOps<T>::gather_state(st, P_.xq, P_.xm, P_.tmpV4, P_.v,
                               reinterpret_cast<uint32_t *>(P_.tmpI),
                               reinterpret_cast<const uint32_t *>(P_.id), P_.fperm, P_.fcell, pm,
                               nf_);
      std::swap(P_.xm, P_.xq);
      std::swap(P_.v, P_.tmpV4);
      std::swap(P_.id, P_.tmpI);
      std::swap(P_.fcell, P_.fperm);
      if (all) {
        for (V4 **f : {&P_.a, &P_.c, &P_.G, &P_.vd})
          gather_swap(*f, P_.tmpV4, pm, nf_, sizeof(V4));
        for (T **f : {&P_.rho, &P_.p, &P_.pr, &P_.s, &P_.AA, &P_.Ap})
          gather_swap(*f, P_.tmpT, pm, nf_, sizeof(T));
        gather_swap(P_.sp, P_.tmpSp, pm, nf_, 9 * sizeof(T));
        gather_swap(P_.gc, P_.tmpSp, pm, nf_, 9 * sizeof(T));
        if (desc.implicit_viscosity == 3) {
          for (V4 **f : {&P_.om, &P_.orhs, &P_.vrhs})
            gather_swap(*f, P_.tmpV4, pm, nf_, sizeof(V4));
          gather_swap(P_.tvg, P_.tmpSp, pm, nf_, 9 * sizeof(T));
          for (T **f : {&P_.prho, &P_.psrho, &P_.pone, &P_.odiag, &P_.vdiag})
            gather_swap(*f, P_.tmpT, pm, nf_, sizeof(T));
        }
      }
    }
    fgrid_dirty_ = false;
    grid_built_ = true;
  }

  void nsearch_update(bool all) {
    ensure_grid();
    if (bgrid_dirty_)
      build_boundary_grid();
    build_fluid_grid(all);
  }

  // Neighborhood::Permute: the host order becomes the slot (grid) order
  void nsearch_permute() {
    if (!grid_built_)
      nsearch_update(true);
    host_permute(res_->stream);
    PRTCL_HIP_CHECK(hipStreamSynchronize(res_->stream));
  }

  // The host order becomes the slot order (id[s] = s), after the gather
  // permutation (new host index s held old host index id[s]) is copied to
  // page-locked memory, where last_permutation() finds it. Used by the
  // procedure and by the permutes inside steps; allocates nothing.
  // the ids stay in HBM (permd) until the host asks for the permutation
  // (last_permutation) or the scene is evicted, so a permute inside a step
  // costs one device pass and no transfer over PCIe
  void host_permute(hipStream_t st) {
    if (nf_ > 0)
      hipLaunchKernelGGL(k_take_ids, dim3(grid_blocks(nf_)), dim3(kBlock), 0, st, P_.id,
                         P_.permd, nf_);
    if (nb_ > 0)
      hipLaunchKernelGGL(k_take_ids, dim3(grid_blocks(nb_)), dim3(kBlock), 0, st, P_.bid,
                         P_.permd + nf_, nb_);
    permuted_ = true;
    perm_on_device_ = true;
  }

  // the last permutation from HBM into its page-locked home
  void fetch_permutation() {
    if (!perm_on_device_)
      return;
    hipStream_t st = res_->stream;
    const size_t n = size_t(nf_) + size_t(nb_);
    if (n > 0)
      PRTCL_HIP_CHECK(hipMemcpyAsync(perm_host_, P_.permd, 4 * n, hipMemcpyDeviceToHost, st));
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
    perm_on_device_ = false;
  }

  void last_permutation(int group, uint32_t *perm) override {
    const int n = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    if (!permuted_) {  // no permute yet: the identity
      for (int i = 0; i < n; ++i)
        perm[i] = static_cast<uint32_t>(i);
      return;
    }
    fetch_permutation();  // resident: evict() fetches it first
    const int32_t *src = perm_host_ + (group == PRTCL_HIP_FLUID ? 0 : nf_);
    for (int i = 0; i < n; ++i)
      perm[i] = static_cast<uint32_t>(src[i]);
  }

  // ------------------------------------------------------------- procedures
  SceneView<T> view() {
    SceneView<T> S{};
    S.nf = nf_;
    S.nb = nb_;
    const T h = static_cast<T>(desc.smoothing_scale);
    S.K.h = h;
    S.K.alpha = kernel_alpha_t<T>();
    S.K.eps = std::sqrt(std::numeric_limits<T>::epsilon());
    S.thr2 = pred_threshold<T>(radius_);
    S.dt = static_cast<T>(desc.time_step);
    S.rho0 = static_cast<T>(desc.rest_density);
    S.omega = static_cast<T>(desc.iisph_relaxation);
    S.mu_f = static_cast<T>(desc.fluid_viscosity);
    S.mu_b = static_cast<T>(desc.boundary_viscosity);
    S.fade = static_cast<T>(desc.fade_duration);
    S.t = static_cast<T>(clock_);
    S.hh001 = (static_cast<T>(0.01) * h) * h;
    S.g0 = static_cast<T>(desc.gravity[0]);
    S.g1 = static_cast<T>(desc.gravity[1]);
    S.g2 = static_cast<T>(desc.gravity[2]);
    S.gc0 = static_cast<T>(desc.gravity_center[0]);
    S.gc1 = static_cast<T>(desc.gravity_center[1]);
    S.gc2 = static_cast<T>(desc.gravity_center[2]);
    S.axial = desc.axial_gravity ? 1 : 0;
    S.grid = gd_;
    S.fstart = G_.fstart;
    S.bstart = G_.bstart;
    S.xm = P_.xm;
    S.v = P_.v;
    S.a = P_.a;
    S.c = P_.c;
    S.G = P_.G;
    S.vd = P_.vd;
    S.xq = P_.xq;
    S.xa = P_.xa;
    S.rho = P_.rho;
    S.p = P_.p;
    S.pr = P_.pr;
    S.s = P_.s;
    S.AA = P_.AA;
    S.Ap = P_.Ap;
    S.sp = P_.sp;
    S.gc = P_.gc;
    S.om = P_.om;
    S.orhs = P_.orhs;
    S.vrhs = P_.vrhs;
    S.tvg = P_.tvg;
    S.prho = P_.prho;
    S.psrho = P_.psrho;
    S.scx = P_.scx;
    S.scr = P_.scr;
    S.scp = P_.scp;
    S.scq = P_.scq;
    S.scy = P_.scy;
    S.pcoef = P_.pcoef;
    S.id = P_.id;
    S.fcell = P_.fcell;
    S.bxV = P_.bxV;
    S.bid = P_.bid;
    S.bcell = P_.bcell;
    S.ctl = P_.ctl;
    S.partials = P_.partials;
    S.cx = P_.cgx;
    S.cr = P_.cgr;
    S.cxp = P_.cgp;
    S.cq = P_.cgq;
    S.cy = P_.cgy;
    S.cb = P_.cgb;
    return S;
  }

  NbrList nbr_list() const { return NbrList{P_.nidx, P_.ninfo, P_.nblk}; }

  void require_grid() {
    if (!grid_built_)
      throw std::logic_error("neighborhood not updated: run the nsearch update first");
  }

  void run(int proc) override {
    ensure_resident();
    hipStream_t st = res_->stream;
    switch (proc) {
      case PRTCL_HIP_P_NSEARCH_UPDATE: nsearch_update(true); break;
      case PRTCL_HIP_P_NSEARCH_PERMUTE: nsearch_permute(); break;
      case PRTCL_HIP_P_AIAST12_COMPUTE_VOLUME:
        if (!grid_mem_ || bgrid_dirty_) {
          ensure_grid();
          build_boundary_grid();
        }
        SphOps<T>::volume(st, view());
        break;
      case PRTCL_HIP_P_DENSITY_COMPUTE_DENSITY: require_grid(); SphOps<T>::density(st, view()); break;
      case PRTCL_HIP_P_GRAVITY_INITIALIZE_ACCELERATION: SphOps<T>::gravity(st, view()); break;
      case PRTCL_HIP_P_GRAVITY_INITIALIZE_AXIAL_ACCELERATION:
        SphOps<T>::gravity_axial(st, view());
        break;
      case PRTCL_HIP_P_GRAVITY_ACCUMULATE_ACCELERATION:
        SphOps<T>::gravity_accumulate(st, view());
        break;
      case PRTCL_HIP_P_VISCOSITY_ACCUMULATE_ACCELERATION: require_grid(); SphOps<T>::viscosity(st, view()); break;
      case PRTCL_HIP_P_SE_INTEGRATE_VELOCITY_WITH_HARD_FADE: SphOps<T>::integrate_v(st, view()); break;
      case PRTCL_HIP_P_SE_INTEGRATE_POSITION:
        SphOps<T>::integrate_x(st, view());
        steps_enqueued_ += nf_ > 0 ? 1u : 0u;
        break;
      case PRTCL_HIP_P_IISPH_SETUP:
        require_grid();
        SphOps<T>::iisph_setup(st, view());
        ctl_op(kCtlSetNprde, 0, nf_);
        break;
      case PRTCL_HIP_P_IISPH_ITERATION_PRESSURE_ACCELERATION:
        require_grid();
        SphOps<T>::iisph_accel(st, view(), false);
        break;
      case PRTCL_HIP_P_IISPH_ITERATION_PRESSURE:
        require_grid();
        SphOps<T>::iisph_pressure(st, view(), false);
        break;
      case PRTCL_HIP_P_WKBB18_COMPUTE_DIAGONAL:
      case PRTCL_HIP_P_WKBB18_ACCUMULATE_ACCELERATION:
      case PRTCL_HIP_P_WKBB18_GC_COMPUTE_DIAGONAL:
      case PRTCL_HIP_P_WKBB18_GC_ACCUMULATE_ACCELERATION:
        require_grid();
        run_wkbb18(proc);
        break;
      case PRTCL_HIP_P_CORRECTION_IDENTITY_GRADIENT_CORRECTION:
        Wkbb18Ops<T>::identity_correction(st, view());
        break;
      case PRTCL_HIP_P_CORRECTION_COMPUTE_GRADIENT_CORRECTION:
        require_grid();
        Wkbb18Ops<T>::correction(st, view(), nullptr);
        break;
      case PRTCL_HIP_P_PT16_SETUP:
      case PRTCL_HIP_P_PT16_SOLVE_VORTICITY_DIFFUSION:
      case PRTCL_HIP_P_PT16_SOLVE_VELOCITY_RECONSTRUCTION:
        require_grid();
        run_pt16(proc);
        break;
      case PRTCL_HIP_P_PT16_VORTICITY_PRESERVATION:
        run_pt16(proc);
        break;
      default: throw std::out_of_range("procedure does not exist");
    }
    PRTCL_HIP_CHECK(hipGetLastError());
  }

  // wkbb18[_gc].compute_diagonal / wkbb18[_gc].accumulate_acceleration;
  // with `fused` the products walk the step's neighbour list
  // with_diagonal (fused steps): the accumulate procedure also computes the
  // diagonal, in the same walk over the list as the CG set-up
  void run_wkbb18(int proc, bool fused = false, bool with_diagonal = false) {
    if (!desc.implicit_viscosity)
      throw std::logic_error("the scene was created without implicit viscosity");
    hipStream_t st = res_->stream;
    const SceneView<T> S = view();
    const NbrList L = nbr_list();
    const NbrList *pl = fused ? &L : nullptr;
    Profiler &pf = res_->prof;
    const bool gc = proc == PRTCL_HIP_P_WKBB18_GC_COMPUTE_DIAGONAL ||
                    proc == PRTCL_HIP_P_WKBB18_GC_ACCUMULATE_ACCELERATION;
    if (proc == PRTCL_HIP_P_WKBB18_COMPUTE_DIAGONAL ||
        proc == PRTCL_HIP_P_WKBB18_GC_COMPUTE_DIAGONAL) {
      int r = pf.begin(st, kStageWkbb18Diag);
      Wkbb18Ops<T>::diagonal(st, S, pl, gc);
      pf.end(st, r);
      return;
    }
    // the PCG of cg_openmp.hpp:32-215, driven like the IISPH loop: the device
    // takes every decision, the host keeps one iteration queued ahead
    const unsigned seq = ++seq_;
    const int max_k = std::max(0, desc.wkbb18_maximum_iterations);
    const bool split = tuner_.current().cg_split;
    int r;
    if (with_diagonal && pl) {
      r = pf.begin(st, kStageWkbb18DiagSetup);
      Wkbb18Ops<T>::diagonal_setup(st, S, *pl, gc, P_.cg_part, desc.wkbb18_maximum_error, max_k,
                                   seq);
    } else {
      r = pf.begin(st, kStagePcgVector);
      Wkbb18Ops<T>::setup(st, S, pl, gc, P_.cg_part, desc.wkbb18_maximum_error, max_k, seq);
    }
    pf.end(st, r);
    const Progress p = cg_loop(seq, max_k, predicted_cg_, [&](int tag) {
      int rp = pf.begin(st, kStagePcgSpmv, tag);
      Wkbb18Ops<T>::iteration_product(st, S, pl, gc, P_.cg_part);
      pf.end(st, rp);
      rp = pf.begin(st, kStagePcgVector, tag);
      Wkbb18Ops<T>::iteration_vectors(st, S, P_.cg_part, split);
      pf.end(st, rp);
    });
    r = pf.begin(st, kStagePcgVector);
    Wkbb18Ops<T>::apply(st, S);
    pf.end(st, r);
    PRTCL_HIP_CHECK(hipGetLastError());
    if (pf.on) {
      PRTCL_HIP_CHECK(hipStreamSynchronize(st));
      pf.resolve(p.iters);
    }
    predicted_cg_ = std::max(1, p.iters);
    last_wkbb_ = p.iters;
  }

  // Iterations of a CG set up under `seq` until the device reports done or
  // the cap (max_k + 1 iterations): `predicted` enqueued at once, then one
  // iteration kept queued ahead while the host polls the progress word.
  template <class F>
  auto cg_loop(unsigned seq, int max_k, int predicted, F &&iteration) {
    int enq = 0;
    auto enqueue = [&](int n) {
      for (int k = 0; k < n; ++k)
        iteration(enq + k);  // the iteration's index tags its profiler records
      enq += n;
    };
    Progress p;
    if (nf_ > 0) {
      enqueue(std::min(std::max(1, std::min(predicted, 16)), max_k + 1));
      for (;;) {
        const int ahead = enq < max_k + 1 ? 1 : 0;
        if (ahead)
          enqueue(1);
        p = wait_progress(seq, enq - ahead, 2);
        if (p.done)
          break;
        if (!ahead) {
          p = wait_progress(seq, enq, 2);
          break;
        }
      }
    }
    return p;
  }

  // pt16.setup / solve_vorticity_diffusion / vorticity_preservation /
  // solve_velocity_reconstruction (pt16.hpp:213-760); each solve is one scalar
  // PCG per vector component with tolerance pt16_maximum_error and cap
  // pt16_maximum_iterations, pt16_iterations = the last component's count
  // with `fused` the fluid sums and the CG products use the step's
  // neighbour list (the pair coefficients are written by the set-up)
  void run_pt16(int proc, bool fused = false) {
    if (desc.implicit_viscosity != 3)
      throw std::logic_error("the scene was created without pt16");
    hipStream_t st = res_->stream;
    const SceneView<T> S = view();
    const NbrList L = nbr_list();
    const NbrList *pl = fused ? &L : nullptr;
    Profiler &pf = res_->prof;
    const T xi = static_cast<T>(pt16_xi_);
    const bool split = tuner_.current().cg_split;
    auto solve = [&](int which) {
      for (int dim = 0; dim < 3; ++dim) {
        const unsigned seq = ++seq_;
        const int max_k = std::max(0, pt16_max_k_);
        int r = pf.begin(st, kStagePcgVector);
        Pt16Ops<T>::cg_setup(st, S, pl, which, dim, P_.cg_part, pt16_tol_, max_k, seq);
        pf.end(st, r);
        const Progress p = cg_loop(seq, max_k, predicted_pt16_[which], [&](int tag) {
          const int rp = pf.begin(st, kStagePcgSpmv, tag);
          Pt16Ops<T>::cg_iteration(st, S, pl, which, P_.cg_part, split);
          pf.end(st, rp);
        });
        r = pf.begin(st, kStagePcgVector);
        Pt16Ops<T>::cg_apply(st, S, which, dim);
        pf.end(st, r);
        PRTCL_HIP_CHECK(hipGetLastError());
        if (pf.on) {
          PRTCL_HIP_CHECK(hipStreamSynchronize(st));
          pf.resolve(p.iters);
        }
        predicted_pt16_[which] = std::max(1, p.iters);
        pt16_iters_ = p.iters;
      }
    };
    int r = pf.begin(st, kStagePt16Setup);
    switch (proc) {
      case PRTCL_HIP_P_PT16_SETUP:
        Pt16Ops<T>::setup(st, S, pl, xi);
        pf.end(st, r);
        break;
      case PRTCL_HIP_P_PT16_SOLVE_VORTICITY_DIFFUSION:
        Pt16Ops<T>::omega_rhs(st, S, pl, xi);
        pf.end(st, r);
        solve(0);
        Pt16Ops<T>::spin(st, S);
        break;
      case PRTCL_HIP_P_PT16_VORTICITY_PRESERVATION:
        Pt16Ops<T>::spin(st, S);
        pf.end(st, r);
        break;
      case PRTCL_HIP_P_PT16_SOLVE_VELOCITY_RECONSTRUCTION:
        Pt16Ops<T>::vrec_rhs(st, S, pl);
        pf.end(st, r);
        solve(1);
        break;
      default: throw std::out_of_range("procedure does not exist");
    }
    PRTCL_HIP_CHECK(hipGetLastError());
  }

  // --------------------------------------------------------------- IISPH loop
  // The convergence decision is taken on the device (finalize kernel), which
  // also publishes (seq, done, iterations, aprde/nprde) to pinned host memory.
  // The host enqueues the predicted number of iterations, then keeps exactly
  // one iteration queued ahead while it polls for the previous ones, so the
  // device never idles on the decision; an iteration enqueued past
  // convergence exits at its first instruction.
  int iisph_solve(double tol, int min_it, int max_it, int *iters, double *rel,
                  bool) override {
    ensure_resident();
    int rc = solve_impl(tol, min_it, max_it, iters, rel, false);
    PRTCL_HIP_CHECK(hipStreamSynchronize(res_->stream));
    return rc;
  }

  struct Progress {
    int iters = 0;
    bool done = false;
    double rel = 0;
  };

  Progress decode(uint64_t w) const {
    Progress p;
    p.iters = static_cast<int>(w & 0x7FFFFFFFull);
    p.done = ((w >> 31) & 1ull) != 0;
    const uint64_t rb = __atomic_load_n(&mirror_host_[1], __ATOMIC_ACQUIRE);
    std::memcpy(&p.rel, &rb, sizeof(double));
    return p;
  }

  // waits until solve `seq` reports done or at least `target` iterations
  Progress wait_progress(unsigned seq, int target, int word = 0) {
    auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0;; ++spin) {
      const uint64_t w = __atomic_load_n(&mirror_host_[word], __ATOMIC_ACQUIRE);
      if ((w >> 32) == seq) {
        Progress p = decode(w);
        if (p.done || p.iters >= target)
          return p;
      }
      if ((spin & 1023u) == 1023u) {
        // the stream drained without the expected word: read the control
        // block itself (the word is only a shortcut)
        if (hipStreamQuery(res_->stream) == hipSuccess) {
          Ctl c = read_ctl();
          Progress p;
          const int it = word == 0 ? c.iters : c.cg_iter;
          const bool dn = word == 0 ? c.done != 0 : c.cg_done != 0;
          p.iters = it;
          p.done = dn || it < target;  // < target: nothing left to wait for
          p.rel = c.rel;
          return p;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300))
          throw HipError{hipErrorLaunchTimeOut, "solver made no progress for 300 s"};
      }
    }
  }

  // IISPH iterations until the device reports convergence or the cap. With
  // `tune`, the fused passes are also timed for the schedule tuner.
  // nprde >= 0 also sets iisph_nprde (one control launch for both)
  // first_done: the fused setup has taken iteration 1 (its pressures and
  // aprde partials are written); its convergence step is the first launch.
  // integrate: the step's fused integration is enqueued right behind the
  // predicted iterations (and again behind any further ones) in its guarded
  // form, so it runs as soon as the device decides convergence instead of
  // after the host has seen the decision; on return it is enqueued.
  // begun != 0: the fused setup launch has begun solve `begun` with these
  // arguments, so no control launch is needed
  int solve_impl(double tol, int min_it, int max_it, int *iters, double *rel, bool fused,
                 const FusedSchedule &sc = FusedSchedule{}, bool tune = false,
                 int nprde = -1, bool first_done = false, bool integrate = false,
                 unsigned begun = 0u) {
    require_grid();
    hipStream_t st = res_->stream;
    const unsigned seq = begun != 0u ? begun : ++seq_;
    // iteration 1 taken by the setup needs its convergence step only where
    // it could end the solve
    const bool first = first_done && nf_ > 0 && max_it > 0;
    const bool preset = first && first_iteration_cannot_stop(min_it, max_it);
    if (begun == 0u)
      ctl_op(preset ? kCtlBeginSolveAfterFirst : kCtlBeginSolve, tol, min_it, max_it,
             static_cast<int>(seq), nprde);
    const SceneView<T> S = view();
    const NbrList L = nbr_list();
    Profiler &pf = res_->prof;
    tprof_.on = tune;
    // iterations enqueued so far (the iteration taken by the setup counts)
    int enq = 0;
    if (first) {
      if (!preset)
        FusedOps<T>::finish_first_iteration(st, S);
      enq = 1;
    }
    // one IISPH iteration (pressure acceleration, pressure, finalize) replayed
    // from a captured graph unless timed or eager
    const bool graph = fused && !sc.eager && !pf.on && !tune;
    // the iteration kernels do not read the clock: a view without it keeps
    // one graph per buffer arrangement (two, alternating step by step)
    SceneView<T> Sg = S;
    Sg.t = T(0);
    const std::string key =
        graph ? graph_key(kGraphIteration | sizeof(T), Sg, L, sc) : std::string();
    auto enqueue = [&](int n) {
      if (graph) {
        for (int k = 0; k < n; ++k)
          graphs_.launch(st, key, [&] {
            FusedOps<T>::accel(st, Sg, L, sc.accel_lds);
            FusedOps<T>::pressure(st, Sg, L, sc.pressure_lds);
          });
        enq += n;
        return;
      }
      for (int k = 0; k < n; ++k) {
        int r = pf.begin(st, kStagePressureAccel, enq + k);
        int q = tprof_.begin(st, kStagePressureAccel, enq + k);
        if (fused)
          FusedOps<T>::accel(st, S, L, sc.accel_lds);
        else
          SphOps<T>::iisph_accel(st, S, true);
        tprof_.end(st, q);
        pf.end(st, r);
        r = pf.begin(st, kStagePressure, enq + k);
        q = tprof_.begin(st, kStagePressure, enq + k);
        if (fused)
          FusedOps<T>::pressure(st, S, L, sc.pressure_lds);
        else
          SphOps<T>::iisph_pressure(st, S, true);
        tprof_.end(st, q);
        pf.end(st, r);
      }
      enq += n;
    };
    Progress p;
    // the guarded integration behind what is enqueued (the caller passes
    // integrate only for fused inviscid steps without per-stage timing)
    if (nf_ > 0 && max_it > 0 && integrate) {
      enqueue(std::max(0, std::min(std::max(min_it, std::min(predicted_iters_, 8)), max_it) - enq));
      FusedOps<T>::integrate(st, S, seq);
      for (;;) {
        p = wait_progress(seq, enq);
        if (p.done)
          break;
        // more iterations than predicted: a growing chunk, each followed by
        // the guarded integration again
        enqueue(std::min(std::max(1, enq / 2), max_it - enq));
        FusedOps<T>::integrate(st, S, seq);
      }
    } else if (nf_ > 0 && max_it > 0) {
      enqueue(std::max(0, std::min(std::max(min_it, std::min(predicted_iters_, 8)), max_it) - enq));
      for (;;) {
        const int ahead = enq < max_it ? 1 : 0;
        if (ahead)
          enqueue(1);
        p = wait_progress(seq, enq - ahead);
        if (p.done)
          break;
        if (!ahead) {
          p = wait_progress(seq, enq);
          break;
        }
      }
    } else {
      p.done = true;
    }
    PRTCL_HIP_CHECK(hipGetLastError());
    if (pf.on) {
      PRTCL_HIP_CHECK(hipStreamSynchronize(st));
      pf.resolve(p.iters);
    }
    if (tune) {
      tprof_.resolve(p.iters);
      tuner_.record(0, tprof_.ms[kStageSetup], static_cast<double>(tprof_.launches[kStageSetup]));
      tuner_.record(1, tprof_.ms[kStagePressureAccel],
                    static_cast<double>(tprof_.launches[kStagePressureAccel]));
      tuner_.record(2, tprof_.ms[kStagePressure],
                    static_cast<double>(tprof_.launches[kStagePressure]));
      tprof_.reset();
      tprof_.on = false;
    }
    // an iteration predicted too many costs a no-op launch, one too few a
    // host round trip: predict the larger of the last two counts
    predicted_iters_ = std::max(1, std::max(p.iters, last_iisph_iters_));
    last_iisph_iters_ = p.iters;
    last_rel_ = p.rel;
    if (iters)
      *iters = p.iters;
    if (rel)
      *rel = p.rel;
    const bool converged = p.iters >= min_it && p.rel < tol;
    return (converged || nf_ == 0 || max_it == 0) ? PRTCL_HIP_OK : PRTCL_HIP_NOT_CONVERGED;
  }

  // ----------------------------------------------------------------- a step
  int step(int k) override {
    ensure_resident();
    hipStream_t st = res_->stream;
    int status = PRTCL_HIP_OK;
    for (int s = 0; s < k; ++s) {
      ensure_grid();
      if (bgrid_dirty_) {
        build_boundary_grid();
        SphOps<T>::volume(st, view());  // boundary volumes follow the boundary grid
      }
      steps_since_permute_ += 1;
      Profiler &pf = res_->prof;
      int r = pf.begin(st, kStageSort);
      build_fluid_grid(false);
      pf.end(st, r);
      if (desc.permute_every > 0 && steps_since_permute_ >= desc.permute_every) {
        host_permute(st);
        steps_since_permute_ = 0;
      }
      const bool fused = desc.step_mode == 0;
      const bool tune = fused && tuner_.trial();
      const FusedSchedule sc = tuner_.current();
      // the fused setup takes IISPH iteration 1 (no neighbours needed at p = 0)
      const bool first_iter =
          fused && !sc.first_split && nf_ > 0 && desc.iisph_max_iterations > 0;
      SolveBegin sb;
      if (fused) {
        // one neighbour pass for density + list + IISPH setup (three with
        // explicit viscosity); the integrated velocity lands in tmpV4
        r = pf.begin(st, kStageSetup);
        tprof_.on = tune;
        const int q = tprof_.begin(st, kStageSetup);
        SceneView<T> Sv = view();
        Sv.first_iter = first_iter ? 1 : 0;
        // the setup also begins the step's IISPH solve (solve_impl's control
        // launch, with the same arguments)
        if (nf_ > 0) {
          sb.tol = desc.iisph_tolerance;
          sb.min_it = desc.iisph_min_iterations;
          sb.max_it = desc.iisph_max_iterations;
          sb.nprde = nf_;
          sb.iters0 = first_iter && first_iteration_cannot_stop(sb.min_it, sb.max_it) ? 1 : 0;
          sb.seq = ++seq_;
        }
        FusedOps<T>::setup(st, Sv, nbr_list(), desc.explicit_viscosity != 0, P_.tmpV4,
                           sc.setup_lds, sb);
        tprof_.end(st, q);
        pf.end(st, r);
        std::swap(P_.v, P_.tmpV4);
      } else {
        const SceneView<T> S = view();
        r = pf.begin(st, kStageDensity);
        SphOps<T>::density(st, S);
        pf.end(st, r);
        r = pf.begin(st, kStageGravityIntegrateV);
        if (desc.axial_gravity)  // example.lua:390-391
          SphOps<T>::gravity_axial(st, S);
        else
          SphOps<T>::gravity(st, S);
        pf.end(st, r);
        if (desc.explicit_viscosity) {
          r = pf.begin(st, kStageViscosity);
          SphOps<T>::viscosity(st, S);
          pf.end(st, r);
        }
        r = pf.begin(st, kStageGravityIntegrateV);
        SphOps<T>::integrate_v(st, S);
        pf.end(st, r);
        r = pf.begin(st, kStageSetup);
        SphOps<T>::iisph_setup(st, S);
        pf.end(st, r);
      }
      int it = 0;
      double rel = 0;
      // the integration rides behind the solve's iterations unless stages
      // are timed
      const bool spec_integrate = fused && !desc.implicit_viscosity && !pf.on && !tune &&
                                  nf_ > 0 && desc.iisph_max_iterations > 0;
      int rc = solve_impl(desc.iisph_tolerance, desc.iisph_min_iterations,
                          desc.iisph_max_iterations, &it, &rel, fused, sc, tune, nf_,
                          first_iter, spec_integrate, sb.seq);
      if (fused)
        tuner_.finish_step();
      if (rc != PRTCL_HIP_OK)
        status = rc;
      const SceneView<T> S = view();
      if (fused && !desc.implicit_viscosity) {
        if (!spec_integrate) {
          r = pf.begin(st, kStageIntegrateX);
          FusedOps<T>::integrate(st, S);
          pf.end(st, r);
        }
        steps_enqueued_ += nf_ > 0 ? 1u : 0u;
      } else {
        if (fused)
          FusedOps<T>::copy_accel(st, S);
        r = pf.begin(st, kStageGravityIntegrateV);
        SphOps<T>::integrate_v(st, S);
        pf.end(st, r);
        if (desc.implicit_viscosity == 1) {
          // example.lua:427-433: diagonal, PCG, integrate velocity again
          if (fused && !sc.wk_split) {
            run_wkbb18(PRTCL_HIP_P_WKBB18_ACCUMULATE_ACCELERATION, fused, true);
          } else {
            run_wkbb18(PRTCL_HIP_P_WKBB18_COMPUTE_DIAGONAL, fused);
            run_wkbb18(PRTCL_HIP_P_WKBB18_ACCUMULATE_ACCELERATION, fused);
          }
          r = pf.begin(st, kStageGravityIntegrateV);
          SphOps<T>::integrate_v(st, S);
          pf.end(st, r);
        } else if (desc.implicit_viscosity == 3) {
          // example.lua:435-445: setup, vorticity diffusion and velocity
          // reconstruction with the solver settings of each (the latter
          // writes the velocity; no integration of v after it)
          run_pt16(PRTCL_HIP_P_PT16_SETUP, fused);
          pt16_tol_ = desc.pt16_vorticity_maximum_error;
          pt16_max_k_ = desc.pt16_vorticity_maximum_iterations;
          run_pt16(PRTCL_HIP_P_PT16_SOLVE_VORTICITY_DIFFUSION, fused);
          pt16_tol_ = desc.pt16_velocity_maximum_error;
          pt16_max_k_ = desc.pt16_velocity_maximum_iterations;
          run_pt16(PRTCL_HIP_P_PT16_SOLVE_VELOCITY_RECONSTRUCTION, fused);
        } else if (desc.implicit_viscosity == 2) {
          // prtcl-lua.bunny.lua:279-281, 321-327: the gradient correction of
          // the step's positions and densities (computed here, right before
          // its only user; x and rho do not change in between), then
          // wkbb18_gc
          r = pf.begin(st, kStageCorrection);
          const NbrList L = nbr_list();
          Wkbb18Ops<T>::correction(st, S, fused ? &L : nullptr);
          pf.end(st, r);
          if (fused && !sc.wk_split) {
            run_wkbb18(PRTCL_HIP_P_WKBB18_GC_ACCUMULATE_ACCELERATION, fused, true);
          } else {
            run_wkbb18(PRTCL_HIP_P_WKBB18_GC_COMPUTE_DIAGONAL, fused);
            run_wkbb18(PRTCL_HIP_P_WKBB18_GC_ACCUMULATE_ACCELERATION, fused);
          }
          r = pf.begin(st, kStageGravityIntegrateV);
          SphOps<T>::integrate_v(st, S);
          pf.end(st, r);
        }
        r = pf.begin(st, kStageIntegrateX);
        SphOps<T>::integrate_x(st, S);
        steps_enqueued_ += nf_ > 0 ? 1u : 0u;
        pf.end(st, r);
      }
      clock_ += static_cast<double>(static_cast<T>(desc.time_step));
      PRTCL_HIP_CHECK(hipGetLastError());
    }
    return status;
  }

  int finish_steps() override {
    if (!arena_)
      return PRTCL_HIP_OK;
    const unsigned want = steps_enqueued_;
    unsigned nonfinite = nonfinite_seen_;
    auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0;; ++spin) {
      const uint64_t w = __atomic_load_n(&mirror_host_[kMirrorStep], __ATOMIC_ACQUIRE);
      if (static_cast<unsigned>(w >> 32) == want) {
        nonfinite = static_cast<unsigned>(w);
        break;
      }
      if ((spin & 1023u) == 1023u) {
        // the stream drained: the control block itself holds the report
        if (hipStreamQuery(res_->stream) == hipSuccess) {
          const Ctl c = read_ctl();
          nonfinite = c.nonfinite_steps;
          break;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300))
          throw HipError{hipErrorLaunchTimeOut, "step made no progress for 300 s"};
      }
    }
    const bool bad = nonfinite != nonfinite_seen_;
    nonfinite_seen_ = nonfinite;
    return bad ? PRTCL_HIP_NONFINITE : PRTCL_HIP_OK;
  }

  // ------------------------------------------------------- batched stepping
  // a scene the batched step can take: fused passes (every viscosity and
  // gravity form of the step)
  bool batchable() const {
    return desc.step_mode == 0 && desc.implicit_viscosity >= 0 &&
           desc.implicit_viscosity <= 3 && nf_ > 0 && desc.iisph_max_iterations > 0 &&
           grid_blocks(nf_) <= kBatchLocalMask;
  }
  // the fused setup takes three launches (FusedOps::setup)
  bool split_setup() const { return desc.explicit_viscosity || desc.axial_gravity; }

  // the grid exists, so a batched step allocates nothing for this scene
  bool stepped_ready() const { return arena_ != nullptr && grid_mem_ != nullptr; }

  // One step of every scene in `v` (all batchable, at most 1 << 12), with one
  // launch per stage over the blocks of all scenes. Per scene the kernels,
  // their inputs and the order of every reduction are those of step(), so
  // each scene's results do not depend on the rest of the batch.
  // With `fixed`, the tables go to the caller's buffers instead of the
  // context's grown ones: no allocation, no graph capture, no timing (the
  // launches are eager and the tuner keeps its current choice).
  static int step_batch(const std::vector<SceneImpl *> &v, CtxResources *res,
                        const BatchTables *fixed = nullptr) {
    const unsigned ns = static_cast<unsigned>(v.size());
    if (ns == 0)
      return PRTCL_HIP_OK;
    hipStream_t st = res->stream;
    Profiler idle;  // never enabled: no events
    Profiler &pf = fixed ? idle : res->prof;
    // boundary grids and volumes (first step after an upload only)
    for (SceneImpl *sc : v) {
      sc->ensure_grid();
      if (sc->bgrid_dirty_) {
        sc->build_boundary_grid();
        SphOps<T>::volume(st, sc->view());
      }
    }
    // block maps: logical block b of the batch (scene order, then local
    // block) is dispatched as hardware block xcd(b), as in the single-scene
    // kernels, so each XCD takes a contiguous run of cell-ordered particles
    // (plist1 / plist3: the blocks of scenes with the one-launch and with
    // the three-launch setup)
    std::vector<uint32_t> plist, tlist, plist1, plist3;
    for (unsigned k = 0; k < ns; ++k) {
      const unsigned nb = grid_blocks(v[k]->nf_);
      for (unsigned b = 0; b < nb; ++b) {
        plist.push_back((k << kBatchLocalBits) | b);
        (v[k]->split_setup() ? plist3 : plist1).push_back((k << kBatchLocalBits) | b);
      }
      const unsigned nt = scan_tiles(v[k]->gd_.ncells);
      for (unsigned b = 0; b < nt; ++b)
        tlist.push_back((k << kBatchLocalBits) | b);
    }
    // the scenes whose step ends with WKBB18 or pt16, with tables and block
    // maps of their own: the WKBB18 scenes all together and, apart, those
    // without and with the gradient correction; the pt16 scenes
    std::vector<unsigned> vsc, psc;  // WKBB18 scenes, pt16 scenes
    std::vector<uint32_t> vlist, vlist_n, vlist_g, ptlist;
    for (unsigned k = 0; k < ns; ++k) {
      if (!v[k]->desc.implicit_viscosity)
        continue;
      if (v[k]->desc.implicit_viscosity == 3) {
        const unsigned kp = static_cast<unsigned>(psc.size());
        psc.push_back(k);
        const unsigned nb = grid_blocks(v[k]->nf_);
        for (unsigned b = 0; b < nb; ++b)
          ptlist.push_back((kp << kBatchLocalBits) | b);
        continue;
      }
      const unsigned kv = static_cast<unsigned>(vsc.size());
      vsc.push_back(k);
      const unsigned nb = grid_blocks(v[k]->nf_);
      const bool gc = v[k]->desc.implicit_viscosity == 2;
      for (unsigned b = 0; b < nb; ++b) {
        vlist.push_back((kv << kBatchLocalBits) | b);
        (gc ? vlist_g : vlist_n).push_back((kv << kBatchLocalBits) | b);
      }
    }
    const unsigned np = static_cast<unsigned>(plist.size());
    const unsigned nt = static_cast<unsigned>(tlist.size());
    const unsigned nvs = static_cast<unsigned>(vsc.size());
    const unsigned nvb = static_cast<unsigned>(vlist.size());
    // the split maps only when some WKBB18 scene has the gradient correction
    const unsigned nvg = static_cast<unsigned>(vlist_g.size());
    const unsigned nvn = nvg > 0 ? static_cast<unsigned>(vlist_n.size()) : 0u;
    const unsigned nps = static_cast<unsigned>(psc.size());
    const unsigned npb = static_cast<unsigned>(ptlist.size());
    // the setup's own block maps only when some scene has the three-launch
    // setup (else pmap serves)
    const unsigned np3 = static_cast<unsigned>(plist3.size());
    const unsigned np1 = np3 > 0 ? static_cast<unsigned>(plist1.size()) : 0u;
    // pinned / device table layout: [sort | setup | solve] entries, the
    // WKBB18 entries, pmap, tmap, vmap, the setup maps, the WKBB18 split
    // maps, the pt16 entries and map
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t tab_bytes = align(sizeof(BatchEntry<T>) * ns);
    const size_t off_visc = 3 * tab_bytes;
    const size_t off_pmap = off_visc + align(sizeof(BatchEntry<T>) * nvs);
    const size_t off_tmap = off_pmap + align(sizeof(uint32_t) * np);
    const size_t off_vmap = off_tmap + align(sizeof(uint32_t) * nt);
    const size_t off_s1map = off_vmap + align(sizeof(uint32_t) * nvb);
    const size_t off_s3map = off_s1map + align(sizeof(uint32_t) * np1);
    const size_t off_vnmap = off_s3map + align(sizeof(uint32_t) * np3);
    const size_t off_vgmap = off_vnmap + align(sizeof(uint32_t) * nvn);
    const size_t off_pt = off_vgmap + align(sizeof(uint32_t) * nvg);
    const size_t off_ptmap = off_pt + align(sizeof(BatchEntry<T>) * nps);
    const size_t total = off_ptmap + align(sizeof(uint32_t) * npb);
    if (fixed && total > fixed->bytes)
      throw TablesTooLarge{total};
    if (!fixed)
      res->ensure_batch(total);
    hipEvent_t copied = fixed ? fixed->copied : res->batch_ev;
    // the previous batch's copy out of the pinned tables must have finished
    PRTCL_HIP_CHECK(hipEventSynchronize(copied));
    char *hb = static_cast<char *>(fixed ? fixed->host : res->batch_host);
    char *db = static_cast<char *>(fixed ? fixed->dev : res->batch_dev);
    auto *t_sort = reinterpret_cast<BatchEntry<T> *>(hb);
    auto *t_setup = reinterpret_cast<BatchEntry<T> *>(hb + tab_bytes);
    auto *t_solve = reinterpret_cast<BatchEntry<T> *>(hb + 2 * tab_bytes);
    auto *t_visc = reinterpret_cast<BatchEntry<T> *>(hb + off_visc);
    uint32_t *pmap = reinterpret_cast<uint32_t *>(hb + off_pmap);
    uint32_t *tmap = reinterpret_cast<uint32_t *>(hb + off_tmap);
    uint32_t *vmap = reinterpret_cast<uint32_t *>(hb + off_vmap);
    for (unsigned b = 0; b < np; ++b)
      pmap[b] = plist[xcd_of(b, np)];
    for (unsigned b = 0; b < nt; ++b)
      tmap[b] = tlist[b];
    for (unsigned b = 0; b < nvb; ++b)
      vmap[b] = vlist[xcd_of(b, nvb)];
    uint32_t *s1map = reinterpret_cast<uint32_t *>(hb + off_s1map);
    uint32_t *s3map = reinterpret_cast<uint32_t *>(hb + off_s3map);
    for (unsigned b = 0; b < np1; ++b)
      s1map[b] = plist1[xcd_of(b, np1)];
    for (unsigned b = 0; b < np3; ++b)
      s3map[b] = plist3[xcd_of(b, np3)];
    uint32_t *vnmap = reinterpret_cast<uint32_t *>(hb + off_vnmap);
    uint32_t *vgmap = reinterpret_cast<uint32_t *>(hb + off_vgmap);
    for (unsigned b = 0; b < nvn; ++b)
      vnmap[b] = vlist_n[xcd_of(b, nvn)];
    for (unsigned b = 0; b < nvg; ++b)
      vgmap[b] = vlist_g[xcd_of(b, nvg)];
    auto *t_pt = reinterpret_cast<BatchEntry<T> *>(hb + off_pt);
    uint32_t *ptmap = reinterpret_cast<uint32_t *>(hb + off_ptmap);
    for (unsigned b = 0; b < npb; ++b)
      ptmap[b] = ptlist[xcd_of(b, npb)];
    // the pointer swaps of build_fluid_grid() and of the setup pass are done
    // on the host while the three tables are filled
    std::vector<char> permute(ns, 0);
    std::vector<unsigned> seqs(ns);
    bool any_first_step = false;  // a scene needs iteration 1's convergence step
    // the setup takes IISPH iteration 1 (batchable scenes have max_it > 0)
    const bool fold_first = !res->batch_tuner[sizeof(T) == 8 ? 1 : 0].current().first_split;
    for (unsigned k = 0; k < ns; ++k) {
      SceneImpl &s = *v[k];
      BatchEntry<T> E{};
      E.S = s.view();
      E.nidx = s.P_.nidx;
      E.ninfo = s.P_.ninfo;
      E.nblk = s.P_.nblk;
      E.count = s.G_.count;
      E.start = s.G_.fstart;
      E.rank = s.P_.frank;
      E.perm = s.P_.fperm;
      E.scan_tmp = s.G_.scan_tmp;
      E.x_dst = s.P_.xq;
      E.v_dst = s.P_.tmpV4;
      E.id_dst = reinterpret_cast<uint32_t *>(s.P_.tmpI);
      E.cell_dst = s.P_.frank;  // the atomic ranks are consumed by then
      E.nblocks = grid_blocks(s.nf_);
      E.ntiles = scan_tiles(s.gd_.ncells);
      t_sort[k] = E;
      std::swap(s.P_.xm, s.P_.xq);
      std::swap(s.P_.v, s.P_.tmpV4);
      std::swap(s.P_.id, s.P_.tmpI);
      std::swap(s.P_.fcell, s.P_.frank);
      s.fgrid_dirty_ = false;
      s.grid_built_ = true;
      s.steps_since_permute_ += 1;
      if (s.desc.permute_every > 0 && s.steps_since_permute_ >= s.desc.permute_every) {
        permute[k] = 1;
        s.steps_since_permute_ = 0;
      }
      E.S = s.view();
      E.S.first_iter = fold_first ? 1 : 0;
      E.vtmp = s.P_.tmpV4;
      E.split_setup = s.split_setup() ? 1 : 0;
      E.explicit_visc = s.desc.explicit_viscosity ? 1 : 0;
      t_setup[k] = E;
      std::swap(s.P_.v, s.P_.tmpV4);
      E.S = s.view();
      E.vtmp = nullptr;
      seqs[k] = ++s.seq_;
      E.tol = s.desc.iisph_tolerance;
      E.min_it = s.desc.iisph_min_iterations;
      E.max_it = s.desc.iisph_max_iterations;
      E.nprde = s.nf_;
      E.seq = seqs[k];
      E.first_preset =
          fold_first && first_iteration_cannot_stop(E.min_it, E.max_it) ? 1 : 0;
      any_first_step = any_first_step || !E.first_preset;
      E.viscous = s.desc.implicit_viscosity ? 1 : 0;
      t_solve[k] = E;
    }
    // WKBB18 after the IISPH solve (the order of run_wkbb18: the CG's
    // sequence number follows the solve's)
    const bool cg_split = res->batch_tuner[sizeof(T) == 8 ? 1 : 0].current().cg_split;
    std::vector<unsigned> cg_seqs(nvs);
    bool any_split = false;
    for (unsigned kv = 0; kv < nvs; ++kv) {
      SceneImpl &s = *v[vsc[kv]];
      BatchEntry<T> E = t_solve[vsc[kv]];
      E.cg_part = s.P_.cg_part;
      E.cg_tol = s.desc.wkbb18_maximum_error;
      E.cg_max_k = std::max(0, static_cast<int>(s.desc.wkbb18_maximum_iterations));
      cg_seqs[kv] = ++s.seq_;
      E.cg_seq = cg_seqs[kv];
      E.cg_fold = !cg_split && E.nblocks <= kFusedCgBlocks ? 1 : 0;
      any_split = any_split || !E.cg_fold;
      t_visc[kv] = E;
    }
    // pt16 after the IISPH solve: six scalar CG solves (vorticity and
    // velocity components) under six consecutive sequence numbers, as
    // run_pt16 takes them
    bool pt_any_split = false;
    std::vector<unsigned> pt_seq0(nps);
    for (unsigned kp = 0; kp < nps; ++kp) {
      SceneImpl &s = *v[psc[kp]];
      BatchEntry<T> E = t_solve[psc[kp]];
      E.cg_part = s.P_.cg_part;
      E.pt16_xi = s.pt16_xi_;
      E.pt16_tol[0] = s.desc.pt16_vorticity_maximum_error;
      E.pt16_max_k[0] = std::max(0, static_cast<int>(s.desc.pt16_vorticity_maximum_iterations));
      E.pt16_tol[1] = s.desc.pt16_velocity_maximum_error;
      E.pt16_max_k[1] = std::max(0, static_cast<int>(s.desc.pt16_velocity_maximum_iterations));
      pt_seq0[kp] = s.seq_ + 1;
      s.seq_ += 6;
      E.cg_seq = pt_seq0[kp];
      E.cg_fold = !cg_split && E.nblocks <= kFusedCgBlocks ? 1 : 0;
      pt_any_split = pt_any_split || !E.cg_fold;
      t_pt[kp] = E;
    }
    // a fixed device buffer may be shared with other streams: the copy into
    // it waits for the kernels of its last user
    if (fixed)
      PRTCL_HIP_CHECK(hipStreamWaitEvent(st, fixed->released, 0));
    PRTCL_HIP_CHECK(hipMemcpyAsync(db, hb, total, hipMemcpyHostToDevice, st));
    PRTCL_HIP_CHECK(hipEventRecord(copied, st));
    const auto *d_sort = reinterpret_cast<const BatchEntry<T> *>(db);
    const auto *d_setup = reinterpret_cast<const BatchEntry<T> *>(db + tab_bytes);
    const auto *d_solve = reinterpret_cast<const BatchEntry<T> *>(db + 2 * tab_bytes);
    const auto *d_pmap = reinterpret_cast<const uint32_t *>(db + off_pmap);
    const auto *d_tmap = reinterpret_cast<const uint32_t *>(db + off_tmap);
    const auto *d_visc = reinterpret_cast<const BatchEntry<T> *>(db + off_visc);
    const auto *d_vmap = reinterpret_cast<const uint32_t *>(db + off_vmap);
    const auto *d_s1map = reinterpret_cast<const uint32_t *>(db + off_s1map);
    const auto *d_s3map = reinterpret_cast<const uint32_t *>(db + off_s3map);
    const auto *d_pt = reinterpret_cast<const BatchEntry<T> *>(db + off_pt);
    const auto *d_ptmap = reinterpret_cast<const uint32_t *>(db + off_ptmap);
    BatchMaps vm;
    vm.all = reinterpret_cast<const uint32_t *>(db + off_vmap);
    vm.n_all = nvb;
    if (nvg > 0) {
      vm.plain = reinterpret_cast<const uint32_t *>(db + off_vnmap);
      vm.n_plain = nvn;
      vm.gc = reinterpret_cast<const uint32_t *>(db + off_vgmap);
      vm.n_gc = nvg;
    } else {
      vm.plain = vm.all;
      vm.n_plain = nvb;
      vm.gc = nullptr;
      vm.n_gc = 0;
    }

    int r = pf.begin(st, kStageSort);
    GridOps<T>::sort_batch(st, d_sort, d_pmap, np, d_tmap, nt, ns);
    pf.end(st, r);
    for (unsigned k = 0; k < ns; ++k)
      if (permute[k])
        v[k]->host_permute(st);
    ScheduleTuner &tn = res->batch_tuner[sizeof(T) == 8 ? 1 : 0];
    const bool tune = !fixed && tn.trial();
    FusedSchedule sc = tn.current();
    if (fixed)
      sc.eager = true;
    Profiler &tp = fixed ? idle : res->batch_prof;
    tp.on = tune;
    r = pf.begin(st, kStageSetup);
    int q = tp.begin(st, kStageSetup);
    if (np3 == 0) {
      FusedOps<T>::setup_batch(st, d_setup, d_pmap, np, sc.setup_lds);
    } else {
      FusedOps<T>::setup_batch(st, d_setup, d_s1map, np1, sc.setup_lds);
      FusedOps<T>::setup_split_batch(st, d_setup, d_s3map, np3, sc.setup_lds);
    }
    tp.end(st, q);
    pf.end(st, r);
    hipLaunchKernelGGL(k_begin_solve_b<T>, dim3(grid_blocks(ns, 64)), dim3(64), 0, st, d_solve,
                       ns);
    // the convergence step of iteration 1, which the setup has taken, for
    // the scenes whose solve it could end
    if (fold_first && any_first_step)
      FusedOps<T>::finish_first_iteration_batch(st, d_solve, ns);

    // IISPH iterations of all scenes; a scene whose solve is done exits its
    // blocks at its flag while the others go on
    int pred = 1, max_it = 0;
    for (SceneImpl *sc : v) {
      pred = std::max(pred, std::max(sc->desc.iisph_min_iterations,
                                     std::min(sc->predicted_iters_, 8)));
      max_it = std::max(max_it, static_cast<int>(sc->desc.iisph_max_iterations));
    }
    int enq = fold_first ? 1 : 0;  // iterations enqueued (iteration 1 by the setup)
    const bool graph = !sc.eager && !pf.on && !tune;
    const std::string key =
        graph ? graph_key(kGraphBatchIteration | sizeof(T), d_solve, d_pmap, np, ns, sc)
              : std::string();
    auto enqueue = [&](int n) {
      if (graph) {
        for (int j = 0; j < n; ++j)
          res->batch_graphs[sizeof(T) == 8 ? 1 : 0].launch(st, key, [&] {
            FusedOps<T>::accel_batch(st, d_solve, d_pmap, np, sc.accel_lds);
            FusedOps<T>::pressure_batch(st, d_solve, d_pmap, np, ns, sc.pressure_lds);
          });
        enq += n;
        return;
      }
      for (int j = 0; j < n; ++j) {
        int q1 = pf.begin(st, kStagePressureAccel, enq + j);
        int q2 = tp.begin(st, kStagePressureAccel, enq + j);
        FusedOps<T>::accel_batch(st, d_solve, d_pmap, np, sc.accel_lds);
        tp.end(st, q2);
        pf.end(st, q1);
        q1 = pf.begin(st, kStagePressure, enq + j);
        q2 = tp.begin(st, kStagePressure, enq + j);
        FusedOps<T>::pressure_batch(st, d_solve, d_pmap, np, ns, sc.pressure_lds);
        tp.end(st, q2);
        pf.end(st, q1);
      }
      enq += n;
    };
    std::vector<Progress> prog(ns);
    std::vector<char> finished(ns, 0);
    enqueue(std::max(0, std::min(pred, max_it) - enq));
    // the integration rides behind the iterations in its guarded form (each
    // scene integrates once its own solve is done) unless stages are timed
    const bool spec = !tune && !pf.on;
    if (spec) {
      FusedOps<T>::integrate_batch(st, d_solve, d_pmap, np, ns, true);
      for (;;) {
        bool all = true;
        for (unsigned k = 0; k < ns; ++k) {
          if (finished[k])
            continue;
          prog[k] = v[k]->wait_progress(seqs[k], enq);
          if (prog[k].done)
            finished[k] = 1;
          else
            all = false;
        }
        if (all)
          break;
        enqueue(std::min(std::max(1, enq / 2), max_it - enq));
        FusedOps<T>::integrate_batch(st, d_solve, d_pmap, np, ns, true);
      }
    }
    for (; !spec;) {
      const int ahead = enq < max_it ? 1 : 0;
      if (ahead)
        enqueue(1);
      bool all = true;
      for (unsigned k = 0; k < ns; ++k) {
        if (finished[k])
          continue;
        prog[k] = v[k]->wait_progress(seqs[k], ahead ? enq - ahead : enq);
        if (prog[k].done || !ahead)
          finished[k] = 1;
        else
          all = false;
      }
      if (all)
        break;
    }
    PRTCL_HIP_CHECK(hipGetLastError());
    if (tune) {
      // the schedules are compared on the iterations every scene of the
      // batch runs (below the smallest min_it): later launches skip the
      // blocks of scenes already done, so their cost depends on the step
      int all_run = 1 << 30;
      for (SceneImpl *sc : v)
        all_run = std::min(all_run, static_cast<int>(sc->desc.iisph_min_iterations));
      tp.resolve(std::max(fold_first ? 2 : 1, all_run));
      tn.record(0, tp.ms[kStageSetup], static_cast<double>(tp.launches[kStageSetup]));
      tn.record(1, tp.ms[kStagePressureAccel],
                static_cast<double>(tp.launches[kStagePressureAccel]));
      tn.record(2, tp.ms[kStagePressure], static_cast<double>(tp.launches[kStagePressure]));
      tp.reset();
      tp.on = false;
    }
    if (!fixed)
      tn.finish_step();
    r = pf.begin(st, kStageIntegrateX);
    if (!spec)
      FusedOps<T>::integrate_batch(st, d_solve, d_pmap, np, ns);
    pf.end(st, r);
    if (pf.on) {
      PRTCL_HIP_CHECK(hipStreamSynchronize(st));
      int it = 0;
      for (const Progress &p : prog)
        it = std::max(it, p.iters);
      pf.resolve(it);
    }
    // the WKBB18 scenes: a = the pressure acceleration, integrate v, the
    // diagonal and the CG set-up in one list walk, the PCG of every scene
    // until each reports done, apply, integrate v and x (the fused step's
    // order for implicit viscosity)
    // The CG iterations of several scenes' solves (scene sc[k] under seq[k],
    // at most cap iterations): the predicted count at once, then one queued
    // ahead while the host polls each scene's progress word
    auto cg_loop_batch = [&](const std::vector<unsigned> &sc, const std::vector<unsigned> &seq,
                             int pred, int cap, const std::function<void(int)> &iteration) {
      std::vector<Progress> out(sc.size());
      int cenq = 0;
      auto cg_enqueue = [&](int n) {
        for (int j = 0; j < n; ++j)
          iteration(cenq + j);
        cenq += n;
      };
      cg_enqueue(std::min(pred, cap));
      std::vector<char> fin(sc.size(), 0);
      for (;;) {
        const int ahead = cenq < cap ? 1 : 0;
        if (ahead)
          cg_enqueue(1);
        bool all = true;
        for (size_t k = 0; k < sc.size(); ++k) {
          if (fin[k])
            continue;
          out[k] = v[sc[k]]->wait_progress(seq[k], cenq - ahead, 2);
          if (out[k].done) {
            fin[k] = 1;
          } else if (!ahead) {
            out[k] = v[sc[k]]->wait_progress(seq[k], cenq, 2);
            fin[k] = 1;
          } else {
            all = false;
          }
        }
        if (all)
          break;
      }
      if (pf.on) {
        PRTCL_HIP_CHECK(hipStreamSynchronize(st));
        int it = 0;
        for (const Progress &p : out)
          it = std::max(it, p.iters);
        pf.resolve(it);
      }
      return out;
    };
    std::vector<Progress> cgp(nvs);
    if (nvs > 0) {
      r = pf.begin(st, kStageGravityIntegrateV);
      SphOps<T>::integrate_v_batch(st, d_visc, d_vmap, nvb, true);
      pf.end(st, r);
      if (nvg > 0) {  // the step's gradient correction (x and rho are final)
        r = pf.begin(st, kStageCorrection);
        Wkbb18Ops<T>::correction_batch(st, d_visc, vm);
        pf.end(st, r);
      }
      r = pf.begin(st, kStageWkbb18DiagSetup);
      Wkbb18Ops<T>::diagonal_setup_batch(st, d_visc, vm, nvs);
      pf.end(st, r);
      int cap = 1, pred = 1;
      for (unsigned kv = 0; kv < nvs; ++kv) {
        const SceneImpl &s = *v[vsc[kv]];
        cap = std::max(cap, std::max(0, static_cast<int>(s.desc.wkbb18_maximum_iterations)) + 1);
        pred = std::max(pred, std::min(std::max(1, s.predicted_cg_), 16));
      }
      cgp = cg_loop_batch(vsc, cg_seqs, pred, cap, [&](int tag) {
        const int q1 = pf.begin(st, kStagePcgSpmv, tag);
        Wkbb18Ops<T>::iteration_batch(st, d_visc, vm, nvs, any_split);
        pf.end(st, q1);
      });
      r = pf.begin(st, kStagePcgVector);
      Wkbb18Ops<T>::apply_batch(st, d_visc, vm);
      pf.end(st, r);
      r = pf.begin(st, kStageGravityIntegrateV);
      SphOps<T>::integrate_v_batch(st, d_visc, d_vmap, nvb, false);
      pf.end(st, r);
      r = pf.begin(st, kStageIntegrateX);
      SphOps<T>::integrate_x_batch(st, d_visc, d_vmap, nvb, nvs);
      pf.end(st, r);
    }
    // the pt16 scenes (the order of the single-scene step): a = the pressure
    // acceleration, integrate v, set-up, vorticity diffusion (three solves),
    // spin, velocity reconstruction (three solves, which write the
    // velocity), integrate x
    std::vector<Progress> ptp(nps);
    std::vector<int> pt_last(nps * 2, 0);
    if (nps > 0) {
      r = pf.begin(st, kStageGravityIntegrateV);
      SphOps<T>::integrate_v_batch(st, d_pt, d_ptmap, npb, true);
      pf.end(st, r);
      r = pf.begin(st, kStagePt16Setup);
      Pt16Ops<T>::setup_batch(st, d_pt, d_ptmap, npb);
      Pt16Ops<T>::omega_rhs_batch(st, d_pt, d_ptmap, npb);
      pf.end(st, r);
      for (int which = 0; which < 2; ++which) {
        if (which == 1) {
          Pt16Ops<T>::spin_batch(st, d_pt, d_ptmap, npb);
          r = pf.begin(st, kStagePt16Setup);
          Pt16Ops<T>::vrec_rhs_batch(st, d_pt, d_ptmap, npb);
          pf.end(st, r);
        }
        int cap = 1, pred = 1;
        for (unsigned kp = 0; kp < nps; ++kp) {
          const SceneImpl &s = *v[psc[kp]];
          const int mk = which == 0 ? s.desc.pt16_vorticity_maximum_iterations
                                    : s.desc.pt16_velocity_maximum_iterations;
          cap = std::max(cap, std::max(0, mk) + 1);
          pred = std::max(pred, std::min(std::max(1, s.predicted_pt16_[which]), 16));
        }
        for (int dim = 0; dim < 3; ++dim) {
          const int solve = 3 * which + dim;
          std::vector<unsigned> seqs_s(nps);
          for (unsigned kp = 0; kp < nps; ++kp)
            seqs_s[kp] = pt_seq0[kp] + static_cast<unsigned>(solve);
          r = pf.begin(st, kStagePcgVector);
          Pt16Ops<T>::cg_setup_batch(st, d_pt, d_ptmap, npb, nps, which, dim, solve);
          pf.end(st, r);
          ptp = cg_loop_batch(psc, seqs_s, pred, cap, [&](int tag) {
            const int q1 = pf.begin(st, kStagePcgSpmv, tag);
            Pt16Ops<T>::cg_iteration_batch(st, d_pt, d_ptmap, npb, nps, which, pt_any_split);
            pf.end(st, q1);
          });
          r = pf.begin(st, kStagePcgVector);
          Pt16Ops<T>::cg_apply_batch(st, d_pt, d_ptmap, npb, which, dim);
          pf.end(st, r);
          for (unsigned kp = 0; kp < nps; ++kp)
            pt_last[2 * kp + which] = ptp[kp].iters;
        }
      }
      r = pf.begin(st, kStageIntegrateX);
      SphOps<T>::integrate_x_batch(st, d_pt, d_ptmap, npb, nps);
      pf.end(st, r);
    }
    if (fixed)  // the last kernel reading the tables is enqueued
      PRTCL_HIP_CHECK(hipEventRecord(fixed->released, st));
    for (SceneImpl *sc : v)
      sc->steps_enqueued_ += sc->nf_ > 0 ? 1u : 0u;
    for (unsigned kv = 0; kv < nvs; ++kv) {
      SceneImpl &s = *v[vsc[kv]];
      s.predicted_cg_ = std::max(1, cgp[kv].iters);
      s.last_wkbb_ = cgp[kv].iters;
    }
    // as run_pt16 leaves them: the predictions from each system's last
    // component, the velocity system's settings and last count
    for (unsigned kp = 0; kp < nps; ++kp) {
      SceneImpl &s = *v[psc[kp]];
      s.predicted_pt16_[0] = std::max(1, pt_last[2 * kp]);
      s.predicted_pt16_[1] = std::max(1, pt_last[2 * kp + 1]);
      s.pt16_iters_ = pt_last[2 * kp + 1];
      s.pt16_tol_ = s.desc.pt16_velocity_maximum_error;
      s.pt16_max_k_ = s.desc.pt16_velocity_maximum_iterations;
    }
    int status = PRTCL_HIP_OK;
    for (unsigned k = 0; k < ns; ++k) {
      SceneImpl &s = *v[k];
      const Progress &p = prog[k];
      s.predicted_iters_ = std::max(1, std::max(p.iters, s.last_iisph_iters_));
      s.last_iisph_iters_ = p.iters;
      s.last_rel_ = p.rel;
      const bool converged = p.iters >= s.desc.iisph_min_iterations &&
                             p.rel < s.desc.iisph_tolerance;
      if (!converged && status == PRTCL_HIP_OK)
        status = PRTCL_HIP_NOT_CONVERGED;
      s.clock_ += static_cast<double>(static_cast<T>(s.desc.time_step));
    }
    PRTCL_HIP_CHECK(hipGetLastError());
    return status;
  }

  // hardware block b of n -> logical block (the XCD remap of the fused kernels)
  static unsigned xcd_of(unsigned b, unsigned n) {
    constexpr unsigned kXcd = 8;
    const unsigned per = n / kXcd, rem = n % kXcd;
    const unsigned x = b % kXcd, k = b / kXcd;
    return x * per + (x < rem ? x : rem) + k;
  }

  // ------------------------------------------------------------ residency
  // Eviction copies the device state byte for byte to host memory and frees
  // the HBM; restoring allocates again, copies back and rebases every device
  // pointer, so the scene continues bit-identically (same slot order, same
  // buffers in use).
  bool resident() const override { return arena_ != nullptr; }

  void evict() override {
    if (!arena_)
      return;
    hipStream_t st = res_->stream;
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
    fetch_permutation();  // while evicted, last_permutation reads the host copy
    graphs_.clear();  // they point into the arena
    host_arena_.resize(arena_bytes_);
    PRTCL_HIP_CHECK(hipMemcpy(host_arena_.data(), arena_, arena_bytes_, hipMemcpyDeviceToHost));
    if (grid_mem_) {
      host_grid_.resize(grid_bytes_);
      PRTCL_HIP_CHECK(hipMemcpy(host_grid_.data(), grid_mem_, grid_bytes_, hipMemcpyDeviceToHost));
    }
    rebase_all(static_cast<char *>(arena_), nullptr, static_cast<char *>(grid_mem_), nullptr);
    PRTCL_HIP_CHECK(hipFree(arena_));
    arena_ = nullptr;
    if (grid_mem_) {
      PRTCL_HIP_CHECK(hipFree(grid_mem_));
      grid_mem_ = nullptr;
      grid_evicted_ = true;
    }
    bytes = 0;
  }

  void ensure_resident() override {
    if (arena_)
      return;
    const uint64_t need = arena_bytes_ + (grid_evicted_ ? grid_bytes_ : 0);
    if (budget_check && !budget_check(need))
      throw std::length_error("restoring the scene exceeds the HBM budget");
    void *a = nullptr, *g = nullptr;
    PRTCL_HIP_CHECK(hipMalloc(&a, arena_bytes_));
    if (grid_evicted_) {
      hipError_t e = hipMalloc(&g, grid_bytes_);
      if (e != hipSuccess) {
        (void)hipFree(a);
        throw HipError{e, "hipMalloc(grid)"};
      }
    }
    PRTCL_HIP_CHECK(hipMemcpy(a, host_arena_.data(), arena_bytes_, hipMemcpyHostToDevice));
    if (g)
      PRTCL_HIP_CHECK(hipMemcpy(g, host_grid_.data(), grid_bytes_, hipMemcpyHostToDevice));
    rebase_all(nullptr, static_cast<char *>(a), nullptr, static_cast<char *>(g));
    arena_ = a;
    grid_mem_ = g;
    grid_evicted_ = false;
    bytes = arena_bytes_ + (g ? grid_bytes_ : 0);
    std::vector<char>().swap(host_arena_);
    std::vector<char>().swap(host_grid_);
  }

  // device bytes this scene still needs before its next step
  uint64_t resident_bytes_needed() const override {
    uint64_t need = arena_ ? 0 : arena_bytes_;
    if (grid_evicted_) {
      need += grid_bytes_;
    } else if (!grid_mem_) {
      Arena A;
      GridPtrs G;
      need += grid_layout(A, G, plan_grid().ncells);
    }
    return need;
  }

  void set_schedule(int mask) override { tuner_.force(mask); }
  int schedule() const override { return tuner_.current().mask(); }

  void step_info(int *it, double *rel, int *wk) override {
    if (it)
      *it = last_iisph_iters_;
    if (rel)
      *rel = last_rel_;
    if (wk)
      *wk = last_wkbb_;
  }

  // --------------------------------------------------------- neighbour export
  void copy_neighbors(int group, int ngroup, uint32_t *offsets, uint32_t *indices,
                      uint64_t *total) override {
    ensure_resident();
    require_grid();
    if ((group != 0 && group != 1) || (ngroup != 0 && ngroup != 1))
      throw std::invalid_argument("group");
    hipStream_t st = res_->stream;
    const int nq = group == PRTCL_HIP_FLUID ? nf_ : nb_;
    const SceneView<T> S = view();
    uint32_t *dcnt = nullptr, *dout = nullptr;
    PRTCL_HIP_CHECK(hipMalloc(&dcnt, sizeof(uint32_t) * (nq + 1)));
    SphOps<T>::count_neighbors(st, S, group, ngroup, dcnt);
    std::vector<uint32_t> cnt(nq);
    if (nq > 0)
      PRTCL_HIP_CHECK(hipMemcpyAsync(cnt.data(), dcnt, 4 * nq, hipMemcpyDeviceToHost, st));
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
    uint64_t tot = 0;
    std::vector<uint32_t> off(nq + 1);
    for (int i = 0; i < nq; ++i) {
      off[i] = static_cast<uint32_t>(tot);
      tot += cnt[i];
    }
    off[nq] = static_cast<uint32_t>(tot);
    if (total)
      *total = tot;
    if (offsets)
      std::copy(off.begin(), off.end(), offsets);
    if (indices && tot > 0) {
      PRTCL_HIP_CHECK(hipMemcpyAsync(dcnt, off.data(), 4 * (nq + 1), hipMemcpyHostToDevice, st));
      PRTCL_HIP_CHECK(hipMalloc(&dout, sizeof(uint32_t) * tot));
      SphOps<T>::write_neighbors(st, S, group, ngroup, dcnt, dout);
      PRTCL_HIP_CHECK(hipMemcpyAsync(indices, dout, 4 * tot, hipMemcpyDeviceToHost, st));
      PRTCL_HIP_CHECK(hipStreamSynchronize(st));
      for (int i = 0; i < nq; ++i)
        std::sort(indices + off[i], indices + off[i + 1]);
    }
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(dcnt);
    if (dout)
      (void)hipFree(dout);
  }

 private:
  struct Ptrs {
    V4 *xm, *v, *a, *c, *G, *vd, *tmpV4, *xq, *xa;
    T *rho, *p, *pr, *s, *AA, *Ap, *tmpT, *sp, *tmpSp, *gc;
    V4 *om, *orhs, *vrhs;
    T *tvg, *prho, *psrho, *pone, *odiag, *vdiag, *scx, *scr, *scp, *scq, *scy, *pcoef;
    int32_t *id, *tmpI;
    uint32_t *fcell, *frank, *fperm, *ninfo;
    uint16_t *nidx;
    BlockRanges *nblk;
    V4 *bxV, *btmpV4;
    int32_t *bid, *btmpI;
    int32_t *permd;  // the ids at the last permute: fluid, then boundary
    uint32_t *bcell, *brank, *bperm;
    double *partials, *cg_part;
    Ctl *ctl;
    V4 *cgx, *cgr, *cgp, *cgq, *cgy, *cgb;
  };
  struct GridPtrs {
    uint32_t *fstart = nullptr, *bstart = nullptr, *count = nullptr, *scan_tmp = nullptr;
  };
  struct BBox {
    double lo[3], hi[3];
    bool valid = false;
  };

  static void layout(Arena &A, Ptrs &P, int nf, int nb, bool pcg, bool pt16) {
    const size_t f = static_cast<size_t>(std::max(nf, 1));
    const size_t b = static_cast<size_t>(std::max(nb, 1));
    for (V4 **q : {&P.v, &P.a, &P.c, &P.G, &P.vd, &P.tmpV4})
      A.take(*q, f);
    // xm and xq swap roles every step; both carry the run padding
    for (V4 **q : {&P.xm, &P.xq})
      A.take(*q, f + kRunPad);
    A.take(P.xa, SceneView<T>::xa_x(f - 1) + 1 + SceneView<T>::kXaGroup);
    for (T **q : {&P.rho, &P.p, &P.pr, &P.s, &P.AA, &P.Ap, &P.tmpT})
      A.take(*q, f);
    A.take(P.sp, 9 * f);
    A.take(P.tmpSp, 9 * f);
    A.take(P.gc, 9 * f);
    A.take(P.id, f);
    A.take(P.tmpI, f);
    A.take(P.fcell, f);
    A.take(P.frank, f);
    A.take(P.fperm, f);
    A.take(P.nidx, static_cast<size_t>(kNbrCap) * f);
    A.take(P.ninfo, f);
    A.take(P.nblk, grid_blocks(static_cast<long>(f)));
    A.take(P.bxV, b);
    A.take(P.btmpV4, b);
    A.take(P.bid, b);
    A.take(P.btmpI, b);
    A.take(P.permd, std::max<size_t>(1, static_cast<size_t>(f) + static_cast<size_t>(b)));
    A.take(P.bcell, b);
    A.take(P.brank, b);
    A.take(P.bperm, b);
    A.take(P.partials, grid_blocks(static_cast<long>(std::max(f, b))) + 1);
    A.take(P.cg_part, 3 * static_cast<size_t>(grid_blocks(static_cast<long>(f))) + 3);
    A.take(P.ctl, 1);
    const size_t nv = pcg ? f : 1;
    for (V4 **q : {&P.cgx, &P.cgr, &P.cgq, &P.cgy, &P.cgb})
      A.take(*q, nv);
    A.take(P.cgp, 2 * nv);  // interleaved (x, m / rho) and p records
    const size_t np = pt16 ? f : 1;
    for (V4 **q : {&P.om, &P.orhs, &P.vrhs})
      A.take(*q, np);
    A.take(P.tvg, 9 * np);
    A.take(P.pcoef, pt16 ? static_cast<size_t>(kNbrCap) * f : 1);
    for (T **q : {&P.prho, &P.psrho, &P.pone, &P.odiag, &P.vdiag, &P.scx, &P.scr, &P.scp,
                  &P.scq, &P.scy})
      A.take(*q, np);
  }

  // every device pointer into the arena (resp. the grid memory)
  template <class F>
  void each_arena_ptr(F &&f) {
    for (V4 **q : {&P_.xm, &P_.v, &P_.a, &P_.c, &P_.G, &P_.vd, &P_.tmpV4, &P_.xq, &P_.xa, &P_.bxV,
                   &P_.btmpV4, &P_.cgx, &P_.cgr, &P_.cgp, &P_.cgq, &P_.cgy, &P_.cgb, &P_.om,
                   &P_.orhs, &P_.vrhs})
      f(reinterpret_cast<void **>(q));
    for (T **q : {&P_.rho, &P_.p, &P_.pr, &P_.s, &P_.AA, &P_.Ap, &P_.tmpT, &P_.sp, &P_.tmpSp,
                  &P_.gc, &P_.tvg, &P_.prho, &P_.psrho, &P_.pone, &P_.odiag, &P_.vdiag,
                  &P_.scx, &P_.scr, &P_.scp, &P_.scq, &P_.scy, &P_.pcoef})
      f(reinterpret_cast<void **>(q));
    for (int32_t **q : {&P_.id, &P_.tmpI, &P_.bid, &P_.btmpI, &P_.permd})
      f(reinterpret_cast<void **>(q));
    for (uint32_t **q : {&P_.fcell, &P_.frank, &P_.fperm, &P_.ninfo, &P_.bcell, &P_.brank,
                         &P_.bperm})
      f(reinterpret_cast<void **>(q));
    f(reinterpret_cast<void **>(&P_.nidx));
    f(reinterpret_cast<void **>(&P_.nblk));
    f(reinterpret_cast<void **>(&P_.partials));
    f(reinterpret_cast<void **>(&P_.cg_part));
    f(reinterpret_cast<void **>(&P_.ctl));
  }
  template <class F>
  void each_grid_ptr(F &&f) {
    for (uint32_t **q : {&G_.fstart, &G_.bstart, &G_.count, &G_.scan_tmp})
      f(reinterpret_cast<void **>(q));
  }
  // pointers relative to old base (or offsets when old is null) -> new base
  // (or offsets when new is null)
  void rebase_all(char *old_a, char *new_a, char *old_g, char *new_g) {
    auto mv = [](void **q, char *from, char *to) {
      const uintptr_t off = reinterpret_cast<uintptr_t>(*q) - reinterpret_cast<uintptr_t>(from);
      *q = reinterpret_cast<void *>(reinterpret_cast<uintptr_t>(to) + off);
    };
    each_arena_ptr([&](void **q) { mv(q, old_a, new_a); });
    if (grid_bytes_ && (old_g || new_g))
      each_grid_ptr([&](void **q) { mv(q, old_g, new_g); });
  }

  void allocate() {
    Arena A;
    layout(A, P_, nf_, nb_, desc.implicit_viscosity != 0, desc.implicit_viscosity == 3);
    PRTCL_HIP_CHECK(hipMalloc(&arena_, A.off));
    bytes = A.off;
    arena_bytes_ = A.off;
    char *base = static_cast<char *>(arena_);
    for (V4 **q : {&P_.xm, &P_.v, &P_.a, &P_.c, &P_.G, &P_.vd, &P_.tmpV4, &P_.xq, &P_.xa, &P_.bxV,
                   &P_.btmpV4,
                   &P_.cgx, &P_.cgr, &P_.cgp, &P_.cgq, &P_.cgy, &P_.cgb, &P_.om, &P_.orhs,
                   &P_.vrhs})
      Arena::rebase(*q, base);
    for (T **q : {&P_.rho, &P_.p, &P_.pr, &P_.s, &P_.AA, &P_.Ap, &P_.tmpT, &P_.sp, &P_.tmpSp,
                  &P_.gc, &P_.tvg, &P_.prho, &P_.psrho, &P_.pone, &P_.odiag, &P_.vdiag,
                  &P_.scx, &P_.scr, &P_.scp, &P_.scq, &P_.scy, &P_.pcoef})
      Arena::rebase(*q, base);
    for (int32_t **q : {&P_.id, &P_.tmpI, &P_.bid, &P_.btmpI, &P_.permd})
      Arena::rebase(*q, base);
    Arena::rebase(P_.nidx, base);
    Arena::rebase(P_.nblk, base);
    for (uint32_t **q : {&P_.fcell, &P_.frank, &P_.fperm, &P_.ninfo, &P_.bcell,
                         &P_.brank, &P_.bperm})
      Arena::rebase(*q, base);
    Arena::rebase(P_.partials, base);
    Arena::rebase(P_.cg_part, base);
    Arena::rebase(P_.ctl, base);
    hipStream_t st = res_->stream;
    // fields start zero-filled (varying_field.hpp:165-167); ids are identity
    PRTCL_HIP_CHECK(hipMemsetAsync(arena_, 0, A.off, st));
    if (nf_ > 0)
      hipLaunchKernelGGL(k_iota, dim3(grid_blocks(nf_)), dim3(kBlock), 0, st, P_.id, nf_);
    if (nb_ > 0)
      hipLaunchKernelGGL(k_iota, dim3(grid_blocks(nb_)), dim3(kBlock), 0, st, P_.bid, nb_);
    // pinned, mapped, coherent progress words of the IISPH solve
    PRTCL_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&mirror_host_), 64,
                                  hipHostMallocMapped | hipHostMallocCoherent));
    // page-locked home of the last host-visible permutation (fluid, then
    // boundary), so the permutes inside steps allocate nothing
    PRTCL_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&perm_host_),
                                  4 * std::max<size_t>(1, size_t(nf_) + size_t(nb_)),
                                  hipHostMallocDefault));
    mirror_host_[0] = mirror_host_[1] = mirror_host_[2] = mirror_host_[kMirrorStep] = 0;
    PRTCL_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&mirror_dev_),
                                            mirror_host_, 0));
    ctl_op(kCtlInit);
    PRTCL_HIP_CHECK(hipGetLastError());
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  }

  CtxResources *res_;
  int nf_ = 0, nb_ = 0;
  double radius_ = 1;
  void *arena_ = nullptr;
  void *grid_mem_ = nullptr;
  uint64_t arena_bytes_ = 0;
  std::vector<char> host_arena_, host_grid_;  // state of an evicted scene
  bool grid_evicted_ = false;
  uint64_t grid_bytes_ = 0;
  Ptrs P_{};
  GridPtrs G_{};
  GridDims gd_{};
  BBox bb_f_, bb_b_;
  bool grid_built_ = false, bgrid_dirty_ = true, fgrid_dirty_ = true;
  double clock_ = 0;
  int steps_since_permute_ = 0;
  ScheduleTuner tuner_;
  Profiler tprof_;     // timing of the tuner's trial steps
  GraphCache graphs_;  // captured IISPH iterations of this scene
  int predicted_iters_ = 2;
  int predicted_cg_ = 4;
  // pt16 uniforms (strain_rate_viscosity, pt16_maximum_error / _iterations,
  // pt16_iterations) and the predicted CG lengths of its two systems
  double pt16_xi_ = 0.0, pt16_tol_ = 1e-3;
  int pt16_max_k_ = 500, pt16_iters_ = 0;
  int predicted_pt16_[2] = {4, 4};
  unsigned seq_ = 0;
  // integrations of the position enqueued (each ends with the end-of-step
  // report) and the non-finite count of the last report read
  unsigned steps_enqueued_ = 0, nonfinite_seen_ = 0;
  uint64_t *mirror_host_ = nullptr;
  unsigned long long *mirror_dev_ = nullptr;
  int last_iisph_iters_ = 0;
  double last_rel_ = 0;
  int last_wkbb_ = 0;
  int32_t *perm_host_ = nullptr;  // last host-visible permutation (page-locked)
  bool permuted_ = false, perm_on_device_ = false;
};

}  // namespace

SceneBase *make_scene(const prtcl_hip_scene_desc &desc, CtxResources *res) {
  if (desc.real_type == 1)
    return new SceneImpl<double>(desc, res);
  return new SceneImpl<float>(desc, res);
}

int step_scenes(const std::vector<SceneBase *> &scenes, int k, CtxResources *res) {
  constexpr size_t kMaxBatch = size_t(1) << (32 - kBatchLocalBits);
  std::vector<SceneImpl<float> *> bf;
  std::vector<SceneImpl<double> *> bd;
  std::vector<SceneBase *> rest;
  for (size_t i = 0; i < scenes.size(); ++i) {
    for (size_t j = 0; j < i; ++j)
      if (scenes[j] == scenes[i])
        throw std::invalid_argument("a scene is listed twice");
    scenes[i]->ensure_resident();
    if (auto *f = dynamic_cast<SceneImpl<float> *>(scenes[i]); f && f->batchable())
      bf.push_back(f);
    else if (auto *d = dynamic_cast<SceneImpl<double> *>(scenes[i]); d && d->batchable())
      bd.push_back(d);
    else
      rest.push_back(scenes[i]);
  }
  // a batch of one is the single-scene step
  if (bf.size() == 1) {
    rest.push_back(bf[0]);
    bf.clear();
  }
  if (bd.size() == 1) {
    rest.push_back(bd[0]);
    bd.clear();
  }
  int status = PRTCL_HIP_OK;
  auto keep = [&](int rc) {
    if (rc != PRTCL_HIP_OK && status == PRTCL_HIP_OK)
      status = rc;
  };
  for (int s = 0; s < k; ++s) {
    for (size_t b = 0; b < bf.size(); b += kMaxBatch) {
      std::vector<SceneImpl<float> *> part(bf.begin() + b,
                                           bf.begin() + std::min(bf.size(), b + kMaxBatch));
      keep(SceneImpl<float>::step_batch(part, res));
    }
    for (size_t b = 0; b < bd.size(); b += kMaxBatch) {
      std::vector<SceneImpl<double> *> part(bd.begin() + b,
                                            bd.begin() + std::min(bd.size(), b + kMaxBatch));
      keep(SceneImpl<double>::step_batch(part, res));
    }
    for (SceneBase *sc : rest)
      keep(sc->step(1));
  }
  // NaN / Inf in a scene's state is an error and outranks the warning of a
  // solve that stopped at its cap
  for (SceneBase *sc : scenes)
    if (sc->finish_steps() == PRTCL_HIP_NONFINITE)
      status = PRTCL_HIP_NONFINITE;
  return status;
}

int step_scenes_noalloc(const std::vector<SceneBase *> &scenes, int k, CtxResources *res,
                        const BatchTables &tables) {
  std::vector<SceneImpl<float> *> bf;
  std::vector<SceneImpl<double> *> bd;
  for (size_t i = 0; i < scenes.size(); ++i) {
    for (size_t j = 0; j < i; ++j)
      if (scenes[j] == scenes[i])
        throw std::invalid_argument("a scene is listed twice");
    SceneBase *sc = scenes[i];
    if (!sc->resident())
      throw std::invalid_argument("a listed scene is evicted (bringing it back allocates)");
    if (auto *f = dynamic_cast<SceneImpl<float> *>(sc); f && f->batchable() && f->stepped_ready())
      bf.push_back(f);
    else if (auto *d = dynamic_cast<SceneImpl<double> *>(sc);
             d && d->batchable() && d->stepped_ready())
      bd.push_back(d);
    else
      throw std::invalid_argument(
          "a listed scene cannot be stepped without allocating: it needs the fused IISPH "
          "step (step mode 0), "
          "and its grid built (by a step, or by neighborhood.update)");
  }
  if (!bf.empty() && !bd.empty())
    throw std::invalid_argument("FP32 and FP64 scenes in one call share the one table buffer");
  int status = PRTCL_HIP_OK;
  for (int s = 0; s < k; ++s) {
    int rc = PRTCL_HIP_OK;
    if (!bf.empty())
      rc = SceneImpl<float>::step_batch(bf, res, &tables);
    if (!bd.empty())
      rc = SceneImpl<double>::step_batch(bd, res, &tables);
    if (rc != PRTCL_HIP_OK && status == PRTCL_HIP_OK)
      status = rc;
  }
  for (SceneBase *sc : scenes)
    if (sc->finish_steps() == PRTCL_HIP_NONFINITE)
      status = PRTCL_HIP_NONFINITE;
  return status;
}

uint64_t scene_particle_bytes(const prtcl_hip_scene_desc &desc) {
  if (desc.real_type == 1)
    return SceneImpl<double>::particle_bytes(desc);
  return SceneImpl<float>::particle_bytes(desc);
}

}  // namespace prtcl_amd
