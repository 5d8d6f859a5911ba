// Host-side owner of one device-resident scene: the HBM arena, the uniform
// grid, the procedure dispatch behind prtcl_hip_run() and the device-driven
// IISPH loop behind prtcl_hip_iisph_solve() / prtcl_hip_step().
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/prtcl_hip.h"
#include "fused.hpp"
#include "scene.hpp"

namespace prtcl_amd {

// Stage timer: hipEvents recorded around kernel groups on the context
// stream, resolved at the next synchronisation point.
enum Stage {
  kStageSort = 0, kStageDensity, kStageGravityIntegrateV, kStageViscosity, kStageSetup,
  kStagePressureAccel, kStagePressure, kStageIntegrateX, kStageWkbb18Diag, kStagePcgSpmv,
  kStagePcgVector, kStageCorrection, kStageWkbb18DiagSetup, kStagePt16Setup, kStageCount
};
const char *stage_name(int s);

struct Profiler {
  bool on = false;
  struct Rec {
    int stage;
    int tag;  // iteration index inside a solve, -1 otherwise
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Rec> recs;
  double ms[kStageCount] = {};
  uint64_t launches[kStageCount] = {};
  hipEvent_t get();
  int begin(hipStream_t st, int stage, int tag = -1);
  void end(hipStream_t st, int rec);
  // accumulate finished records; records of solve iterations with tag >=
  // valid_iters are dropped (launches that exited at the convergence flag)
  void resolve(int valid_iters = 1 << 30);
  void reset();
};

// Captured hipGraphs of launch sequences, keyed by the bytes of their
// arguments and replayed with one hipGraphLaunch each. The sequences must not
// contain memsets (graph memset nodes are avoided on purpose) and must read
// any per-launch data from device memory, not from captured host values that
// change between replays. A key holds only argument bytes, so one cache must
// serve one launch sequence of one element type: the FP32 and the FP64
// batched iterations take the same kinds of arguments (table and block-map
// pointers, counts) and their keys can coincide.
class GraphCache {
 public:
  GraphCache() = default;
  GraphCache(const GraphCache &) = delete;
  GraphCache &operator=(const GraphCache &) = delete;
  ~GraphCache() { clear(); }
  // drops every graph (the caller makes sure none is in flight)
  void clear();
  // replays the graph of `key`, capturing enqueue() on `st` first if absent
  void launch(hipStream_t st, const std::string &key, const std::function<void()> &enqueue);

 private:
  struct Entry {
    std::string key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
  };
  std::vector<Entry> entries_;
  uint64_t clock_ = 0;
  static constexpr size_t kMaxEntries = 8;
};

// the key of a launch sequence: a tag naming the sequence and its element
// type, then the bytes of its arguments
template <class... A>
std::string graph_key(uint32_t tag, const A &...a) {
  std::string k(reinterpret_cast<const char *>(&tag), sizeof(tag));
  (k.append(reinterpret_cast<const char *>(&a), sizeof(A)), ...);
  return k;
}

// Shared per-context resources a scene borrows.
struct CtxResources {
  Profiler prof;
  hipStream_t stream = nullptr;
  // device staging for host transfers, grown on demand
  void *staging = nullptr;
  size_t staging_bytes = 0;
  // pinned host words for asynchronous flag read-back
  int32_t *pinned = nullptr;  // 64 words
  hipEvent_t ev[2] = {nullptr, nullptr};
  void *ensure_staging(size_t bytes);
  // tables of batched steps: a pinned host copy written per step, its device
  // copy, and an event marking the end of the last host -> device copy
  void *batch_host = nullptr;
  void *batch_dev = nullptr;
  size_t batch_bytes = 0;
  hipEvent_t batch_ev = nullptr;
  // schedule of the batched fused passes (FP32, FP64) and its trial timing
  ScheduleTuner batch_tuner[2];
  Profiler batch_prof;
  // captured batched IISPH iterations (FP32, FP64): one cache per type
  GraphCache batch_graphs[2];
  void ensure_batch(size_t bytes);
  void release();
};

class SceneBase {
 public:
  virtual ~SceneBase() = default;
  prtcl_hip_scene_desc desc{};
  uint64_t bytes = 0;  // device bytes held
  std::string err;
  // admission hook: returns false when `extra` more bytes exceed the budget
  std::function<bool(uint64_t)> budget_check;

  virtual void upload(int group, int field, const void *host, size_t count) = 0;
  virtual void download(int group, int field, void *host, size_t count) = 0;
  virtual void run(int proc) = 0;
  virtual double get_global(int g) = 0;
  virtual void set_global(int g, double v) = 0;
  virtual int iisph_solve(double tol, int min_it, int max_it, int *iters, double *rel,
                          bool wait) = 0;
  virtual int step(int k) = 0;
  // Waits until every step enqueued so far has finished on the device and
  // returns PRTCL_HIP_NONFINITE when one of them, since the last call,
  // produced a non-finite position or speed (else PRTCL_HIP_OK).
  virtual int finish_steps() = 0;
  virtual void copy_neighbors(int group, int ngroup, uint32_t *offsets, uint32_t *indices,
                              uint64_t *total) = 0;
  virtual void last_permutation(int group, uint32_t *perm) = 0;
  virtual void step_info(int *iisph_iters, double *rel, int *wkbb18_iters) = 0;
  virtual uint64_t grid_bytes_for_domain() const = 0;
  // fused-pass schedule: mask >= 0 forces it (bit 0 setup, 1 pressure
  // acceleration, 2 pressure in LDS), -1 returns to the tuner
  virtual void set_schedule(int mask) = 0;
  virtual int schedule() const = 0;
  // residency: evict() moves the device state to host memory and frees the
  // HBM; ensure_resident() brings it back (throws std::length_error when the
  // budget does not allow it)
  virtual bool resident() const = 0;
  virtual void evict() = 0;
  virtual void ensure_resident() = 0;
  virtual uint64_t resident_bytes_needed() const = 0;

  // checkpoint support. Components of a field of a group, 0 when the scene
  // does not hold it.
  virtual int field_components(int group, int field) = 0;
  // download() without the synchronisation: the values land in `pinned`
  // (page-locked host memory) once the context stream reaches this point
  virtual void download_async(int group, int field, void *pinned, size_t count) = 0;
  // the field in host order into device memory `dev` (count values of the
  // real type), enqueued on the context stream
  virtual void gather_device(int group, int field, void *dev, size_t count) = 0;
  // slot of every particle of a group, in host order
  virtual void slots(int group, int32_t *slot_of_host) = 0;
  // puts host particle h in slot slot_of_host[h] (a permutation of 0..n-1);
  // fields uploaded afterwards land in those slots, and the first neighbour
  // sort breaks ties in that order, so a scene saved with its slots is
  // continued bit-identically
  virtual void set_slots(int group, const int32_t *slot_of_host) = 0;
  // the step clock (current_time is this value rounded to the real type)
  // and the steps since the last host-visible permute
  virtual double exact_clock() const = 0;
  virtual void set_exact_clock(double t) = 0;
  virtual int steps_since_permute() const = 0;
  virtual void set_steps_since_permute(int k) = 0;
};

// Caller-provided home of a batched step's tables: `host` (page-locked) is
// filled and copied to `dev` (device memory of at least `bytes`). `copied`
// marks the end of the last copy out of `host`; `released` the end of the
// last kernel that read `dev` (on whichever stream used it last): the next
// copy into `dev` waits for it.
struct BatchTables {
  void *host = nullptr;
  void *dev = nullptr;
  size_t bytes = 0;
  hipEvent_t copied = nullptr;
  hipEvent_t released = nullptr;
};
// thrown when a batch's tables exceed BatchTables::bytes
struct TablesTooLarge {
  size_t bytes;
};

SceneBase *make_scene(const prtcl_hip_scene_desc &desc, CtxResources *res);
// k whole steps of every listed scene; scenes of the fused path without
// explicit or implicit viscosity are advanced together (one launch per stage
// for all of them), the others one after the other. Returns the first
// non-OK status of a scene.
int step_scenes(const std::vector<SceneBase *> &scenes, int k, CtxResources *res);
// The same steps with every batched table in `tables` and nothing allocated:
// every scene must be resident, take the batched step (fused IISPH without
// viscosity or axial gravity) and have its grid; all FP32 or all FP64.
// Throws std::invalid_argument otherwise, TablesTooLarge when the tables do
// not fit.
int step_scenes_noalloc(const std::vector<SceneBase *> &scenes, int k, CtxResources *res,
                        const BatchTables &tables);
uint64_t scene_particle_bytes(const prtcl_hip_scene_desc &desc);

}  // namespace prtcl_amd
