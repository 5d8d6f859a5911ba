// Launchers of the fused step kernels (fused.hip).
#pragma once

#include "nbrlist.hpp"

namespace prtcl_amd {

// Schedule of the three neighbour passes of the fused step: the neighbours'
// fields are gathered from global memory (through L2) or staged slab by slab
// in LDS first. Both schedules visit the same neighbours in the same order
// and give bit-identical results; which one is faster depends on the scene.
struct FusedSchedule {
  bool setup_lds = false;
  bool accel_lds = false;
  bool pressure_lds = false;
  // each IISPH iteration is replayed from a captured hipGraph unless eager
  bool eager = false;
  // the CG iterations (WKBB18, pt16) launch their one-block alpha and beta
  // kernels even where the vector kernels can derive the scalars themselves
  bool cg_split = false;
  // the WKBB18 diagonal and the CG set-up of a fused step in two list walks
  // instead of one
  bool wk_split = false;
  // IISPH iteration 1 in its own two neighbour passes instead of folded into
  // the setup pass (the fold needs no neighbours: every pressure starts at 0)
  bool first_split = false;
  int mask() const {
    return (setup_lds ? 1 : 0) | (accel_lds ? 2 : 0) | (pressure_lds ? 4 : 0) | (eager ? 8 : 0) |
           (cg_split ? 16 : 0) | (wk_split ? 32 : 0) | (first_split ? 64 : 0);
  }
  static FusedSchedule from_mask(int m) {
    FusedSchedule s;
    s.setup_lds = (m & 1) != 0;
    s.accel_lds = (m & 2) != 0;
    s.pressure_lds = (m & 4) != 0;
    s.eager = (m & 8) != 0;
    s.cg_split = (m & 16) != 0;
    s.wk_split = (m & 32) != 0;
    s.first_split = (m & 64) != 0;
    return s;
  }
};

// Picks each pass's schedule from timed trial steps: after one warm-up step,
// kTrials steps alternate the all-global and the all-LDS schedule; per pass
// the variant with the lower mean launch time wins. A forced mask (>= 0)
// bypasses the trials.
struct ScheduleTuner {
  static constexpr int kTrials = 4;
  int forced = -1;
  int steps = 0;
  double ms[3][2] = {};
  double launches[3][2] = {};
  // graph replay of the iteration measured no faster than eager launches on
  // this stack: eager unless a forced mask clears bit 3
  FusedSchedule chosen = FusedSchedule::from_mask(8);
  bool trial() const { return forced < 0 && steps >= 1 && steps <= kTrials; }
  int trial_variant() const { return (steps - 1) & 1; }
  FusedSchedule current() const {
    if (forced >= 0)
      return FusedSchedule::from_mask(forced);
    if (trial())  // timed with events between the kernels: eager launches
      return FusedSchedule::from_mask(trial_variant() ? 7 | 8 : 8);
    return chosen;
  }
  // pass: 0 setup, 1 pressure acceleration, 2 pressure
  void record(int pass, double total_ms, double n) {
    ms[pass][trial_variant()] += total_ms;
    launches[pass][trial_variant()] += n;
  }
  void finish_step() {
    if (forced >= 0)
      return;
    if (trial() && steps == kTrials) {
      bool lds[3];
      for (int p = 0; p < 3; ++p) {
        const double g = launches[p][0] > 0 ? ms[p][0] / launches[p][0] : 0.0;
        const double l = launches[p][1] > 0 ? ms[p][1] / launches[p][1] : 1e30;
        lds[p] = l < g;
      }
      chosen.setup_lds = lds[0];
      chosen.accel_lds = lds[1];
      chosen.pressure_lds = lds[2];
      chosen.eager = true;
    }
    ++steps;
  }
  void force(int mask) {
    forced = mask;
    if (mask < 0) {  // restart the trials
      steps = 0;
      for (auto &r : ms)
        r[0] = r[1] = 0;
      for (auto &r : launches)
        r[0] = r[1] = 0;
    }
  }
};

template <class T>
struct FusedOps {
  // density, neighbour list, IISPH c / AA / s, integrated velocity into vtmp
  // B.seq != 0: the first kernel also begins solve B (one launch fewer)
  static void setup(hipStream_t st, const SceneView<T> &S, const NbrList &L,
                    bool explicit_viscosity, vec4_t<T> *vtmp, bool lds,
                    const SolveBegin &B = SolveBegin{});
  // the pressure acceleration goes to the second half of S.xa only
  static void accel(hipStream_t st, const SceneView<T> &S, const NbrList &L, bool lds);
  static void pressure(hipStream_t st, const SceneView<T> &S, const NbrList &L, bool lds);
  // the convergence step of IISPH iteration 1, whose pressures and aprde
  // partials the setup (with S.first_iter) has written (batched: for the
  // scenes without first_preset)
  static void finish_first_iteration(hipStream_t st, const SceneView<T> &S);
  static void finish_first_iteration_batch(hipStream_t st, const BatchEntry<T> *tab,
                                           unsigned nscenes);
  // v += dt a H, x += dt v, with a from S.xa, also stored to S.a; seq > 0:
  // enqueued ahead of the host's knowledge that IISPH solve seq is done, it
  // integrates only once that solve is done, and only once (integrate_now)
  static void integrate(hipStream_t st, const SceneView<T> &S, unsigned seq = 0);
  // S.a from S.xa (before a procedure kernel reads the acceleration)
  static void copy_accel(hipStream_t st, const SceneView<T> &S);
  // the same passes over every scene of a batch (pmap: block map)
  static void setup_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                          unsigned npblocks, bool lds);
  // the three setup launches of scenes with explicit viscosity (entry flag
  // explicit_visc) or axial gravity, over the blocks of `map`
  static void setup_split_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *map,
                                unsigned nblocks, bool lds);
  static void accel_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                          unsigned npblocks, bool lds);
  static void pressure_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                             unsigned npblocks, unsigned nscenes, bool lds);
  // guarded: each scene as integrate() with seq = its solve's sequence
  static void integrate_batch(hipStream_t st, const BatchEntry<T> *tab, const uint32_t *pmap,
                              unsigned npblocks, unsigned nscenes, bool guarded = false);
};

}  // namespace prtcl_amd
