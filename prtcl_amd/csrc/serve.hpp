// Request loop of the serving path: requests ask to advance a resident scene
// by k steps; every tick advances each scene that has pending work by one
// step, all of them in one batched step (continuous batching at step
// granularity: a request submitted between two ticks joins at the next one,
// and a finished request leaves without waiting for the others). Scenes whose
// device state was evicted are brought back when they have work and the HBM
// budget allows it, evicting idle scenes (no pending work, least recently
// stepped first) to make room; a scene that still does not fit waits.
#pragma once

#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <vector>

#include "scene_impl.hpp"

namespace prtcl_amd {

class Server {
 public:
  struct Hooks {
    std::function<SceneBase *(int32_t)> find;          // throws for an unknown id
    std::function<std::vector<int32_t>()> scene_ids;   // every scene of the context
    std::function<uint64_t()> bytes_in_use;
    uint64_t budget = 0;                               // 0: unlimited
  };

  int64_t submit(int32_t scene, int32_t k_steps);
  // one step of every scene with pending work; returns the number of scenes
  // stepped and the first non-OK step status in *status
  int tick(const Hooks &h, CtxResources *res, int *status);
  // completed tickets since the last call (up to cap; the rest stay queued)
  int poll(int64_t *tickets, int cap);
  void pending(int64_t *requests, int64_t *steps) const;
  // a destroyed scene's requests are dropped
  void forget(int32_t scene);
  // evicts idle scenes (no pending work, least recently stepped first, never
  // `keep`) until `need` more bytes fit the budget; false if they cannot
  bool make_room(const Hooks &h, uint64_t need, int32_t keep = -1);

 private:
  struct Request {
    int64_t ticket;
    int32_t remaining;
  };
  std::map<int32_t, std::deque<Request>> queue_;  // per scene, FIFO
  std::deque<int64_t> done_;
  std::map<int32_t, uint64_t> last_used_;
  uint64_t clock_ = 0;
  int64_t next_ticket_ = 1;

};

}  // namespace prtcl_amd
