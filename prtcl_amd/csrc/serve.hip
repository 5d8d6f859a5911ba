// Request loop of the serving path (serve.hpp).
#include "serve.hpp"

#include <algorithm>
#include <stdexcept>
#include <utility>

namespace prtcl_amd {

int64_t Server::submit(int32_t scene, int32_t k_steps) {
  if (k_steps < 0)
    throw std::invalid_argument("k_steps must be >= 0");
  const int64_t t = next_ticket_++;
  if (k_steps == 0)
    done_.push_back(t);
  else
    queue_[scene].push_back(Request{t, k_steps});
  return t;
}

bool Server::make_room(const Hooks &h, uint64_t need, int32_t keep) {
  if (h.budget == 0 || h.bytes_in_use() + need <= h.budget)
    return true;
  // idle resident scenes (no pending work), least recently stepped first
  std::vector<std::pair<uint64_t, int32_t>> idle;
  for (int32_t o : h.scene_ids()) {
    if (o == keep)
      continue;
    auto q = queue_.find(o);
    if (q != queue_.end() && !q->second.empty())
      continue;
    SceneBase *os = h.find(o);
    if (os->resident() && os->bytes > 0) {
      auto lu = last_used_.find(o);
      idle.emplace_back(lu == last_used_.end() ? 0 : lu->second, o);
    }
  }
  std::sort(idle.begin(), idle.end());
  for (const auto &e : idle) {
    if (h.bytes_in_use() + need <= h.budget)
      break;
    h.find(e.second)->evict();
  }
  return h.bytes_in_use() + need <= h.budget;
}

int Server::tick(const Hooks &h, CtxResources *res, int *status) {
  // scenes with work, oldest pending request first (first come, first
  // admitted when HBM is short)
  std::vector<std::pair<int64_t, int32_t>> order;
  for (auto &kv : queue_)
    if (!kv.second.empty())
      order.emplace_back(kv.second.front().ticket, kv.first);
  std::sort(order.begin(), order.end());
  std::vector<SceneBase *> batch;
  std::vector<int32_t> ids;
  for (const auto &o : order) {
    SceneBase *sc = h.find(o.second);
    // a scene that cannot be brought back within the budget waits
    if (!make_room(h, sc->resident_bytes_needed(), o.second))
      continue;
    sc->ensure_resident();
    batch.push_back(sc);
    ids.push_back(o.second);
  }
  *status = PRTCL_HIP_OK;
  if (batch.empty())
    return 0;
  *status = step_scenes(batch, 1, res);
  // a request is complete when its last step has finished on the device
  PRTCL_HIP_CHECK(hipStreamSynchronize(res->stream));
  ++clock_;
  for (int32_t id : ids) {
    last_used_[id] = clock_;
    auto &q = queue_[id];
    if (--q.front().remaining == 0) {
      done_.push_back(q.front().ticket);
      q.pop_front();
    }
  }
  return static_cast<int>(batch.size());
}

int Server::poll(int64_t *tickets, int cap) {
  int n = 0;
  while (n < cap && !done_.empty()) {
    tickets[n++] = done_.front();
    done_.pop_front();
  }
  return n;
}

void Server::pending(int64_t *requests, int64_t *steps) const {
  int64_t r = 0, s = 0;
  for (const auto &kv : queue_)
    for (const Request &q : kv.second) {
      ++r;
      s += q.remaining;
    }
  if (requests)
    *requests = r;
  if (steps)
    *steps = s;
}

void Server::forget(int32_t scene) {
  queue_.erase(scene);
  last_used_.erase(scene);
}

}  // namespace prtcl_amd
