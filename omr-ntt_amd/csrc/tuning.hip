// tuning.hip -- schedule table filled by the autotuner (host code).
#include <map>
#include <mutex>
#include <utility>

#include "internal.h"

namespace omr {

static std::mutex g_tune_mu;
static std::map<std::pair<int, uint32_t>, Schedule> g_tuned;

bool tuned_schedule(int f, uint32_t logN, Schedule &s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find({f, logN});
  if (it == g_tuned.end()) return false;
  s = it->second;
  return true;
}

void set_tuned_schedule(int f, uint32_t logN, const Schedule &s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[{f, logN}] = s;
}

void clear_tuned_schedules() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.clear();
}

}  // namespace omr
