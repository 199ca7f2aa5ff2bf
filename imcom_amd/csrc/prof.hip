// HIP-event stage timing for the solve path (see Prof in kernels.h).
#include "common.h"
#include "kernels.h"

namespace imcom {

static constexpr size_t kNoToken = (size_t)-1;

hipError_t Prof::begin(int slot, int need_level, size_t* token) {
  (void)slot;
  *token = kNoToken;
  if (level < need_level) return hipSuccess;
  while (pool.size() < used + 2) {
    hipEvent_t e;
    IMCOM_HIP_TRY(hipEventCreate(&e));
    pool.push_back(e);
  }
  *token = used;
  used += 2;
  return hipEventRecord(pool[*token], st);
}

hipError_t Prof::end(int slot, size_t token) {
  if (token == kNoToken) return hipSuccess;
  IMCOM_HIP_TRY(hipEventRecord(pool[token + 1], st));
  pending.emplace_back(slot, token);
  return hipSuccess;
}

hipError_t Prof::events(int slot, int need_level, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (level < need_level) return hipSuccess;
  while (pool.size() < used + 2) {
    hipEvent_t e;
    IMCOM_HIP_TRY(hipEventCreate(&e));
    pool.push_back(e);
  }
  *e0 = pool[used];
  *e1 = pool[used + 1];
  pending.emplace_back(slot, used);
  used += 2;
  return hipSuccess;
}

hipError_t Prof::collect() {
  if (pending.empty()) return hipSuccess;
  IMCOM_HIP_TRY(hipStreamSynchronize(st));
  for (auto& p : pending) {
    float t = 0.0f;
    IMCOM_HIP_TRY(hipEventElapsedTime(&t, pool[p.second], pool[p.second + 1]));
    ms[p.first] += t;
    count[p.first] += 1;
  }
  pending.clear();
  used = 0;
  return hipSuccess;
}

void Prof::reset() {
  for (int i = 0; i < PS_NSLOTS; ++i) {
    ms[i] = 0.0;
    count[i] = 0;
  }
  pending.clear();
  used = 0;
}

void Prof::destroy() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  pool.clear();
  reset();
}

}  // namespace imcom
