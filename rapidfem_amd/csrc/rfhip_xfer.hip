// Large host <-> device copies of pageable host buffers through a reused
// ring of two page-locked slots per device: the DMA engine moves one slot
// while host threads copy the other to or from the caller's buffer (which
// also spreads the first-touch page faults of a fresh destination over the
// threads). Copies below kStageMin go through hipMemcpyAsync directly.
// Measured on a 311 MB download into fresh memory: pageable hipMemcpy
// 4.1 GB/s, the ring 20 GB/s with 8 threads and 27.5 GB/s with 16 (the
// host copy into fresh pages is the limit; pinned DMA alone does 57 GB/s).
#include <algorithm>
#include <map>
#include <mutex>
#include <thread>

#include "rfhip_impl.hpp"

namespace rfhip {
namespace {

constexpr size_t kSlot = size_t(64) << 20;      // bytes per pinned slot
constexpr size_t kStageMin = size_t(16) << 20;  // smaller copies skip the ring
constexpr size_t kThreadSlice = size_t(4) << 20;

struct Ring {
  std::mutex mu;
  HostPinned slot[2];
  hipEvent_t ev[2] = {nullptr, nullptr};
};

Ring &ring_for_current_device() {
  static std::mutex mu;
  static std::map<int, Ring *> rings;  // held until the process ends
  int dev = 0;
  RF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  Ring *&r = rings[dev];
  if (!r) {
    auto *nr = new Ring;
    for (int k = 0; k < 2; ++k) {
      nr->slot[k].alloc(kSlot);
      RF_HIP(hipEventCreateWithFlags(&nr->ev[k], hipEventDisableTiming));
    }
    r = nr;
  }
  return *r;
}

// memcpy with up to 16 host threads, slices of at least kThreadSlice bytes.
void host_copy(void *dst, const void *src, size_t bytes) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>({16, hw, std::max<size_t>(1, bytes / kThreadSlice)});
  if (nt <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t per = (bytes + nt - 1) / nt;
  auto slice = [=](size_t a, size_t b) {
    std::memcpy(static_cast<char *>(dst) + a, static_cast<const char *>(src) + a, b - a);
  };
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  for (size_t t = 1; t < nt; ++t) {
    const size_t a = t * per, b = std::min(bytes, a + per);
    if (a >= b) continue;
    try {
      th.emplace_back(slice, a, b);
    } catch (...) {
      slice(a, b);  // no thread available: this slice on the calling thread
    }
  }
  slice(0, std::min(bytes, per));
  for (auto &x : th) x.join();
}

}  // namespace

void copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes < kStageMin) {
    RF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    RF_HIP(hipStreamSynchronize(s));
    return;
  }
  Ring &r = ring_for_current_device();
  std::lock_guard<std::mutex> g(r.mu);
  const size_t nch = (bytes + kSlot - 1) / kSlot;
  auto len = [&](size_t k) { return std::min(kSlot, bytes - k * kSlot); };
  auto issue = [&](size_t k) {
    const int sl = (int)(k & 1);
    // a host-to-device copy of an earlier call may still read this slot
    RF_HIP(hipEventSynchronize(r.ev[sl]));
    RF_HIP(hipMemcpyAsync(r.slot[sl].p, static_cast<const char *>(src) + k * kSlot, len(k), hipMemcpyDeviceToHost,
                          s));
    RF_HIP(hipEventRecord(r.ev[sl], s));
  };
  issue(0);
  for (size_t k = 0; k < nch; ++k) {
    // chunk k + 1 lands in the slot whose chunk k - 1 the host has copied out
    if (k + 1 < nch) issue(k + 1);
    RF_HIP(hipEventSynchronize(r.ev[k & 1]));
    host_copy(static_cast<char *>(dst) + k * kSlot, r.slot[k & 1].p, len(k));
  }
}

void copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes < kStageMin) {
    RF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  Ring &r = ring_for_current_device();
  std::lock_guard<std::mutex> g(r.mu);
  const size_t nch = (bytes + kSlot - 1) / kSlot;
  for (size_t k = 0; k < nch; ++k) {
    const int sl = (int)(k & 1);
    const size_t n = std::min(kSlot, bytes - k * kSlot);
    // the copy of chunk k - 2 (or of an earlier call) has finished reading the slot
    RF_HIP(hipEventSynchronize(r.ev[sl]));
    host_copy(r.slot[sl].p, static_cast<const char *>(src) + k * kSlot, n);
    RF_HIP(hipMemcpyAsync(static_cast<char *>(dst) + k * kSlot, r.slot[sl].p, n, hipMemcpyHostToDevice, s));
    RF_HIP(hipEventRecord(r.ev[sl], s));
  }
}

}  // namespace rfhip
