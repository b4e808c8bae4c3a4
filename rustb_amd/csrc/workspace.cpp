// Call workspaces: per device, a few slots, each a HIP stream and one device
// block that grows to the largest workspace a call on it has needed.  A call
// leases a slot, carves its buffers from the block with a bump pointer and
// returns the slot when it ends; once a slot's block is large enough, calls
// on it allocate nothing.  (Stream-ordered pool allocations were used before:
// memory freed on one call's stream was not reused by the next call's stream,
// so every call past the first few mapped tens of GB anew.)
#include <condition_variable>
#include <cstdlib>
#include <map>
#include <mutex>

#include "rtb_internal.hpp"

namespace rtb {

struct Slot {
  int device = 0;
  hipStream_t stream = nullptr;
  void* block = nullptr;  // the reusable block
  size_t cap = 0;
  std::vector<void*> extra;  // blocks added during a call that outgrew `block`
  size_t need = 0;           // bytes (aligned) the current call has taken
  std::vector<hipEvent_t> events;  // timing events, created once, reused
};

namespace {

constexpr size_t kAlign = 256;
constexpr size_t kGrain = size_t(64) << 20;  // block sizes are multiples of this

struct DeviceSlots {
  std::vector<Slot*> all;
  std::vector<Slot*> idle;
};

std::mutex g_mu;
std::condition_variable g_cv;
std::map<int, DeviceSlots> g_slots;

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

int max_concurrent_calls() {
  if (const char* env = std::getenv("RTB_MAX_CONCURRENT_CALLS")) {
    const int v = std::atoi(env);
    if (v >= 1) return v;
  }
  return 4;
}

Lease::Lease() : rc_(RTB_OK) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    rc_ = fail(RTB_ERR_DEVICE, "no HIP device");
    return;
  }
  const size_t limit = static_cast<size_t>(max_concurrent_calls());
  std::unique_lock<std::mutex> lock(g_mu);
  DeviceSlots& ds = g_slots[dev];
  g_cv.wait(lock, [&] { return !ds.idle.empty() || ds.all.size() < limit; });
  if (!ds.idle.empty()) {
    s_ = ds.idle.back();
    ds.idle.pop_back();
  } else {
    auto* s = new Slot();
    s->device = dev;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
      delete s;
      rc_ = fail(RTB_ERR_DEVICE, "stream creation failed");
      return;
    }
    ds.all.push_back(s);
    s_ = s;
  }
  s_->need = 0;
}

Lease::~Lease() {
  if (!s_) return;
  (void)hipStreamSynchronize(s_->stream);
  if (!s_->extra.empty()) {
    // the call outgrew the block: replace block and extras by one block of
    // the call's total, so that the next call of this size allocates nothing
    for (void* p : s_->extra) (void)hipFree(p);
    s_->extra.clear();
    if (s_->block) (void)hipFree(s_->block);
    s_->block = nullptr;
    s_->cap = 0;
    const size_t want = round_up(s_->need, kGrain);
    if (hipMalloc(&s_->block, want) == hipSuccess) {
      s_->cap = want;
    } else {
      s_->block = nullptr;
      (void)hipGetLastError();
    }
  }
  {
    std::lock_guard<std::mutex> lock(g_mu);
    g_slots[s_->device].idle.push_back(s_);
  }
  g_cv.notify_one();
}

hipStream_t Lease::stream() const { return s_ ? s_->stream : nullptr; }

int Lease::event(size_t i, hipEvent_t* out) {
  if (!s_) return fail(RTB_ERR_DEVICE, "no workspace slot");
  while (s_->events.size() <= i) {
    hipEvent_t e;
    RTB_HIP(hipEventCreate(&e));
    s_->events.push_back(e);
  }
  *out = s_->events[i];
  return RTB_OK;
}

Work::Work(Lease& l) : slot_(l.slot()) {
  if (slot_) {
    base_ = static_cast<char*>(slot_->block);
    cap_ = slot_->cap;
  }
}

Work::Work(void* base, size_t cap) : base_(static_cast<char*>(base)), cap_(cap) {}

int Work::take(size_t bytes, void** out) {
  bytes = round_up(bytes == 0 ? 1 : bytes, kAlign);
  *out = nullptr;
  if (off_ + bytes > cap_) {
    if (!slot_) {
      set_error("workspace of %zu bytes exceeded (%zu taken, %zu more requested)", cap_, off_, bytes);
      return RTB_ERR_UNSUPPORTED;
    }
    // grow: a new block for this and later requests of the call
    const size_t sz = round_up(bytes, kGrain);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed: %s", sz, hipGetErrorString(e));
      return RTB_ERR_DEVICE;
    }
    slot_->extra.push_back(p);
    base_ = static_cast<char*>(p);
    cap_ = sz;
    off_ = 0;
  }
  *out = base_ ? base_ + off_ : nullptr;  // no base: only counting bytes
  off_ += bytes;
  if (slot_) slot_->need += bytes;
  used_ += bytes;
  return RTB_OK;
}

}  // namespace rtb

using namespace rtb;

extern "C" {

int rtb_workspace_release(void) {
  int dev = 0;
  RTB_HIP(hipGetDevice(&dev));
  std::vector<Slot*> drop;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    DeviceSlots& ds = g_slots[dev];
    drop.swap(ds.idle);
    std::vector<Slot*> keep;
    for (Slot* s : ds.all) {
      bool idle = false;
      for (Slot* d : drop) idle = idle || d == s;
      if (!idle) keep.push_back(s);
    }
    ds.all.swap(keep);
  }
  g_cv.notify_all();
  for (Slot* s : drop) {
    (void)hipStreamSynchronize(s->stream);
    if (s->block) (void)hipFree(s->block);
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s->stream);
    delete s;
  }
  return RTB_OK;
}

int rtb_workspace_bytes(size_t* bytes, int* slots) {
  if (!bytes) return fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  int dev = 0;
  RTB_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  size_t b = 0;
  auto it = g_slots.find(dev);
  int n = 0;
  if (it != g_slots.end()) {
    n = static_cast<int>(it->second.all.size());
    for (Slot* s : it->second.all) b += s->cap;
  }
  *bytes = b;
  if (slots) *slots = n;
  return RTB_OK;
}

}  // extern "C"
