// rcg_solver.cpp -- handles, the device engine, the Rayleigh-Chebyshev state
// machine and the extern "C" entry points of librcg.so.
//
// The state machine is RayleighChebyshevLM::getMinIntervalEigenSystem_Base
// (RayleighChebyshevLM.h:817-1776): the same shift/degree rules, stop
// conditions, locking, guard and shrink logic, driven over device blocks.
// What runs where:
//   device : Lanczos vectors, Chebyshev filter steps, Gram matrices, block
//            rotations and deflation, norms and residuals, start vectors;
//   host   : scalars of the state machine, the b x b Cholesky and Jacobi.
// One stream per context; the host waits once per Gram/projection result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/rcg_c_api.h"
#include "rcg_cpoly.hpp"
#include "rcg_dense.hpp"
#include "rcg_internal.hpp"

using rcg::Error;
using cplx = std::complex<double>;

inline double conj_s(double x) { return x; }
inline cplx conj_s(cplx x) { return std::conj(x); }

// Waits for an event.  hipEventSynchronize busy-waits even for events made
// with hipEventBlockingSync, and dozens of server threads waiting that way
// take the host cores the solves need; a polite wait queries the event,
// yields a few times, then sleeps in short steps.
inline void wait_event(hipEvent_t e, bool polite) {
  if (!polite) {
    RCG_HIP(hipEventSynchronize(e));
    return;
  }
  for (int k = 0;; ++k) {
    const hipError_t r = hipEventQuery(e);
    if (r == hipSuccess) return;
    (void)hipGetLastError();  // hipErrorNotReady is not an error
    if (r != hipErrorNotReady) RCG_HIP(r);
    if (k < 8) std::this_thread::yield();
    else std::this_thread::sleep_for(std::chrono::microseconds(k < 200 ? 20 : 200));
  }
}

// ------------------------------------------------------------------ memory
// Cached device and pinned-host allocations.  hipFree and hipHostFree
// synchronise the whole device, so a solve that frees its buffers would wait
// for every other solve in flight (the server runs many at once).  Buffers
// bound to a stream are returned to a free list instead, in size classes of
// 2^k * {1, 1.25, 1.5, 1.75}, after that stream has drained; later requests
// of the same class (on any thread) reuse them.  At most kMaxCached bytes
// per device (and of pinned memory) are kept; rcg_memory_trim frees them.
class MemPool {
 public:
  static MemPool& get() {
    static MemPool* p = new MemPool;  // never destroyed: no HIP calls at exit
    return *p;
  }
  static size_t size_class(size_t b) {
    if (b <= 4096) return 4096;
    const int k = 63 - __builtin_clzll((unsigned long long)(b - 1));  // 2^k < b <= 2^(k+1)
    const size_t base = size_t(1) << k, step = base / 4;
    return base + ((b - base + step - 1) / step) * step;
  }
  // device < 0: pinned host memory
  void* alloc(int device, size_t bytes, size_t* cap) {
    *cap = size_class(bytes);
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = free_.find({device, *cap});
      if (it != free_.end() && !it->second.empty()) {
        void* q = it->second.back();
        it->second.pop_back();
        cached_[device] -= *cap;
        return q;
      }
    }
    void* q = nullptr;
    hipError_t e = raw_alloc(device, &q, *cap);
    if (e != hipSuccess) {  // out of memory: give the cached buffers back and retry once
      (void)hipGetLastError();
      trim(device);
      e = raw_alloc(device, &q, *cap);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(RCG_ERR_ALLOC, std::string(device < 0 ? "hipHostMalloc" : "hipMalloc") +
                                       " of " + std::to_string(*cap) + " bytes failed: " +
                                       hipGetErrorString(e));
      }
    }
    return q;
  }
  void release(int device, void* q, size_t cap) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (cached_[device] + cap <= kMaxCached) {
        free_[{device, cap}].push_back(q);
        cached_[device] += cap;
        return;
      }
    }
    raw_free(device, q);
  }
  // frees the cached buffers of one device (device < 0: pinned), or of all
  // (device == kAll)
  void trim(int device = kAll) {
    std::vector<std::pair<int, void*>> out;
    {
      std::lock_guard<std::mutex> g(mu_);
      for (auto it = free_.begin(); it != free_.end();) {
        if (device == kAll || it->first.first == device) {
          for (void* q : it->second) out.push_back({it->first.first, q});
          cached_[it->first.first] = 0;
          it = free_.erase(it);
        } else {
          ++it;
        }
      }
    }
    for (auto& d : out) raw_free(d.first, d.second);
  }
  static constexpr int kAll = -1000;

 private:
  static constexpr size_t kMaxCached = size_t(24) << 30;
  static hipError_t raw_alloc(int device, void** q, size_t b) {
    if (device < 0) return hipHostMalloc(q, b, hipHostMallocDefault);
    return hipMalloc(q, b);
  }
  static void raw_free(int device, void* q) {
    if (device < 0) (void)hipHostFree(q);
    else (void)hipFree(q);
  }
  std::mutex mu_;
  std::map<std::pair<int, size_t>, std::vector<void*>> free_;
  std::map<int, size_t> cached_;
};

// A device buffer.  Bound to a stream (use_stream), it comes from and goes
// back to the cache; the stream is drained before the memory is handed on.
// Unbound, it is a plain hipMalloc / hipFree allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int dev = 0;
  bool pooled = false;
  hipStream_t stream = nullptr;
  DevBuf() = default;
  explicit DevBuf(hipStream_t s) { use_stream(s); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void use_stream(hipStream_t s) {
    pooled = true;
    stream = s;
  }
  void release() {
    if (p) {
      if (pooled) {
        (void)hipStreamSynchronize(stream);
        MemPool::get().release(dev, p, bytes);
      } else {
        (void)hipFree(p);
      }
    }
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (b == 0) return;
    if (pooled) {
      RCG_HIP(hipGetDevice(&dev));
      p = MemPool::get().alloc(dev, b, &bytes);
      return;
    }
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Error(RCG_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(b) +
                                     " bytes failed: " + hipGetErrorString(e));
    }
    bytes = b;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Pinned host staging; cached like DevBuf when bound to a stream.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool pooled = false;
  hipStream_t stream = nullptr;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { release(); }
  void use_stream(hipStream_t s) {
    pooled = true;
    stream = s;
  }
  void release() {
    if (p) {
      if (pooled) {
        (void)hipStreamSynchronize(stream);
        MemPool::get().release(-1, p, bytes);
      } else {
        (void)hipHostFree(p);
      }
    }
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (pooled) {
      p = MemPool::get().alloc(-1, b, &bytes);
      return;
    }
    RCG_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
    bytes = b;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Pinned host memory that reduction kernels write their results into, read on
// the host once the stream has been waited for: coherent, so the results need
// no device-to-host copy (three copies fewer per inner iteration).  Kept for
// the context's lifetime and only grown.
struct CohBuf {
  void* h = nullptr;
  void* d = nullptr;  // the same memory as the device addresses it
  size_t bytes = 0;
  CohBuf() = default;
  CohBuf(const CohBuf&) = delete;
  CohBuf& operator=(const CohBuf&) = delete;
  ~CohBuf() {
    if (h) (void)hipHostFree(h);
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (h) {
      RCG_HIP(hipHostFree(h));
      h = d = nullptr;
      bytes = 0;
    }
    RCG_HIP(hipHostMalloc(&h, b, hipHostMallocCoherent | hipHostMallocMapped));
    RCG_HIP(hipHostGetDevicePointer(&d, h, 0));
    bytes = b;
  }
};

// ----------------------------------------------------------------- handles
struct rcg_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // set for server workers: waits block the thread on this event instead of
  // spinning, so many workers share the host cores
  hipEvent_t blocking_ev = nullptr;
  DevBuf ws;      // reduction partials
  DevBuf small;   // b x b matrices and b-vectors on the device
  HostBuf pinned; // host staging of the same
  CohBuf results; // reduction results the host reads (not released with the scratch)
  void bind_scratch() {
    ws.use_stream(stream);
    small.use_stream(stream);
    pinned.use_stream(stream);
  }
  void release_scratch() {
    ws.release();
    small.release();
    pinned.release();
  }
};

struct rcg_block_s {
  rcg_context_s* ctx = nullptr;
  int64_t n = 0, cols = 0, ld = 0;
  int dtype = rcg::F64;
  DevBuf mem;
  void* d() const { return mem.p; }
};

// A candidate visiting order of the CSR kernels' row blocks.
struct SchedCand {
  std::string name;
  DevBuf order;
  rcg::CsrSched sc;
};

struct rcg_operator_s {
  rcg_context_s* ctx = nullptr;
  int64_t n = 0, nnz = 0;
  int dtype = rcg::F64;
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const double* val = nullptr;
  DevBuf own_rowptr, own_col, own_val;
  int64_t gx = 0, gy = 0, gz = 0;  // structured-grid hint (rows x + gx(y + gy z))
  std::vector<std::unique_ptr<SchedCand>> scheds;
  struct Tuned {
    int panel;
    int sched;
  };
  std::map<int64_t, Tuned> tuned;  // autotuned filter shape per (kind, ld, columns)
  DevBuf val32;                    // FP32 copy of val for the mixed-precision filter
  // dense symmetric operator (kind 1): n x n, row-major = column-major
  std::mutex lazy_mu;              // guards the lazily built members below and above
  int kind = 0;                    // 0 CSR, 1 dense, 2 caller's apply
  const double* dense = nullptr;
  int64_t lda = 0;
  DevBuf own_dense;
  // caller's operator (kind 2): out = alpha A in + beta out on column-major blocks
  rcg_apply_fn apply_fn = nullptr;
  void* user = nullptr;
  bool is_dense() const { return kind == 1; }
  bool is_csr() const { return kind == 0; }
  bool is_callback() const { return kind == 2; }
};

namespace {

thread_local std::string g_last_error;

}  // namespace

void rcg::set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    return RCG_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return RCG_ERR_ALLOC;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return RCG_ERR_INVALID;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) throw Error(RCG_ERR_INVALID, msg);
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    RCG_HIP(hipGetDevice(&prev));
    if (prev != dev) RCG_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

int64_t block_ld(int64_t cols) { return std::max<int64_t>(4, rcg::round_up(cols, 4)); }

// ---------------------------------------------- no-allocation workspace
// rcg_residuals_noalloc keeps its reduction partials, Ritz values and sums
// in `scratch`, a file-scope device array of kNoAllocDoubles (16 MiB) that
// each device holds for the life of the process.  In host code `scratch` is
// only a stand-in for the symbol: its device address, one per device, comes
// from hipGetSymbolAddress (made once per device when a context is created
// there, with that device current) and only that address is used.
constexpr size_t kNoAllocDoubles = size_t(2) << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current.
double* noalloc_scratch(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices)
    throw Error(RCG_ERR_UNSUPPORTED, "device index beyond the no-allocation workspace table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    RCG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(RCG_ERR_HIP, "hipGetSymbolAddress returned no address for the workspace");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[device];
}

double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------- filter batcher
// Collects the filter sweeps of concurrent solves (each on its own stream and
// host thread) and runs them together: for every step j one launch per
// launch class covers the row blocks of all requests whose degree reaches j.
// A solve submits once its stream has drained (or, with `ready`, the
// batcher's stream waits on that event); the batcher records `done` after the
// request's last step.  Up to kDepth batches are in flight on the batcher's stream,
// each with its own slot of task tables, so the next batch is gathered and
// enqueued while the previous one runs; with nothing in flight the batcher
// waits up to window_us for more requests.
struct FilterReq {
  rcg::BatchDesc desc{};
  int cls = 0;
  bool f32 = false, src = false, cx = false;
  int jstart = 1;
  double step_bytes = 0;  // algorithmic bytes of one step (instrumentation)
  hipEvent_t ready = nullptr, done = nullptr;
  bool enqueued = false;
};

class FilterBatcher {
  static constexpr int kDepth = 2;            // batches in flight
  static constexpr int kSlots = kDepth + 1;  // task-table slots
  struct Slot {
    DevBuf ddesc, dtask;
    HostBuf hdesc, htask;
    hipEvent_t end = nullptr;  // the slot's last batch has finished
    bool used = false;
  };

 public:
  FilterBatcher(int device, int window_us) : device_(device), window_us_(window_us) {
    RCG_HIP(hipSetDevice(device_));
    RCG_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    for (Slot& sl : slots_) {
      sl.hdesc.use_stream(st_);
      sl.htask.use_stream(st_);
      sl.ddesc.use_stream(st_);
      sl.dtask.use_stream(st_);
      RCG_HIP(hipEventCreateWithFlags(&sl.end, hipEventDisableTiming | hipEventBlockingSync));
    }
    thr_ = std::thread([this] { loop(); });
  }
  ~FilterBatcher() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    if (thr_.joinable()) thr_.join();
    (void)hipStreamSynchronize(st_);
    {
      std::lock_guard<std::mutex> g(prof_mu_);
      harvest_locked(true);
      for (hipEvent_t e : ev_free_) (void)hipEventDestroy(e);
      ev_free_.clear();
    }
    for (Slot& sl : slots_) {  // bound to st_
      sl.ddesc.release();
      sl.dtask.release();
      sl.hdesc.release();
      sl.htask.release();
      (void)hipEventDestroy(sl.end);
    }
    (void)hipStreamDestroy(st_);
  }
  // Solves that may submit; with nothing in flight the batcher launches as
  // soon as all of them wait.
  void attach() {
    std::lock_guard<std::mutex> g(mu_);
    ++members_;
  }
  void detach() {
    {
      std::lock_guard<std::mutex> g(mu_);
      --members_;
    }
    cv_.notify_all();
  }
  // Blocks until the request's steps are enqueued (not until they finish).
  void submit(FilterReq* r) {
    std::unique_lock<std::mutex> g(mu_);
    if (!error_.empty()) throw Error(RCG_ERR_HIP, error_);
    pending_.push_back(r);
    cv_.notify_all();
    cv_done_.wait(g, [&] { return r->enqueued || !error_.empty(); });
    if (!r->enqueued) {
      // r lives in the caller's frame, which the throw ends: it must not
      // stay queued (the loop never runs a batch once error_ is set)
      pending_.erase(std::remove(pending_.begin(), pending_.end(), r), pending_.end());
      throw Error(RCG_ERR_HIP, error_);
    }
  }
  int64_t batches() const { return batches_; }
  int64_t requests() const { return requests_; }
  // out[4]: device seconds spent in the timed batches (first launch to last),
  // kernel launches, algorithmic bytes of the launched steps, timed batches
  void profile(double* out) {
    std::lock_guard<std::mutex> g(prof_mu_);
    harvest_locked(false);
    out[0] = busy_s_;
    out[1] = (double)launches_;
    out[2] = bytes_;
    out[3] = (double)timed_;
  }

 private:
  // drops finished batches from the in-flight list (the oldest first)
  void retire_finished() {
    while (!inflight_.empty() && hipEventQuery(slots_[inflight_.front()].end) == hipSuccess)
      inflight_.pop_front();
    (void)hipGetLastError();
  }

  void loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> g(mu_);
    int next_slot = 0;
    while (true) {
      cv_.wait(g, [&] { return stop_ || !pending_.empty(); });
      if (stop_ && pending_.empty()) return;
      retire_finished();
      if ((int)inflight_.size() >= kDepth) {
        // the device has work queued: wait for the oldest batch while more
        // requests gather
        const hipEvent_t e = slots_[inflight_.front()].end;
        g.unlock();
        wait_event(e, true);
        g.lock();
        continue;
      }
      if (inflight_.empty()) {
        // the device is idle: gather until every attached solve waits here,
        // or the window closes
        const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(window_us_);
        cv_.wait_until(g, until, [&] { return stop_ || (int)pending_.size() >= members_; });
      }
      if (!error_.empty()) {
        // a failed launch leaves the stream's state unknown: no further
        // batches; the waiting solves see error_ and throw
        pending_.clear();
        cv_done_.notify_all();
        continue;
      }
      std::vector<FilterReq*> batch;
      batch.swap(pending_);
      const int slot = next_slot;
      next_slot = (next_slot + 1) % kSlots;
      g.unlock();
      std::string err;
      try {
        run(batch, slots_[slot]);
      } catch (const std::exception& e) {
        err = e.what();
      }
      g.lock();
      if (err.empty()) inflight_.push_back(slot);
      if (!err.empty()) error_ = err;
      for (FilterReq* r : batch) r->enqueued = err.empty();
      ++batches_;
      requests_ += (int64_t)batch.size();
      cv_done_.notify_all();
    }
  }

  void run(std::vector<FilterReq*>& batch, Slot& sl) {
    for (FilterReq* r : batch)
      if (r->ready) RCG_HIP(hipStreamWaitEvent(st_, r->ready, 0));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    {
      std::lock_guard<std::mutex> g(prof_mu_);
      harvest_locked(false);
      for (hipEvent_t* e : {&t0, &t1}) {
        if (!ev_free_.empty()) {
          *e = ev_free_.back();
          ev_free_.pop_back();
        } else {
          RCG_HIP(hipEventCreate(e));
        }
      }
    }
    // groups of one launch kind, requests by degree descending
    std::stable_sort(batch.begin(), batch.end(), [](const FilterReq* a, const FilterReq* b) {
      const int ka = (a->f32 ? 1000 : 0) + (a->cx ? 100 : 0) + a->cls;
      const int kb = (b->f32 ? 1000 : 0) + (b->cx ? 100 : 0) + b->cls;
      if (ka != kb) return ka < kb;
      return a->desc.m > b->desc.m;
    });
    std::vector<rcg::BatchTask> tasks;
    std::vector<int> first_task(batch.size() + 1, 0);
    for (size_t q = 0; q < batch.size(); ++q) {
      first_task[q] = (int)tasks.size();
      const rcg::BatchDesc& d = batch[q]->desc;
      // what the batched kernel assumes of each request: row blocks within
      // the LDS row-pointer stage, the launch class of its width, a degree
      // and start step the sweep covers, both recurrence buffers
      const FilterReq* r = batch[q];
      if (d.n <= 0 || d.rows_per_blk <= 0 || d.rows_per_blk > rcg::kBatchMaxRows ||
          d.ncv <= 0 || d.tail < 1 || d.ld < 4 * d.ncv / (r->f32 ? 1 : 2) || d.m < 1 ||
          r->jstart < 1 || (r->f32 && r->jstart < 2) || r->cls != rcg::batch_class(d.ncv) ||
          !d.rowptr || !d.col || !d.val || !d.buf[0] || !d.buf[1] ||
          (r->src && (!d.R || !d.kap)))
        throw Error(RCG_ERR_INVALID, "malformed batched filter request");
      if ((int64_t)tasks.size() + (d.n + d.rows_per_blk - 1) / d.rows_per_blk > INT32_MAX)
        throw Error(RCG_ERR_INVALID, "too many row blocks in one filter batch");
      const int nb = (int)((d.n + d.rows_per_blk - 1) / d.rows_per_blk);
      for (int b = 0; b < nb; ++b) tasks.push_back({(int)q, b});
    }
    first_task[batch.size()] = (int)tasks.size();
    const size_t db = sizeof(rcg::BatchDesc) * batch.size();
    const size_t tb = sizeof(rcg::BatchTask) * tasks.size();
    // the slot's previous batch (kSlots batches ago) has finished: at most
    // kDepth are in flight
    if (sl.used) wait_event(sl.end, true);
    sl.used = true;
    sl.hdesc.ensure(db);
    sl.htask.ensure(tb);
    sl.ddesc.ensure(db);
    sl.dtask.ensure(tb);
    for (size_t q = 0; q < batch.size(); ++q) sl.hdesc.as<rcg::BatchDesc>()[q] = batch[q]->desc;
    std::memcpy(sl.htask.p, tasks.data(), tb);
    RCG_HIP(hipMemcpyAsync(sl.ddesc.p, sl.hdesc.p, db, hipMemcpyHostToDevice, st_));
    RCG_HIP(hipMemcpyAsync(sl.dtask.p, sl.htask.p, tb, hipMemcpyHostToDevice, st_));
    const rcg::BatchTask* dt = sl.dtask.as<rcg::BatchTask>();
    const rcg::BatchDesc* dd = sl.ddesc.as<rcg::BatchDesc>();
    // groups of one launch kind; within a group, requests by degree
    // descending, so step j covers a prefix of the group
    struct Group {
      size_t g0, end;
      int jstart;
    };
    std::vector<Group> groups;
    int mmax = 0, jmin = 1 << 30;
    for (size_t g0 = 0; g0 < batch.size();) {
      size_t g1 = g0;
      const FilterReq* h = batch[g0];
      while (g1 < batch.size() && batch[g1]->f32 == h->f32 && batch[g1]->cx == h->cx &&
             batch[g1]->cls == h->cls)
        ++g1;
      groups.push_back({g0, g1, h->jstart});
      mmax = std::max(mmax, h->desc.m);
      jmin = std::min(jmin, h->jstart);
      g0 = g1;
    }
    // a request's `done` is recorded right after its own last step, so a
    // short sweep does not wait for the longest one in its batch
    for (FilterReq* r : batch)
      if (r->desc.m < r->jstart) RCG_HIP(hipEventRecord(r->done, st_));
    RCG_HIP(hipEventRecord(t0, st_));
    int64_t nl = 0;
    double nb = 0;
    for (int j = jmin; j <= mmax; ++j) {
      for (Group& gr : groups) {
        if (j < gr.jstart) continue;
        while (gr.end > gr.g0 && batch[gr.end - 1]->desc.m < j) --gr.end;
        if (gr.end == gr.g0) continue;
        const FilterReq* h = batch[gr.g0];
        const int nt = first_task[gr.end] - first_task[gr.g0];
        rcg::launch_cheb_batch(dd, dt + first_task[gr.g0], nt, j, h->f32, h->src, h->cx, h->cls,
                               st_);
        ++nl;
        for (size_t q = gr.g0; q < gr.end; ++q) nb += batch[q]->step_bytes;
      }
      for (FilterReq* r : batch)
        if (r->desc.m == j && j >= r->jstart) RCG_HIP(hipEventRecord(r->done, st_));
    }
    RCG_HIP(hipEventRecord(t1, st_));
    RCG_HIP(hipEventRecord(sl.end, st_));
    std::lock_guard<std::mutex> g(prof_mu_);
    timing_.push_back({t0, t1});
    launches_ += nl;
    bytes_ += nb;
  }

  // adds the intervals of finished batches (all of them when `wait`)
  void harvest_locked(bool wait) {
    while (!timing_.empty()) {
      auto pr = timing_.front();
      if (wait) (void)hipEventSynchronize(pr.second);
      if (hipEventQuery(pr.second) != hipSuccess) break;
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
        busy_s_ += 1e-3 * ms;
        ++timed_;
      }
      (void)hipGetLastError();
      ev_free_.push_back(pr.first);
      ev_free_.push_back(pr.second);
      timing_.pop_front();
    }
  }

  int device_;
  int window_us_;
  hipStream_t st_ = nullptr;
  std::thread thr_;
  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::vector<FilterReq*> pending_;
  int members_ = 0;
  bool stop_ = false;
  std::string error_;
  int64_t batches_ = 0, requests_ = 0;
  Slot slots_[kSlots];
  std::deque<int> inflight_;  // slots of enqueued, unfinished batches
  std::mutex prof_mu_;
  std::deque<std::pair<hipEvent_t, hipEvent_t>> timing_;
  std::vector<hipEvent_t> ev_free_;
  double busy_s_ = 0, bytes_ = 0;
  int64_t launches_ = 0, timed_ = 0;
};

// ------------------------------------------------------------------ engine
// Device operations of the solve over attached row-interleaved buffers:
// V (the subspace, ld), T (second recurrence / rotation buffer, ld),
// W (A V, ld) and Y (found vectors, ldy).
class Engine {
 public:
  // Complex operators (dtype C128): every block column is a complex column
  // stored as two interleaved real columns; ld and the attached buffers are
  // in real elements, and widths passed to the methods count complex columns.
  // cx_blocks: complex blocks with no operator (orthonormalisation, deflation,
  // rotation of C128 blocks); with an operator its dtype decides.
  Engine(rcg_context_s* ctx, rcg_operator_s* op, int64_t n, bool cx_blocks = false)
      : ctx_(ctx), op_(op), n_(n) {
    st_ = ctx->stream;
    cx_ = op ? op->dtype == RCG_C128 : cx_blocks;
    cw_ = cx_ ? 2 : 1;
    coef_.use_stream(st_);
    hcoef_.use_stream(st_);
    cb_in_.use_stream(st_);
    cb_out_.use_stream(st_);
  }
  bool cx() const { return cx_; }
  int cw() const { return cw_; }
  rcg_context_s* context() const { return ctx_; }

  // Size the context's workspaces for widths up to bmax and kmax.
  void reserve(int64_t bmax, int64_t kmax) {
    bb_ = std::max<int64_t>({bmax, kmax, 4});
    size_t e = rcg::gram_ws_bound(n_, (int)bb_);
    // ortho_project's paired Gram products (at most kFoldCols real columns)
    const int fold = (int)std::min<int64_t>(kFoldCols, cw_ * bb_);
    e = std::max(e, 2 * rcg::gram_ws_elems<double>(n_, fold, fold));
    e = std::max(e, rcg::colred_ws_elems(n_, 2 * (int)bb_));
    e = std::max(e, rcg::lz_ws_elems(n_));
    ctx_->ws.ensure(e * sizeof(double));
    ws_elems_ = ctx_->ws.bytes / sizeof(double);
    const size_t sm = sizeof(double) * (size_t)(2 * bb_ * bb_ + 4 * bb_ + 64);
    ctx_->small.ensure(sm);
    ctx_->pinned.ensure(sm);
    const size_t rw = (size_t)(cw_ * bb_);
    ctx_->results.ensure(sizeof(double) * (2 * rw * rw + 4 * rw + 64));
  }
  void attach(double* V, double* T, double* W, int64_t ld, double* Y, int64_t ldy) {
    buf_[0] = V;
    buf_[1] = T;
    iv_ = 0;
    W_ = W;
    ld_ = ld;
    Y_ = Y;
    ldy_ = ldy;
  }

  double* V() const { return buf_[iv_]; }
  double* T() const { return buf_[1 - iv_]; }
  double* W() const { return W_; }
  double* Y() const { return Y_; }
  bool v_is_first() const { return iv_ == 0; }
  void swapVT() { iv_ = 1 - iv_; }
  int64_t ld() const { return ld_; }
  int64_t ldy() const { return ldy_; }
  hipStream_t stream() const { return st_; }
  void sync() {
    if (ctx_->blocking_ev) {
      RCG_HIP(hipEventRecord(ctx_->blocking_ev, st_));
      wait_event(ctx_->blocking_ev, true);
    } else {
      RCG_HIP(hipStreamSynchronize(st_));
    }
  }

  // Scratch layout (device `small` and host `pinned` alike):
  //   [0, bb^2)          matrix A (Gram results, uploaded factors)
  //   [bb^2, 2bb^2)      matrix B
  //   [2bb^2, +bb)       vector: column sums
  //   [2bb^2+bb, +bb)    vector: uploaded eigenvalues
  //   [2bb^2+2bb, +bb)   vector: downloaded sums
  double* dmatA() { return ctx_->small.as<double>(); }
  double* dsums() { return ctx_->small.as<double>() + 2 * bb_ * bb_; }
  double* dtheta() { return dsums() + bb_; }
  double* hmatA() { return ctx_->pinned.as<double>(); }
  // reduction results written by the device into host memory (hres) at dres
  double* hres() { return static_cast<double*>(ctx_->results.h); }
  double* dres() { return static_cast<double*>(ctx_->results.d); }
  double* hmatB() { return ctx_->pinned.as<double>() + bb_ * bb_; }
  double* hvecUp() { return ctx_->pinned.as<double>() + 2 * bb_ * bb_ + bb_; }
  double* hlz() { return ctx_->pinned.as<double>() + 2 * bb_ * bb_ + 3 * bb_; }

  double* ws() { return ctx_->ws.as<double>(); }

  // Candidate row-block orders: contiguous blocks with and without the XCD
  // remap and, for operators with a structured-grid hint, x-lines visited in
  // (y, z) tiles so that the lines a stencil row gathers from are in flight
  // together.
  void build_schedules() {
    if (!op_->scheds.empty()) return;
    auto add = [&](const std::string& name, std::vector<int> order, int64_t rows, int remap) {
      auto c = std::make_unique<SchedCand>();
      c->name = name;
      c->sc.remap = remap;
      if (!order.empty()) {
        c->order.ensure(sizeof(int) * order.size());
        RCG_HIP(hipMemcpy(c->order.p, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
        c->sc.order = c->order.as<int>();
        c->sc.rows_per_blk = rows;
        c->sc.nblk = (int64_t)order.size();
      }
      op_->scheds.push_back(std::move(c));
    };
    add("rows", {}, 0, 1);
    add("rows-noremap", {}, 0, 0);
    const int64_t gx = op_->gx, gy = op_->gy, gz = op_->gz;
    if (gx >= 64 && gx <= 512 && gy >= 4 && gz >= 4 && gx * gy * gz == n_) {
      const int tiles[][2] = {{8, 8}, {16, 16}, {32, 8}, {16, 4}, {4, 16}};
      for (const auto& t : tiles) {
        const int ty = t[0], tz = t[1];
        std::vector<int> order;
        order.reserve(gy * gz);
        for (int64_t z0 = 0; z0 < gz; z0 += tz)
          for (int64_t y0 = 0; y0 < gy; y0 += ty)
            for (int64_t z = z0; z < std::min(gz, z0 + tz); ++z)
              for (int64_t y = y0; y < std::min(gy, y0 + ty); ++y) order.push_back((int)(y + gy * z));
        for (int remap = 0; remap < 2; ++remap)
          add("tile" + std::to_string(ty) + "x" + std::to_string(tz) + (remap ? "" : "-noremap"),
              order, gx, remap);
      }
    }
  }

  // Filter shape for b columns: the fastest row-block order, then the fastest
  // column-panel width {b, 32, 16} under it, timed with hipEvents on this
  // operator (3 Chebyshev steps after a warm-up step) and cached on it per
  // step kind (0: FP64 step, 1: FP32 correction step).  `launch(panel, sched)`
  // runs one step on scratch buffers.  RCG_FILTER_PANEL / RCG_FILTER_SCHED
  // override the choice.
  template <class Launch>
  rcg_operator_s::Tuned tuned_for(int64_t b, int kind, Launch&& launch) {
    std::lock_guard<std::mutex> lk(op_->lazy_mu);
    build_schedules();
    const int64_t key = (int64_t)kind * 1000000000000LL + ld_ * 100000 + b;
    auto it = op_->tuned.find(key);
    rcg_operator_s::Tuned t{(int)b, 0};
    if (it != op_->tuned.end()) {
      t = it->second;
    } else if (n_ * b >= (int64_t)1 << 22) {  // only large blocks are worth timing
      hipEvent_t e0, e1;
      RCG_HIP(hipEventCreate(&e0));
      RCG_HIP(hipEventCreate(&e1));
      auto time_it = [&](int panel, int sched) {
        const rcg::CsrSched* sc = &op_->scheds[sched]->sc;
        launch(panel, sc);
        RCG_HIP(hipEventRecord(e0, st_));
        for (int r = 0; r < 3; ++r) launch(panel, sc);
        RCG_HIP(hipEventRecord(e1, st_));
        RCG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        RCG_HIP(hipEventElapsedTime(&ms, e0, e1));
        return ms;
      };
      float best = 1e30f;
      for (int sidx = 0; sidx < (int)op_->scheds.size(); ++sidx) {
        const float ms = time_it((int)b, sidx);
        if (ms < best) {
          best = ms;
          t = {(int)b, sidx};
        }
      }
      for (int cand : {32, 16}) {
        if (cand >= b) continue;
        const float ms = time_it(cand, t.sched);
        if (ms < best) {
          best = ms;
          t.panel = cand;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      op_->tuned[key] = t;
    }
    if (const char* env = std::getenv("RCG_FILTER_PANEL")) {
      const int v = std::atoi(env);
      if (v > 0) t.panel = v;
    }
    if (const char* env = std::getenv("RCG_FILTER_SCHED")) {
      int pick = -1;
      for (int sidx = 0; sidx < (int)op_->scheds.size(); ++sidx)
        if (op_->scheds[sidx]->name == env) pick = sidx;
      if (pick < 0)
        throw Error(RCG_ERR_INVALID, std::string("RCG_FILTER_SCHED names no schedule of this "
                                                 "operator: ") + env);
      t.sched = pick;
    }
    return t;
  }
  // b counts real columns here.  Timing writes Chebyshev steps into T, so an
  // engine without a T block (the piecewise residual / apply entry points)
  // never tunes: it uses the cached choice or the contiguous default.
  rcg_operator_s::Tuned tuned_for(int64_t b) {
    if (!T() || !V()) {
      std::lock_guard<std::mutex> lk(op_->lazy_mu);
      build_schedules();
      auto it = op_->tuned.find((int64_t)(cx_ ? 2 : 0) * 1000000000000LL + ld_ * 100000 + b);
      return it != op_->tuned.end() ? it->second : rcg_operator_s::Tuned{(int)b, 0};
    }
    return tuned_for(b, cx_ ? 2 : 0, [&](int panel, const rcg::CsrSched* sc) {
      rcg::launch_cheb_step<double>(n_, op_->rowptr, op_->col, op_->val, V(), T(), T(), (int)ld_,
                                    (int)b, 0.0, 1.0, 0.0, false, st_, panel, sc, cx_);
    });
  }
  // W[:, :b] = A V[:, :b] (b complex columns when complex).
  void apply(const double* X, double* out, int64_t b) {
    if (op_->is_callback()) {
      const double* ax = cb_product(X, b);
      rcg::launch_layout(n_, (int)b, cw_, out, ld_, const_cast<double*>(ax), n_, false, st_);
      return;
    }
    if (op_->is_dense()) {
      rcg::launch_dense_step<double>(n_, op_->dense, op_->lda, X, nullptr, out, (int)ld_, (int)b,
                                     1.0, 0.0, 0.0, rcg::kStepApply, nullptr, nullptr, st_);
      return;
    }
    rcg::launch_csr_apply<double>(n_, op_->rowptr, op_->col, op_->val, X, out, (int)ld_,
                                  (int)(cw_ * b), 1.0, 0.0, st_, sched_for(cw_ * b), cx_);
  }
  void copy_cols(const double* src, int64_t lds, int64_t s0, double* dst, int64_t ldd, int64_t d0,
                 int64_t cnt) {
    rcg::launch_copy_cols<double>(n_, src, (int)lds, (int)(cw_ * s0), dst, (int)ldd,
                                  (int)(cw_ * d0), (int)(cw_ * cnt), st_);
  }
  const rcg::CsrSched* sched_for(int64_t b) {
    if (!op_ || batcher_ || n_ * b < (int64_t)1 << 22) return nullptr;
    return &op_->scheds[tuned_for(b).sched]->sc;
  }

  // V <- P_m(A')^rep V over b columns (LanczosCpolyOperatorLM.h:192-304):
  // X1 = g2(g1 A X0 + X0), Xj = g2(g1 A X_{j-1} + X_{j-1}) - X_{j-2}, out = Xm/(m+1).
  void filter(int64_t b, long m, long rep, double lmax, double shift) {
    if (op_->is_dense()) return filter_dense(b, m, rep, lmax, shift);
    if (op_->is_callback()) return filter_callback(b, m, rep, lmax, shift);
    if (batcher_) return filter_batched(b, m, rep, lmax, shift);
    const auto c = rcg::cpoly::coefficients(m, lmax, shift);
    const int64_t rb = cw_ * b;  // real columns
    const auto tn = tuned_for(rb);
    const rcg::CsrSched* sc = &op_->scheds[tn.sched]->sc;
    last_panel_ = tn.panel;
    last_sched_ = op_->scheds[tn.sched]->name;
    const double csr = double(op_->nnz) * (4.0 + 8.0 * cw_) + 4.0 * double(n_ + 1);
    const double blk = 8.0 * double(n_) * double(rb);
    step_begin(rep * m * ((rb + tn.panel - 1) / tn.panel),
               double(rep) * (csr + 2.0 * blk + double(m - 1) * (csr + 3.0 * blk)));
    for (long r = 0; r < rep; ++r) {
      double* X0 = V();
      double* X1 = T();
      const double s1 = (m == 1) ? 0.5 : 1.0;
      rcg::launch_cheb_step<double>(n_, op_->rowptr, op_->col, op_->val, X0, nullptr, X1,
                                    (int)ld_, (int)rb, c.g1, c.g2, s1, true, st_, tn.panel, sc,
                                    cx_);
      for (long j = 2; j <= m; ++j) {
        const double s = (j == m) ? 1.0 / double(m + 1) : 1.0;
        rcg::launch_cheb_step<double>(n_, op_->rowptr, op_->col, op_->val, X1, X0, X0,
                                      (int)ld_, (int)rb, c.g1, c.g2, s, false, st_, tn.panel, sc,
                                      cx_);
        std::swap(X0, X1);
      }
      if (X1 != V()) swapVT();
    }
    step_end();
  }

  // Serving: the steps run in the shared batcher.  The two sides are
  // ordered by host waits, not by device-side event waits: a stream waiting
  // on an event holds up the other streams that share its hardware queue.
  void batch_submit(const rcg::BatchDesc& d, bool f32, bool cx, int jstart) {
    if (!ev_done_)
      RCG_HIP(hipEventCreateWithFlags(&ev_done_, hipEventDisableTiming | hipEventBlockingSync));
    sync();  // the block is ready
    FilterReq r;
    r.desc = d;
    r.f32 = f32;
    r.src = f32;
    r.cx = cx;
    r.jstart = jstart;
    r.cls = rcg::batch_class(d.ncv);
    {
      const double vs = (f32 ? 4.0 : 8.0) * (cx ? 2.0 : 1.0);  // bytes per matrix value
      const double xs = f32 ? 4.0 : 8.0;                       // bytes per block element
      const double cols = (double)last_panel_;
      r.step_bytes = (double)op_->nnz * (vs + 4.0) + 4.0 * (double)(n_ + 1) +
                     (double)n_ * cols * xs * (f32 ? 4.0 : 3.0);
    }
    r.done = ev_done_;
    batcher_->submit(&r);
    wait_event(ev_done_, true);
  }
  rcg::BatchDesc batch_desc(long m, double g1, double g2, int64_t rb, int w) {
    rcg::BatchDesc d{};
    d.rowptr = op_->rowptr;
    d.col = op_->col;
    d.n = n_;
    d.ld = (int)ld_;
    d.ncv = (int)((rb + w - 1) / w);
    d.tail = (int)(rb - (int64_t)(d.ncv - 1) * w);
    d.m = (int)m;
    d.rows_per_blk = rcg::batch_rows(rcg::batch_class(d.ncv));
    d.g1 = g1;
    d.g2 = g2;
    return d;
  }
  void filter_batched(int64_t b, long m, long rep, double lmax, double shift) {
    const auto c = rcg::cpoly::coefficients(m, lmax, shift);
    const int64_t rb = cw_ * b;
    last_panel_ = (int)rb;
    last_sched_ = "batched";
    for (long r = 0; r < rep; ++r) {
      rcg::BatchDesc d = batch_desc(m, c.g1, c.g2, rb, 2);
      d.val = op_->val;
      d.buf[0] = V();
      d.buf[1] = T();
      batch_submit(d, false, cx_, 1);
      if (m % 2 == 1) swapVT();
    }
  }
  void set_batcher(FilterBatcher* b) { batcher_ = b; }

  // The same recurrence with the dense MFMA step.
  void filter_dense(int64_t b, long m, long rep, double lmax, double shift) {
    const auto c = rcg::cpoly::coefficients(m, lmax, shift);
    last_panel_ = (int)b;
    last_sched_ = "dense";
    const double nn = double(n_) * double(n_);
    step_begin(rep * m, double(rep) * double(m) * (8.0 * nn + 24.0 * double(n_) * double(b)));
    for (long r = 0; r < rep; ++r) {
      double* X0 = V();
      double* X1 = T();
      const double s1 = (m == 1) ? 0.5 : 1.0;
      rcg::launch_dense_step<double>(n_, op_->dense, op_->lda, X0, nullptr, X1, (int)ld_, (int)b,
                                     c.g1, c.g2, s1, rcg::kStepFirst, nullptr, nullptr, st_);
      for (long j = 2; j <= m; ++j) {
        const double s = (j == m) ? 1.0 / double(m + 1) : 1.0;
        rcg::launch_dense_step<double>(n_, op_->dense, op_->lda, X1, X0, X0, (int)ld_, (int)b,
                                       c.g1, c.g2, s, rcg::kStepCheb, nullptr, nullptr, st_);
        std::swap(X0, X1);
      }
      if (X1 != V()) swapVT();
    }
    step_end();
  }

  // The caller's operator (LanczosCpolyOperatorLMCuda.h:236,245, its
  // Otype::apply(in, out, alpha, beta)): its product on the column-major
  // copy of the block, then the rest of the step in one pass back to the
  // block layout.  One layout change, the caller's apply and one combine per
  // degree; the step's bytes depend on the caller's product (not counted).
  void filter_callback(int64_t b, long m, long rep, double lmax, double shift) {
    const auto c = rcg::cpoly::coefficients(m, lmax, shift);
    last_panel_ = (int)b;
    last_sched_ = "caller";
    step_begin(rep * m, 0.0);
    for (long r = 0; r < rep; ++r) {
      double* X0 = V();
      double* X1 = T();
      const double s1 = (m == 1) ? 0.5 : 1.0;
      rcg::launch_cheb_combine(n_, (int)b, cw_, cb_product(X0, b), n_, X0, nullptr, X1, ld_, c.g1,
                               c.g2, s1, st_);
      for (long j = 2; j <= m; ++j) {
        const double s = (j == m) ? 1.0 / double(m + 1) : 1.0;
        rcg::launch_cheb_combine(n_, (int)b, cw_, cb_product(X1, b), n_, X1, X0, X0, ld_, c.g1,
                                 c.g2, s, st_);
        std::swap(X0, X1);
      }
      if (X1 != V()) swapVT();
    }
    step_end();
  }
  // A X over b columns of X (rows layout, ld_) through the caller's apply:
  // the column-major product (n x b, ld n), valid until the next call.
  const double* cb_product(const double* X, int64_t b) {
    const size_t bytes = sizeof(double) * (size_t)cw_ * (size_t)n_ * (size_t)b;
    cb_in_.ensure(bytes);
    cb_out_.ensure(bytes);
    rcg::launch_layout(n_, (int)b, cw_, const_cast<double*>(X), ld_, cb_in_.as<double>(), n_, true,
                       st_);
    call_user(cb_in_.p, cb_out_.p, b, n_, 1.0, 0.0);
    return cb_out_.as<double>();
  }
  // out = alpha A X + beta out over b columns (rows layout, ld_) through the
  // caller's apply (rcg_operator_apply on a caller's operator)
  void cb_apply_ab(const double* X, double* out, int64_t b, double alpha, double beta) {
    const size_t bytes = sizeof(double) * (size_t)cw_ * (size_t)n_ * (size_t)b;
    cb_in_.ensure(bytes);
    cb_out_.ensure(bytes);
    rcg::launch_layout(n_, (int)b, cw_, const_cast<double*>(X), ld_, cb_in_.as<double>(), n_, true,
                       st_);
    if (beta != 0.0)
      rcg::launch_layout(n_, (int)b, cw_, out, ld_, cb_out_.as<double>(), n_, true, st_);
    call_user(cb_in_.p, cb_out_.p, b, n_, alpha, beta);
    rcg::launch_layout(n_, (int)b, cw_, out, ld_, cb_out_.as<double>(), n_, false, st_);
  }
  void call_user(const void* in, void* out, int64_t cols, int64_t ld, double alpha, double beta) {
    const int rc = op_->apply_fn(op_->user, in, out, n_, cols, ld, alpha, beta, (void*)st_);
    if (rc != 0)
      throw Error(RCG_ERR_INVALID, "the operator's apply callback returned " + std::to_string(rc));
  }

  // Mixed-precision filter (the correction recurrence of
  // launch_cheb_step_src): V <- P_m(A')^rep V over b columns with FP32 steps
  // and FP32 CSR values.  theta[c] / ritz[c]: the Ritz value of column c and
  // whether column c is a Ritz vector (others are filtered directly in FP32).
  // Needs m >= 2 and the FP32 buffers attached with attach_mixed.  Complex
  // columns run as their two real columns with the same theta.
  void filter_mixed(int64_t b, long m, long rep, double lmax, double shift,
                    const std::vector<double>& theta_c, const std::vector<char>& ritz_c) {
    if (op_->is_callback())
      throw Error(RCG_ERR_UNSUPPORTED, "mixed-precision filtering needs the operator's values "
                                       "(a CSR or dense operator), not a caller's apply");
    if (m < 2) return filter(b, m, rep, lmax, shift);
    if (!E_[0]) throw Error(RCG_ERR_INVALID, "mixed-precision buffers are not attached");
    const int64_t rb = cw_ * b;
    std::vector<double> theta(rb);
    std::vector<char> ritz(rb);
    for (int64_t c = 0; c < rb; ++c) {
      theta[c] = theta_c[c / cw_];
      ritz[c] = ritz_c[c / cw_];
    }
    const auto c = rcg::cpoly::coefficients(m, lmax, shift);
    const double g12 = c.g1 * c.g2;
    // coefficient table: cp (theta, flag) [2 rb doubles], um [ld doubles],
    // kap rows j = 0..m [ld floats each]
    const size_t nd = 2 * (size_t)rb + (size_t)ld_;
    const size_t bytes = sizeof(double) * nd + sizeof(float) * (size_t)(m + 1) * (size_t)ld_;
    // the pinned table of the previous call is no longer in flight
    if (coef_ev_) wait_event(coef_ev_, ctx_->blocking_ev != nullptr);
    else RCG_HIP(hipEventCreateWithFlags(&coef_ev_, hipEventDisableTiming));
    coef_.ensure(bytes);
    hcoef_.ensure(bytes);
    double* hd = hcoef_.as<double>();
    float* hk = reinterpret_cast<float*>(hd + nd);
    std::fill(hd, hd + nd, 0.0);
    std::fill(hk, hk + (size_t)(m + 1) * (size_t)ld_, 0.0f);
    for (int64_t col = 0; col < rb; ++col) {
      if (!ritz[col]) continue;
      hd[col] = theta[col];
      hd[rb + col] = 1.0;
      const double l = c.g2 * (1.0 + c.g1 * theta[col]);
      double u0 = 1.0, u1 = l;  // u_{j-2}, u_{j-1}
      for (long j = 2; j <= m; ++j) {
        hk[(size_t)j * ld_ + col] = float(g12 * u1);
        const double u = l * u1 - u0;
        u0 = u1;
        u1 = u;
      }
      hd[2 * rb + col] = u1;  // u_m
    }
    RCG_HIP(hipMemcpyAsync(coef_.p, hcoef_.p, bytes, hipMemcpyHostToDevice, st_));
    RCG_HIP(hipEventRecord(coef_ev_, st_));
    const double* dcp = coef_.as<double>();
    const double* dum = dcp + 2 * rb;
    const float* dk = reinterpret_cast<const float*>(dcp + nd);
    const float* val32 = values_f32();
    const float g1f = float(c.g1), g2f = float(c.g2);
    const bool dn = op_->is_dense();
    rcg_operator_s::Tuned tn{(int)rb, 0};
    if (!dn && !batcher_)
      tn = tuned_for(rb, cx_ ? 3 : 1, [&](int panel, const rcg::CsrSched* sc) {
        rcg::launch_cheb_step_src(n_, op_->rowptr, op_->col, val32, E_[0], E_[1], E_[1], R32_,
                                  dk + 2 * ld_, (int)ld_, (int)rb, g1f, g2f, 1.0f, st_, panel, sc,
                                  cx_);
      });
    const rcg::CsrSched* sc = (dn || batcher_) ? nullptr : &op_->scheds[tn.sched]->sc;
    last_panel_ = tn.panel;
    last_sched_ = dn ? "dense" : (batcher_ ? "batched" : op_->scheds[tn.sched]->name);
    const double csr = dn ? 4.0 * double(n_) * double(n_)
                          : double(op_->nnz) * (4.0 + 4.0 * cw_) + 4.0 * double(n_ + 1);
    const double blk = 4.0 * double(n_) * double(rb);
    // CSR operators: W = A V and the start in one pass (W is not stored);
    // RCG_MIXED_UNFUSED=1 runs the apply and the start kernel separately
    // (bit-identical results).
    const char* unf = std::getenv("RCG_MIXED_UNFUSED");
    const bool fused = !dn && !(unf && unf[0] == '1');
    for (long r = 0; r < rep; ++r) {
      if (fused) {
        rcg::MixedArgs ms;
        ms.cp = dcp;
        ms.e0 = E_[0];
        ms.e1 = E_[1];
        ms.r = R32_;
        ms.e1c = g12;
        ms.g1 = c.g1;
        ms.g2 = c.g2;
        ms.nc = (int)rb;
        rcg::launch_csr_apply_start(n_, op_->rowptr, op_->col, op_->val, V(), (int)ld_, (int)rb,
                                    ms, st_, sched_for(rb), cx_);
      } else {
        apply(V(), W_, b);
        rcg::launch_mixed_start(n_, V(), W_, (int)ld_, (int)rb, dcp, g12, c.g1, c.g2, E_[0],
                                E_[1], R32_, st_);
      }
      // the recombination V <- (um V + E_m) / (m + 1): in the last FP32 step's
      // epilogue for a single CSR operator (FP32 step at 128^3 x 64 with the
      // filter's ends, 561 -> 546 us; RCG_MIXED_UNFUSED=1: a kernel of its
      // own), after the batched steps a kernel of its own
      rcg::MixedArgs end;
      end.v_end = V();
      end.um = dum;
      end.s_end = 1.0 / double(m + 1);
      if (batcher_ && !dn) {
        rcg::BatchDesc d = batch_desc(m, c.g1, c.g2, rb, 4);
        d.val = val32;
        d.buf[0] = E_[0];
        d.buf[1] = E_[1];
        d.R = R32_;
        d.kap = dk;
        batch_submit(d, true, cx_, 2);
        rcg::launch_mixed_combine(n_, V(), (int)ld_, (int)rb, dum, end.s_end, E_[m % 2], st_);
        continue;
      }
      float* X0 = E_[0];
      float* X1 = E_[1];
      step_begin((m - 1) * ((rb + tn.panel - 1) / tn.panel), double(m - 1) * (csr + 4.0 * blk));
      for (long j = 2; j <= m; ++j) {
        if (dn) {
          rcg::launch_dense_step<float>(n_, val32, op_->lda, X1, X0, X0, (int)ld_, (int)rb, g1f,
                                        g2f, 1.0f, rcg::kStepSrc, R32_, dk + (size_t)j * ld_, st_);
        } else if (j == m && fused) {
          rcg::launch_cheb_step_src_end(n_, op_->rowptr, op_->col, val32, X1, X0, R32_,
                                        dk + (size_t)j * ld_, (int)ld_, (int)rb, g1f, g2f, end,
                                        st_, tn.panel, sc, cx_);
        } else {
          rcg::launch_cheb_step_src(n_, op_->rowptr, op_->col, val32, X1, X0, X0, R32_,
                                    dk + (size_t)j * ld_, (int)ld_, (int)rb, g1f, g2f, 1.0f, st_,
                                    tn.panel, sc, cx_);
        }
        std::swap(X0, X1);
      }
      step_end();
      if (dn || !fused)
        rcg::launch_mixed_combine(n_, V(), (int)ld_, (int)rb, dum, end.s_end, X1, st_);
    }
  }
  void attach_mixed(float* E0, float* E1, float* R) {
    E_[0] = E0;
    E_[1] = E1;
    R32_ = R;
  }
  const float* values_f32() {
    std::lock_guard<std::mutex> lk(op_->lazy_mu);
    if (!op_->val32.p) {
      // dense operators hold their FP64 matrix in `dense` (lda x n), CSR ones in `val`
      const double* src = op_->is_dense() ? op_->dense : op_->val;
      const int64_t nv = op_->is_dense() ? op_->lda * n_ : cw_ * op_->nnz;
      if (!src) throw Error(RCG_ERR_INVALID, "operator has no values to convert");
      op_->val32.ensure(sizeof(float) * (size_t)std::max<int64_t>(nv, 1));
      rcg::launch_to_f32(nv, src, op_->val32.as<float>(), st_);
    }
    return op_->val32.as<float>();
  }

  // Device time of the Chebyshev step launches (hipEvents around each run
  // of steps), their count and algorithmic bytes; read once the stream has
  // passed the end event.
  void step_begin(long launches, double bytes) {
    if (!fev_[0]) {
      RCG_HIP(hipEventCreate(&fev_[0]));
      RCG_HIP(hipEventCreate(&fev_[1]));
    }
    collect_step_time();
    RCG_HIP(hipEventRecord(fev_[0], st_));
    pend_launches_ = launches;
    pend_bytes_ = bytes;
  }
  void step_end() {
    RCG_HIP(hipEventRecord(fev_[1], st_));
    step_pending_ = true;
  }
  void collect_step_time() {
    if (!step_pending_) return;
    RCG_HIP(hipEventSynchronize(fev_[1]));
    float ms = 0.f;
    RCG_HIP(hipEventElapsedTime(&ms, fev_[0], fev_[1]));
    step_seconds_ += 1.0e-3 * ms;
    step_launches_ += pend_launches_;
    step_bytes_ += pend_bytes_;
    step_pending_ = false;
  }
  double step_seconds() { collect_step_time(); return step_seconds_; }
  double step_bytes() { collect_step_time(); return step_bytes_; }
  long step_launches() { collect_step_time(); return step_launches_; }
  ~Engine() {
    for (auto& e : fev_)
      if (e) (void)hipEventDestroy(e);
    if (ev_done_) (void)hipEventDestroy(ev_done_);
    if (coef_ev_) (void)hipEventDestroy(coef_ev_);
  }
  int last_panel() const { return last_panel_; }
  const std::string& last_sched() const { return last_sched_; }

  // host[bx x by] = X^H Y (column-major); widths in columns of S.  Complex:
  // the real Gram of the interleaved blocks (2bx x 2by) is combined on the host.
  // lower_only: only the entries a >= c of the result are meaningful.
  template <class S>
  void gram_to_host(const double* X, int64_t ldx, int64_t bx, const double* Yp, int64_t ldy,
                    int64_t by, S* host, bool lower_only = false) {
    constexpr int w = std::is_same<S, cplx>::value ? 2 : 1;
    const int64_t rx = w * bx, ry = w * by;
    rcg::launch_gram<double>(n_, X, (int)ldx, (int)rx, Yp, (int)ldy, (int)ry, dres(), ws(),
                             ws_elems_, st_, lower_only);
    sync();
    const double* g = hres();
    if constexpr (w == 1) {
      std::memcpy(host, g, sizeof(double) * bx * by);
    } else {
      for (int64_t c = 0; c < by; ++c)
        for (int64_t a = 0; a < bx; ++a) {
          const double g00 = g[2 * a + 2 * c * rx], g10 = g[2 * a + 1 + 2 * c * rx];
          const double g01 = g[2 * a + (2 * c + 1) * rx], g11 = g[2 * a + 1 + (2 * c + 1) * rx];
          host[a + c * bx] = cplx(g00 + g11, g01 - g10);
        }
    }
  }

  // V <- V C, C host b x b column-major; the result lands in T, then swaps.
  // Complex C enters the real product through its 2b x 2b real embedding.
  template <class S>
  void right_multiply(int64_t b, const S* C) {
    if constexpr (std::is_same<S, double>::value) {
      std::memcpy(hmatB(), C, sizeof(double) * b * b);
    } else {
      const int64_t r = 2 * b;
      double* e = hmatB();
      for (int64_t c = 0; c < b; ++c)
        for (int64_t k = 0; k < b; ++k) {
          const cplx v = C[k + c * b];
          e[2 * k + 2 * c * r] = v.real();
          e[2 * k + 1 + 2 * c * r] = -v.imag();
          e[2 * k + (2 * c + 1) * r] = v.imag();
          e[2 * k + 1 + (2 * c + 1) * r] = v.real();
        }
    }
    const int64_t rb = cw_ * b;
    RCG_HIP(hipMemcpyAsync(dmatA(), hmatB(), sizeof(double) * rb * rb, hipMemcpyHostToDevice, st_));
    rcg::launch_tall_gemm<double>(n_, V(), (int)ld_, (int)rb, dmatA(), (int)rb, (int)rb, 1.0, 0.0,
                                  nullptr, T(), (int)ld_, st_);
    swapVT();
  }

  // One CholQR pass on V[:, :b]: G = V^H V scaled to unit diagonal,
  // G = R^H R, V <- V D R^{-1}.  A failed factorisation retries with the
  // shift 11(nb + b(b+1)) u ||G|| (shifted CholeskyQR); returns false then.
  template <class S>
  bool cholqr_pass(int64_t b, bool allow_shift) {
    std::vector<S> G(b * b);
    gram_to_host<S>(V(), ld_, b, V(), ld_, b, G.data());
    std::vector<double> d(b);
    for (int64_t j = 0; j < b; ++j) {
      const double g = std::real(G[j + j * b]);
      d[j] = (g > 0.0 && std::isfinite(g)) ? 1.0 / std::sqrt(g) : 1.0;
    }
    for (int64_t j = 0; j < b; ++j)
      for (int64_t i = 0; i < b; ++i) G[i + j * b] *= d[i] * d[j];
    std::vector<S> R, Ri;
    bool clean = rcg::dense::cholesky_upper(G.data(), b, R);
    if (!clean) {
      if (!allow_shift) throw Error(RCG_ERR_NUMERIC, "CholQR breakdown after a shifted pass");
      double gn = 0.0;
      for (const S& x : G) gn += std::norm(x);
      const double s = 11.0 * (double(n_) * b + double(b) * (b + 1)) * 1.1102230246251565e-16 * std::sqrt(gn);
      for (int64_t j = 0; j < b; ++j) G[j + j * b] += s;
      if (!rcg::dense::cholesky_upper(G.data(), b, R))
        throw Error(RCG_ERR_NUMERIC, "shifted CholQR breakdown (block is numerically rank deficient)");
    }
    rcg::dense::upper_inverse(R, b, Ri);
    for (int64_t j = 0; j < b; ++j)
      for (int64_t i = 0; i < b; ++i) Ri[i + j * b] *= d[i];
    right_multiply<S>(b, Ri.data());
    return clean;
  }

  // orthogonalize (RayleighChebyshevLM.h:1782-1837): CholQR2, plus a third
  // pass when the first needed a shift.
  void orthonormalize(int64_t b) {
    fold_b_ = 0;
    if (b <= 0) return;
    if (cx_) return orthonormalize_t<cplx>(b);
    orthonormalize_t<double>(b);
  }
  template <class S>
  void orthonormalize_t(int64_t b) {
    const bool clean = cholqr_pass<S>(b, true);
    cholqr_pass<S>(b, true);
    if (!clean) cholqr_pass<S>(b, false);
  }


  // OrthogonalizeAtoB (RayleighChebyshevLM.h:1844-1917) against Y[:, :found]:
  // V -= Y (Y^H V), twice.  The normalisation of the reference is implied by
  // the CholQR that follows in the solve.
  void deflate(int64_t b, int64_t found) {
    if (found <= 0 || b <= 0) return;
    const int64_t rf = cw_ * found, rb = cw_ * b;
    for (int pass = 0; pass < 2; ++pass) {
      rcg::launch_gram<double>(n_, Y_, (int)ldy_, (int)rf, V(), (int)ld_, (int)rb, dmatA(), ws(),
                               ws_elems_, st_);
      if (cx_) rcg::launch_herm_embed(dmatA(), (int)found, (int)b, st_);
      rcg::launch_tall_gemm<double>(n_, Y_, (int)ldy_, (int)rf, dmatA(), (int)rf, (int)rb, -1.0,
                                    1.0, V(), V(), (int)ld_, st_);
    }
  }

  void normalize(int64_t b) {
    const int64_t rb = cw_ * b;
    rcg::launch_col_sumsq<double>(n_, V(), (int)ld_, (int)rb, dsums(), ws(), ws_elems_, st_);
    if (cx_) rcg::launch_pair_sums(dsums(), (int)b, st_);
    rcg::launch_normalize_cols<double>(n_, V(), (int)ld_, (int)rb, dsums(), st_);
  }

  // formVtAV (RayleighChebyshevLM.h:1928-1969): W = A V, H(j,i) = V_j^H W_i
  // for j >= i, mirrored (conjugated).
  template <class S>
  void project(int64_t b, std::vector<S>& H) {
    fold_b_ = 0;
    apply(V(), W_, b);
    H.assign(b * b, S(0.0));
    gram_to_host<S>(V(), ld_, b, W_, ld_, b, H.data(), true);
    for (int64_t i = 0; i < b; ++i) {
      H[i + i * b] = std::real(H[i + i * b]);
      for (int64_t j = i + 1; j < b; ++j) H[i + j * b] = conj_s(H[j + i * b]);
    }
  }

  // Complex Gram of interleaved blocks from the real Gram g of the 2bx x 2by
  // real columns (column-major, leading dimension 2bx).
  template <class S>
  static void combine_gram(const double* g, int64_t bx, int64_t by, S* host) {
    if constexpr (std::is_same<S, double>::value) {
      std::memcpy(host, g, sizeof(double) * bx * by);
    } else {
      const int64_t rx = 2 * bx;
      for (int64_t c = 0; c < by; ++c)
        for (int64_t a = 0; a < bx; ++a) {
          const double g00 = g[2 * a + 2 * c * rx], g10 = g[2 * a + 1 + 2 * c * rx];
          const double g01 = g[2 * a + (2 * c + 1) * rx], g11 = g[2 * a + 1 + (2 * c + 1) * rx];
          host[a + c * bx] = cplx(g00 + g11, g01 - g10);
        }
    }
  }

  // orthogonalize + formVtAV (RayleighChebyshevLM.h:1782-1837, 1928-1969) with
  // CholQR2's second pass folded into the projection, for blocks of at most 64
  // real columns (the host b^3 products then cost less than the device block
  // product they replace).  After one CholQR pass V1: W = A V1; G = V1^H V1 and
  // H1 = V1^H W come back with one host wait; G = R^H R, and
  // H = R^{-H} H1 R^{-1} is the projection on V1 R^{-1}, the basis the second
  // pass would have formed.  R^{-1} is kept and multiplied into the Ritz
  // rotation (rotate_residuals), so the block is rotated once.  Wider blocks,
  // a first pass that needed a shift, or a G that does not factor take the
  // plain passes.
  template <class S>
  void ortho_project(int64_t b, std::vector<S>& H) {
    fold_b_ = 0;
    if (b <= 0) return;
    if (cw_ * b > kFoldCols) {
      orthonormalize_t<S>(b);
      project<S>(b, H);
      return;
    }
    if (!cholqr_pass<S>(b, true)) {
      cholqr_pass<S>(b, true);
      cholqr_pass<S>(b, false);
      project<S>(b, H);
      return;
    }
    apply(V(), W_, b);
    const int64_t rb = cw_ * b;
    // G = V^H V and H1 = V^H W right after it, so that one copy brings both
    // back.  Below 2^22 block entries both products go in one paired launch
    // (the same values; plain solves of served-size requests 568 -> 586
    // eigenpairs/s); above, the paired kernel's one wave per SIMD costs more
    // than the launch it saves (128^3 x 64 solve 1987 -> 2003 ms).
    // RCG_GRAM_UNPAIRED=1 always runs them separately.
    const char* unp = std::getenv("RCG_GRAM_UNPAIRED");
    if ((unp && unp[0] == '1') || n_ * rb >= (int64_t(1) << 22)) {
      rcg::launch_gram<double>(n_, V(), (int)ld_, (int)rb, V(), (int)ld_, (int)rb, dres(), ws(),
                               ws_elems_, st_);
      rcg::launch_gram<double>(n_, V(), (int)ld_, (int)rb, W_, (int)ld_, (int)rb,
                               dres() + rb * rb, ws(), ws_elems_, st_, true);
    } else {
      rcg::launch_gram_pair(n_, V(), (int)ld_, W_, (int)ld_, (int)rb, dres(), ws(), ws_elems_,
                            st_);
    }
    sync();
    std::vector<S> G(b * b), H1(b * b);
    combine_gram<S>(hres(), b, b, G.data());
    combine_gram<S>(hres() + rb * rb, b, b, H1.data());
    std::vector<S> R, Ri;
    if (!rcg::dense::cholesky_upper(G.data(), b, R)) {
      // V1 is left as it is: the plain second (and third) pass, then project
      if (!cholqr_pass<S>(b, true)) cholqr_pass<S>(b, false);
      project<S>(b, H);
      return;
    }
    rcg::dense::upper_inverse(R, b, Ri);
    for (int64_t i = 0; i < b; ++i) {
      H1[i + i * b] = std::real(H1[i + i * b]);
      for (int64_t j = i + 1; j < b; ++j) H1[i + j * b] = conj_s(H1[j + i * b]);
    }
    // T = H1 Ri, H = Ri^H T (Ri upper triangular)
    std::vector<S> T(b * b, S(0.0));
    for (int64_t j = 0; j < b; ++j)
      for (int64_t k = 0; k <= j; ++k) {
        const S r = Ri[k + j * b];
        for (int64_t i = 0; i < b; ++i) T[i + j * b] += H1[i + k * b] * r;
      }
    H.assign(b * b, S(0.0));
    for (int64_t j = 0; j < b; ++j)
      for (int64_t i = 0; i < b; ++i) {
        S t = S(0.0);
        for (int64_t k = 0; k <= i; ++k) t += conj_s(Ri[k + i * b]) * T[k + j * b];
        H[i + j * b] = t;
      }
    for (int64_t i = 0; i < b; ++i) {
      H[i + i * b] = std::real(H[i + i * b]);
      for (int64_t j = i + 1; j < b; ++j) {
        const S m = 0.5 * (H[i + j * b] + conj_s(H[j + i * b]));
        H[i + j * b] = m;
        H[j + i * b] = conj_s(m);
      }
    }
    fold<S>().swap(Ri);
    fold_b_ = b;
  }

  // createEigenVectorsAndResiduals (RayleighChebyshevLM.h:1984-2087):
  // V <- V Y; W = A V; r_c = ||theta_c V_c - W_c|| / ||V_c||, c < count.  With V
  // orthonormal and Y orthogonal the columns of V Y have unit norm to
  // round-off; the residuals are divided by the computed norms instead of
  // rescaling the block (the filter is linear and CholQR normalises), and the
  // columns handed out as eigenvectors are scaled at lock time.
  // After ortho_project the block is V1 and the rotation is R^{-1} Y.
  template <class S>
  void rotate_residuals(int64_t b, const std::vector<S>& Yh, const std::vector<double>& th,
                        int64_t count, std::vector<double>& res) {
    if (fold_b_ == b && b > 0) {
      const std::vector<S>& Ri = fold<S>();
      std::vector<S> C(b * b, S(0.0));
      for (int64_t j = 0; j < b; ++j)
        for (int64_t k = 0; k < b; ++k) {
          const S y = Yh[k + j * b];
          for (int64_t i = 0; i <= k; ++i) C[i + j * b] += Ri[i + k * b] * y;
        }
      fold_b_ = 0;
      right_multiply<S>(b, C.data());
    } else {
      fold_b_ = 0;
      right_multiply<S>(b, Yh.data());
    }
    residuals_normalized(count, th.data(), res);
  }

  // r_c = ||theta_c V_c - (A V)_c|| / ||V_c||, c < count, and ||V_c||^2 kept
  // on the device (dsums() + cw*count) for scale_found.
  void residuals_normalized(int64_t count, const double* th, std::vector<double>& res) {
    res.assign(count, 0.0);
    if (count <= 0) return;
    apply(V(), W_, count);
    const int64_t rc = cw_ * count;
    if (rc <= rcg::kThetaArgs) {
      // the Ritz values travel with the launch
      rcg::ThetaArgs ta{};
      for (int64_t c = 0; c < rc; ++c) ta.v[c] = th[c / cw_];
      rcg::launch_resid_norm_sumsq_args(n_, V(), W_, (int)ld_, (int)rc, ta, dres(), ws(),
                                        ws_elems_, st_);
    } else {
      for (int64_t c = 0; c < rc; ++c) hvecUp()[c] = th[c / cw_];
      RCG_HIP(hipMemcpyAsync(dtheta(), hvecUp(), sizeof(double) * rc, hipMemcpyHostToDevice, st_));
      rcg::launch_resid_norm_sumsq<double>(n_, V(), W_, (int)ld_, (int)rc, dtheta(), dres(), ws(),
                                           ws_elems_, st_);
    }
    sync();
    const double* sums = hres();
    last_norms_.assign(count, 1.0);
    for (int64_t c = 0; c < count; ++c) {
      double t = 0.0, q = 0.0;
      for (int e = 0; e < cw_; ++e) {
        t += sums[cw_ * c + e];
        q += sums[rc + cw_ * c + e];
      }
      last_norms_[c] = q > 0.0 ? std::sqrt(q) : 1.0;
      res[c] = std::sqrt(std::abs(t)) / last_norms_[c];
    }
  }
  // The columns [c0, c0+cnt) of `dst` (ld ldd) came from V's first columns at
  // the last residual computation: scale them to unit norm.
  void scale_found(double* dst, int64_t ldd, int64_t c0, int64_t cnt) {
    if (cnt <= 0) return;
    const int64_t rc = cw_ * cnt;
    for (int64_t c = 0; c < rc; ++c) {
      const double nrm = c / cw_ < (int64_t)last_norms_.size() ? last_norms_[c / cw_] : 1.0;
      hvecUp()[c] = nrm * nrm;
    }
    RCG_HIP(hipMemcpyAsync(dtheta(), hvecUp(), sizeof(double) * rc, hipMemcpyHostToDevice, st_));
    rcg::launch_normalize_cols<double>(n_, dst + cw_ * c0, (int)ldd, (int)rc, dtheta(), st_);
    sync();
  }

  void residuals(int64_t count, const double* th, std::vector<double>& res) {
    res.assign(count, 0.0);
    if (count <= 0) return;
    apply(V(), W_, count);
    const int64_t rc = cw_ * count;
    for (int64_t c = 0; c < rc; ++c) hvecUp()[c] = th[c / cw_];
    RCG_HIP(hipMemcpyAsync(dtheta(), hvecUp(), sizeof(double) * rc, hipMemcpyHostToDevice, st_));
    rcg::launch_residual_sumsq<double>(n_, V(), W_, (int)ld_, (int)rc, dtheta(), dres(), ws(),
                                       ws_elems_, st_);
    sync();
    const double* sums = hres();
    for (int64_t c = 0; c < count; ++c) {
      double t = 0.0;
      for (int e = 0; e < cw_; ++e) t += sums[cw_ * c + e];
      res[c] = std::sqrt(std::abs(t));
    }
  }

  // ArandomizeOpClassLMCuda::randomize semantics: column j of the range
  // takes draws counter + j*n + i (complex: counter + 2(j*n + i) + {0, 1}).
  void randomize(int64_t c0, int64_t cnt, uint64_t seed, uint64_t* counter) {
    if (cx_) {
      rcg::launch_randomize_cplx(n_, V(), (int)ld_, (int)c0, (int)cnt, seed, *counter, st_);
      *counter += 2 * (uint64_t)cnt * (uint64_t)n_;
      return;
    }
    rcg::launch_randomize_cols<double>(n_, V(), (int)ld_, (int)c0, (int)cnt, seed, *counter, st_);
    *counter += (uint64_t)cnt * (uint64_t)n_;
  }

  // LanczosMaxMinFinder::getMinMaxEigenvalues (LanczosMaxMinFinder.h:248-544)
  // with the vectors, alpha and beta on the device; every 5 steps the host
  // bisects the tridiagonal for its extreme eigenvalues.
  void lanczos(double errorTolerance, bool maxFlag, bool minFlag, long itmax, uint64_t seed,
               uint64_t* counter, double* mnOut, double* mxOut, long* itOut) {
    const double smallTol = 1.0e-10;
    auto relErr = [&](double v, double t) {
      return (std::abs(v) * t > smallTol) ? std::abs(v) : smallTol / t;
    };
    if (errorTolerance > 0.01) errorTolerance = 0.01;
    const double tol = errorTolerance < smallTol ? smallTol : errorTolerance;
    // complex vectors are 2n interleaved reals: the start draws, norms,
    // scalings and the real-alpha axpy act on them as real vectors
    const int64_t nr = cw_ * n_;
    // real CSR operators run the fused step on a ping-pong pair (v, w), (v2, w2)
    const bool fused = !cx_ && op_->is_csr();
    DevBuf vecs(st_), ab(st_);
    vecs.ensure(sizeof(double) * (size_t)nr * (fused ? 4 : 2));
    ab.ensure(sizeof(double) * (size_t)(3 * itmax + 8));
    double* v = vecs.as<double>();
    double* w = v + nr;
    double* alpha = ab.as<double>();
    double* beta = alpha + itmax;
    double* alpha_im = beta + itmax;
    double* scratch = alpha_im + itmax;
    unsigned* cnt = reinterpret_cast<unsigned*>(scratch + 4);
    RCG_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), st_));
    rcg::launch_randomize_cols<double>(nr, v, 1, 0, 1, seed, *counter, st_);
    *counter += (uint64_t)nr;
    rcg::launch_lz_start(nr, v, w, ws(), ws_elems_, scratch, st_);
    double* pv[2] = {v, fused ? w + nr : nullptr};
    double* pw[2] = {w, fused ? w + 2 * nr : nullptr};
    auto step = [&](long k) {
      if (fused) {
        const int c = (int)(k & 1);
        rcg::launch_lz_step(n_, op_->rowptr, op_->col, op_->val, pv[c], pw[c], pv[1 - c],
                            pw[1 - c], k == 0 ? nullptr : beta + (k - 1), alpha + k, beta + k, cnt,
                            ws(), ws_elems_, st_);
        return;
      }
      if (k != 0) rcg::launch_lz_swap(nr, v, w, beta + (k - 1), st_);
      if (op_->is_dense()) {
        rcg::launch_dense_step<double>(n_, op_->dense, op_->lda, w, nullptr, v, 1, 1, 1.0, 1.0,
                                       0.0, rcg::kStepApplyBeta, nullptr, nullptr, st_);
        rcg::launch_lz_dot(n_, w, v, alpha + k, ws(), ws_elems_, st_);
      } else if (op_->is_callback()) {
        // v = A w + v: an n-vector is a one-column block in either layout
        call_user(w, v, 1, n_, 1.0, 1.0);
        if (cx_)
          rcg::launch_lz_dot_c(n_, w, v, alpha + k, alpha_im + k, ws(), ws_elems_, st_);
        else
          rcg::launch_lz_dot(n_, w, v, alpha + k, ws(), ws_elems_, st_);
      } else {
        rcg::launch_lz_apply_dot_c(n_, op_->rowptr, op_->col, op_->val, w, v, alpha + k,
                                   alpha_im + k, ws(), ws_elems_, st_);
      }
      rcg::launch_lz_axpy_norm(nr, w, v, alpha + k, beta + k, ws(), ws_elems_, st_);
    };
    // Steps go out in chunks of 5 (the convergence test's period); each full
    // chunk's alpha/beta are copied to one of two host slots behind an
    // event.  The next chunk is issued before the host waits for the
    // current one, so the device keeps stepping while the host tests; steps
    // past the stopping point only move the internal vectors.
    hipEvent_t cev[2] = {nullptr, nullptr};
    const unsigned evf = hipEventDisableTiming | (ctx_->blocking_ev ? hipEventBlockingSync : 0u);
    for (auto& e : cev) RCG_HIP(hipEventCreateWithFlags(&e, evf));
    struct EvGuard {
      hipEvent_t* e;
      ~EvGuard() {
        for (int i = 0; i < 2; ++i)
          if (e[i]) (void)hipEventDestroy(e[i]);
      }
    } evg{cev};
    long issued = 0;
    auto issue_chunk = [&]() {
      const long k0 = issued, k1 = std::min(itmax, k0 + 5);
      for (long q = k0; q < k1; ++q) step(q);
      issued = k1;
      if (k1 - k0 == 5) {
        double* h = hlz() + 24 * ((k0 / 5) & 1);
        RCG_HIP(hipMemcpyAsync(h, alpha + k0, sizeof(double) * 5, hipMemcpyDeviceToHost, st_));
        RCG_HIP(hipMemcpyAsync(h + 8, beta + k0, sizeof(double) * 5, hipMemcpyDeviceToHost, st_));
        if (cx_)
          RCG_HIP(hipMemcpyAsync(h + 16, alpha_im + k0, sizeof(double) * 5, hipMemcpyDeviceToHost,
                                 st_));
        RCG_HIP(hipEventRecord(cev[(k0 / 5) & 1], st_));
      }
    };
    std::vector<double> ha(itmax), hb(itmax);
    double minDiff = 1.0e10, maxDiff = 1.0e10, lmin = 0.0, lmax = 0.0;
    double rfMax = 9.0, rfMin = 9.0, rateMax = 1.0, rateMin = 1.0;
    double oldMax = 0.0, oldMin = 0.0, dMaxA = 0.0, dMinA = 0.0, dMaxB = 0.0, dMinB = 0.0;
    int diffFlag = 0, rateFlag = 0;
    long k = 0;
    bool done = false;
    if (itmax > 0) issue_chunk();
    while (k < itmax && !done) {
      const long k0 = k;
      if (issued < itmax) issue_chunk();  // the next chunk, ahead of the test
      k = std::min(itmax, k0 + 5);
      if (k - k0 == 5) {
        const double* h = hlz() + 24 * ((k0 / 5) & 1);
        wait_event(cev[(k0 / 5) & 1], ctx_->blocking_ev != nullptr);
        for (int q = 0; q < 5; ++q) {
          ha[k - 5 + q] = h[q];
          hb[k - 5 + q] = h[8 + q];
          // LanczosMaxMinFinder.h:350-356: a non-real <w, Aw> means A is not Hermitian
          if (cx_ && std::abs(h[16 + q]) > (1.0 + std::abs(h[q])) * 1.0e-12)
            throw Error(RCG_ERR_NOT_HERMITIAN, "Lanczos max/min finder: operator is not Hermitian");
        }
        std::vector<double> D(ha.begin(), ha.begin() + k), E(hb.begin(), hb.begin() + (k - 1));
        lmax = rcg::dense::tridiag_eigenvalue(D, E, k - 1, 1.0e-14);
        lmin = rcg::dense::tridiag_eigenvalue(D, E, 0, 1.0e-14);
        if (diffFlag == 0) {
          oldMax = lmax;
          oldMin = lmin;
          diffFlag = 1;
        } else {
          maxDiff = std::abs(lmax - oldMax) / relErr(lmax, tol);
          minDiff = std::abs(lmin - oldMin) / relErr(lmin, tol);
          if (rateFlag == 0) {
            dMaxA = std::abs(lmax - oldMax);
            dMinA = std::abs(lmin - oldMin);
            rateFlag = 1;
          } else {
            if (maxDiff > 1.0e-09) {
              dMaxB = std::abs(lmax - oldMax);
              rateMax = dMaxB / dMaxA;
            } else {
              rateMax = 0.95;
            }
            if (minDiff > 1.0e-09) {
              dMinB = std::abs(lmin - oldMin);
              rateMin = dMinB / dMinA;
            } else {
              rateMin = 0.95;
            }
            rfMax = 9.0;
            rfMin = 9.0;
            if (rateMax < 0.90) rfMax = rateMax / (1.0 - rateMax);
            if (rateMin < 0.90) rfMin = rateMin / (1.0 - rateMin);
            dMaxA = dMaxB;
            dMinA = dMinB;
          }
          oldMax = lmax;
          oldMin = lmin;
          minDiff *= rfMin;
          maxDiff *= rfMax;
        }
        if (maxFlag && minFlag) done = (minDiff < tol && maxDiff < tol);
        else if (maxFlag) done = maxDiff < tol;
        else if (minFlag) done = minDiff < tol;
      }
    }
    sync();
    if (lmin > lmax) std::swap(lmin, lmax);
    *mnOut = lmin;
    *mxOut = lmax;
    *itOut = k;
  }

 private:
  rcg_context_s* ctx_;
  rcg_operator_s* op_;
  int64_t n_;
  hipStream_t st_;
  double* buf_[2] = {nullptr, nullptr};
  int iv_ = 0;
  double* W_ = nullptr;
  double* Y_ = nullptr;
  int64_t ld_ = 0, ldy_ = 0, bb_ = 4;
  size_t ws_elems_ = 0;
  bool cx_ = false;
  int cw_ = 1;
  int last_panel_ = 0;
  std::string last_sched_;
  float* E_[2] = {nullptr, nullptr};  // FP32 correction blocks (mixed filter)
  float* R32_ = nullptr;              // FP32 residual block (mixed filter)
  std::vector<double> last_norms_;    // column norms of V at the last residuals
  // the second CholQR factor's inverse kept by ortho_project for the rotation
  static constexpr int64_t kFoldCols = 64;
  int64_t fold_b_ = 0;
  std::vector<double> fold_r_;
  std::vector<cplx> fold_c_;
  template <class S>
  std::vector<S>& fold() {
    if constexpr (std::is_same<S, double>::value) return fold_r_;
    else return fold_c_;
  }
  DevBuf coef_;                       // mixed filter coefficient table
  DevBuf cb_in_, cb_out_;             // column-major blocks of a caller's operator
  hipEvent_t coef_ev_ = nullptr;      // its upload from hcoef_ has finished
  HostBuf hcoef_;
  hipEvent_t fev_[2] = {nullptr, nullptr};
  FilterBatcher* batcher_ = nullptr;
  hipEvent_t ev_done_ = nullptr;  // the batcher has run this engine's steps
  bool step_pending_ = false;
  long pend_launches_ = 0, step_launches_ = 0;
  double pend_bytes_ = 0.0, step_bytes_ = 0.0, step_seconds_ = 0.0;
};

// ------------------------------------------------------------- the solver
struct TraceRow {
  double v[10];
};

// Stream time per phase of a solve.  An event is recorded on the solve's
// stream where a phase begins; the time between consecutive events goes to
// the phase that began at the first of them, so it includes the stream's
// idle time while the host works inside the phase (the small eigensolve
// counts to Rayleigh-Ritz, a served solve's wait for its shared filter batch
// to the filter).  Finished intervals are folded whenever a new phase
// begins; the solve's host waits keep the list of pending events short.
class PhaseClock {
 public:
  enum { kLanczos = 0, kFilter, kOrtho, kRR, kOther, kCount };
  explicit PhaseClock(hipStream_t st) : st_(st) {}
  PhaseClock(const PhaseClock&) = delete;
  PhaseClock& operator=(const PhaseClock&) = delete;
  ~PhaseClock() {
    for (const Mark& m : marks_) (void)hipEventDestroy(m.ev);
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }
  void mark(int phase) {
    fold(false);
    hipEvent_t e;
    if (!pool_.empty()) {
      e = pool_.back();
      pool_.pop_back();
    } else {
      RCG_HIP(hipEventCreate(&e));
    }
    RCG_HIP(hipEventRecord(e, st_));
    marks_.push_back({e, phase});
  }
  // closes the last phase and folds every interval (waits for the stream)
  void finish() {
    mark(kOther);
    if (!marks_.empty()) wait_event(marks_.back().ev, true);
    fold(true);
  }
  double seconds(int phase) const { return sec_[phase]; }

 private:
  struct Mark {
    hipEvent_t ev;
    int phase;
  };
  void fold(bool done) {
    while (marks_.size() >= 2) {
      if (!done) {
        const hipError_t r = hipEventQuery(marks_[1].ev);
        if (r != hipSuccess) {
          (void)hipGetLastError();
          break;
        }
      }
      float ms = 0.f;
      RCG_HIP(hipEventElapsedTime(&ms, marks_[0].ev, marks_[1].ev));
      sec_[marks_[0].phase] += 1.0e-3 * ms;
      pool_.push_back(marks_[0].ev);
      marks_.pop_front();
    }
  }
  hipStream_t st_;
  std::deque<Mark> marks_;
  std::vector<hipEvent_t> pool_;
  double sec_[kCount] = {};
};

class Solver {
 public:
  Solver(rcg_operator_s* op, const rcg_solve_params& p, rcg_context_s* ctx = nullptr,
         FilterBatcher* batcher = nullptr)
      : op_(op), p_(p), eng_(ctx ? ctx : op->ctx, op, op->n), clock_(eng_.stream()) {
    eng_.set_batcher(batcher);
    for (DevBuf* b : {&Vb_, &Tb_, &Wb_, &Yb_, &Mb_}) b->use_stream(eng_.stream());
  }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  static double rel(double val, double tol) {
    return (std::abs(val) * tol > 1.0e-11) ? std::abs(val) : 1.0e-11 / tol;
  }

  void record(long pass, long it, long deg, double sb, double sh, double ls, double mr, double md,
              double ratio, double gap) {
    trace_.push_back({{double(pass), double(it), double(deg), sb, sh, ls, mr, md, ratio, gap}});
  }

  long run(rcg_solve_result* R, std::vector<double>& eigValues, std::vector<double>& eigRes) {
    if (eng_.cx()) return run_t<cplx>(R, eigValues, eigRes);
    return run_t<double>(R, eigValues, eigRes);
  }

  template <class S>
  long run_t(rcg_solve_result* R, std::vector<double>& eigValues, std::vector<double>& eigRes) {
    const double t0 = wall();
    const int64_t n = op_->n;
    const long N = (long)n;
    long inc = (long)p_.inc, buf = (long)p_.buf;
    need(inc >= 0 && buf >= 0 && inc + buf >= 1, "subspace increment + buffer must be >= 1");
    long subspaceSize = inc + buf;
    long maxDim = (long)p_.k;
    if (subspaceSize > N) {
      if (inc < N) {
        buf = N - inc;
        subspaceSize = N;
      } else {
        subspaceSize = N;
        inc = N;
        buf = 0;
      }
      maxDim = N;
    }
    // leading dimensions in real elements (2 per complex entry)
    const int cw = eng_.cw();
    const int64_t ld = cw * block_ld(subspaceSize);
    const int64_t ycols = std::max<long>(std::min<long>(maxDim, N), subspaceSize);
    const int64_t ldy = cw * block_ld(ycols);
    const size_t blk = sizeof(double) * (size_t)n * (size_t)ld;
    Vb_.ensure(blk);
    Tb_.ensure(blk);
    Wb_.ensure(blk);
    Yb_.ensure(sizeof(double) * (size_t)n * (size_t)ldy);
    hipStream_t st = eng_.stream();
    RCG_HIP(hipMemsetAsync(Vb_.p, 0, blk, st));
    RCG_HIP(hipMemsetAsync(Tb_.p, 0, blk, st));
    RCG_HIP(hipMemsetAsync(Wb_.p, 0, blk, st));
    eng_.reserve(ld, ldy);
    eng_.attach(Vb_.as<double>(), Tb_.as<double>(), Wb_.as<double>(), ld, Yb_.as<double>(), ldy);
    const bool mixed = p_.filter_dtype == RCG_F32;
    if (mixed) {  // E_{j-1}, E_{j-2} and R of the FP32 correction recurrence
      const size_t fb = sizeof(float) * (size_t)n * (size_t)ld;
      Mb_.ensure(3 * fb);
      RCG_HIP(hipMemsetAsync(Mb_.p, 0, 3 * fb, st));
      float* m0 = Mb_.as<float>();
      eng_.attach_mixed(m0, m0 + (size_t)n * ld, m0 + 2 * (size_t)n * ld);
    }

    const long maxInner = p_.fixed_iteration ? 1 : (p_.max_inner > 0 ? (long)p_.max_inner : 10000);
    const long degMax = p_.deg_max > 0 ? (long)p_.deg_max : 200;
    const long rep = p_.rep > 0 ? (long)p_.rep : 1;
    const int stop = p_.stop == 0 ? RCG_STOP_COMBINATION : p_.stop;
    const double subspaceTol = p_.tol < 1.0e-11 ? 1.0e-11 : p_.tol;

    // spectral bounds (getMinEigenSystem, LM:712-738; interval, LM:745-782)
    double minEigValue, maxEigValue, lambdaMax;
    clock_.mark(PhaseClock::kLanczos);
    if (p_.use_bounds) {  // caller's bounds, no Lanczos (RayleighChebyshev.h:709-723, 812-821)
      minEigValue = p_.min_eig;
      maxEigValue = p_.max_eig;
      lambdaMax = p_.interval ? p_.lambda_max : maxEigValue;
    } else {
      double mn, mx;
      long it;
      eng_.lanczos(p_.maxmin_tol > 0 ? p_.maxmin_tol : 1.0e-6, true, p_.interval != 0,
                   p_.lanczos_itmax > 0 ? (long)p_.lanczos_itmax : 10000, p_.seed, &counter_, &mn,
                   &mx, &it);
      R->lanczos_iters = it;
      R->min_est = mn;
      R->max_est = mx;
      if (p_.interval && p_.lambda_max < mn) {
        R->found = 0;
        clock_.finish();
        R->t_lanczos = clock_.seconds(PhaseClock::kLanczos);
        R->t_total = wall() - t0;
        return 0;
      }
      if (std::abs(mx - mn) > 1.0e-12) mx += 0.001 * std::abs(mx - mn);
      minEigValue = mn;
      maxEigValue = mx;
      lambdaMax = p_.interval ? p_.lambda_max : mx;
    }
    const bool intervalFlag = p_.interval != 0;

    long foundSize = 0;
    // Ritz value of each block column at filter time, and whether the column
    // is a Ritz vector (the mixed filter's correction needs both)
    std::vector<double> fTheta(subspaceSize, 0.0);
    std::vector<char> fRitz(subspaceSize, 0);
    std::vector<double> theta(subspaceSize, 0.0), oldEigs(subspaceSize, 0.0),
        eigDiffs(subspaceSize, 1.0), oldEigDiffs(subspaceSize, 1.0), subspaceResiduals,
        residualHistory;
    std::vector<S> H, Yh;

    // start vectors (LM:948-985)
    clock_.mark(PhaseClock::kOrtho);
    if (start_ && start_cols_ > 0) {
      const long have = std::min<long>((long)start_cols_, subspaceSize);
      eng_.copy_cols(start_, cw * start_ld_, 0, eng_.V(), ld, 0, have);
      if (have < subspaceSize) eng_.randomize(have, subspaceSize - have, p_.seed, &counter_);
    } else {
      eng_.randomize(0, subspaceSize, p_.seed, &counter_);
    }

    if (N == subspaceSize) {  // quick return (LM:1005-1036)
      eng_.orthonormalize(subspaceSize);
      eng_.project(subspaceSize, H);
      rcg::dense::small_eigensystem(H.data(), subspaceSize, theta, Yh);
      eng_.rotate_residuals(subspaceSize, Yh, theta, subspaceSize, eigRes);
      eng_.copy_cols(eng_.V(), ld, 0, eng_.Y(), ldy, 0, subspaceSize);
      eng_.scale_found(eng_.Y(), ldy, 0, subspaceSize);
      eigValues = theta;
      found_ = subspaceSize;
      R->found = subspaceSize;
      clock_.finish();
      R->t_lanczos = clock_.seconds(PhaseClock::kLanczos);
      R->t_ortho = clock_.seconds(PhaseClock::kOrtho);
      R->t_total = wall() - t0;
      return subspaceSize;
    }

    bool completedBasis = false, exitFlag = false;
    long starDegree = 0, starDegreeSave = 0, pass = 0;
    double starBound = 0.0, starBoundSave = 0.0, shift = 0.0, maxEigDiff = 0.0,
           maxResidual = 0.0, maxGap = 0.0, eigDiffRatio = 0.0;
    long applyCountCumulative = 0;
    double fmaxres = 0.0, fmaxdiff = 0.0, fmaxgap = 0.0;

    while (!exitFlag) {
      for (long k = buf; k < subspaceSize; ++k) oldEigs[k] = oldEigs[buf];
      if (applyCountCumulative > 0) {
        eng_.randomize(buf, subspaceSize - buf, p_.seed, &counter_);
        for (long k = buf; k < subspaceSize; ++k) fRitz[k] = 0;
      }
      clock_.mark(PhaseClock::kOrtho);
      eng_.orthonormalize(subspaceSize);
      count_ortho_ += 2;

      double lambdaStar = maxEigValue;
      eigDiffRatio = 1.0;
      long innerLoopCount = 0;
      starDegreeSave = 0;
      starBoundSave = 0.0;
      double stopCheckValue = subspaceTol + 1.0e10;
      long applyCount = 0;
      residualHistory.clear();
      maxGap = maxResidual = maxEigDiff = 0.0;
      long oscillationCount = 0;

      while (stopCheckValue > subspaceTol && innerLoopCount < maxInner) {
        // filter parameters (LM:1135-1206)
        shift = -minEigValue;
        starDegreeSave = starDegree;
        starBoundSave = starBound;
        starDegree = 0;
        const double eMin = std::abs(minEigValue - lambdaStar) / rel(minEigValue, subspaceTol);
        const double eMax = std::abs(maxEigValue - lambdaStar) / rel(maxEigValue, subspaceTol);
        if (eMin < subspaceTol && eMax < subspaceTol) {
          starDegree = 1;
          starBound = maxEigValue;
        } else if (!(maxEigDiff > subspaceTol) && innerLoopCount > 3 && eigDiffRatio < .2) {
          starDegree = starDegreeSave;
          starBound = starBoundSave;
        }
        if (starDegree == 0)
          rcg::cpoly::star_degree_and_bound(shift, maxEigValue, lambdaStar, degMax, &starDegree,
                                            &starBound);
        if (starDegree < starDegreeSave) {
          starDegree = starDegreeSave;
          starBound = starBoundSave;
        }
        clock_.mark(PhaseClock::kFilter);
        if (!completedBasis) {
          if (mixed)
            eng_.filter_mixed(subspaceSize, starDegree, rep, starBound, shift, fTheta, fRitz);
          else
            eng_.filter(subspaceSize, starDegree, rep, starBound, shift);
          degree_sum_ += starDegree * rep;
        }
        applyCount += 1;
        count_apply_ += starDegree * subspaceSize;

        clock_.mark(PhaseClock::kOrtho);
        eng_.deflate(subspaceSize, foundSize);
        count_ortho_ += 1;

        // orthogonalize and project (the CholQR passes count as Rayleigh-Ritz)
        clock_.mark(PhaseClock::kRR);
        eng_.ortho_project(subspaceSize, H);
        count_apply_ += subspaceSize;
        const double te = wall();
        rcg::dense::small_eigensystem(H.data(), subspaceSize, theta, Yh);
        t_eig_ += wall() - te;
        const long residualCheckCount = (foundSize + subspaceSize < N) ? inc : subspaceSize;
        eng_.rotate_residuals(subspaceSize, Yh, theta, residualCheckCount, subspaceResiduals);
        fTheta.assign(theta.begin(), theta.end());
        fRitz.assign(subspaceSize, 1);
        count_apply_ += residualCheckCount;
        count_eig_ += 1;

        // convergence bookkeeping (LM:1285-1457)
        maxResidual = 0.0;
        for (size_t k = 0; k < subspaceResiduals.size(); ++k)
          maxResidual = std::max(maxResidual, subspaceResiduals[k] / rel(theta[k], subspaceTol));
        residualHistory.push_back(maxResidual);
        const long checkSize = intervalFlag ? inc + 1 : inc;
        for (long i = 0; i < checkSize; ++i) oldEigDiffs[i] = eigDiffs[i];
        maxEigDiff = 0.0;
        for (long i = 0; i < checkSize; ++i) {
          const double d = std::abs(theta[i] - oldEigs[i]) / rel(oldEigs[i], subspaceTol);
          eigDiffs[i] = d;
          maxEigDiff = d > maxEigDiff ? d : maxEigDiff;
        }
        for (long i = 0; i < subspaceSize; ++i) oldEigs[i] = theta[i];
        long diffCount = 0;
        eigDiffRatio = 0.0;
        if (maxEigDiff > subspaceTol) {
          for (long i = 0; i < checkSize; ++i)
            if (std::abs(oldEigDiffs[i]) > subspaceTol / 10.0) {
              eigDiffRatio += std::abs(eigDiffs[i] / oldEigDiffs[i]);
              diffCount++;
            }
          eigDiffRatio = diffCount > 0 ? eigDiffRatio / double(diffCount) : 1.0;
        }
        const double spectralRange = std::abs(lambdaMax - minEigValue);
        maxGap = 0.0;
        for (long i = 1; i < checkSize; ++i)
          maxGap = std::max(maxGap, std::abs(theta[i] - theta[i - 1]) / spectralRange);
        record(pass, innerLoopCount, starDegree, starBound, shift, lambdaStar, maxResidual,
               maxEigDiff, eigDiffRatio, maxGap);
        if (p_.verbose)
          std::printf("%-5ld : Degree %-3ld  Residual Max: %-10.5g  Eig Diff Max: %-10.5g  Eig Conv Factor: %-10.5g Max Gap %-10.5g\n",
                      innerLoopCount, starDegree, maxResidual, maxEigDiff, eigDiffRatio, maxGap);

        if (stop == RCG_STOP_RESIDUAL_ONLY) stopCheckValue = maxResidual;
        else if (stop == RCG_STOP_EIGENVALUE_ONLY) stopCheckValue = maxEigDiff;
        else {
          stopCheckValue = maxEigDiff;
          if (maxResidual > std::sqrt(subspaceTol)) stopCheckValue = maxResidual;
        }
        if (inc == 0) stopCheckValue = 0.0;
        if (stop == RCG_STOP_RESIDUAL_ONLY && maxResidual < std::sqrt(subspaceTol)) {
          const long r = (long)residualHistory.size() - 1;
          if (r > 3) {
            const double dA = (residualHistory[r - 3] - 2.0 * residualHistory[r - 2] + residualHistory[r - 1]) / std::abs(maxResidual);
            const double dB = (residualHistory[r - 2] - 2.0 * residualHistory[r - 1] + residualHistory[r]) / std::abs(maxResidual);
            if (dA * dB < 0.0 && ++oscillationCount > 5) stopCheckValue = 0.0;
          }
        }
        lambdaStar = theta[subspaceSize - 1];
        minEigValue = minEigValue < theta[0] ? minEigValue : theta[0];
        innerLoopCount++;
      }
      fmaxres = std::max(fmaxres, maxResidual);
      fmaxdiff = std::max(fmaxdiff, maxEigDiff);
      fmaxgap = std::max(fmaxgap, maxGap);
      applyCountCumulative += applyCount;

      // lock found pairs (LM:1486-1543)
      long foundCount = 0;
      const long checkIndexCount = (foundSize + subspaceSize < N) ? inc : subspaceSize;
      for (long i = 0; i < checkIndexCount; ++i)
        if ((theta[i] - lambdaMax) / rel(lambdaMax, subspaceTol) < subspaceTol) foundCount++;
      if (foundCount + foundSize >= maxDim) {
        foundCount = maxDim - foundSize;
        exitFlag = true;
      }
      if (foundCount + foundSize >= N) {
        foundCount = N - foundSize;
        exitFlag = true;
      }
      if (foundCount > 0) {
        eng_.copy_cols(eng_.V(), ld, 0, eng_.Y(), ldy, foundSize, foundCount);
        eng_.scale_found(eng_.Y(), ldy, foundSize, foundCount);
        for (long i = 0; i < foundCount; ++i) {
          eigValues.push_back(theta[i]);
          eigRes.push_back(subspaceResiduals[i]);
        }
        foundSize += foundCount;
      }
      if (!exitFlag && foundCount > 0) {  // shuffle (LM:1549-1555)
        eng_.copy_cols(eng_.V(), ld, foundCount, eng_.T(), ld, 0, subspaceSize - foundCount);
        eng_.swapVT();
        for (long i = 0; i < subspaceSize; ++i) {
          const bool kept = i + foundCount < subspaceSize;
          fTheta[i] = kept ? theta[i + foundCount] : 0.0;
          fRitz[i] = kept ? fRitz[i + foundCount] : 0;
        }
      }
      const double gv = buf > 0 ? theta[subspaceSize - buf] : lambdaMax;  // guard (LM:1559-1587)
      guard_ = gv;
      if (!intervalFlag) {
        if (gv > lambdaMax) exitFlag = true;
      } else if ((gv - lambdaMax) / rel(lambdaMax, subspaceTol) >= 10.0 * subspaceTol) {
        exitFlag = true;
      }
      if (foundSize > 0) minEigValue = eigValues[foundSize - 1];  // LM:1593
      starDegree = 1;
      starBound = maxEigValue;
      if (!exitFlag && foundSize + subspaceSize >= N) {  // shrink (LM:1605-1638)
        if (foundSize + inc >= N) {
          buf = 0;
          inc = N - foundSize;
          subspaceSize = inc;
          completedBasis = true;
          eng_.randomize(0, subspaceSize, p_.seed, &counter_);
          fRitz.assign(subspaceSize, 0);
        } else {
          buf = N - (foundSize + inc);
          subspaceSize = inc + buf;
        }
        theta.resize(subspaceSize, 0.0);
        fTheta.resize(subspaceSize, 0.0);
        fRitz.resize(subspaceSize, 0);
      }
      ++pass;
    }

    // non-monotone repair (LM:1664-1707), or the requested FP64 refinement
    bool nonMonotone = false;
    for (long i = 0; i + 1 < foundSize; ++i)
      if (eigValues[i] > eigValues[i + 1]) nonMonotone = true;
    if (nonMonotone || ((p_.refine || mixed) && foundSize > 0)) {
      clock_.mark(PhaseClock::kRR);
      repair<S>(foundSize, eigValues, eigRes);
      count_apply_ += 2 * foundSize;
      count_eig_ += 1;
    }
    clock_.mark(PhaseClock::kOther);
    if (intervalFlag) {  // trim (LM:1715-1732)
      long fc = 0;
      for (double e : eigValues)
        if ((e - lambdaMax) / rel(lambdaMax, subspaceTol) < subspaceTol) fc++;
      if (fc < (long)eigValues.size()) {
        eigValues.resize(fc);
        eigRes.resize(fc);
        foundSize = fc;
      }
    }
    eng_.sync();
    found_ = foundSize;
    R->found = foundSize;
    R->guard = guard_;
    R->count_ortho = count_ortho_;
    R->count_apply = count_apply_;
    R->count_eig = count_eig_;
    R->final_maxres = fmaxres;
    R->final_maxdiff = fmaxdiff;
    R->final_maxgap = fmaxgap;
    clock_.finish();
    R->t_total = wall() - t0;
    R->t_lanczos = clock_.seconds(PhaseClock::kLanczos);
    R->t_filter = clock_.seconds(PhaseClock::kFilter);
    R->t_ortho = clock_.seconds(PhaseClock::kOrtho);
    R->t_rr = clock_.seconds(PhaseClock::kRR);
    R->t_eig = t_eig_;
    R->filter_degree_sum = degree_sum_;
    R->t_filter_device = eng_.step_seconds();
    R->filter_bytes = eng_.step_bytes();
    R->filter_launches = eng_.step_launches();
    R->filter_panel = eng_.last_panel();
    return foundSize;
  }

  // Rayleigh-Ritz over the found vectors Y[:, :f] (LM:1672-1707): the
  // projection, its eigensystem, the rotated and normalized vectors and
  // their residuals, with the Ritz values stored as the eigenvalues.
  template <class S>
  void repair(long f, std::vector<double>& eigValues, std::vector<double>& eigRes) {
    const int64_t n = op_->n;
    const int64_t ldy = eng_.ldy();
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)ldy;
    DevBuf AY(eng_.stream()), T2(eng_.stream());
    AY.ensure(bytes);
    T2.ensure(bytes);
    Engine e2(eng_.context(), op_, n);
    e2.reserve(ldy, ldy);
    e2.attach(eng_.Y(), T2.as<double>(), AY.as<double>(), ldy, eng_.Y(), ldy);
    std::vector<S> H, Z;
    std::vector<double> th, r;
    e2.project(f, H);
    const double te = wall();
    rcg::dense::small_eigensystem(H.data(), f, th, Z);
    t_eig_ += wall() - te;
    e2.rotate_residuals(f, Z, th, f, r);
    if (!e2.v_is_first()) e2.copy_cols(e2.V(), ldy, 0, eng_.Y(), ldy, 0, f);
    e2.scale_found(eng_.Y(), ldy, 0, f);
    e2.sync();
    eigValues.assign(th.begin(), th.end());
    eigRes = r;
  }

  void export_vectors(rcg_block_s* out) {
    if (!out || found_ <= 0) return;
    const long cnt = std::min<long>(found_, (long)out->cols);
    eng_.copy_cols(eng_.Y(), eng_.ldy(), 0, static_cast<double*>(out->d()), eng_.cw() * out->ld, 0,
                   cnt);
    eng_.sync();
  }

  std::vector<TraceRow> trace_;
  uint64_t counter_ = 0;
  const double* start_ = nullptr;
  int64_t start_ld_ = 0, start_cols_ = 0;

 private:
  rcg_operator_s* op_;
  rcg_solve_params p_;
  Engine eng_;
  DevBuf Vb_, Tb_, Wb_, Yb_, Mb_;
  long found_ = 0;
  PhaseClock clock_;
  long count_ortho_ = 0, count_apply_ = 0, count_eig_ = 0, degree_sum_ = 0;
  double guard_ = 0.0, t_eig_ = 0.0;
};

// allow_cx: a complex operator with complex (C128) blocks is accepted too.
// ------------------------------------------------------------------ server
// Continuous batching of lowest-eigenpair solves.  Requests queue in arrival
// order; a worker admits the queue head when a worker slot is free and the
// head's device footprint fits the HBM budget next to the running solves.
// Each worker has its own stream and scratch (a context of its own) and runs
// the solve's state machine; the filter sweeps of all running solves go
// through one FilterBatcher, so solves that arrive while others run share
// their step launches.  Eigenvectors stay on the device in a server-owned
// block until the ticket is released.
struct Ticket {
  int64_t id = 0;
  rcg_operator_s* op = nullptr;
  rcg_solve_params p{};
  int64_t footprint = 0;
  int64_t resident = 0;  // bytes of evecs, held against the budget until release
  int state = 0;  // 0 queued, 1 running, 2 done
  double t_submit = 0, t_admit = 0, t_done = 0;
  rcg_solve_result res{};
  std::vector<double> evals, resid;
  rcg_block_s* evecs = nullptr;
  int status = RCG_OK;
  std::string error;
  Ticket* start = nullptr;  // warm start: a finished ticket whose vectors start this solve
  int pins = 0;             // running solves that start from this ticket's vectors
};

int64_t solve_footprint(const rcg_operator_s* op, const rcg_solve_params& p) {
  const int64_t n = op->n;
  const int cw = op->dtype == RCG_C128 ? 2 : 1;
  int64_t sub = std::min<int64_t>(p.inc + p.buf, n);
  const int64_t ld = cw * block_ld(sub);
  const int64_t ldy = cw * block_ld(std::max<int64_t>(std::min<int64_t>(p.k, n), sub));
  int64_t bytes = 8 * n * (3 * ld + 2 * ldy);  // V, T, W, Y and the caller's copy of Y
  if (p.filter_dtype == RCG_F32) bytes += 4 * 3 * n * ld;
  if (p.refine || p.filter_dtype == RCG_F32) bytes += 8 * 2 * n * ldy;  // refinement blocks
  return bytes;
}

}  // namespace

struct rcg_server_s {
  rcg_context_s* ctx = nullptr;
  // used: footprints of the running solves plus the resident eigenvectors of
  // finished, unreleased tickets
  int64_t budget = 0, used = 0, peak = 0;
  int running = 0;
  int max_active = 1;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Ticket*> queue;
  std::map<int64_t, std::unique_ptr<Ticket>> tickets;
  int64_t next_id = 1;
  bool stop = false;
  std::unique_ptr<FilterBatcher> batcher;
  std::vector<std::unique_ptr<rcg_context_s>> wctx;
  std::vector<std::thread> workers;
  int64_t completed = 0;
};

namespace {

double now_s() { return wall(); }

void server_worker(rcg_server_s* S, int w) {
  (void)hipSetDevice(S->ctx->device);
  rcg_context_s* ctx = S->wctx[w].get();
  while (true) {
    Ticket* t = nullptr;
    {
      std::unique_lock<std::mutex> g(S->mu);
      S->cv.wait(g, [&] {
        // a request that does not fit next to the others runs once no other
        // solve is running (resident vectors alone cannot hold it back forever)
        return S->stop || (!S->queue.empty() && (S->used + S->queue.front()->footprint <= S->budget ||
                                                 S->running == 0));
      });
      if (S->stop) return;
      t = S->queue.front();
      S->queue.pop_front();
      S->used += t->footprint;
      ++S->running;
      S->peak = std::max(S->peak, S->used);
      t->state = 1;
      t->t_admit = now_s();
      S->batcher->attach();
    }
    try {
      Solver s(t->op, t->p, ctx, S->batcher.get());
      if (t->start) {  // resident vectors of a finished request (LM:957-985)
        s.start_ = static_cast<const double*>(t->start->evecs->d());
        s.start_ld_ = t->start->evecs->ld;
        s.start_cols_ = t->p.nonrandom_start;
      }
      std::vector<double> ev, rr;
      s.run(&t->res, ev, rr);
      t->res.trace_len = (int64_t)s.trace_.size();
      t->res.rng_counter = s.counter_;
      t->evals = ev;
      t->resid = rr;
      if (t->res.found > 0) {
        auto b = std::make_unique<rcg_block_s>();
        b->ctx = ctx;  // filled on the worker stream, then handed to the server context
        b->n = t->op->n;
        b->cols = t->res.found;
        b->ld = block_ld(b->cols);
        b->dtype = t->op->dtype;
        b->mem.use_stream(ctx->stream);
        b->mem.ensure(rcg::dtype_size(b->dtype) * (size_t)b->n * (size_t)b->ld);
        s.export_vectors(b.get());
        b->ctx = S->ctx;
        b->mem.stream = S->ctx->stream;  // the caller's work on it goes there from now
        t->evecs = b.release();
      }
      RCG_HIP(hipEventRecord(ctx->blocking_ev, ctx->stream));
      wait_event(ctx->blocking_ev, true);
    } catch (const Error& e) {
      t->status = e.code;
      t->error = e.what();
    } catch (const std::exception& e) {
      t->status = RCG_ERR_INVALID;
      t->error = e.what();
    }
    S->batcher->detach();
    {
      std::lock_guard<std::mutex> g(S->mu);
      if (t->start) {
        --t->start->pins;
        t->start = nullptr;
      }
      S->used -= t->footprint;
      --S->running;
      if (t->evecs) {
        t->resident = (int64_t)t->evecs->mem.bytes;
        S->used += t->resident;
        S->peak = std::max(S->peak, S->used);
      }
      t->state = 2;
      t->t_done = now_s();
      ++S->completed;
    }
    S->cv.notify_all();
  }
}

void check_pair(rcg_operator op, rcg_block a, rcg_block b, int64_t cols, bool allow_cx = false) {
  need(op && a && b, "NULL argument");
  need(a->n == op->n && b->n == op->n, "block row count differs from the operator");
  need(a->ld == b->ld, "blocks must share the leading dimension");
  need(cols >= 0 && cols <= a->cols && cols <= b->cols, "column count exceeds the block");
  if (allow_cx && op->dtype == RCG_C128) {
    need(a->dtype == RCG_C128 && b->dtype == RCG_C128, "a complex operator takes C128 blocks");
  } else {
    need(a->dtype == RCG_F64 && b->dtype == RCG_F64, "FP64 blocks expected");
    need(op->dtype == RCG_F64, "this entry point takes real FP64 operators");
  }
  need(a->ctx == op->ctx && b->ctx == op->ctx, "handles belong to different contexts");
}

// leading dimension of a block in real elements
int64_t real_ld(const rcg_block_s* b) { return (b->dtype == RCG_C128 ? 2 : 1) * b->ld; }

// Engine over caller blocks for the piecewise entry points (C128 blocks run
// as their interleaved real columns).
struct Piece {
  Engine eng;
  Piece(rcg_block V, rcg_block T, rcg_operator_s* op, int64_t kmax = 4)
      : eng(V->ctx, op, V->n, V->dtype == RCG_C128) {
    eng.reserve(real_ld(V), std::max<int64_t>(kmax, 4));
    eng.attach(static_cast<double*>(V->d()), T ? static_cast<double*>(T->d()) : nullptr, nullptr,
               real_ld(V), nullptr, 0);
  }
  // make sure the result sits in the caller's V block
  void settle(rcg_block V, int64_t cols) {
    if (!eng.v_is_first())
      rcg::launch_copy_cols<double>(V->n, eng.V(), (int)real_ld(V), 0, static_cast<double*>(V->d()),
                                    (int)real_ld(V), 0, (int)(eng.cw() * cols), eng.stream());
    eng.sync();
  }
};

// X and tmp of one shape and dtype (FP64, or C128 for complex blocks).
void check_block_pair(rcg_block X, rcg_block tmp, int64_t cols) {
  need(X && tmp && X != tmp, "X and tmp must be distinct blocks");
  need(X->n == tmp->n && X->ld == tmp->ld && X->ctx == tmp->ctx, "X and tmp differ in shape");
  need(cols >= 1 && cols <= X->cols && cols <= tmp->cols, "bad column count");
  need((X->dtype == RCG_F64 || X->dtype == RCG_C128) && tmp->dtype == X->dtype,
       "FP64 or C128 blocks of one dtype expected");
}

}  // namespace

// =================================================================== C ABI
extern "C" {

int rcg_get_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return RCG_ERR_INVALID;
  std::snprintf(buf, len, "%s", g_last_error.c_str());
  return RCG_OK;
}

const char* rcg_version(void) { return "rcg 0.1 gfx950"; }

int rcg_device_count(int* count) {
  return guarded([&] {
    need(count != nullptr, "count is NULL");
    RCG_HIP(hipGetDeviceCount(count));
  });
}

int rcg_context_create(int device, void* hip_stream, rcg_context* out) {
  return guarded([&] {
    need(out != nullptr, "out is NULL");
    int ndev = 0;
    RCG_HIP(hipGetDeviceCount(&ndev));
    need(device >= 0 && device < ndev, "device index out of range");
    auto c = std::make_unique<rcg_context_s>();
    c->device = device;
    DeviceGuard g(device);
    if (hip_stream) {
      c->stream = static_cast<hipStream_t>(hip_stream);
    } else {
      RCG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->own_stream = true;
    }
    c->bind_scratch();
    (void)noalloc_scratch(device);
    *out = c.release();
  });
}

int rcg_context_destroy(rcg_context ctx) {
  return guarded([&] {
    if (!ctx) return;
    DeviceGuard g(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->release_scratch();  // while the stream they are bound to exists
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int rcg_memory_trim(void) {
  return guarded([&] { MemPool::get().trim(); });
}

int rcg_context_synchronize(rcg_context ctx) {
  return guarded([&] {
    need(ctx != nullptr, "context is NULL");
    DeviceGuard g(ctx->device);
    RCG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int rcg_block_create(rcg_context ctx, int64_t n, int64_t cols, int dtype, rcg_block* out) {
  return guarded([&] {
    need(ctx && out, "NULL argument");
    need(n > 0 && cols > 0, "block dimensions must be positive");
    need(dtype == RCG_F64 || dtype == RCG_F32 || dtype == RCG_C128, "blocks are FP64, FP32 or C128");
    DeviceGuard g(ctx->device);
    auto b = std::make_unique<rcg_block_s>();
    b->ctx = ctx;
    b->n = n;
    b->cols = cols;
    b->ld = block_ld(cols);
    b->dtype = dtype;
    const size_t bytes = rcg::dtype_size(dtype) * (size_t)n * (size_t)b->ld;
    b->mem.ensure(bytes);
    RCG_HIP(hipMemsetAsync(b->mem.p, 0, bytes, ctx->stream));
    RCG_HIP(hipStreamSynchronize(ctx->stream));
    *out = b.release();
  });
}

int rcg_block_destroy(rcg_block b) {
  return guarded([&] {
    if (!b) return;
    DeviceGuard g(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
  });
}

namespace {
// Host column-major <-> device row-interleaved through a host staging copy.
void transfer(rcg_block b, void* host, int64_t ld_host, int64_t cols, bool up) {
  need(b && host, "NULL argument");
  need(cols >= 0 && cols <= b->cols, "column count exceeds the block");
  need(ld_host >= b->n, "host leading dimension smaller than n");
  DeviceGuard g(b->ctx->device);
  const size_t es = rcg::dtype_size(b->dtype);
  const size_t bytes = es * (size_t)b->n * (size_t)b->ld;
  std::vector<char> rows(bytes);
  char* h = static_cast<char*>(host);
  hipStream_t st = b->ctx->stream;
  if (up) {
    RCG_HIP(hipMemcpyAsync(rows.data(), b->mem.p, bytes, hipMemcpyDeviceToHost, st));
    RCG_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < b->n; ++i)
      for (int64_t c = 0; c < cols; ++c)
        std::memcpy(&rows[es * (i * b->ld + c)], h + es * (c * ld_host + i), es);
    RCG_HIP(hipMemcpyAsync(b->mem.p, rows.data(), bytes, hipMemcpyHostToDevice, st));
    RCG_HIP(hipStreamSynchronize(st));
  } else {
    RCG_HIP(hipMemcpyAsync(rows.data(), b->mem.p, bytes, hipMemcpyDeviceToHost, st));
    RCG_HIP(hipStreamSynchronize(st));
    for (int64_t c = 0; c < cols; ++c)
      for (int64_t i = 0; i < b->n; ++i)
        std::memcpy(h + es * (c * ld_host + i), &rows[es * (i * b->ld + c)], es);
  }
}
}  // namespace

int rcg_block_upload(rcg_block b, const void* host, int64_t ld_host, int64_t cols) {
  return guarded([&] { transfer(b, const_cast<void*>(host), ld_host, cols, true); });
}

int rcg_block_download(rcg_block b, void* host, int64_t ld_host, int64_t cols) {
  return guarded([&] { transfer(b, host, ld_host, cols, false); });
}

int rcg_block_copy(rcg_block dst, rcg_block src) {
  return guarded([&] {
    need(dst && src, "NULL argument");
    need(dst->n == src->n && dst->ld == src->ld && dst->dtype == src->dtype,
         "blocks differ in shape or type");
    DeviceGuard g(dst->ctx->device);
    RCG_HIP(hipMemcpyAsync(dst->mem.p, src->mem.p,
                           rcg::dtype_size(src->dtype) * (size_t)src->n * (size_t)src->ld,
                           hipMemcpyDeviceToDevice, dst->ctx->stream));
    RCG_HIP(hipStreamSynchronize(dst->ctx->stream));
  });
}

int rcg_block_randomize(rcg_block b, int64_t col0, int64_t ncols, uint64_t seed,
                        uint64_t* counter) {
  return guarded([&] {
    need(b && counter, "NULL argument");
    need(col0 >= 0 && ncols >= 0 && col0 + ncols <= b->cols, "column range exceeds the block");
    DeviceGuard g(b->ctx->device);
    if (b->dtype == RCG_C128) {
      rcg::launch_randomize_cplx(b->n, static_cast<double*>(b->d()), (int)(2 * b->ld), (int)col0,
                                 (int)ncols, seed, *counter, b->ctx->stream);
      *counter += 2 * (uint64_t)ncols * (uint64_t)b->n;
    } else {
      if (b->dtype == RCG_F64)
        rcg::launch_randomize_cols<double>(b->n, static_cast<double*>(b->d()), (int)b->ld,
                                           (int)col0, (int)ncols, seed, *counter, b->ctx->stream);
      else
        rcg::launch_randomize_cols<float>(b->n, static_cast<float*>(b->d()), (int)b->ld, (int)col0,
                                          (int)ncols, seed, *counter, b->ctx->stream);
      *counter += (uint64_t)ncols * (uint64_t)b->n;
    }
    RCG_HIP(hipStreamSynchronize(b->ctx->stream));
  });
}

int rcg_block_info(rcg_block b, int64_t* n, int64_t* cols, int64_t* ld, int* dtype,
                   void** device_ptr) {
  return guarded([&] {
    need(b != nullptr, "block is NULL");
    if (n) *n = b->n;
    if (cols) *cols = b->cols;
    if (ld) *ld = b->ld;
    if (dtype) *dtype = b->dtype;
    if (device_ptr) *device_ptr = b->d();
  });
}

int rcg_operator_csr_create(rcg_context ctx, int64_t n, int64_t nnz, const int32_t* rowptr,
                            const int32_t* col, const void* val, int val_dtype, int mem,
                            rcg_operator* out) {
  return guarded([&] {
    need(ctx && rowptr && col && val && out, "NULL argument");
    need(n > 0 && nnz >= 0, "bad operator size");
    need(nnz <= INT32_MAX && n <= INT32_MAX, "sizes exceed the int32 index range");
    need(val_dtype == RCG_F64 || val_dtype == RCG_C128, "CSR values must be FP64 or C128");
    need(mem == RCG_MEM_HOST || mem == RCG_MEM_DEVICE, "bad memory flag");
    DeviceGuard g(ctx->device);
    auto o = std::make_unique<rcg_operator_s>();
    o->ctx = ctx;
    o->n = n;
    o->nnz = nnz;
    o->dtype = val_dtype;
    if (mem == RCG_MEM_DEVICE) {
      o->rowptr = rowptr;
      o->col = col;
      o->val = static_cast<const double*>(val);
    } else {
      // validated on the host before anything reaches a kernel
      need(rowptr[0] == 0 && rowptr[n] == nnz, "rowptr must start at 0 and end at nnz");
      for (int64_t i = 0; i < n; ++i) need(rowptr[i] <= rowptr[i + 1], "rowptr must be non-decreasing");
      for (int64_t p = 0; p < nnz; ++p) need(col[p] >= 0 && col[p] < n, "column index out of range");
      o->own_rowptr.ensure(sizeof(int) * (size_t)(n + 1));
      o->own_col.ensure(sizeof(int) * (size_t)std::max<int64_t>(nnz, 1));
      const size_t vs = rcg::dtype_size(val_dtype);
      o->own_val.ensure(vs * (size_t)std::max<int64_t>(nnz, 1));
      RCG_HIP(hipMemcpy(o->own_rowptr.p, rowptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice));
      RCG_HIP(hipMemcpy(o->own_col.p, col, sizeof(int) * nnz, hipMemcpyHostToDevice));
      RCG_HIP(hipMemcpy(o->own_val.p, val, vs * nnz, hipMemcpyHostToDevice));
      o->rowptr = o->own_rowptr.as<int>();
      o->col = o->own_col.as<int>();
      o->val = o->own_val.as<double>();
    }
    *out = o.release();
  });
}

int rcg_operator_csr_share(rcg_operator base, const void* val, int mem, rcg_operator* out) {
  return guarded([&] {
    need(base && val && out, "NULL argument");
    need(base->is_csr(), "the pattern comes from a CSR operator");
    need(mem == RCG_MEM_HOST || mem == RCG_MEM_DEVICE, "bad memory flag");
    DeviceGuard g(base->ctx->device);
    auto o = std::make_unique<rcg_operator_s>();
    o->ctx = base->ctx;
    o->n = base->n;
    o->nnz = base->nnz;
    o->dtype = base->dtype;
    o->rowptr = base->rowptr;
    o->col = base->col;
    o->gx = base->gx;
    o->gy = base->gy;
    o->gz = base->gz;
    if (mem == RCG_MEM_DEVICE) {
      o->val = static_cast<const double*>(val);
    } else {
      const size_t vs = rcg::dtype_size(base->dtype);
      o->own_val.ensure(vs * (size_t)std::max<int64_t>(base->nnz, 1));
      RCG_HIP(hipMemcpy(o->own_val.p, val, vs * base->nnz, hipMemcpyHostToDevice));
      o->val = o->own_val.as<double>();
    }
    *out = o.release();
  });
}

int rcg_operator_dense_create(rcg_context ctx, int64_t n, const double* A, int64_t lda, int mem,
                              rcg_operator* out) {
  return guarded([&] {
    need(ctx && A && out, "NULL argument");
    need(n > 0 && lda >= n, "bad dense operator size");
    need(mem == RCG_MEM_HOST || mem == RCG_MEM_DEVICE, "bad memory flag");
    DeviceGuard g(ctx->device);
    auto o = std::make_unique<rcg_operator_s>();
    o->ctx = ctx;
    o->n = n;
    o->nnz = n * n;
    o->dtype = RCG_F64;
    o->kind = 1;
    if (mem == RCG_MEM_DEVICE) {
      o->dense = A;
      o->lda = lda;
    } else {
      o->own_dense.ensure(sizeof(double) * (size_t)n * (size_t)n);
      RCG_HIP(hipMemcpy2D(o->own_dense.p, sizeof(double) * n, A, sizeof(double) * lda,
                          sizeof(double) * n, n, hipMemcpyHostToDevice));
      o->dense = o->own_dense.as<double>();
      o->lda = n;
    }
    *out = o.release();
  });
}

int rcg_operator_callback_create(rcg_context ctx, int64_t n, int dtype, rcg_apply_fn apply,
                                 void* user, rcg_operator* out) {
  return guarded([&] {
    need(ctx && apply && out, "NULL argument");
    need(n > 0, "bad operator size");
    need(dtype == RCG_F64 || dtype == RCG_C128, "a caller's operator is RCG_F64 or RCG_C128");
    auto o = std::make_unique<rcg_operator_s>();
    o->ctx = ctx;
    o->n = n;
    o->nnz = 0;
    o->dtype = dtype;
    o->kind = 2;
    o->apply_fn = apply;
    o->user = user;
    *out = o.release();
  });
}

int rcg_operator_set_grid(rcg_operator op, int64_t nx, int64_t ny, int64_t nz) {
  return guarded([&] {
    need(op != nullptr, "operator is NULL");
    need(op->is_csr(), "a grid hint applies to CSR operators");
    need(nx >= 1 && ny >= 1 && nz >= 1 && nx * ny * nz == op->n, "grid does not match the operator size");
    DeviceGuard g(op->ctx->device);
    RCG_HIP(hipStreamSynchronize(op->ctx->stream));
    op->gx = nx;
    op->gy = ny;
    op->gz = nz;
    op->scheds.clear();
    op->tuned.clear();
  });
}

int rcg_operator_destroy(rcg_operator op) {
  return guarded([&] {
    if (!op) return;
    DeviceGuard g(op->ctx->device);
    (void)hipStreamSynchronize(op->ctx->stream);
    delete op;
  });
}

int rcg_operator_apply(rcg_operator op, rcg_block in, rcg_block out, double alpha, double beta,
                       int64_t cols) {
  return guarded([&] {
    check_pair(op, in, out, cols, true);
    need(in != out, "in and out must differ");
    DeviceGuard g(op->ctx->device);
    const bool cx = op->dtype == RCG_C128;
    if (op->is_callback()) {
      Engine e(op->ctx, op, op->n);
      e.attach(nullptr, nullptr, nullptr, real_ld(in), nullptr, 0);
      e.cb_apply_ab(static_cast<const double*>(in->d()), static_cast<double*>(out->d()), cols,
                    alpha, beta);
    } else if (op->is_dense())
      rcg::launch_dense_step<double>(op->n, op->dense, op->lda, static_cast<const double*>(in->d()),
                                     nullptr, static_cast<double*>(out->d()), (int)in->ld,
                                     (int)cols, alpha, beta, 0.0,
                                     beta == 0.0 ? rcg::kStepApply : rcg::kStepApplyBeta, nullptr,
                                     nullptr, op->ctx->stream);
    else
      rcg::launch_csr_apply<double>(op->n, op->rowptr, op->col, op->val,
                                    static_cast<const double*>(in->d()), static_cast<double*>(out->d()),
                                    (int)real_ld(in), (int)((cx ? 2 : 1) * cols), alpha, beta,
                                    op->ctx->stream, nullptr, cx);
    RCG_HIP(hipStreamSynchronize(op->ctx->stream));
  });
}

int rcg_cheb_filter(rcg_operator op, rcg_block X, rcg_block tmp, int64_t cols, int64_t degree,
                    int64_t repetition, double lambda_max, double shift) {
  return guarded([&] {
    check_pair(op, X, tmp, cols, true);
    need(X != tmp, "X and tmp must differ");
    need(degree >= 1 && repetition >= 1, "degree and repetition must be >= 1");
    DeviceGuard g(op->ctx->device);
    Piece pc(X, tmp, op);
    pc.eng.filter(cols, (long)degree, (long)repetition, lambda_max, shift);
    pc.settle(X, cols);
  });
}

int rcg_cheb_filter_batch(int64_t count, const rcg_operator* ops, const rcg_block* X,
                          const rcg_block* tmp, const int64_t* cols, const int64_t* degree,
                          const double* lambda_max, const double* shift) {
  return guarded([&] {
    need(count >= 1 && ops && X && tmp && cols && degree && lambda_max && shift, "bad arguments");
    rcg_context_s* ctx = ops[0] ? ops[0]->ctx : nullptr;
    need(ctx != nullptr, "NULL operator");
    DeviceGuard g(ctx->device);
    // descriptors in the order the launches want: one launch per step and
    // launch class over the requests whose degree reaches the step
    struct Req {
      rcg::BatchDesc d;
      int cls;
      bool cx;
      int64_t q;
    };
    std::vector<Req> reqs;
    for (int64_t q = 0; q < count; ++q) {
      rcg_operator op = ops[q];
      check_pair(op, X[q], tmp[q], cols[q], true);
      need(X[q] != tmp[q], "X and tmp must differ");
      need(op->is_csr(), "the batched filter takes CSR operators");
      need(op->ctx == ctx, "operators of one context");
      need(degree[q] >= 1 && degree[q] <= INT32_MAX, "degree must be >= 1");
      if (cols[q] == 0) continue;
      const bool cx = op->dtype == RCG_C128;
      const int64_t rb = (cx ? 2 : 1) * cols[q];
      const auto c = rcg::cpoly::coefficients((long)degree[q], lambda_max[q], shift[q]);
      rcg::BatchDesc d{};
      d.rowptr = op->rowptr;
      d.col = op->col;
      d.val = op->val;
      d.buf[0] = X[q]->d();
      d.buf[1] = tmp[q]->d();
      d.n = op->n;
      d.ld = (int)real_ld(X[q]);
      d.ncv = (int)((rb + 1) / 2);
      d.tail = (int)(rb - (int64_t)(d.ncv - 1) * 2);
      d.m = (int)degree[q];
      const int cls = rcg::batch_class(d.ncv);
      d.rows_per_blk = rcg::batch_rows(cls);
      d.g1 = c.g1;
      d.g2 = c.g2;
      reqs.push_back({d, cls, cx, q});
    }
    if (reqs.empty()) return;
    std::stable_sort(reqs.begin(), reqs.end(), [](const Req& a, const Req& b) {
      const int ka = (a.cx ? 100 : 0) + a.cls, kb = (b.cx ? 100 : 0) + b.cls;
      if (ka != kb) return ka < kb;
      return a.d.m > b.d.m;
    });
    std::vector<rcg::BatchDesc> descs;
    std::vector<rcg::BatchTask> tasks;
    std::vector<int> first(reqs.size() + 1, 0);
    for (size_t i = 0; i < reqs.size(); ++i) {
      descs.push_back(reqs[i].d);
      first[i] = (int)tasks.size();
      const int64_t nb = (reqs[i].d.n + reqs[i].d.rows_per_blk - 1) / reqs[i].d.rows_per_blk;
      need((int64_t)tasks.size() + nb <= INT32_MAX, "too many row blocks");
      for (int64_t b = 0; b < nb; ++b) tasks.push_back({(int)i, (int)b});
    }
    first[reqs.size()] = (int)tasks.size();
    const hipStream_t st = ctx->stream;
    DevBuf dd(st), dt(st);
    dd.ensure(sizeof(rcg::BatchDesc) * descs.size());
    dt.ensure(sizeof(rcg::BatchTask) * tasks.size());
    RCG_HIP(hipMemcpyAsync(dd.p, descs.data(), sizeof(rcg::BatchDesc) * descs.size(),
                           hipMemcpyHostToDevice, st));
    RCG_HIP(hipMemcpyAsync(dt.p, tasks.data(), sizeof(rcg::BatchTask) * tasks.size(),
                           hipMemcpyHostToDevice, st));
    const int mmax = reqs[0].d.m > 0 ? std::max_element(reqs.begin(), reqs.end(), [](const Req& a,
                                           const Req& b) { return a.d.m < b.d.m; })->d.m : 0;
    for (int j = 1; j <= mmax; ++j)
      for (size_t g0 = 0; g0 < reqs.size();) {
        size_t g1 = g0;
        while (g1 < reqs.size() && reqs[g1].cx == reqs[g0].cx && reqs[g1].cls == reqs[g0].cls) ++g1;
        size_t e = g0;  // degree descending within the group: a prefix reaches step j
        while (e < g1 && reqs[e].d.m >= j) ++e;
        if (e > g0)
          rcg::launch_cheb_batch(dd.as<rcg::BatchDesc>(), dt.as<rcg::BatchTask>() + first[g0],
                                 first[e] - first[g0], j, false, false, reqs[g0].cx, reqs[g0].cls,
                                 st);
        g0 = g1;
      }
    // an odd degree leaves X_m in tmp
    for (const Req& r : reqs)
      if (r.d.m % 2 == 1) {
        rcg_block b = X[r.q];
        rcg::launch_copy_cols<double>(b->n, static_cast<const double*>(tmp[r.q]->d()),
                                      (int)real_ld(b), 0, static_cast<double*>(b->d()),
                                      (int)real_ld(b), 0, (int)((r.cx ? 2 : 1) * cols[r.q]), st);
      }
    RCG_HIP(hipStreamSynchronize(st));
  });
}

int rcg_cheb_filter_mixed(rcg_operator op, rcg_block X, rcg_block tmp, int64_t cols,
                          int64_t degree, int64_t repetition, double lambda_max, double shift,
                          const double* theta, const int32_t* ritz) {
  return guarded([&] {
    check_pair(op, X, tmp, cols, true);
    need(X != tmp, "X and tmp must differ");
    need(degree >= 1 && repetition >= 1, "degree and repetition must be >= 1");
    need(!ritz || theta, "Ritz flags need Ritz values");
    DeviceGuard g(op->ctx->device);
    Piece pc(X, tmp, op);
    // tmp doubles as the FP64 A*X block of the correction start
    const int64_t rld = real_ld(X);
    pc.eng.attach(static_cast<double*>(X->d()), static_cast<double*>(tmp->d()),
                  static_cast<double*>(tmp->d()), rld, nullptr, 0);
    const size_t fb = sizeof(float) * (size_t)X->n * (size_t)rld;
    DevBuf mb(pc.eng.stream());
    mb.ensure(3 * fb);
    RCG_HIP(hipMemsetAsync(mb.p, 0, 3 * fb, pc.eng.stream()));
    float* m0 = mb.as<float>();
    pc.eng.attach_mixed(m0, m0 + (size_t)X->n * rld, m0 + 2 * (size_t)X->n * rld);
    std::vector<double> th(cols, 0.0);
    std::vector<char> rz(cols, 0);
    for (int64_t c = 0; c < cols; ++c) {
      if (theta) th[c] = theta[c];
      if (ritz) rz[c] = ritz[c] != 0;
    }
    pc.eng.filter_mixed(cols, (long)degree, (long)repetition, lambda_max, shift, th, rz);
    pc.settle(X, cols);
  });
}

int rcg_orthonormalize(rcg_block X, rcg_block tmp, int64_t cols) {
  return guarded([&] {
    check_block_pair(X, tmp, cols);
    DeviceGuard g(X->ctx->device);
    Piece pc(X, tmp, nullptr);
    pc.eng.orthonormalize(cols);
    pc.settle(X, cols);
  });
}

int rcg_deflate(rcg_block X, int64_t cols, rcg_block Y, int64_t found) {
  return guarded([&] {
    need(X && Y && X != Y, "bad arguments");
    need(X->n == Y->n && X->ctx == Y->ctx, "X and Y differ in rows or context");
    need((X->dtype == RCG_F64 || X->dtype == RCG_C128) && Y->dtype == X->dtype,
         "FP64 or C128 blocks of one dtype expected");
    need(cols >= 1 && cols <= X->cols && found >= 0 && found <= Y->cols, "bad column counts");
    DeviceGuard g(X->ctx->device);
    Engine e(X->ctx, nullptr, X->n, X->dtype == RCG_C128);
    e.reserve(std::max(real_ld(X), real_ld(Y)), std::max(real_ld(X), real_ld(Y)));
    e.attach(static_cast<double*>(X->d()), nullptr, nullptr, real_ld(X),
             static_cast<double*>(Y->d()), real_ld(Y));
    e.deflate(cols, found);
    e.normalize(cols);
    e.sync();
  });
}

int rcg_project(rcg_operator op, rcg_block V, rcg_block W, int64_t cols, double* H_host) {
  return guarded([&] {
    check_pair(op, V, W, cols, true);
    need(V != W && H_host && cols >= 1, "bad arguments");
    DeviceGuard g(op->ctx->device);
    Engine e(op->ctx, op, op->n);
    e.reserve(real_ld(V), 4);
    e.attach(static_cast<double*>(V->d()), nullptr, static_cast<double*>(W->d()), real_ld(V),
             nullptr, 0);
    if (e.cx()) {
      std::vector<cplx> H;
      e.project(cols, H);
      std::memcpy(H_host, H.data(), sizeof(cplx) * cols * cols);
    } else {
      std::vector<double> H;
      e.project(cols, H);
      std::memcpy(H_host, H.data(), sizeof(double) * cols * cols);
    }
  });
}

int rcg_rotate(rcg_block V, rcg_block tmp, int64_t cols, const double* Y_host) {
  return guarded([&] {
    need(Y_host != nullptr, "bad arguments");
    check_block_pair(V, tmp, cols);
    DeviceGuard g(V->ctx->device);
    Piece pc(V, tmp, nullptr);
    if (pc.eng.cx()) pc.eng.right_multiply(cols, reinterpret_cast<const cplx*>(Y_host));
    else pc.eng.right_multiply(cols, Y_host);
    pc.eng.normalize(cols);
    pc.settle(V, cols);
  });
}

int rcg_small_eigensystem(int64_t n, const double* H, double* evals, double* evecs) {
  return guarded([&] {
    need(n >= 1 && H && evals && evecs, "bad arguments");
    std::vector<double> w, Z;
    rcg::dense::small_eigensystem(H, n, w, Z);
    std::memcpy(evals, w.data(), sizeof(double) * n);
    std::memcpy(evecs, Z.data(), sizeof(double) * n * n);
  });
}

int rcg_residuals(rcg_operator op, rcg_block V, rcg_block W, int64_t count, const double* theta,
                  double* res_host) {
  return guarded([&] {
    check_pair(op, V, W, count, true);
    need(V != W && theta && res_host, "bad arguments");
    DeviceGuard g(op->ctx->device);
    Engine e(op->ctx, op, op->n);
    e.reserve(real_ld(V), 4);
    e.attach(static_cast<double*>(V->d()), nullptr, static_cast<double*>(W->d()), real_ld(V),
             nullptr, 0);
    std::vector<double> r;
    e.residuals(count, theta, r);
    std::memcpy(res_host, r.data(), sizeof(double) * count);
  });
}

int rcg_residuals_noalloc(rcg_operator op, rcg_block V, rcg_block W, int64_t count,
                          const double* theta, double* res_host) {
  return guarded([&] {
    check_pair(op, V, W, count, true);
    need(V != W && theta && res_host, "bad arguments");
    if (op->is_callback())
      throw Error(RCG_ERR_UNSUPPORTED, "rcg_residuals_noalloc takes CSR and dense operators");
    if (count == 0) return;
    DeviceGuard g(op->ctx->device);
    double* base = noalloc_scratch(op->ctx->device);
    const int cw = op->dtype == RCG_C128 ? 2 : 1;
    const int64_t rc = cw * count;
    const int64_t ld = real_ld(V);
    const size_t part = rcg::colred_ws_elems(op->n, (int)rc);
    if ((size_t)(3 * rc) + part > kNoAllocDoubles)
      throw Error(RCG_ERR_INVALID,
                  "rcg_residuals_noalloc: " + std::to_string(3 * rc + (int64_t)part) +
                      " workspace doubles needed (3 per real column plus one per column and "
                      "512 rows), the no-allocation workspace holds " +
                      std::to_string(kNoAllocDoubles));
    // every region below is written in full before it is read: nothing
    // left by an earlier call is used
    double* th = base;           // theta per real column
    double* sums = base + rc;    // per real column
    double* res = base + 2 * rc; // per column
    double* ws = base + 3 * rc;  // partials
    const hipStream_t st = op->ctx->stream;
    if (cw == 1) {
      RCG_HIP(hipMemcpyAsync(th, theta, sizeof(double) * count, hipMemcpyHostToDevice, st));
    } else {
      for (int e = 0; e < cw; ++e)
        RCG_HIP(hipMemcpy2DAsync(th + e, sizeof(double) * cw, theta, sizeof(double),
                                 sizeof(double), (size_t)count, hipMemcpyHostToDevice, st));
    }
    const double* v = static_cast<const double*>(V->d());
    double* w = static_cast<double*>(W->d());
    if (op->is_dense())
      rcg::launch_dense_step<double>(op->n, op->dense, op->lda, v, nullptr, w, (int)ld, (int)count,
                                     1.0, 0.0, 0.0, rcg::kStepApply, nullptr, nullptr, st);
    else
      rcg::launch_csr_apply<double>(op->n, op->rowptr, op->col, op->val, v, w, (int)ld, (int)rc,
                                    1.0, 0.0, st, nullptr, cw == 2);
    rcg::launch_residual_sumsq<double>(op->n, v, w, (int)ld, (int)rc, th, sums, ws, part, st);
    rcg::launch_group_sqrt(sums, cw, (int)count, res, st);
    RCG_HIP(hipMemcpyAsync(res_host, res, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    RCG_HIP(hipStreamSynchronize(st));
  });
}

int rcg_lanczos_bounds(rcg_operator op, double tol, int stop_flags, int64_t itmax, uint64_t seed,
                       uint64_t* counter, double* min_eig, double* max_eig, int64_t* iters) {
  return guarded([&] {
    need(op && counter && min_eig && max_eig && iters, "NULL argument");
    need(itmax >= 5, "itmax must be >= 5");
    DeviceGuard g(op->ctx->device);
    Engine e(op->ctx, op, op->n);
    e.reserve(4, 4);
    double mn, mx;
    long it;
    e.lanczos(tol, (stop_flags & 1) != 0, (stop_flags & 2) != 0, (long)itmax, seed, counter, &mn,
              &mx, &it);
    *min_eig = mn;
    *max_eig = mx;
    *iters = it;
  });
}

int rcg_solve_lowest(rcg_operator op, const rcg_solve_params* params, rcg_solve_result* result,
                     double* evals, double* res, int64_t cap, rcg_block evecs, double* trace,
                     int64_t trace_cap) {
  return guarded([&] {
    need(op && params && result && evals && res, "NULL argument");
    need(params->k >= 1, "eigCount must be >= 1");
    need(params->filter_dtype == RCG_F64 || params->filter_dtype == RCG_F32,
         "filter_dtype must be RCG_F64 or RCG_F32");
    if (op->is_callback() && params->filter_dtype == RCG_F32)
      throw Error(RCG_ERR_UNSUPPORTED, "mixed-precision filtering needs the operator's values "
                                       "(a CSR or dense operator), not a caller's apply");
    if (evecs)
      need(evecs->n == op->n && evecs->dtype == op->dtype && evecs->ctx == op->ctx,
           "evecs block mismatch (rows, dtype of the operator, context)");
    DeviceGuard g(op->ctx->device);
    std::memset(result, 0, sizeof(*result));
    Solver s(op, *params);
    if (params->nonrandom_start > 0) {
      need(evecs != nullptr, "a non-random start needs the evecs block");
      need(params->nonrandom_start <= evecs->cols, "more start vectors than evecs columns");
      s.start_ = static_cast<const double*>(evecs->d());
      s.start_ld_ = evecs->ld;
      s.start_cols_ = params->nonrandom_start;
    }
    std::vector<double> ev, rr;
    const long f = s.run(result, ev, rr);
    result->trace_len = (int64_t)s.trace_.size();
    result->rng_counter = s.counter_;
    for (long i = 0; i < f && i < cap; ++i) {
      evals[i] = ev[i];
      res[i] = rr[i];
    }
    if (trace)
      for (int64_t t = 0; t < (int64_t)s.trace_.size() && t < trace_cap; ++t)
        std::memcpy(trace + 10 * t, s.trace_[t].v, sizeof(double) * 10);
    if (evecs) s.export_vectors(evecs);
  });
}

int64_t rcg_sizeof_solve_params(void) { return (int64_t)sizeof(rcg_solve_params); }
int64_t rcg_sizeof_solve_result(void) { return (int64_t)sizeof(rcg_solve_result); }

int rcg_server_create(rcg_context ctx, int64_t hbm_budget, int32_t max_active, int32_t window_us,
                      rcg_server* out) {
  return guarded([&] {
    need(ctx && out, "NULL argument");
    need(hbm_budget > 0 && max_active >= 1 && max_active <= 1024 && window_us >= 0,
         "bad server limits");
    DeviceGuard g(ctx->device);
    auto S = std::make_unique<rcg_server_s>();
    S->ctx = ctx;
    S->budget = hbm_budget;
    S->max_active = max_active;
    S->batcher = std::make_unique<FilterBatcher>(ctx->device, window_us);
    for (int w = 0; w < max_active; ++w) {
      auto c = std::make_unique<rcg_context_s>();
      c->device = ctx->device;
      RCG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->own_stream = true;
      RCG_HIP(hipEventCreateWithFlags(&c->blocking_ev,
                                      hipEventBlockingSync | hipEventDisableTiming));
      c->bind_scratch();
      S->wctx.push_back(std::move(c));
    }
    rcg_server_s* raw = S.get();
    for (int w = 0; w < max_active; ++w) S->workers.emplace_back(server_worker, raw, w);
    *out = S.release();
  });
}

namespace {
void server_enqueue(rcg_server s, rcg_operator op, const rcg_solve_params* params,
                    int64_t start_ticket, int64_t* ticket) {
  need(s && op && params && ticket, "NULL argument");
  need(params->k >= 1, "eigCount must be >= 1");
  need(params->inc >= 0 && params->buf >= 0 && params->inc + params->buf >= 1,
       "subspace increment + buffer must be >= 1");
  need(params->filter_dtype == RCG_F64 || params->filter_dtype == RCG_F32,
       "filter_dtype must be RCG_F64 or RCG_F32");
  need(params->nonrandom_start == 0,
       "served solves start from random vectors or, with rcg_server_submit_warm, from a "
       "finished ticket's vectors");
  need(op->ctx->device == s->ctx->device, "operator on another device");
  if (op->is_callback())
    throw Error(RCG_ERR_UNSUPPORTED, "the server takes CSR and dense operators (a caller's apply "
                                     "would run on the server's worker threads)");
  auto t = std::make_unique<Ticket>();
  t->op = op;
  t->p = *params;
  t->footprint = solve_footprint(op, *params);
  t->t_submit = now_s();
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (start_ticket > 0) {
      auto it = s->tickets.find(start_ticket);
      need(it != s->tickets.end(), "unknown start ticket");
      Ticket* st = it->second.get();
      need(st->state == 2 && st->status == RCG_OK && st->evecs != nullptr,
           "the start ticket has not finished with eigenvectors");
      need(st->evecs->n == op->n && st->evecs->dtype == op->dtype,
           "the start ticket's vectors do not match the operator (rows, dtype)");
      t->start = st;
      t->p.nonrandom_start = (int32_t)std::min<int64_t>(
          st->evecs->cols, std::max<int64_t>(params->k, params->inc + params->buf));
      ++st->pins;
    }
    t->id = s->next_id++;
    *ticket = t->id;
    s->queue.push_back(t.get());
    s->tickets[t->id] = std::move(t);
  }
  s->cv.notify_all();
}
}  // namespace

int rcg_server_submit(rcg_server s, rcg_operator op, const rcg_solve_params* params,
                      int64_t* ticket) {
  return guarded([&] { server_enqueue(s, op, params, 0, ticket); });
}

int rcg_server_submit_warm(rcg_server s, rcg_operator op, const rcg_solve_params* params,
                           int64_t start_ticket, int64_t* ticket) {
  return guarded([&] {
    need(start_ticket > 0, "bad start ticket");
    server_enqueue(s, op, params, start_ticket, ticket);
  });
}

int rcg_server_wait(rcg_server s, int64_t ticket, rcg_solve_result* result, double* evals,
                    double* res, int64_t cap, double* times) {
  int status = RCG_OK;
  std::string err;
  int rc = guarded([&] {
    need(s && result, "NULL argument");
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->tickets.find(ticket);
    need(it != s->tickets.end(), "unknown ticket");
    Ticket* t = it->second.get();
    s->cv.wait(g, [&] { return t->state == 2; });
    *result = t->res;
    for (int64_t i = 0; i < (int64_t)t->evals.size() && i < cap; ++i) {
      if (evals) evals[i] = t->evals[i];
      if (res) res[i] = t->resid[i];
    }
    if (times) {
      times[0] = t->t_submit;
      times[1] = t->t_admit;
      times[2] = t->t_done;
    }
    status = t->status;
    err = t->error;
  });
  if (rc != RCG_OK) return rc;
  if (status != RCG_OK) g_last_error = err;
  return status;
}

int rcg_server_vectors(rcg_server s, int64_t ticket, rcg_block* evecs) {
  return guarded([&] {
    need(s && evecs, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->tickets.find(ticket);
    need(it != s->tickets.end(), "unknown ticket");
    need(it->second->state == 2, "the solve has not finished");
    *evecs = it->second->evecs;
  });
}

int rcg_server_release(rcg_server s, int64_t ticket) {
  return guarded([&] {
    need(s != nullptr, "NULL argument");
    std::unique_ptr<Ticket> t;
    {
      std::lock_guard<std::mutex> g(s->mu);
      auto it = s->tickets.find(ticket);
      need(it != s->tickets.end(), "unknown ticket");
      need(it->second->state == 2, "the solve has not finished");
      need(it->second->pins == 0, "the ticket's vectors are the start of a queued or running solve");
      t = std::move(it->second);
      s->tickets.erase(it);
      s->used -= t->resident;
    }
    s->cv.notify_all();  // the freed bytes may admit the queue head
    if (t->evecs) delete t->evecs;
  });
}

int rcg_server_stats(rcg_server s, int64_t* stats) {
  return guarded([&] {
    need(s && stats, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    stats[0] = s->completed;
    stats[1] = s->batcher->batches();
    stats[2] = s->batcher->requests();
    stats[3] = s->used;
    stats[4] = s->peak;
    stats[5] = (int64_t)s->queue.size();
  });
}

int rcg_server_filter_profile(rcg_server s, double* out) {
  return guarded([&] {
    need(s && out, "NULL argument");
    s->batcher->profile(out);
  });
}

int rcg_server_destroy(rcg_server s) {
  return guarded([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stop = true;
    }
    s->cv.notify_all();
    for (auto& th : s->workers)
      if (th.joinable()) th.join();
    s->batcher.reset();
    for (auto& kv : s->tickets)
      if (kv.second->evecs) delete kv.second->evecs;
    for (auto& c : s->wctx) {
      (void)hipStreamSynchronize(c->stream);
      c->release_scratch();
      (void)hipEventDestroy(c->blocking_ev);
      (void)hipStreamDestroy(c->stream);
    }
    delete s;
  });
}

}  // extern "C"
