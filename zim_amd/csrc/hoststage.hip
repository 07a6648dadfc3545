// Host-memory transfers of the host-pointer entry points (mem ==
// PSGC_MEM_HOST), through pinned staging buffers.
//
// Caller buffers are pageable (a cv::Mat, a numpy array). Copied directly,
// pageable memory moves at about half the PCIe rate of pinned memory (30 and
// 18 GB/s against 57 GB/s host to device and back, measured on the MI355X
// host). Each context keeps two pinned chunks: a chunk of the caller's data
// is copied into one by several host threads (about 110 GB/s together) while
// the DMA of the other runs, so the transfer goes at the pinned rate. When
// pinned memory cannot be had, the copies fall back to direct ones.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "psgc_internal.h"

namespace psgc {
namespace {

constexpr size_t kPinChunk = (size_t)32 << 20;  // bytes per pinned chunk
constexpr size_t kSplitMin = (size_t)1 << 20;   // below this one thread copies

// A fixed set of worker threads for the host-side copies; run(n, fn) calls
// fn(0) .. fn(n - 1) on them and the caller and returns when all are done.
// One job at a time (concurrent callers take turns). The workers live until
// the process ends.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* p = new CopyPool();  // never destroyed: workers outlive static teardown
    return *p;
  }
  int threads() const { return (int)workers_ + 1; }
  void run(int n, const std::function<void(int)>& fn) {
    std::lock_guard<std::mutex> turn(submit_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      next_ = 0;
      count_ = n;
      left_ = n;
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return left_ == 0; });
    job_ = nullptr;
  }

 private:
  CopyPool() {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned want = std::min(7u, hw > 1 ? hw - 1 : 0u);
    // a thread that cannot be started leaves fewer workers (the caller
    // always takes part)
    for (unsigned i = 0; i < want; ++i) {
      try {
        std::thread([this] { loop(); }).detach();
        ++workers_;
      } catch (...) {
        break;
      }
    }
  }
  // takes tasks of the current job until none is left
  void work() {
    for (;;) {
      int i;
      const std::function<void(int)>* f;
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!job_ || next_ >= count_) return;
        i = next_++;
        f = job_;
      }
      (*f)(i);
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_.notify_all();
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      work();
    }
  }
  unsigned workers_ = 0;
  std::mutex submit_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int next_ = 0, count_ = 0, left_ = 0;
  uint64_t gen_ = 0;
};

// dst[r * dpitch .. + width) = src[r * spitch .. + width) for r < rows, the
// rows split over the pool (a flat copy when both pitches equal width).
void par_copy2d(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width,
                size_t rows) {
  const size_t total = width * rows;
  if (dpitch == width && spitch == width) {
    if (total < kSplitMin) {
      std::memcpy(dst, src, total);
      return;
    }
    CopyPool& pool = CopyPool::get();
    const int n = pool.threads();
    const size_t part = (total + n - 1) / n;
    pool.run(n, [&](int i) {
      const size_t b = std::min(total, (size_t)i * part), e = std::min(total, b + part);
      if (e > b) std::memcpy(dst + b, src + b, e - b);
    });
    return;
  }
  auto rows_of = [&](size_t r0, size_t r1) {
    for (size_t r = r0; r < r1; ++r) std::memcpy(dst + r * dpitch, src + r * spitch, width);
  };
  if (total < kSplitMin) {
    rows_of(0, rows);
    return;
  }
  CopyPool& pool = CopyPool::get();
  const int n = pool.threads();
  const size_t part = (rows + n - 1) / n;
  pool.run(n, [&](int i) {
    const size_t b = std::min(rows, (size_t)i * part);
    rows_of(b, std::min(rows, b + part));
  });
}

// The context's two pinned chunks and their events (allocated on first use);
// false when pinned memory is not available.
bool pinned_ready(psgc_ctx* ctx) {
  if (ctx->pin_failed) return false;
  if (ctx->pin[0]) return true;
  for (int k = 0; k < 2; ++k) {
    if (hipHostMalloc(&ctx->pin[k], kPinChunk, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->pin_ev[k], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      host_stage_release(ctx);
      ctx->pin_failed = true;
      return false;
    }
    ctx->pin_busy[k] = false;
  }
  return true;
}

// Waits until no DMA uses pinned chunk k.
int pin_wait(psgc_ctx* ctx, int k) {
  if (ctx->pin_busy[k]) {
    PSGC_HIP(hipEventSynchronize(ctx->pin_ev[k]));
    ctx->pin_busy[k] = false;
  }
  return PSGC_OK;
}

}  // namespace

void host_stage_release(psgc_ctx* ctx) {
  for (int k = 0; k < 2; ++k) {
    if (ctx->pin_ev[k]) {
      hipEventSynchronize(ctx->pin_ev[k]);
      hipEventDestroy(ctx->pin_ev[k]);
      ctx->pin_ev[k] = nullptr;
    }
    if (ctx->pin[k]) {
      hipHostFree(ctx->pin[k]);
      ctx->pin[k] = nullptr;
    }
    ctx->pin_busy[k] = false;
  }
}

// Host rows (pitch spitch) to a dense device block (pitch width), on the
// context's stream. Returns once the caller's memory has been read; the
// device copy may still be in flight (ordered before later work on the
// stream).
int host_to_device_2d(psgc_ctx* ctx, void* dst, const void* src, size_t spitch, size_t width,
                      size_t rows) {
  const size_t total = width * rows;
  if (total == 0) return PSGC_OK;
  if (!pinned_ready(ctx)) {
    if (spitch == width)
      PSGC_HIP(hipMemcpyAsync(dst, src, total, hipMemcpyHostToDevice, ctx->stream));
    else
      PSGC_HIP(hipMemcpy2DAsync(dst, width, src, spitch, width, rows, hipMemcpyHostToDevice,
                                ctx->stream));
    return PSGC_OK;
  }
  // chunks of whole rows (a row wider than a chunk is split flat when dense)
  const bool dense = spitch == width;
  const size_t rows_per = std::max<size_t>(1, kPinChunk / width);
  if (!dense && width > kPinChunk) {
    PSGC_HIP(hipMemcpy2DAsync(dst, width, src, spitch, width, rows, hipMemcpyHostToDevice,
                              ctx->stream));
    return PSGC_OK;
  }
  size_t done = 0;
  for (int c = 0; done < total; ++c) {
    const int k = c & 1;
    size_t len;
    PSGC_TRY(pin_wait(ctx, k));
    if (dense) {
      len = std::min(kPinChunk, total - done);
      par_copy2d((char*)ctx->pin[k], len, (const char*)src + done, len, len, 1);
    } else {
      const size_t r0 = done / width, nr = std::min(rows_per, rows - r0);
      len = nr * width;
      par_copy2d((char*)ctx->pin[k], width, (const char*)src + r0 * spitch, spitch, width, nr);
    }
    PSGC_HIP(hipMemcpyAsync((char*)dst + done, ctx->pin[k], len, hipMemcpyHostToDevice,
                            ctx->stream));
    PSGC_HIP(hipEventRecord(ctx->pin_ev[k], ctx->stream));
    ctx->pin_busy[k] = true;
    done += len;
  }
  return PSGC_OK;
}

int host_to_device(psgc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return host_to_device_2d(ctx, dst, src, bytes, bytes, 1);
}

// Device block to host memory, after the work before it on the context's
// stream; returns when the data is in the caller's memory.
int device_to_host(psgc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return PSGC_OK;
  if (!pinned_ready(ctx)) {
    PSGC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PSGC_HIP(hipStreamSynchronize(ctx->stream));
    return PSGC_OK;
  }
  const int nch = (int)((bytes + kPinChunk - 1) / kPinChunk);
  auto issue = [&](int c) -> int {
    const int k = c & 1;
    const size_t off = (size_t)c * kPinChunk, len = std::min(kPinChunk, bytes - off);
    PSGC_TRY(pin_wait(ctx, k));
    PSGC_HIP(hipMemcpyAsync(ctx->pin[k], (const char*)src + off, len, hipMemcpyDeviceToHost,
                            ctx->stream));
    PSGC_HIP(hipEventRecord(ctx->pin_ev[k], ctx->stream));
    ctx->pin_busy[k] = true;
    return PSGC_OK;
  };
  PSGC_TRY(issue(0));
  for (int c = 0; c < nch; ++c) {
    // the next chunk's DMA runs while this one is copied out
    if (c + 1 < nch) PSGC_TRY(issue(c + 1));
    const int k = c & 1;
    PSGC_TRY(pin_wait(ctx, k));
    const size_t off = (size_t)c * kPinChunk, len = std::min(kPinChunk, bytes - off);
    par_copy2d((char*)dst + off, len, (const char*)ctx->pin[k], len, len, 1);
  }
  return PSGC_OK;
}

}  // namespace psgc
