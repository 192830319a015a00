// Host -> device upload of edge lists through pinned staging buffers.
//
// The caller's arrays are ordinary (pageable) host memory.  They are copied
// chunk by chunk into one of two pinned buffers by a few host threads, and
// each filled buffer is sent to the device with asynchronous copies on a copy
// stream; while the DMA engine drains one buffer the threads fill the other.
// A consumer of each chunk (the CSR build's canonicalisation) can run on the
// compute stream as soon as that chunk has arrived.  Pieces (one per graph of
// a batch) are packed back to back, so a batch of thousands of small graphs
// costs a few large copies instead of three small ones per graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.h"

namespace stc {

namespace {

#define STC_HIP_TRY(expr)            \
  do {                               \
    hipError_t e_ = (expr);          \
    if (e_ != hipSuccess) return e_; \
  } while (0)

// Lines [a, b) of the concatenated pieces (pre = prefix sums of piece sizes)
// to the staging arrays, where line `base` lands at index 0.
void copy_lines(const std::vector<EdgePiece> &pc, const std::vector<uint64_t> &pre, uint64_t a,
                uint64_t b, uint64_t base, uint32_t *hu, uint32_t *hv, int8_t *hs) {
  if (a >= b) return;
  size_t i = (size_t)(std::upper_bound(pre.begin(), pre.end(), a) - pre.begin()) - 1;
  for (uint64_t x = a; x < b; i++) {
    const uint64_t in = x - pre[i];
    const uint64_t take = std::min(b, pre[i + 1]) - x;
    if (take) {
      std::memcpy(hu + (x - base), pc[i].u + in, take * 4);
      std::memcpy(hv + (x - base), pc[i].v + in, take * 4);
      std::memcpy(hs + (x - base), pc[i].s + in, take);
    }
    x += take;
  }
}

}  // namespace

hipError_t UploadRing::ensure() {
  for (int b = 0; b < 2; b++) {
    // (no pre-touch: writing the fresh pinned pages once cost about 55 ms on an
    // MI355X host and saved 8 ms of the first upload)
    if (!host[b]) STC_HIP_TRY(hipHostMalloc(&host[b], kStageLines * 9, hipHostMallocDefault));
    if (!done[b]) STC_HIP_TRY(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
  }
  return hipSuccess;
}

void UploadRing::release() {
  for (int b = 0; b < 2; b++) {
    if (host[b]) (void)hipHostFree(host[b]);
    if (done[b]) (void)hipEventDestroy(done[b]);
    host[b] = nullptr;
    done[b] = nullptr;
  }
}

hipError_t UploadRing::upload(const std::vector<EdgePiece> &pc, uint32_t *du, uint32_t *dv,
                              int8_t *ds, hipStream_t st, hipStream_t ks, const ChunkFn &on_chunk) {
  std::vector<uint64_t> pre(pc.size() + 1, 0);
  for (size_t i = 0; i < pc.size(); i++) pre[i + 1] = pre[i] + pc[i].m;
  const uint64_t total = pre.back();
  if (total == 0) return hipSuccess;
  STC_HIP_TRY(ensure());
  const uint64_t nchunks = (total + kStageLines - 1) / kStageLines;
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned nthreads = std::max(1u, std::min(8u, hw ? hw : 1u));
  // chunk i goes to buffer i & 1; `go` is the last chunk the workers may
  // fill, `filled` counts (chunk, thread) parts done
  std::mutex mu;
  std::condition_variable cv;
  int64_t go = -1;
  uint64_t filled = 0;
  bool quit = false;
  auto part = [&](uint64_t i, unsigned t, uint64_t &a, uint64_t &b) {
    const uint64_t c0 = i * kStageLines, c1 = std::min(total, c0 + kStageLines);
    const uint64_t len = c1 - c0;
    a = c0 + len * t / nthreads;
    b = c0 + len * (t + 1) / nthreads;
  };
  auto worker = [&](unsigned t) {
    for (uint64_t i = 0; i < nchunks; i++) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return quit || go >= (int64_t)i; });
        if (quit) return;
      }
      uint64_t a, b;
      part(i, t, a, b);
      char *h = (char *)host[i & 1];
      copy_lines(pc, pre, a, b, i * kStageLines, (uint32_t *)h,
                 (uint32_t *)(h + kStageLines * 4), (int8_t *)(h + kStageLines * 8));
      {
        std::lock_guard<std::mutex> lk(mu);
        filled++;
      }
      cv.notify_all();
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthreads; t++) pool.emplace_back(worker, t);
  hipError_t e = hipSuccess;
  for (uint64_t i = 0; i < nchunks && e == hipSuccess; i++) {
    // the buffer's previous chunk (i - 2) must have left before it is refilled
    if (i >= 2) e = hipEventSynchronize(done[i & 1]);
    if (e != hipSuccess) break;
    {
      std::lock_guard<std::mutex> lk(mu);
      go = (int64_t)i;
    }
    cv.notify_all();
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return filled >= (i + 1) * nthreads; });
    }
    const uint64_t c0 = i * kStageLines, len = std::min(total, c0 + kStageLines) - c0;
    const char *h = (const char *)host[i & 1];
    e = hipMemcpyAsync(du + c0, h, len * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dv + c0, h + kStageLines * 4, len * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ds + c0, h + kStageLines * 8, len, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(done[i & 1], st);
    if (e == hipSuccess && on_chunk) {
      // the chunk's consumer runs on ks once its copies are done
      if (ks != st) e = hipStreamWaitEvent(ks, done[i & 1], 0);
      if (e == hipSuccess) e = on_chunk(c0, c0 + len);
    }
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    quit = true;
  }
  cv.notify_all();
  for (auto &th : pool) th.join();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

}  // namespace stc
