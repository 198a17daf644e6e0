/*
 * pool.h -- per-device memory pool for every allocation of the library.
 *
 * hipMalloc / hipFree of large HBM blocks cost tens of milliseconds per
 * gigabyte and stall unpredictably; a job allocates its table and several
 * per-phase lists on every run.  The pool takes memory from HIP in chunks
 * (one large chunk when the caller reserves up front) and hands out
 * 256-byte-aligned blocks by best fit, coalescing neighbours on release, so
 * a steady-state run makes no HIP allocator calls.  Fully free chunks go
 * back to HIP only on trim, or when a new chunk does not fit otherwise.
 *
 * A block is released to the pool only once the work that used it is
 * finished: callers synchronise the stream that last touched the block
 * before they free it.  All methods are thread-safe.
 */
#ifndef BRUNO_POOL_H_
#define BRUNO_POOL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace br {

class DevicePool {
 public:
  static constexpr uint64_t ALIGN = 256;
  static constexpr uint64_t GROW_MIN = 1ULL << 30;  /* chunk for small requests */
  static constexpr uint64_t CHUNK_ROUND = 2ULL << 20;

  static DevicePool& current() {
    static DevicePool pools[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
      (void)hipGetLastError();
      dev = 0;
    }
    return pools[dev];
  }

  hipError_t alloc(void** out, uint64_t bytes) {
    const uint64_t need = round_up(std::max<uint64_t>(bytes, 1), ALIGN);
    std::lock_guard<std::mutex> g(mu_);
    if (take(out, need)) return hipSuccess;
    uint64_t want = round_up(std::max(need, GROW_MIN), CHUNK_ROUND);
    hipError_t e = add_chunk(want);
    if (e != hipSuccess && want > need) e = add_chunk(round_up(need, CHUNK_ROUND));
    if (e != hipSuccess) {
      trim_locked();
      e = add_chunk(round_up(need, CHUNK_ROUND));
    }
    if (e != hipSuccess) {
      *out = nullptr;
      return e;
    }
    return take(out, need) ? hipSuccess : hipErrorOutOfMemory;
  }

  /* p must come from alloc and its users must be finished */
  void release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    auto it = used_.find((char*)p);
    if (it == used_.end()) return;
    Blk b = it->second;
    char* a = it->first;
    used_.erase(it);
    in_use_ -= b.size;
    /* coalesce with the following and the preceding free block of the same chunk */
    auto nx = free_.lower_bound(a);
    if (nx != free_.end() && nx->first == a + b.size && nx->second.chunk == b.chunk) {
      b.size += nx->second.size;
      free_.erase(nx);
    }
    auto pv = free_.lower_bound(a);
    if (pv != free_.begin()) {
      --pv;
      if (pv->first + pv->second.size == a && pv->second.chunk == b.chunk) {
        pv->second.size += b.size;
        return;
      }
    }
    free_[a] = b;
  }

  /* make sure one free block of at least `bytes` exists */
  hipError_t reserve(uint64_t bytes) {
    const uint64_t need = round_up(std::max<uint64_t>(bytes, 1), CHUNK_ROUND);
    std::lock_guard<std::mutex> g(mu_);
    for (auto& f : free_)
      if (f.second.size >= need) return hipSuccess;
    hipError_t e = add_chunk(need);
    if (e != hipSuccess) {
      trim_locked();
      e = add_chunk(need);
    }
    return e;
  }

  void trim() {
    std::lock_guard<std::mutex> g(mu_);
    trim_locked();
  }

  void stats(uint64_t* reserved, uint64_t* in_use, uint64_t* largest_free) {
    std::lock_guard<std::mutex> g(mu_);
    uint64_t lf = 0;
    for (auto& f : free_) lf = std::max(lf, f.second.size);
    if (reserved) *reserved = reserved_;
    if (in_use) *in_use = in_use_;
    if (largest_free) *largest_free = lf;
  }

 private:
  struct Blk {
    uint64_t size;
    int chunk;
  };
  struct Chunk {
    char* base;
    uint64_t size;
  };

  static uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

  bool take(void** out, uint64_t need) {
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second.size >= need && (best == free_.end() || it->second.size < best->second.size)) best = it;
    if (best == free_.end()) return false;
    char* a = best->first;
    Blk b = best->second;
    free_.erase(best);
    if (b.size > need) free_[a + need] = Blk{b.size - need, b.chunk};
    used_[a] = Blk{need, b.chunk};
    in_use_ += need;
    *out = a;
    return true;
  }

  hipError_t add_chunk(uint64_t size) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, size);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return e;
    }
    int id = next_chunk_++;
    chunks_[id] = Chunk{(char*)p, size};
    free_[(char*)p] = Blk{size, id};
    reserved_ += size;
    return hipSuccess;
  }

  void trim_locked() {
    for (auto it = chunks_.begin(); it != chunks_.end();) {
      auto f = free_.find(it->second.base);
      if (f != free_.end() && f->second.chunk == it->first && f->second.size == it->second.size) {
        free_.erase(f);
        (void)hipFree(it->second.base);
        reserved_ -= it->second.size;
        it = chunks_.erase(it);
      } else {
        ++it;
      }
    }
  }

  std::mutex mu_;
  std::map<int, Chunk> chunks_;
  std::map<char*, Blk> free_;
  std::unordered_map<char*, Blk> used_;
  int next_chunk_ = 0;
  uint64_t reserved_ = 0, in_use_ = 0;
};

}  // namespace br

#endif
