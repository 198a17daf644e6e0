/*
 * fastq.hip -- FASTQ records to concatenated reads, on the device.
 *
 * Replaces the reader under parse_file_data (compact_dbg_io.hpp:165-217: the
 * file bytes go to kmerind's FASTQ parser, one sequence per 4-line record,
 * compact_dbg_freq_construct_clean.cpp:210-219).  The file's bytes are in
 * device memory; the output is what bruno_parse_count takes: the sequence
 * lines back to back and n_reads + 1 byte offsets.  Reads with a non-ACGT
 * base are not removed here: the parse kernel drops them, as
 * NFilterSequencesIterator does.
 *
 * Format accepted: records of exactly four lines (@header, sequence,
 * +[header], quality); '\n' or "\r\n" line ends; the last line may lack its
 * newline; blank lines after the last record are ignored.  Anything else
 * (a header not starting with '@', a separator not starting with '+', text
 * after the last complete record) is BRUNO_BAD_INPUT.
 *
 * Three passes: newlines per 16 KB chunk; a prefix sum gives every chunk the
 * number of the first line it contains; each chunk, staged in LDS, is walked
 * by 256 threads over 64-byte runs (a block scan of the per-run newline
 * counts numbers every line), which records where each sequence line starts
 * and ends and checks the record structure.  A prefix sum of the sequence
 * lengths gives the offsets, and one wave per read copies its bases.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/bruno_hip.h"
#include "pool.h"

namespace {

constexpr int FQ_BLOCK = 256;
constexpr uint64_t FQ_RUN = 64;                     /* bytes per thread */
constexpr uint64_t FQ_CHUNK = FQ_BLOCK * FQ_RUN;    /* bytes per workgroup */

/* number of '\n' bytes in a 32-bit word: x has a zero byte exactly where w
 * has a newline; the high bit of ((x & 7F..) + 7F..) | x is clear exactly
 * for zero bytes */
__device__ __forceinline__ uint32_t newlines4(uint32_t w) {
  const uint32_t x = w ^ 0x0A0A0A0Au;
  const uint32_t t = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
  return (uint32_t)__popc(t);
}

/* whole 16-byte vectors when the chunk is aligned and complete, bytes
 * otherwise */
__global__ __launch_bounds__(FQ_BLOCK) void k_fq_count(const uint8_t* s, uint64_t n, unsigned long long* nl) {
  const uint64_t c0 = (uint64_t)blockIdx.x * FQ_CHUNK;
  const uint64_t len = n - c0 < FQ_CHUNK ? n - c0 : FQ_CHUNK;
  unsigned long long c = 0;
  if (len == FQ_CHUNK && (((uintptr_t)(s + c0)) & 15) == 0) {
    const uint4* v = (const uint4*)(s + c0);
    for (uint64_t i = threadIdx.x; i < FQ_CHUNK / 16; i += FQ_BLOCK) {
      const uint4 q = v[i];
      c += newlines4(q.x) + newlines4(q.y) + newlines4(q.z) + newlines4(q.w);
    }
  } else {
    for (uint64_t i = threadIdx.x; i < len; i += FQ_BLOCK) c += (s[c0 + i] == '\n');
  }
  typedef hipcub::BlockReduce<unsigned long long, FQ_BLOCK> BR;
  __shared__ typename BR::TempStorage tmp;
  unsigned long long t = BR(tmp).Sum(c);
  if (threadIdx.x == 0) nl[blockIdx.x] = t;
}

struct FqOut {
  uint64_t* seq_start;
  uint64_t* seq_end;
  uint64_t n_records;   /* complete records R */
  uint64_t n_lines;     /* lines T, the last one possibly without newline */
  uint32_t* err;
};

/* A line ends at p (a newline, or the end of the file for the last line);
 * line is its number. */
__device__ __forceinline__ void fq_line_end(const uint8_t* s, uint64_t n, uint64_t p, uint64_t line, const FqOut& o) {
  const uint64_t R = o.n_records;
  if (line < 4 * R) {
    const uint64_t r = line >> 2;
    const uint32_t m = (uint32_t)(line & 3);
    if (m == 0) o.seq_start[r] = p + 1;
    if (m == 1) o.seq_end[r] = (p > 0 && s[p - 1] == '\r') ? p - 1 : p;
  }
  /* the line after this one must start as its role says */
  const uint64_t next = line + 1;
  if (next < 4 * R && p + 1 < n) {
    const uint8_t c = s[p + 1];
    if ((next & 3) == 0 && c != '@') atomicOr(o.err, 1u);
    if ((next & 3) == 2 && c != '+') atomicOr(o.err, 1u);
  }
}

/* Each thread walks its 64-byte run from registers: four 16-byte loads
 * (neighbouring threads' runs share cache lines, so the chunk is read from
 * memory once), a SWAR newline count, a block scan for the line number of
 * the run's first byte, then one pass over the 64 bytes. */
__global__ __launch_bounds__(FQ_BLOCK) void k_fq_mark(const uint8_t* s, uint64_t n, const unsigned long long* base,
                                                      FqOut o) {
  typedef hipcub::BlockScan<unsigned long long, FQ_BLOCK> BS;
  __shared__ typename BS::TempStorage tmp;
  const uint64_t c0 = (uint64_t)blockIdx.x * FQ_CHUNK;
  const uint64_t len = n - c0 < FQ_CHUNK ? n - c0 : FQ_CHUNK;
  const uint64_t r0 = (uint64_t)threadIdx.x * FQ_RUN;
  const uint64_t r1 = r0 + FQ_RUN < len ? r0 + FQ_RUN : len;
  uint32_t w[FQ_RUN / 4];
  const bool whole = r1 - r0 == FQ_RUN && r0 < r1 && (((uintptr_t)(s + c0 + r0)) & 15) == 0;
  if (whole) {
    const uint4* v = (const uint4*)(s + c0 + r0);
#pragma unroll
    for (int q = 0; q < (int)(FQ_RUN / 16); ++q) {
      const uint4 x = v[q];
      w[4 * q] = x.x;
      w[4 * q + 1] = x.y;
      w[4 * q + 2] = x.z;
      w[4 * q + 3] = x.w;
    }
  } else {
    /* partial or unaligned run: bytes, padded with a non-newline */
#pragma unroll
    for (int q = 0; q < (int)(FQ_RUN / 4); ++q) {
      uint32_t x = 0;
      for (int b = 0; b < 4; ++b) {
        const uint64_t i = r0 + 4 * q + b;
        const uint32_t c = i < r1 ? s[c0 + i] : 0u;
        x |= c << (8 * b);
      }
      w[q] = x;
    }
  }
  unsigned long long cnt = 0;
#pragma unroll
  for (int q = 0; q < (int)(FQ_RUN / 4); ++q) cnt += newlines4(w[q]);
  unsigned long long pre = 0;
  BS(tmp).ExclusiveSum(cnt, pre);
  uint64_t line = base[blockIdx.x] + pre;
  const uint64_t R = o.n_records;
  const uint64_t nb = r1 > r0 ? r1 - r0 : 0;
#pragma unroll
  for (int q = 0; q < (int)(FQ_RUN / 4); ++q) {
    for (int b = 0; b < 4; ++b) {
      const uint64_t i = 4 * (uint64_t)q + b;
      if (i >= nb) break;
      const uint8_t c = (uint8_t)(w[q] >> (8 * b));
      if (line >= 4 * R && c != '\n' && c != '\r' && c != ' ') atomicOr(o.err, 2u); /* text after the last record */
      if (c == '\n') {
        fq_line_end(s, n, c0 + r0 + i, line, o);
        ++line;
      }
    }
  }
  /* the last line without a newline ends at the end of the file */
  if (nb && c0 + r1 == n && s[n - 1] != '\n') fq_line_end(s, n, n, line, o);
  if (blockIdx.x == 0 && threadIdx.x == 0 && R > 0 && s[0] != '@') atomicOr(o.err, 1u);
}

__global__ void k_fq_len(const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t* len) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) len[i] = b[i] >= a[i] ? b[i] - a[i] : 0;
}

/* one wave per read */
__global__ __launch_bounds__(256) void k_fq_copy(const uint8_t* s, const uint64_t* start, const uint64_t* off,
                                                 uint64_t n_reads, uint8_t* out) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t r = wave; r < n_reads; r += nw) {
    const uint8_t* src = s + start[r];
    uint8_t* dst = out + off[r];
    const uint64_t L = off[r + 1] - off[r];
    for (uint64_t i = lane; i < L; i += 64) dst[i] = src[i];
  }
}

/* out[0] = 0, out[i+1] = in[0] + ... + in[i], with the scan's temporary
 * storage from mem */
template <typename Mem>
hipError_t scan_u64(const uint64_t* in, uint64_t* out_plus1, uint64_t n, hipStream_t st, Mem& mem) {
  hipError_t e = hipMemsetAsync(out_plus1, 0, 8, st);
  if (e != hipSuccess || n == 0) return e;
  size_t tmp = 0;
  e = hipcub::DeviceScan::InclusiveSum(nullptr, tmp, in, out_plus1 + 1, (int)n, st);
  if (e != hipSuccess) return e;
  uint8_t* t = nullptr;
  if (!mem.get(&t, tmp + 16)) return hipErrorOutOfMemory;
  return hipcub::DeviceScan::InclusiveSum(t, tmp, in, out_plus1 + 1, (int)n, st);
}

/* Scratch of one call from the memory pool, released when the call returns
 * (after its stream has finished with it). */
struct Blocks {
  std::vector<void*> v;
  hipStream_t st;
  static constexpr bruno_status SHORT = BRUNO_OOM;
  explicit Blocks(hipStream_t s) : st(s) {}
  ~Blocks() {
    if (!v.empty()) (void)hipStreamSynchronize(st);
    for (void* p : v) br::DevicePool::current().release(p);
  }
  template <typename T>
  bool get(T** p, uint64_t bytes) {
    void* q = nullptr;
    if (br::DevicePool::current().alloc(&q, std::max<uint64_t>(bytes, 16)) != hipSuccess) return false;
    v.push_back(q);
    *p = (T*)q;
    return true;
  }
  bool reads_fit(uint64_t) const { return true; }
};

/* Scratch of the allocation-free split: consecutive 256-byte-aligned pieces
 * of one fixed device array (the caller's device's copy of `scratch`).
 * Nothing is allocated or released. */
constexpr uint64_t FQ_NOALLOC_MAX_BYTES = 1ULL << 30;   /* FASTQ bytes per call */
constexpr uint64_t FQ_NOALLOC_MAX_READS = 1ULL << 21;   /* records per call     */
constexpr uint64_t FQ_SCRATCH_BYTES = 80ULL << 20;
static_assert(FQ_SCRATCH_BYTES >= 2 * (FQ_NOALLOC_MAX_BYTES / FQ_CHUNK + 2) * 8 + 256 + 4 * (FQ_NOALLOC_MAX_READS + 1) * 8 +
                                      (8ULL << 20),
              "scratch too small for its stated limits");
struct Bump {
  uint8_t* base;
  uint64_t cap, used = 0;
  static constexpr bruno_status SHORT = BRUNO_CAPACITY;
  Bump(void* b, uint64_t c) : base((uint8_t*)b), cap(c) {}
  template <typename T>
  bool get(T** p, uint64_t bytes) {
    const uint64_t need = (std::max<uint64_t>(bytes, 16) + 255) & ~255ULL;
    if (need > cap - used) return false;
    *p = (T*)(base + used);
    used += need;
    return true;
  }
  bool reads_fit(uint64_t R) const { return R <= FQ_NOALLOC_MAX_READS; }
};

/* One pass.  out_bases / out_offsets are either given (with capacities) or
 * obtained from alloc(bytes) once the sizes are known. */
template <typename Mem, typename Alloc>
bruno_status fastq_split_impl(Mem& mem, const uint8_t* s, uint64_t bytes, char** out_bases, uint64_t** out_offs,
                              uint64_t cap_reads, uint64_t cap_bases, bool have_out, bool copy, Alloc alloc,
                              uint64_t* n_reads, uint64_t* n_bases, hipStream_t st,
                              uint64_t** out_start = nullptr, uint64_t** out_end = nullptr) {
  *n_reads = *n_bases = 0;
  if (bytes == 0) {
    if (copy && !have_out) {
      if (!(*out_offs = (uint64_t*)alloc(8)) || !(*out_bases = (char*)alloc(16))) return BRUNO_OOM;
    }
    if (*out_offs && hipMemsetAsync(*out_offs, 0, 8, st) != hipSuccess) return BRUNO_INTERNAL;
    return hipStreamSynchronize(st) == hipSuccess ? BRUNO_OK : BRUNO_INTERNAL;
  }
  const uint64_t nchunks = (bytes + FQ_CHUNK - 1) / FQ_CHUNK;
  unsigned long long *nl = nullptr, *base = nullptr;
  uint32_t* err = nullptr;
  if (!mem.get(&nl, nchunks * 8) || !mem.get(&base, (nchunks + 1) * 8) || !mem.get(&err, 4)) return Mem::SHORT;
  hipLaunchKernelGGL(k_fq_count, dim3((unsigned)nchunks), dim3(FQ_BLOCK), 0, st, s, bytes, nl);
  if (scan_u64((const uint64_t*)nl, (uint64_t*)base, nchunks, st, mem) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  unsigned long long total_nl = 0;
  uint8_t last = 0;
  if (hipMemcpyAsync(&total_nl, base + nchunks, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&last, s + bytes - 1, 1, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemsetAsync(err, 0, 4, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  const uint64_t T = total_nl + (last != '\n' ? 1 : 0);
  const uint64_t R = T / 4;
  if (!mem.reads_fit(R)) return Mem::SHORT;
  FqOut o{};
  o.n_records = R;
  o.n_lines = T;
  o.err = err;
  uint64_t* lens = nullptr;
  if (out_start) {
    /* ranges only: the sequence lines stay where they are in the file */
    if (!(o.seq_start = (uint64_t*)alloc(std::max<uint64_t>(1, R) * 8)) ||
        !(o.seq_end = (uint64_t*)alloc(std::max<uint64_t>(1, R) * 8)))
      return BRUNO_OOM;
    *out_start = o.seq_start;
    *out_end = o.seq_end;
  } else if (!mem.get(&o.seq_start, R * 8) || !mem.get(&o.seq_end, R * 8)) {
    return Mem::SHORT;
  }
  if (!mem.get(&lens, R * 8)) return Mem::SHORT;
  uint64_t* offs = nullptr;
  if (have_out) {
    if (R > cap_reads) return BRUNO_CAPACITY;
    offs = *out_offs;
  } else if (copy) {
    if (!(offs = (uint64_t*)alloc((R + 1) * 8))) return BRUNO_OOM;
    *out_offs = offs;
  } else if (!mem.get(&offs, (R + 1) * 8)) {
    return Mem::SHORT;
  }
  hipLaunchKernelGGL(k_fq_mark, dim3((unsigned)nchunks), dim3(FQ_BLOCK), 0, st, s, bytes,
                     (const unsigned long long*)base, o);
  if (R)
    hipLaunchKernelGGL(k_fq_len, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, st, o.seq_start, o.seq_end, R, lens);
  if (scan_u64(lens, offs, R, st, mem) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  uint32_t herr = 0;
  uint64_t nb = 0;
  if (hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&nb, offs + R, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  if (herr) return BRUNO_BAD_INPUT;
  *n_reads = R;
  *n_bases = nb;
  if (!copy) return BRUNO_OK;
  char* ob = nullptr;
  if (have_out) {
    if (nb > cap_bases) return BRUNO_CAPACITY;
    ob = *out_bases;
  } else {
    if (!(ob = (char*)alloc(nb + 1))) return BRUNO_OOM;
    *out_bases = ob;
  }
  if (R) {
    const uint64_t waves = std::min<uint64_t>(R, 65536);
    hipLaunchKernelGGL(k_fq_copy, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, s, o.seq_start, offs, R,
                       (uint8_t*)ob);
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  return BRUNO_OK;
}

}  // namespace

extern "C" bruno_status bruno_fastq_split(const char* d_fastq, uint64_t bytes, char* d_out_bases,
                                          uint64_t* d_out_offsets, uint64_t* n_reads, uint64_t* n_bases,
                                          void* stream) {
  if (!n_reads || !n_bases || (bytes && !d_fastq)) return BRUNO_BAD_INPUT;
  if ((d_out_bases == nullptr) != (d_out_offsets == nullptr)) return BRUNO_BAD_INPUT;
  const bool have = d_out_bases != nullptr;
  char* ob = d_out_bases;
  uint64_t* oo = d_out_offsets;
  const uint64_t cr = *n_reads, cb = *n_bases;
  Blocks mem((hipStream_t)stream);
  return fastq_split_impl(mem, (const uint8_t*)d_fastq, bytes, &ob, &oo, cr, cb, have, have,
                          [](uint64_t) -> void* { return nullptr; }, n_reads, n_bases, (hipStream_t)stream);
}

/* The device's copy of the allocation-free split's scratch (every device
 * that loads this library holds one until the process ends). */
static __device__ __align__(256) uint8_t scratch[FQ_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

/* Address of the current device's scratch, looked up once per device with
 * that device current (a host-side use of the array's name is not a device
 * address). */
bruno_status scratch_of_current_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return BRUNO_NO_DEVICE;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return BRUNO_INTERNAL;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return BRUNO_OK;
}
}  // namespace

extern "C" bruno_status bruno_fastq_split_noalloc(const char* d_fastq, uint64_t bytes, char* d_out_bases,
                                                  uint64_t* d_out_offsets, uint64_t* n_reads, uint64_t* n_bases,
                                                  void* stream) {
  if (!n_reads || !n_bases || (bytes && !d_fastq)) return BRUNO_BAD_INPUT;
  if ((d_out_bases == nullptr) != (d_out_offsets == nullptr)) return BRUNO_BAD_INPUT;
  if (bytes > FQ_NOALLOC_MAX_BYTES) return BRUNO_CAPACITY;
  void* sp = nullptr;
  bruno_status s = scratch_of_current_device(&sp);
  if (s != BRUNO_OK) return s;
  const bool have = d_out_bases != nullptr;
  char* ob = d_out_bases;
  uint64_t* oo = d_out_offsets;
  const uint64_t cr = *n_reads, cb = *n_bases;
  Bump mem(sp, FQ_SCRATCH_BYTES);
  s = fastq_split_impl(mem, (const uint8_t*)d_fastq, bytes, &ob, &oo, cr, cb, have, have,
                       [](uint64_t) -> void* { return nullptr; }, n_reads, n_bases, (hipStream_t)stream);
  /* the next call may reuse the scratch at once, from any stream */
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    (void)hipGetLastError();
    if (s == BRUNO_OK) s = BRUNO_INTERNAL;
  }
  return s;
}

/* Record ranges only: start / end byte of every sequence line in the file,
 * in arrays from the caller's allocator (no copy; the parse kernel reads
 * the sequences where they are; used by bruno_parse_count_fastq). */
bruno_status bruno_fastq_ranges(const char* d_fastq, uint64_t bytes, void* (*alloc)(void*, uint64_t), void* ctx,
                                uint64_t** starts, uint64_t** ends, uint64_t* n_reads, uint64_t* n_bases,
                                void* stream) {
  if (!n_reads || !n_bases || !starts || !ends || !alloc || (bytes && !d_fastq)) return BRUNO_BAD_INPUT;
  *starts = *ends = nullptr;
  char* ob = nullptr;
  uint64_t* oo = nullptr;
  if (bytes == 0) {
    *n_reads = *n_bases = 0;
    return BRUNO_OK;
  }
  Blocks mem((hipStream_t)stream);
  return fastq_split_impl(mem, (const uint8_t*)d_fastq, bytes, &ob, &oo, 0, 0, false, false,
                          [&](uint64_t b) { return alloc(ctx, b); }, n_reads, n_bases, (hipStream_t)stream, starts,
                          ends);
}

/* Split with outputs allocated by the caller's allocator once their sizes
 * are known (one pass). */
bruno_status bruno_fastq_split_alloc(const char* d_fastq, uint64_t bytes, void* (*alloc)(void*, uint64_t), void* ctx,
                                     char** out_bases, uint64_t** out_offsets, uint64_t* n_reads, uint64_t* n_bases,
                                     void* stream) {
  if (!n_reads || !n_bases || !out_bases || !out_offsets || !alloc || (bytes && !d_fastq)) return BRUNO_BAD_INPUT;
  *out_bases = nullptr;
  *out_offsets = nullptr;
  Blocks mem((hipStream_t)stream);
  return fastq_split_impl(mem, (const uint8_t*)d_fastq, bytes, out_bases, out_offsets, 0, 0, false, true,
                          [&](uint64_t b) { return alloc(ctx, b); }, n_reads, n_bases, (hipStream_t)stream);
}
