/*
 * regions.h -- counting by table region (plain jobs, k <= 63).
 *
 * The table is 2^rbits regions of rspan slots (one region fills a build
 * workgroup's LDS); a k-mer's probe window never leaves the region its
 * hash's top rbits bits pick (see Probe in table.h).  Counting then needs no
 * HBM atomics on the table:
 *
 *   1. k_parse_records: one read per lane, k-mer and reverse complement
 *      rolled along the read; each instance becomes a record (8 bytes,
 *      (hash bits 0..58) << 5 | biedge index; k > 31 adds the k-mer's high
 *      word) in one of 32 buckets picked by hash bits 59..63, so a record's
 *      top bits are its region inside the bucket.  A workgroup stages
 *      records per bucket in LDS and appends runs of >= FLUSH with one
 *      atomic per chunk.
 *   2. two multisplit passes (partition.h) group each bucket's records by
 *      region and note where every region's records start.
 *   3. k_region_build: a workgroup takes regions one by one; per region it
 *      loads the slots into LDS (or starts from EMPTY when the table holds
 *      nothing yet), adds the region's records with LDS atomics, and writes
 *      the region back with coalesced stores.  Entries beyond a slot's NSUB
 *      sub-entries go to the HBM side table exactly as in the atomic count.
 *
 * The table contents equal those of the atomic count up to the order of
 * keys inside a probe run, which no result depends on.
 */
#ifndef BRUNO_REGIONS_H_
#define BRUNO_REGIONS_H_

#include "kernels.h"
#include "partition.h"

namespace br {

constexpr int REC_BUCKET_BITS = 5;
constexpr int REC_BUCKETS = 1 << REC_BUCKET_BITS;
/* Region sizes: 2560 slots (80 KB of LDS, two 512-thread builds per CU) or
 * 1280 (40 KB, four 256-thread builds per CU); the region_slots knob picks
 * one when a table is allocated. */
constexpr uint32_t REGION_MAX_SLOTS = 2560;
constexpr uint32_t REGION_SMALL_SLOTS = 1280;
/* the same LDS for 64-byte slots (k > 31) */
constexpr uint32_t REGION_WIDE_SLOTS = 1280;
constexpr uint32_t REGION_WIDE_SMALL_SLOTS = 640;
constexpr uint64_t REGION_BUILD_CU_SLOTS = 512; /* build workgroups resident on the device (x2 for small) */
/* record parser (k <= 31): workgroup size, positions between appends,
 * staged records per bucket, append threshold (16-byte records, k > 31:
 * a quarter of the positions, half the stage and half the threshold) */
constexpr int REC_BS = 512, REC_PF = 8, REC_STAGE = 192, REC_FLUSH = 32;
constexpr int REGION_ITEMS = 8;              /* records per thread loaded at once by the build */

/* biedge byte (nibbles 0, 1, 2, 4, 8 per side) -> index into BIEDGE25 */
__device__ __forceinline__ uint32_t biedge_index(uint8_t e) {
  const uint32_t l = e >> 4, r = e & 15;
  const uint32_t il = l ? 1 + __builtin_ctz(l) : 0, ir = r ? 1 + __builtin_ctz(r) : 0;
  return il * 5 + ir;
}

/* Records of a bucket fill chunks of `chunk` slots taken from the front of
 * the bucket's area with one atomic each (a workgroup keeps one open chunk
 * per bucket and pads it with skipped records at the end); records that
 * find their workgroup's stage full (reads longer than the stage) are
 * appended one by one from the back. */
struct RecArgs {
  void* recs;                /* REC_BUCKETS x bcap records (RecOf<KT>::T) */
  uint64_t bcap;
  unsigned long long* front; /* per bucket: records in chunks taken (> bcap - back: overflow) */
  unsigned long long* back;  /* per bucket: records appended from the back */
  unsigned long long* stat;  /* [0] instances, [1] dropped reads */
  uint32_t chunk;            /* power of two, >= REC_STAGE */
};

/* Upper bound of the instances of a batch of reads (reads with non-ACGT
 * bases included). */
__global__ __launch_bounds__(BLOCK) void k_count_instances(ParseArgs a, unsigned long long* out) {
  unsigned long long n = 0;
  for (uint64_t r = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; r < a.n_reads; r += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t off = a.starts ? a.starts[r] : a.offsets[r];
    const uint64_t L = (a.starts ? a.ends[r] : a.offsets[r + 1]) - off;
    if (L >= (uint64_t)a.K + 2) n += L - a.K + 1;
  }
  block_add(n, out);
}

/* A k-mer instance's record: the key word (hash << 5) | biedge index, and
 * for k > 31 the k-mer's high word (with the hash it gives the low word
 * back, see kmer_unhash_lo). */
template <typename KT> __device__ __forceinline__ typename RecOf<KT>::T make_rec(uint64_t key, KT km);
template <> __device__ __forceinline__ uint64_t make_rec<uint64_t>(uint64_t key, uint64_t) { return key; }
template <> __device__ __forceinline__ ulonglong2 make_rec<u128>(uint64_t key, u128 km) {
  return make_ulonglong2(key, (uint64_t)(km >> 64));
}

template <typename R>
__device__ __forceinline__ void rec_put(const RecArgs& ra, uint32_t b, uint64_t at, const R& rec) {
  if (at < ra.bcap) ((R*)ra.recs)[(uint64_t)b * ra.bcap + at] = rec;
}

/* chunk padding: a record the first multisplit pass drops (code 31) */
template <typename R> __device__ __forceinline__ R rec_filler();
template <> __device__ __forceinline__ uint64_t rec_filler<uint64_t>() { return REC_CODE_FILL; }
template <> __device__ __forceinline__ ulonglong2 rec_filler<ulonglong2>() { return make_ulonglong2(REC_CODE_FILL, 0); }

/* Parse into bucketed instance records, one read per lane.  A lane rolls
 * its read's k-mer and reverse complement along the read one base at a
 * time (a shift and an or each), so an instance costs its hash and its
 * record, not a window extraction and a reverse complement.  The lanes
 * step through their reads in aligned 16-byte blocks, all lanes of the
 * workgroup at the same block index (bases before a read's start or past
 * its end are masked off), loading the next block while the current one
 * is parsed; every PF positions the workgroup appends its staged records
 * (about BS x PF / 32 per bucket), so the loop runs to the longest read of
 * the workgroup's BS reads.  A read with a non-ACGT base is dropped whole
 * (NFilterSequencesIterator, compact_dbg_freq_construct_clean.cpp:210-213),
 * found by a first pass over its blocks. */
template <typename KT, int BS, int PF, int STAGE, int FLUSH>
__global__ __launch_bounds__(BS) void k_parse_records(ParseArgs a, RecArgs ra) {
  typedef typename RecOf<KT>::T R;
  __shared__ R stg[REC_BUCKETS][STAGE];
  __shared__ uint32_t cnt[REC_BUCKETS];
  __shared__ unsigned long long ch_base[REC_BUCKETS]; /* open chunk (~0: none) */
  __shared__ uint32_t ch_fill[REC_BUCKETS];
  __shared__ uint32_t fl_n1[REC_BUCKETS], fl_n2[REC_BUCKETS];
  __shared__ unsigned long long fl_d1[REC_BUCKETS], fl_d2[REC_BUCKETS];
  __shared__ uint32_t blocks_max;
  constexpr int WPB = BS / 64;
  static_assert(16 % PF == 0, "flush points split a block evenly");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int K = a.K;
  const uint32_t C = ra.chunk;
  kops<KT> ops(K);
  unsigned long long my_inst = 0, my_drop = 0;
  if (threadIdx.x < REC_BUCKETS) {
    cnt[threadIdx.x] = 0;
    ch_base[threadIdx.x] = ~0ULL;
    ch_fill[threadIdx.x] = 0;
  }
  __syncthreads();
  /* append every bucket holding >= at staged records to its open chunk
   * (the rest to a new chunk), then empty them */
  auto flush = [&](uint32_t at) {
    if (threadIdx.x < REC_BUCKETS) {
      const uint32_t t = threadIdx.x;
      const uint32_t c = cnt[t];
      uint32_t n = c < (uint32_t)STAGE ? c : (uint32_t)STAGE;
      if (n < at) n = 0;
      const unsigned long long base = ch_base[t];
      const uint32_t fill = ch_fill[t];
      const uint32_t room = base == ~0ULL ? 0u : C - fill;
      const uint32_t m1 = n < room ? n : room, m2 = n - m1;
      fl_n1[t] = m1;
      fl_d1[t] = base + fill;
      fl_n2[t] = m2;
      if (m2) {
        const unsigned long long nb = atomicAdd(&ra.front[t], (unsigned long long)C);
        fl_d2[t] = nb;
        ch_base[t] = nb;
        ch_fill[t] = m2;
      } else {
        ch_fill[t] = fill + m1;
      }
      if (n) cnt[t] = 0;
    }
    lds_barrier();
    for (int b = w; b < REC_BUCKETS; b += WPB) {
      const uint32_t m1 = fl_n1[b], m2 = fl_n2[b];
      const unsigned long long d1 = fl_d1[b], d2 = fl_d2[b];
      for (uint32_t i = lane; i < m1 + m2; i += 64) rec_put(ra, b, i < m1 ? d1 + i : d2 + (i - m1), stg[b][i]);
    }
    lds_barrier();
  };
  const int sh = 2 * (K - 1);
  const KT kmask = ops.mask;
  for (uint64_t r0 = (uint64_t)blockIdx.x * BS; r0 < a.n_reads; r0 += (uint64_t)gridDim.x * BS) {
    const uint64_t r = r0 + threadIdx.x;
    uint64_t lo = 0, L = 0;
    if (r < a.n_reads) {
      lo = a.starts ? a.starts[r] : a.offsets[r];
      L = (a.starts ? a.ends[r] : a.offsets[r + 1]) - lo;
    }
    const uint8_t* s = a.bases + lo;
    const uint4* blk = (const uint4*)((uintptr_t)s & ~(uintptr_t)15);
    const int64_t mis = (int64_t)((uintptr_t)s & 15);
    /* the blocks holding the read; an aligned block that holds a byte of
     * the read never crosses a page the read does not touch */
    const uint64_t nblk = L ? (uint64_t)(mis + (int64_t)L + 15) / 16 : 0;
    bool bad = false;
    for (uint64_t t = 0; t < nblk; ++t) {
      const uint4 v = blk[t];
      const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int64_t j = (int64_t)(16 * t) + q - mis;
        if (j >= 0 && j < (int64_t)L) bad |= base_code((uint8_t)(wv[q >> 2] >> (8 * (q & 3)))) < 0;
      }
    }
    if (bad) ++my_drop;
    /* reads shorter than the parser window k + 2 give nothing
     * (debruijn_biedge_loader.hpp:464-465) */
    const uint64_t Lp = (bad || L < (uint64_t)K + 2) ? 0 : L;
    const uint64_t nb = Lp ? (uint64_t)(mis + (int64_t)Lp + 16) / 16 : 0; /* through position Lp */
    /* every thread has read the last round's maximum before it is reset */
    __syncthreads();
    if (threadIdx.x == 0) blocks_max = 0;
    __syncthreads();
    if (nb) atomicMax(&blocks_max, (uint32_t)nb);
    __syncthreads();
    const uint32_t nbw = blocks_max;
    /* fw: the last k bases read, rv its reverse complement; gone: the base
     * that left fw at the previous step.  At position j >= k the instance
     * starting at j - k is complete: its out base is base j (none at
     * j == L) and its in base the one that left at step j - 1 (none for
     * the first instance). */
    KT fw = 0, rv = 0;
    uint32_t gone = 0;
    uint4 cur = nb ? blk[0] : make_uint4(0, 0, 0, 0);
    for (uint32_t t = 0; t < nbw; ++t) {
      const uint4 nxt = (t + 1 < nb && t + 1 < nblk) ? blk[t + 1] : make_uint4(0, 0, 0, 0);
      const uint32_t wv[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int64_t j = (int64_t)(16 * t) + q - mis;
        if (j >= 0 && j <= (int64_t)Lp) {
          const uint32_t x = ((wv[q >> 2] >> (8 * (q & 3))) & 0xFFu) | 0x20u;
          const uint32_t c = ((x >> 1) ^ (x >> 2)) & 3u; /* A/C/G/T, either case */
          const bool in_read = j < (int64_t)Lp;
          if (j >= K) {
            const bool flip = !(fw < rv);
            const KT km = flip ? rv : fw;
            const uint32_t has_in = j > K, has_out = in_read;
            const uint32_t il = !has_in ? 0u : flip ? 0u : 1u + gone;
            const uint32_t ir = !has_out ? 0u : flip ? 0u : 1u + c;
            /* biedge_rc: the in side becomes the complement of the out
             * base and the other way round */
            const uint32_t il2 = flip && has_out ? 4u - c : il;
            const uint32_t ir2 = flip && has_in ? 4u - gone : ir;
            const uint64_t h = kmer_hash(km);
            const uint32_t b = (uint32_t)(h >> (64 - REC_BUCKET_BITS));
            const uint64_t key = (h << REC_BUCKET_BITS) | (il2 * 5u + ir2);
            const R rec = make_rec<KT>(key, km);
            const uint32_t pos = atomicAdd(&cnt[b], 1u);
            if (pos < (uint32_t)STAGE) {
              stg[b][pos] = rec;
            } else {
              /* stage full: one record from the back */
              const unsigned long long k = atomicAdd(&ra.back[b], 1ULL);
              if (k < ra.bcap) rec_put(ra, b, ra.bcap - 1 - k, rec);
            }
            ++my_inst;
          }
          if (in_read) {
            gone = (uint32_t)(fw >> sh) & 3u;
            fw = ((fw << 2) | (KT)c) & kmask;
            rv = (rv >> 2) | ((KT)(3u - c) << sh);
          }
        }
        if (q % PF == PF - 1) {
          lds_barrier();
          flush(FLUSH);
        }
      }
      cur = nxt;
    }
  }
  flush(1);
  /* pad the open chunks */
  for (int b = w; b < REC_BUCKETS; b += WPB) {
    const unsigned long long base = ch_base[b];
    if (base == ~0ULL) continue;
    for (uint32_t i = ch_fill[b] + lane; i < C; i += 64) rec_put(ra, b, base + i, rec_filler<R>());
  }
  const unsigned long long si = wave_sum(my_inst), sd = wave_sum(my_drop);
  if (lane == 0) {
    if (si) atomicAdd(&ra.stat[0], si);
    if (sd) atomicAdd(&ra.stat[1], sd);
  }
}

struct RegionArgs {
  void* table;               /* Slot<KT> */
  uint64_t span;             /* slots per region */
  uint32_t rbits;            /* region bits (the bucket's 5 and rb below them) */
  uint64_t n_regions;        /* 2^rbits */
  uint64_t q0, q1;           /* this launch: regions [q0, q1) */
  const void* recs;          /* records (RecOf<KT>::T) grouped by region */
  const uint64_t* region_start; /* [n_regions + 1]: region q holds recs[region_start[q], region_start[q + 1]) */
  int valid;                 /* the table holds earlier counts (else start from EMPTY) */
  Spill* spill;
  uint64_t spcap;
  /* side-table entries are queued, not added inline: wave w of the grid
   * appends (slot << 5 | biedge index) to spq[w * spq_cap ...] and leaves
   * its count in spq_n[w]; k_spill_flush adds them after the build (an
   * inline add is a chain of dependent HBM atomics that the whole
   * workgroup waits for at the next barrier) */
  uint64_t* spq;
  uint32_t* spq_n;
  uint32_t spq_cap;
  Counters* ctr;
  /* filter pre-pass (optional): per region `stride` words of candidate
   * bits (occupied, no spilled entries, k-mer count >= t1lo when
   * thresholding) and of spilled-slot bits, bit i of word w = slot 64 w + i */
  uint64_t* cbits;
  uint64_t* sbits;
  uint64_t stride;
  int thresholding;
  uint64_t t1lo;
  /* the filter fused into the build (fuse != 0, with the bits): every
   * candidate slot gets local_insert_by_freqencies and its branch / chain
   * class in LDS before the write-back, and its node record; the candidate
   * bits then keep only what a wave could not place (a full node list).
   * List positions come in chunks of FUSE_CHUNK per wave; wave w leaves
   * its open chunks' unused positions marked LIST_HOLE and described in
   * holes[2 w + list] = start << 7 | count. */
  int fuse; /* the kernel's FUSE parameter */
  int K;
  Thresh th;
  NodeLists nl;
  uint64_t* holes;
};

constexpr uint32_t FUSE_CHUNK = 64;
constexpr uint64_t LIST_HOLE = ~0ULL;

__device__ __forceinline__ uint64_t lds_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

/* A region in LDS, one row per slot word (row 0 the key / meta word, then
 * the two-word key's words, then the sub-entries): a wave's lanes touch
 * random slots, and with 32- or 64-byte slots side by side every key word
 * would sit in one of a few banks; in rows, consecutive slots' words are
 * consecutive and random slots spread over all banks.  Set-up and
 * write-back move the region's bytes in table order, 16 per lane, so every
 * load or store instruction covers whole cache lines (lane i takes piece
 * i % P of slot i / P, P = 16-byte pieces per slot). */
template <typename KT, uint32_t SLOTS>
struct LdsRegion {
  static constexpr int W = sizeof(Slot<KT>) / 8;
  static constexpr int P = W / 2;
  static constexpr int SUB0 = sizeof(KT) == 8 ? 1 : 3; /* first sub-entry row */
  uint64_t w[W][SLOTS];
  __device__ __forceinline__ uint64_t* flag(uint32_t s) { return &w[0][s]; }
  __device__ __forceinline__ uint64_t* sub(int j, uint32_t s) { return &w[SUB0 + j][s]; }
  /* slots [0, span) from / to their table copy t, BS threads */
  template <int BS>
  __device__ __forceinline__ void load(uint32_t span, const Slot<KT>* t) {
    const ulonglong2* v = (const ulonglong2*)t;
    for (uint32_t i = threadIdx.x; i < span * P; i += BS) {
      const ulonglong2 x = v[i];
      const uint32_t s = i / P, q = i % P;
      w[2 * q][s] = x.x;
      w[2 * q + 1][s] = x.y;
    }
  }
  template <int BS>
  __device__ __forceinline__ void clear(uint32_t span) {
    for (uint32_t s = threadIdx.x; s < span; s += BS)
#pragma unroll
      for (int q = 0; q < W; ++q) w[q][s] = EMPTY64;
  }
  template <int BS>
  __device__ __forceinline__ void store(uint32_t span, Slot<KT>* t) const {
    typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
    u64x2* v = (u64x2*)t;
    for (uint32_t i = threadIdx.x; i < span * P; i += BS) {
      const uint32_t s = i / P, q = i % P;
      u64x2 x;
      x.x = w[2 * q][s];
      x.y = w[2 * q + 1][s];
      /* streaming: the region is not read again during the count */
      __builtin_nontemporal_store(x, v + i);
    }
  }
};

/* Find or claim k-mer v in the LDS region L of span slots from home slot
 * s; returns the slot, or ~0u when the region is full.  *claimed as in
 * SlotOps::find_or_insert. */
template <uint32_t SLOTS>
__device__ __forceinline__ uint32_t lds_claim(LdsRegion<uint64_t, SLOTS>& L, uint32_t span, uint32_t s, uint64_t v,
                                              int* claimed) {
  for (uint32_t probe = 0; probe < span; ++probe) {
    uint64_t k0 = lds_load(L.flag(s));
    if (k0 == EMPTY64) {
      const uint64_t old = atomicCAS((unsigned long long*)L.flag(s), EMPTY64, (unsigned long long)v);
      if (old == EMPTY64) {
        *claimed = 1;
        return s;
      }
      k0 = old;
    }
    if ((k0 & KEY_BITS) == v) return s;
    s = (s + 1 == span) ? 0 : s + 1;
  }
  return ~0u;
}
/* Two-word keys: meta is claimed by CAS (EMPTY -> LOCK), the words are
 * stored and meta published with a release store, all in one block ahead
 * of any wait, so a lane of the claiming wave never waits for its own
 * wave; readers wait (bounded) while meta is LOCK. */
template <uint32_t SLOTS>
__device__ __forceinline__ uint32_t lds_claim(LdsRegion<u128, SLOTS>& L, uint32_t span, uint32_t s, u128 v,
                                              int* claimed) {
  const uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
  for (uint32_t probe = 0; probe < span; ++probe) {
    uint64_t m = __hip_atomic_load(L.flag(s), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (m == EMPTY64) {
      const uint64_t old = atomicCAS((unsigned long long*)L.flag(s), EMPTY64, (unsigned long long)LOCK64);
      if (old == EMPTY64) {
        L.w[1][s] = lo;
        L.w[2][s] = hi;
        __hip_atomic_store(L.flag(s), (uint64_t)0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        *claimed = 1;
        return s;
      }
      m = old;
    }
    for (uint32_t spin = 0; m == LOCK64 && spin < MAX_SPIN; ++spin)
      m = __hip_atomic_load(L.flag(s), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (m == LOCK64) return ~0u;
    if (L.w[1][s] == lo && L.w[2][s] == hi) return s;
    s = (s + 1 == span) ? 0 : s + 1;
  }
  return ~0u;
}

/* add_k2 on LDS slot s; gs is the slot's table index (side-table key).
 * Returns 1 for a new k2 entry in the slot, 0 for a count added, and
 * LDS_SPILL when the entry belongs in the side table (the slot is marked
 * as having spilled entries; the caller queues the entry). */
constexpr int LDS_SPILL = 2;
template <typename KT, uint32_t SLOTS>
__device__ __forceinline__ int lds_add_k2(LdsRegion<KT, SLOTS>& L, uint32_t s, uint8_t e) {
  const uint64_t code = (uint64_t)e + 1;
#pragma unroll
  for (int j = 0; j < Slot<KT>::NSUB; ++j) {
    uint64_t* q = L.sub(j, s);
    uint64_t v = lds_load(q);
    if (v == EMPTY64) {
      const uint64_t old = atomicCAS((unsigned long long*)q, EMPTY64, (unsigned long long)((code << 48) | 1));
      if (old == EMPTY64) return 1;
      v = old;
    }
    if ((v >> 48) == code) {
      atomicAdd((unsigned long long*)q, 1ULL);
      return 0;
    }
  }
  uint64_t* fw = L.flag(s);
  if (!(lds_load(fw) & F_SPILL)) atomicOr((unsigned long long*)fw, (unsigned long long)F_SPILL);
  return LDS_SPILL;
}

/* an empty use of a register: the compiler waits for its load here */
__device__ __forceinline__ void consume(uint64_t v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void consume(const ulonglong2& v) {
  asm volatile("" ::"v"(v.x));
  asm volatile("" ::"v"(v.y));
}

/* the k-mer of a record of region q */
__device__ __forceinline__ uint64_t rec_hash(const RegionArgs& g, uint64_t q, uint64_t key) {
  /* the bucket bits (the region's top 5 bits), then the record's */
  return ((q >> (g.rbits - REC_BUCKET_BITS)) << (64 - REC_BUCKET_BITS)) | (key >> REC_BUCKET_BITS);
}
__device__ __forceinline__ uint64_t rec_kmer(uint64_t h, uint64_t) { return kmer_unhash(h); }
__device__ __forceinline__ u128 rec_kmer(uint64_t h, const ulonglong2& r) { return kmer_unhash(h, r.y); }

/* The biedge byte of biedge index i = il * 5 + ir (nibble 0 for side index
 * 0, else 1 << (index - 1)), computed: a table lookup would be a global
 * load per record on the build's critical path. */
__device__ __forceinline__ uint32_t biedge_of_index(uint32_t i) {
  const uint32_t il = (i * 52u) >> 8, ir = i - 5u * il; /* i / 5 for i < 25 */
  return (((1u << il) >> 1) << 4) | ((1u << ir) >> 1);
}

/* Add one record of region q (table slots from base) to the LDS region.
 * Returns the side-table entry to queue (slot << 5 | biedge index), or
 * NO_SPILL. */
constexpr uint64_t NO_SPILL = ~0ULL;
template <typename KT, uint32_t SLOTS, typename R>
__device__ __forceinline__ uint64_t region_add(const RegionArgs& g, LdsRegion<KT, SLOTS>& L, uint32_t span,
                                               uint64_t q, uint64_t base, const R& rec, unsigned long long& my_new,
                                               unsigned long long& my_k2) {
  const uint64_t key = rec_key(rec);
  const uint32_t code = (uint32_t)(key & 31);
  const uint64_t h = rec_hash(g, q, key);
  const KT km = rec_kmer(h, rec);
  int claimed = 0;
  const uint32_t s = lds_claim(L, span, (uint32_t)__umul64hi(h << g.rbits, g.span), km, &claimed);
  if (s == ~0u) {
    atomicOr(&g.ctr->err, ERR_TABLE_FULL);
    return NO_SPILL;
  }
  my_new += claimed;
  const int r = lds_add_k2<KT, SLOTS>(L, s, (uint8_t)biedge_of_index(code));
  if (r == LDS_SPILL) return ((base + s) << 5) | code;
  my_k2 += r;
  return NO_SPILL;
}

/* Queue side-table entries from converged code: one ballot per record
 * slot of the tile; qn (this wave's queue fill) stays wave-uniform.  A full
 * queue adds the entry inline. */
__device__ __forceinline__ void spill_queue(const RegionArgs& g, uint64_t* q, uint32_t& qn, uint64_t e,
                                            unsigned long long& my_k2) {
  const uint64_t m = __ballot(e != NO_SPILL);
  if (!m) return;
  const uint32_t lane = __lane_id();
  const uint32_t at = qn + (uint32_t)__popcll(m & ((1ULL << lane) - 1));
  if (e != NO_SPILL) {
    if (at < g.spq_cap) {
      q[at] = e;
    } else {
      const uint64_t bq = biedge_of_index((uint32_t)(e & 31)) + 1;
      my_k2 += spill_add(g.spill, g.spcap, e >> 5, bq, &g.ctr->err);
    }
  }
  qn += (uint32_t)__popcll(m);
}

/* ------------------------------------------------------------------ */
/* the filter fused into the build                                      */
/* ------------------------------------------------------------------ */
template <uint32_t SLOTS>
__device__ __forceinline__ uint64_t lds_kmer(LdsRegion<uint64_t, SLOTS>& L, uint32_t s) {
  return *L.flag(s) & KEY_BITS;
}
template <uint32_t SLOTS>
__device__ __forceinline__ u128 lds_kmer(LdsRegion<u128, SLOTS>& L, uint32_t s) {
  return ((u128)L.w[2][s] << 64) | L.w[1][s];
}

/* The filter pre-pass's choice for LDS slot i: a candidate (occupied, no
 * side-table entries, not yet a node, k-mer count >= t1lo when
 * thresholding), or *spilled for a slot with side-table entries. */
template <typename KT, uint32_t SLOTS>
__device__ __forceinline__ bool lds_candidate(const RegionArgs& g, LdsRegion<KT, SLOTS>& L, uint32_t i,
                                              bool& spilled) {
  const uint64_t fw = *L.flag(i);
  spilled = fw != EMPTY64 && (fw & F_SPILL);
  if (fw == EMPTY64 || (fw & (F_SPILL | F_PRESENT))) return false;
  if (!g.thresholding) return true;
  uint64_t tot = 0;
#pragma unroll
  for (int j = 0; j < Slot<KT>::NSUB; ++j) {
    const uint64_t v = *L.sub(j, i);
    tot += v != EMPTY64 ? (v & 0xFFFFFFFFFFFFULL) : 0;
  }
  return tot >= g.t1lo;
}

/* A wave's open chunk of a node list: positions [base, base + size), of
 * which the first `used` are given out.  Wave-uniform. */
struct ListChunk {
  uint64_t base;
  uint32_t used;
};

/* Consecutive list positions for the lanes with take set, called by the
 * whole wave: from the open chunk, then from a new chunk of `size` >= 64
 * (one atomic) when it runs out. */
__device__ __forceinline__ uint64_t chunk_take(ListChunk& ch, bool take, unsigned long long* counter,
                                               uint32_t size = FUSE_CHUNK) {
  const uint64_t m = __ballot(take);
  const uint32_t n = (uint32_t)__popcll(m);
  const uint32_t lane = __lane_id();
  const uint32_t rank = (uint32_t)__popcll(m & ((1ULL << lane) - 1));
  const uint32_t room = size - ch.used;
  uint64_t pos = ch.base + ch.used + rank;
  if (n > room) {
    uint64_t nb = 0;
    if (lane == 0) nb = atomicAdd(counter, (unsigned long long)size);
    nb = readlane64(nb, 0);
    if (rank >= room) pos = nb + (rank - room);
    ch.base = nb;
    ch.used = n - room;
  } else {
    ch.used += n;
  }
  return pos;
}

/* The fused filter for queued slots of the region whose table slots start
 * at base (lane l holds LDS slot qs when active): the slot's k2 entries
 * from LDS, local_insert_by_freqencies (debruijn_graph_map.hpp:1351-1609;
 * the k+1-mer edge filter, or none, as the job asks), IsBranchPoint /
 * IsChainNode (debruijn_graph_filters.hpp:142-220), then the node record
 * with its list position and F_PRESENT written into the LDS slot and the
 * slot appended to its list.  Called by the whole wave; true (for the
 * wave) when a node found its list full: that slot stays a candidate for
 * the filter pass. */
template <typename KT, uint32_t SLOTS>
__device__ __forceinline__ bool fuse_nodes(const RegionArgs& g, LdsRegion<KT, SLOTS>& L, uint64_t base, uint32_t qs,
                                           bool active, ListChunk (&ch)[2]) {
  constexpr int N = Slot<KT>::NSUB;
  const kops<KT> ops(g.K);
  bool kept = false;
  NodeRec rec;
  rec.cls = 0;
  if (active) {
    K2L<N> Lk;
    Lk.n = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const uint64_t v = *L.sub(j, qs);
      const bool ok = v != EMPTY64;
      Lk.e[j] = ok ? (uint8_t)((v >> 48) - 1) : 0;
      Lk.c[j] = ok ? (uint32_t)(v & 0xFFFFFFFFFFFFULL) : 0;
      Lk.n += ok;
    }
    const KT k = lds_kmer(L, qs);
    const KmerFacts f = kmer_facts(ops, k);
    uint16_t c[9];
    if (g.thresholding == 2) {
      kept = edge_counts<KT, N>(ops, k, Lk, g.th, c);
    } else if (g.thresholding) {
      kept = p2_counts_slot<KT, N>(ops, f, Lk, g.th, c);
    } else {
      p1_counts(Lk, c);
      kept = true;
    }
    if (kept) rec = make_node_facts<KT>(ops, f, c);
  }
  const bool chn = kept && rec.cls == CLS_CHAIN, brn = kept && rec.cls == CLS_BRANCH;
  const uint64_t pc = chunk_take(ch[0], chn, &g.ctr->n_chain);
  const uint64_t pb = chunk_take(ch[1], brn, &g.ctr->n_branch);
  bool full = false;
  if (kept) {
    const uint64_t pos = chn ? pc : pb;
    if (pos < g.nl.cap && pos <= 0x7FFFFFFFULL) {
      rec.idx = (uint32_t)pos;
      uint64_t w[3];
      __builtin_memcpy(w, &rec, sizeof(rec));
#pragma unroll
      for (int j = 0; j < 3; ++j) *L.sub(j, qs) = w[j];
      *L.flag(qs) |= F_PRESENT;
      (chn ? g.nl.chain_slot : g.nl.branch_slot)[pos] = base + qs;
    } else {
      full = true;
    }
  }
  return __ballot(full) != 0;
}

/* Side-table entries the region builds queued, one workgroup per queue. */
__global__ __launch_bounds__(BLOCK) void k_spill_flush(const uint64_t* spq, const uint32_t* spq_n, uint32_t spq_cap,
                                                        Spill* spill, uint64_t spcap, Counters* ctr) {
  const uint32_t n = min(spq_n[blockIdx.x], spq_cap);
  const uint64_t* q = spq + (uint64_t)blockIdx.x * spq_cap;
  unsigned long long mine = 0;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
    const uint64_t e = q[i];
    mine += spill_add(spill, spcap, e >> 5, biedge_of_index((uint32_t)(e & 31)) + 1, &ctr->err);
  }
  block_add(mine, &ctr->n_k2);
}

/* Build regions [q0, q1), workgroup i taking an even share [qa, qb) of them.
 * Per region: set up the slots in LDS (EMPTY, or the table's slots when an
 * earlier count wrote them; a region without records is then left as it
 * is), add the region's records (REGION_ITEMS per thread at a time), write
 * the filter pre-pass bits when asked, and write the region back (16-byte
 * stores, a slot per lane).  The next region's first records are loaded
 * before this region's records are added, so their latency hides behind
 * the LDS work; barriers order LDS only, so one region's write-back
 * overlaps the next region's set-up. */
template <typename KT, uint32_t SLOTS, int BS, bool FUSE>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_region_build(RegionArgs g) {
  typedef typename RecOf<KT>::T R;
  /* records per thread per tile: a 64-byte-slot region holds half the
   * slots, so half as many (and the registers stay without spills) */
  constexpr int REGION_ITEMS = sizeof(R) == 8 ? br::REGION_ITEMS : br::REGION_ITEMS / 2;
  __shared__ LdsRegion<KT, SLOTS> L;
  const uint32_t span = (uint32_t)g.span;
  const uint64_t nreg = g.n_regions;
  const uint64_t nlaunch = g.q1 - g.q0;
  const uint64_t qa = g.q0 + (uint64_t)blockIdx.x * nlaunch / gridDim.x,
                 qb = g.q0 + ((uint64_t)blockIdx.x + 1) * nlaunch / gridDim.x;
  const R* recs = (const R*)g.recs;
  unsigned long long my_new = 0, my_k2 = 0;
  constexpr uint64_t TILE = (uint64_t)REGION_ITEMS * BS;
  /* region_start[q] for the workgroup's regions, 64 at a time: lane l of
   * every wave holds entry rs_base + l and the next 64 are loaded ahead,
   * so a region's bounds are two lane reads, not a dependent load */
  const int lane = threadIdx.x & 63;
  uint64_t rs_base = qa;
  auto rs_load = [&](uint64_t b) -> uint64_t { return b + lane <= nreg ? g.region_start[b + lane] : 0; };
  uint64_t rs = rs_load(qa), rs_nx = rs_load(qa + 64);
  auto rstart = [&](uint64_t q) -> uint64_t { /* q never decreases */
    while (q >= rs_base + 64) {
      rs = rs_nx;
      rs_base += 64;
      rs_nx = rs_load(rs_base + 64);
    }
    return readlane64(rs, (int)(q - rs_base));
  };
  /* this wave's side-table queue */
  const uint64_t wave_id = (uint64_t)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  uint64_t* spq = g.spq + wave_id * g.spq_cap;
  uint32_t qn = 0;
  /* this wave's open chunks of the chain and branch lists (fused filter) */
  ListChunk ch[2] = {{0, FUSE_CHUNK}, {0, FUSE_CHUNK}};
  /* first tile of the next region with records, loaded ahead */
  R rn[REGION_ITEMS];
  uint64_t nq = qa, nlo = 0, nhi = 0;
  auto ahead = [&](uint64_t from) {
    nq = from;
    for (; nq < qb; ++nq) {
      nlo = rstart(nq);
      nhi = rstart(nq + 1);
      if (!(g.valid && nlo == nhi)) break; /* no records: the region stays as it is */
    }
    if (nq < qb) {
#pragma unroll
      for (int k = 0; k < REGION_ITEMS; ++k) {
        const uint64_t i = nlo + (uint64_t)k * BS + threadIdx.x;
        rn[k] = i < nhi ? recs[i] : rec_zero<R>();
      }
    }
  };
  ahead(qa);
  while (nq < qb) {
    const uint64_t q = nq, lo = nlo, hi = nhi;
    R rr[REGION_ITEMS];
#pragma unroll
    for (int k = 0; k < REGION_ITEMS; ++k) rr[k] = rn[k];
    const uint64_t base = q * g.span;
    Slot<KT>* T = (Slot<KT>*)g.table + base;
    if (g.valid) {
      L.template load<BS>(span, T);
    } else {
      L.template clear<BS>(span);
    }
    ahead(q + 1);
    lds_barrier();
    /* the first tile, loaded ahead; a region longer than one tile loads
     * the rest here (in registers of their own, so the first tile's adds
     * wait for nothing) */
#pragma unroll
    for (int k = 0; k < REGION_ITEMS; ++k) {
      uint64_t e = NO_SPILL;
      if (lo + (uint64_t)k * BS + threadIdx.x < hi) e = region_add<KT, SLOTS, R>(g, L, span, q, base, rr[k], my_new, my_k2);
      spill_queue(g, spq, qn, e, my_k2);
    }
    for (uint64_t i0 = lo + TILE; i0 < hi; i0 += TILE) {
      R rt[REGION_ITEMS];
#pragma unroll
      for (int k = 0; k < REGION_ITEMS; ++k) {
        const uint64_t i = i0 + (uint64_t)k * BS + threadIdx.x;
        rt[k] = i < hi ? recs[i] : rec_zero<R>();
      }
#pragma unroll
      for (int k = 0; k < REGION_ITEMS; ++k) {
        uint64_t e = NO_SPILL;
        if (i0 + (uint64_t)k * BS + threadIdx.x < hi)
          e = region_add<KT, SLOTS, R>(g, L, span, q, base, rt[k], my_new, my_k2);
        spill_queue(g, spq, qn, e, my_k2);
      }
    }
    /* the next region's records (loaded since this region began) are
     * waited for here, before the write-back's stores: stores and loads
     * share one in-order counter, so waiting for them after the stores
     * would wait for the stores too */
#pragma unroll
    for (int k = 0; k < REGION_ITEMS; ++k) consume(rn[k]);
    lds_barrier();
    bool left = false; /* a candidate of this wave's slots found its list full */
    if constexpr (FUSE) {
      /* each wave queues the candidates of its slots (lane l holds the
       * l-th) and filters them 64 at a time, so every lane of a pass has
       * a slot */
      const uint32_t lane = __lane_id();
      uint32_t fs = 0, fn = 0; /* this lane's queued slot, the queue's fill */
      for (uint32_t i0 = 0; i0 < span; i0 += BS) {
        const uint32_t i = i0 + threadIdx.x;
        bool sp = false;
        const bool c = i < span && lds_candidate(g, L, i, sp);
        const uint64_t m = __ballot(c);
        const uint32_t n = (uint32_t)__popcll(m);
        if (n == 0) continue;
        const uint32_t rank = (uint32_t)__popcll(m & ((1ULL << lane) - 1));
        /* a permutation of the lanes: the candidates to lanes fn + rank,
         * the other lanes after them (mod 64) */
        const uint32_t to = (c ? fn + rank : fn + n + lane - rank) & 63;
        const uint32_t v = (uint32_t)__builtin_amdgcn_ds_permute((int)(to << 2), (int)i);
        if (fn + n < 64) {
          if (((lane - fn) & 63) < n) fs = v;
          fn += n;
        } else {
          if (lane >= fn) fs = v;
          left |= fuse_nodes<KT, SLOTS>(g, L, base, fs, true, ch);
          fs = v; /* lanes below the new fill hold the rest */
          fn = fn + n - 64;
        }
      }
      if (fn) left |= fuse_nodes<KT, SLOTS>(g, L, base, fs, lane < fn, ch);
    }
    if (g.cbits) {
      /* the filter scan's choice for every slot of the region, one
       * ballot per 64 slots; with the filter fused, only what a full
       * list left */
      for (uint32_t i0 = 0; i0 < span; i0 += BS) {
        const uint32_t i = i0 + threadIdx.x;
        bool c = false, sp = false;
        if (i < span) c = lds_candidate(g, L, i, sp);
        if (FUSE && !left) c = false;
        const uint64_t bc = __ballot(c), bsp = __ballot(sp);
        const uint32_t w0 = i0 + (threadIdx.x & ~63u);
        if ((threadIdx.x & 63) == 0 && w0 < span) {
          g.cbits[q * g.stride + w0 / 64] = bc;
          g.sbits[q * g.stride + w0 / 64] = bsp;
        }
      }
    }
    if (FUSE) lds_barrier(); /* the node records are in LDS */
    L.template store<BS>(span, T);
    lds_barrier();
  }
  const unsigned long long sn = wave_sum(my_new), sk = wave_sum(my_k2);
  if ((threadIdx.x & 63) == 0) {
    if (sn) atomicAdd(&g.ctr->n_distinct, sn);
    if (sk) atomicAdd(&g.ctr->n_k2, sk);
    g.spq_n[wave_id] = qn;
  }
  if constexpr (FUSE) {
    /* the open chunks' unused positions */
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const uint64_t start = ch[l].base + ch[l].used;
      const uint32_t cnt = FUSE_CHUNK - ch[l].used;
      if (lane < cnt && start + lane < g.nl.cap) (l ? g.nl.branch_slot : g.nl.chain_slot)[start + lane] = LIST_HOLE;
      if (lane == 0) g.holes[2 * wave_id + l] = (start << 7) | cnt;
    }
  }
}

}  // namespace br

#endif
