/*
 * regions.h -- counting by table region (k <= 31, plain jobs).
 *
 * The table is 2^rbits regions of rspan <= REGION_MAX_SLOTS slots; a k-mer's
 * probe window never leaves the region its hash's top rbits bits pick (see
 * Probe in table.h).  Counting then needs no HBM atomics on the table:
 *
 *   1. k_parse_records: the parser (same instances as k_parse) writes one
 *      8-byte record per instance, (hash bits 0..58) << 5 | biedge index,
 *      rotated left so the region bits below the bucket bits come lowest,
 *      into one of 32 buckets picked by hash bits 59..63.  A workgroup stages
 *      records per bucket in LDS and appends runs of >= 128 with one atomic.
 *   2. a radix sort of each bucket on its low (region) bits.
 *      (The region bits are put lowest because the library radix sort
 *      returned wrong orders for bit ranges starting above bit 0 on more
 *      than a block of keys.)
 *   3. k_region_build: a workgroup walks the sorted records of a run of
 *      regions; per region it loads the slots into LDS (or starts from EMPTY
 *      when the table holds nothing yet), adds the region's records with LDS
 *      atomics, and writes the region back with coalesced stores.  Entries beyond a slot's NSUB sub-entries go to
 *      the HBM side table exactly as in the atomic count.
 *
 * The table contents equal those of the atomic count up to the order of
 * keys inside a probe run, which no result depends on.
 */
#ifndef BRUNO_REGIONS_H_
#define BRUNO_REGIONS_H_

#include "kernels.h"

namespace br {

constexpr int REC_BUCKET_BITS = 5;
constexpr int REC_BUCKETS = 1 << REC_BUCKET_BITS;
/* Region sizes: 2560 slots (80 KB of LDS, two 512-thread builds per CU) or
 * 1280 (40 KB, four 256-thread builds per CU); the region_slots knob picks
 * one when a table is allocated. */
constexpr uint32_t REGION_MAX_SLOTS = 2560;
constexpr uint32_t REGION_SMALL_SLOTS = 1280;
constexpr uint64_t REGION_BUILD_CU_SLOTS = 512; /* build workgroups resident on the device (x2 for small) */
/* record parser: workgroup size, most reads per wave between appends (the
 * count picks fewer for longer reads, so a round's records fit the stage),
 * staged records per bucket, append threshold */
constexpr int REC_BS = 512, REC_RPW = 4, REC_STAGE = 192, REC_FLUSH = 32;
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
  uint64_t* recs;            /* REC_BUCKETS x bcap records */
  uint64_t bcap;
  unsigned long long* front; /* per bucket: records in chunks taken (> bcap - back: overflow) */
  unsigned long long* back;  /* per bucket: records appended from the back */
  unsigned long long* stat;  /* [0] instances, [1] dropped reads */
  uint32_t chunk;            /* power of two, >= REC_STAGE */
  int rot;                   /* region bits inside a bucket (1..59): the record rotation */
  uint32_t rpw;              /* reads per wave between appends (1..REC_RPW) */
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

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

__device__ __forceinline__ void rec_put(const RecArgs& ra, uint32_t b, uint64_t at, uint64_t rec) {
  if (at < ra.bcap) ra.recs[(uint64_t)b * ra.bcap + at] = rec;
}

/* a record the build skips (code 31), its region bits spread by position */
__device__ __forceinline__ uint64_t rec_filler(const RecArgs& ra, uint64_t at) {
  return rotl64((fmix64(at) << REC_BUCKET_BITS) | 31u, ra.rot);
}

/* Parse (k <= 31) into bucketed instance records.  The read loop is
 * workgroup-uniform (wave w of a workgroup takes read r0 + w) so staged
 * buckets can be appended between reads. */
template <int BS, int RPW, int STAGE, int FLUSH>
__global__ __launch_bounds__(BS) void k_parse_records(ParseArgs a, RecArgs ra) {
  __shared__ uint64_t stg[REC_BUCKETS][STAGE];
  __shared__ uint32_t cnt[REC_BUCKETS];
  __shared__ unsigned long long ch_base[REC_BUCKETS]; /* open chunk (~0: none) */
  __shared__ uint32_t ch_fill[REC_BUCKETS];
  __shared__ uint32_t fl_n1[REC_BUCKETS], fl_n2[REC_BUCKETS];
  __shared__ unsigned long long fl_d1[REC_BUCKETS], fl_d2[REC_BUCKETS];
  constexpr int WPB = BS / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int K = a.K;
  const uint32_t C = ra.chunk;
  kops<uint64_t> ops(K);
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
  const uint32_t rpw = ra.rpw;
  for (uint64_t r0 = (uint64_t)blockIdx.x * WPB * rpw; r0 < a.n_reads; r0 += (uint64_t)gridDim.x * WPB * rpw) {
    for (uint32_t jr = 0; jr < rpw; ++jr) {
    const uint64_t r = r0 + (uint64_t)jr * WPB + w;
    if (r < a.n_reads) {
      const uint64_t off = a.starts ? a.starts[r] : a.offsets[r];
      const uint64_t L = (a.starts ? a.ends[r] : a.offsets[r + 1]) - off;
      const uint8_t* s = a.bases + off;
      /* NFilterSequencesIterator (compact_dbg_freq_construct_clean.cpp:210-213) */
      int bad = 0;
      for (uint64_t i = lane; i < L; i += 64) bad |= (base_code(s[i]) < 0);
      if (__any(bad)) {
        if (lane == 0) ++my_drop;
      } else if (L >= (uint64_t)K + 2) {
        for_each_instance<uint64_t>(s, L, K, lane, ops, [&](uint64_t, uint64_t km, uint8_t e) {
          const uint64_t h = kmer_hash(km);
          const uint32_t b = (uint32_t)(h >> (64 - REC_BUCKET_BITS));
          const uint64_t rec = rotl64((h << REC_BUCKET_BITS) | biedge_index(e), ra.rot);
          const uint32_t pos = atomicAdd(&cnt[b], 1u);
          if (pos < (uint32_t)STAGE) {
            stg[b][pos] = rec;
          } else {
            /* stage full (a long read): one record from the back */
            const unsigned long long k = atomicAdd(&ra.back[b], 1ULL);
            if (k < ra.bcap) rec_put(ra, b, ra.bcap - 1 - k, rec);
          }
          ++my_inst;
        });
      }
    }
    }
    lds_barrier();
    flush(FLUSH);
  }
  flush(1);
  /* pad the open chunks */
  for (int b = w; b < REC_BUCKETS; b += WPB) {
    const unsigned long long base = ch_base[b];
    if (base == ~0ULL) continue;
    for (uint32_t i = ch_fill[b] + lane; i < C; i += 64) rec_put(ra, b, base + i, rec_filler(ra, base + i));
  }
  const unsigned long long si = wave_sum(my_inst), sd = wave_sum(my_drop);
  if (lane == 0) {
    if (si) atomicAdd(&ra.stat[0], si);
    if (sd) atomicAdd(&ra.stat[1], sd);
  }
}

struct RegionArgs {
  Slot<uint64_t>* table;
  uint64_t span;             /* slots per region */
  uint32_t rbits;
  uint32_t bucket;
  uint64_t n_regions;        /* regions per bucket */
  const uint64_t* recs;      /* the bucket's records sorted by region */
  uint64_t n;                /* records of the bucket */
  int valid;                 /* the table holds earlier counts (else start from EMPTY) */
  Spill* spill;
  uint64_t spcap;
  Counters* ctr;
  /* filter pre-pass (optional): per region `stride` words of candidate
   * bits (occupied, no spilled entries, k-mer count >= t1lo when
   * thresholding) and of spilled-slot bits, bit i of word w = slot 64 w + i */
  uint64_t* cbits;
  uint64_t* sbits;
  uint64_t stride;
  int thresholding;
  uint64_t t1lo;
};

__device__ __forceinline__ uint64_t lds_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

/* Find or claim k-mer v in the LDS region L of span slots from home slot s;
 * returns the slot, or ~0u when the region is full.  *claimed as in
 * SlotOps::find_or_insert. */
__device__ __forceinline__ uint32_t lds_claim(Slot<uint64_t>* L, uint32_t span, uint32_t s, uint64_t v,
                                              int* claimed) {
  for (uint32_t probe = 0; probe < span; ++probe) {
    uint64_t k0 = lds_load(&L[s].key);
    if (k0 == EMPTY64) {
      const uint64_t old = atomicCAS((unsigned long long*)&L[s].key, EMPTY64, (unsigned long long)v);
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

/* add_k2 on an LDS slot; gs is the slot's table index (side-table key) */
__device__ __forceinline__ int lds_add_k2(Slot<uint64_t>* p, uint8_t e, uint64_t gs, Spill* sp, uint64_t spcap,
                                          uint32_t* err) {
  const uint64_t code = (uint64_t)e + 1;
#pragma unroll
  for (int j = 0; j < Slot<uint64_t>::NSUB; ++j) {
    uint64_t v = lds_load(&p->sub[j]);
    if (v == EMPTY64) {
      const uint64_t old = atomicCAS((unsigned long long*)&p->sub[j], EMPTY64, (unsigned long long)((code << 48) | 1));
      if (old == EMPTY64) return 1;
      v = old;
    }
    if ((v >> 48) == code) {
      atomicAdd((unsigned long long*)&p->sub[j], 1ULL);
      return 0;
    }
  }
  if (!(lds_load(&p->key) & F_SPILL)) atomicOr((unsigned long long*)&p->key, (unsigned long long)F_SPILL);
  return spill_add(sp, spcap, (gs << 9) | code, err);
}

/* Add one record to the LDS region. */
__device__ __forceinline__ void region_add(const RegionArgs& g, Slot<uint64_t>* L, uint32_t span, int rot,
                                           uint64_t base, uint64_t rrec, unsigned long long& my_new,
                                           unsigned long long& my_k2) {
  const uint64_t rec = rotl64(rrec, 64 - rot);
  const uint32_t code = (uint32_t)(rec & 31);
  if (code >= 25) return;
  const uint64_t h = ((uint64_t)g.bucket << (64 - REC_BUCKET_BITS)) | (rec >> REC_BUCKET_BITS);
  const uint64_t km = kmer_unhash(h);
  int claimed = 0;
  const uint32_t s = lds_claim(L, span, (uint32_t)__umul64hi(h << g.rbits, g.span), km, &claimed);
  if (s == ~0u) {
    atomicOr(&g.ctr->err, ERR_TABLE_FULL);
    return;
  }
  my_new += claimed;
  my_k2 += lds_add_k2(L + s, BIEDGE25[code], base + s, g.spill, g.spcap, &g.ctr->err);
}

/* Workgroup-wide count of pred, through one LDS word the caller is not
 * using (the build's slot array between regions: two workgroups of 80 KB
 * fill a CU's LDS, so no other LDS is available). */
__device__ __forceinline__ uint32_t block_count(bool pred, uint32_t* word) {
  if (threadIdx.x == 0) *word = 0;
  lds_barrier();
  const uint64_t bal = __ballot(pred);
  if ((threadIdx.x & 63) == 0 && bal) atomicAdd(word, (uint32_t)__popcll(bal));
  lds_barrier();
  const uint32_t c = *word;
  lds_barrier();
  return c;
}

/* First index of the sorted records whose region (rec & m) is >= target,
 * found by the whole workgroup: each round samples one record per thread,
 * evenly
 * spaced records and keeps the gap holding the boundary. */
template <int BS>
__device__ __forceinline__ uint64_t block_lower_bound(const uint64_t* recs, uint64_t n, uint64_t m, uint64_t target,
                                                      uint32_t* word) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t step = (hi - lo + BS - 1) / BS;
    const uint64_t pt = lo + (uint64_t)threadIdx.x * step;
    const uint32_t c = block_count(pt < hi && (recs[pt] & m) < target, word);
    if (c == 0) return lo;
    const uint64_t last_less = lo + (uint64_t)(c - 1) * step, first_ge = lo + (uint64_t)c * step;
    lo = last_less + 1;
    if (first_ge < hi) hi = first_ge;
  }
  return lo;
}

/* Build regions [qa, qb) of one bucket in LDS, workgroup i taking an even
 * share of the bucket's regions.  Its records are contiguous in the sorted
 * bucket: it finds the first with a binary search, then walks tiles of
 * BS * REGION_ITEMS records (all loads of a tile in flight at once, the
 * next tile loading while one is processed),
 * closing a region when the records move past it.  Regions without
 * records are written EMPTY when the table holds nothing yet and are
 * skipped otherwise.  Barriers order LDS only, so the write-back of one
 * region overlaps the next. */
template <uint32_t SLOTS, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_region_build(RegionArgs g) {
  __shared__ Slot<uint64_t> L[SLOTS];
  const uint32_t span = (uint32_t)g.span;
  const uint32_t words = span * 4;
  const int rot = (int)(g.rbits - REC_BUCKET_BITS);
  const uint64_t nreg = g.n_regions, m = nreg - 1, n = g.n;
  const uint64_t qa = (uint64_t)blockIdx.x * nreg / gridDim.x, qb = ((uint64_t)blockIdx.x + 1) * nreg / gridDim.x;
  unsigned long long my_new = 0, my_k2 = 0;
  uint64_t* Lw = (uint64_t*)L;
  uint64_t q = qa;
  uint64_t p = qa < qb ? block_lower_bound<BS>(g.recs, n, m, qa, (uint32_t*)L) : n;
  bool open = false; /* LDS holds region q */
  /* tile t is processed while tile t + 1 loads */
  constexpr uint64_t TILE = (uint64_t)REGION_ITEMS * BS;
  uint64_t rr[REGION_ITEMS];
#pragma unroll
  for (int k = 0; k < REGION_ITEMS; ++k) {
    const uint64_t i = p + (uint64_t)k * BS + threadIdx.x;
    rr[k] = i < n ? g.recs[i] : ~0ULL;
  }
  uint64_t tile_end = p + TILE < n ? p + TILE : n;
  /* region of the tile's last record; past the bucket's end: none */
  uint64_t last = tile_end < n ? (g.recs[tile_end - 1] & m) : nreg;
  while (q < qb) {
    uint64_t rn[REGION_ITEMS];
#pragma unroll
    for (int k = 0; k < REGION_ITEMS; ++k) {
      const uint64_t i = tile_end + (uint64_t)k * BS + threadIdx.x;
      rn[k] = i < n ? g.recs[i] : ~0ULL;
    }
    const uint64_t next_end = tile_end + TILE < n ? tile_end + TILE : n;
    const uint64_t next_last = next_end < n ? (g.recs[next_end - 1] & m) : nreg;
    while (q < qb && q <= last) {
      if (!open) {
        if (g.valid) {
          bool has = false;
#pragma unroll
          for (int k = 0; k < REGION_ITEMS; ++k) has |= rr[k] != ~0ULL && (rr[k] & m) == q;
          if (!block_count(has, (uint32_t*)L)) { /* no records: the region stays as it is */
            ++q;
            continue;
          }
        }
        ulonglong2* T2 = (ulonglong2*)(g.table + (((uint64_t)g.bucket << rot) | q) * g.span);
        if (g.valid) {
          for (uint32_t i = threadIdx.x; i < words / 2; i += BS) ((ulonglong2*)Lw)[i] = T2[i];
        } else {
          for (uint32_t i = threadIdx.x; i < words; i += BS) Lw[i] = EMPTY64;
        }
        lds_barrier();
        open = true;
      }
      const uint64_t base = (((uint64_t)g.bucket << rot) | q) * g.span;
#pragma unroll
      for (int k = 0; k < REGION_ITEMS; ++k)
        if (rr[k] != ~0ULL && (rr[k] & m) == q) region_add(g, L, span, rot, base, rr[k], my_new, my_k2);
      if (q == last) break; /* the region continues in the next tile */
      lds_barrier();
      if (g.cbits) {
        /* the filter scan's choice for every slot of the region, one
         * ballot per 64 slots */
        const uint64_t region = ((uint64_t)g.bucket << rot) | q;
        for (uint32_t i0 = 0; i0 < span; i0 += BS) {
          const uint32_t i = i0 + threadIdx.x;
          bool c = false, sp = false;
          if (i < span) {
            const uint64_t key = L[i].key;
            if (key != EMPTY64) {
              if (key & F_SPILL) {
                sp = true;
              } else if (!g.thresholding) {
                c = true;
              } else {
                uint64_t tot = 0;
#pragma unroll
                for (int j = 0; j < Slot<uint64_t>::NSUB; ++j) {
                  const uint64_t v = L[i].sub[j];
                  tot += v != EMPTY64 ? (v & 0xFFFFFFFFFFFFULL) : 0;
                }
                c = tot >= g.t1lo;
              }
            }
          }
          const uint64_t bc = __ballot(c), bsp = __ballot(sp);
          const uint32_t w0 = i0 + (threadIdx.x & ~63u);
          if ((threadIdx.x & 63) == 0 && w0 < span) {
            g.cbits[region * g.stride + w0 / 64] = bc;
            g.sbits[region * g.stride + w0 / 64] = bsp;
          }
        }
      }
      ulonglong2* T2 = (ulonglong2*)(g.table + base);
      for (uint32_t i = threadIdx.x; i < words / 2; i += BS) T2[i] = ((const ulonglong2*)Lw)[i];
      lds_barrier();
      open = false;
      ++q;
    }
#pragma unroll
    for (int k = 0; k < REGION_ITEMS; ++k) rr[k] = rn[k];
    tile_end = next_end;
    last = next_last;
  }
  const unsigned long long sn = wave_sum(my_new), sk = wave_sum(my_k2);
  if ((threadIdx.x & 63) == 0) {
    if (sn) atomicAdd(&g.ctr->n_distinct, sn);
    if (sk) atomicAdd(&g.ctr->n_k2, sk);
  }
}

}  // namespace br

#endif
