/*
 * partition.h -- grouping the count's instance records by table region.
 *
 * The record parser (k_parse_records, regions.h) leaves one 8-byte record
 * per k-mer instance, rec = (hash << 5) | biedge index, in 32 buckets picked
 * by the hash's top 5 bits.  The next rb hash bits (the record's top rb
 * bits) name the region inside the bucket.  Two multisplit passes, both
 * hand-written and each one read and one write of the records, group them
 * by region so that a region build reads each region's records as one
 * contiguous run:
 *
 *   k_split1   all buckets at once, by the top b1 region bits (up to 128
 *              sub-buckets per bucket).  Each sub-bucket gets room for its
 *              share of the bucket plus a margin (hash bits are uniform, so a
 *              sub-bucket holds 1/2^b1 of its bucket up to sampling noise);
 *              a workgroup sorts a tile of records by digit in LDS and
 *              reserves each digit's run with one atomic per tile.  A
 *              sub-bucket that would overflow its room (a k-mer in a large
 *              share of all instances) is reported, and the pass runs again
 *              with room from exact counts (k_split_count).  The chunk
 *              padding the parser left behind (code 31) is dropped.
 *   k_split2   one workgroup per sub-bucket (persistent), by the remaining
 *              b2 <= 9 region bits: a count pass over the sub-bucket, then
 *              the same LDS tile sort with workgroup-local cursors, into a
 *              compact output.  It also writes the first record of every
 *              region.
 *
 * Both passes load the next tile while sorting and writing the current one.
 *
 * Neither pass keeps the order of records inside a digit, which nothing
 * downstream depends on (a region's slots come out the same up to the
 * order of keys inside a probe run).
 */
#ifndef BRUNO_PARTITION_H_
#define BRUNO_PARTITION_H_

#include "kernels.h"

namespace br {

constexpr int SPLIT_BS = 512;
constexpr int SPLIT_ITEMS = 8;
constexpr int SPLIT_TILE = SPLIT_BS * SPLIT_ITEMS;  /* records per tile */
constexpr int SPLIT1_MAX_BITS = 7;
constexpr int SPLIT2_MAX_BITS = 9;
constexpr uint32_t REC_CODE_FILL = 31;              /* chunk padding, never a biedge index (0..24) */

/* A record: one word (k <= 31) or the key word and the k-mer's high word
 * (k <= 63).  The key word, (hash << 5) | biedge index, is what the passes
 * split on. */
template <typename KT> struct RecOf;
template <> struct RecOf<uint64_t> { typedef uint64_t T; };
template <> struct RecOf<u128> { typedef ulonglong2 T; };
__device__ __forceinline__ uint64_t rec_key(uint64_t r) { return r; }
__device__ __forceinline__ uint64_t rec_key(const ulonglong2& r) { return r.x; }
template <typename R> __device__ __forceinline__ R rec_zero();
template <> __device__ __forceinline__ uint64_t rec_zero<uint64_t>() { return 0; }
template <> __device__ __forceinline__ ulonglong2 rec_zero<ulonglong2>() { return make_ulonglong2(0, 0); }

/* One segment of a bucket's record area: [begin, end) holds records of
 * bucket `bucket`; its tiles are numbered from first_tile. */
struct SplitSeg {
  uint64_t begin, end, first_tile;
  uint32_t bucket, pad;
};
constexpr int SPLIT_MAX_SEGS = 64;

/* Workgroups go to the device's XCDs (each with its own L2) in turn:
 * workgroup w runs on XCD w % SPLIT_XCDS. */
constexpr int SPLIT_XCDS = 8;

struct Split1Args {
  const void* in;                 /* records (RecOf<KT>::T) */
  SplitSeg seg[SPLIT_MAX_SEGS];   /* grouped by bucket % SPLIT_XCDS */
  int n_seg;
  uint64_t n_tiles;
  uint64_t xcd_tiles[SPLIT_XCDS + 1]; /* tiles of XCD x's buckets: [xcd_tiles[x], xcd_tiles[x + 1]) */
  int b1;                         /* digit bits (0: one sub-bucket per bucket) */
  unsigned long long* cursor;     /* [32 << b1]: next free position of each sub-bucket */
  const uint64_t* limit;          /* [32 << b1]: end of each sub-bucket's room */
  void* out;
  unsigned long long* overflow;   /* set when a sub-bucket ran out of room */
};

/* Exclusive prefix of cnt[0..n) into start[0..n) (n <= 512), computed by
 * the first wave; returns nothing, the caller synchronises. */
__device__ __forceinline__ void wave_exclusive_scan(const uint32_t* cnt, uint32_t* start, int n) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const int per = (n + 63) / 64;
    uint32_t mine = 0;
    for (int i = 0; i < per; ++i) {
      const int q = lane * per + i;
      mine += q < n ? cnt[q] : 0;
    }
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    uint32_t run = incl - mine;
    for (int i = 0; i < per; ++i) {
      const int q = lane * per + i;
      if (q < n) {
        start[q] = run;
        run += cnt[q];
      }
    }
  }
}

/* The LDS tile sort both passes share: the tile's records (ITEMS per
 * thread, `ok` false for holes and padding) are ranked by digit with LDS
 * atomics and scattered into buf in digit order; reserve(d, count) (one
 * thread per digit) fixes where digit d's records of this tile go, and
 * store(d, j, rec) writes the j-th of them, so each digit's run is written
 * to consecutive places. */
template <typename R, int ITEMS, typename DigitF, typename ReserveF, typename StoreF>
__device__ __forceinline__ void tile_split(const R (&r)[ITEMS], const bool (&ok)[ITEMS], int nbins, R* buf,
                                           uint32_t* cnt, uint32_t* start, DigitF digit, ReserveF reserve,
                                           StoreF store) {
  for (int i = threadIdx.x; i < nbins; i += SPLIT_BS) cnt[i] = 0;
  __syncthreads();
  uint32_t rank[ITEMS], d[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    d[k] = ok[k] ? digit(rec_key(r[k])) : 0;
    rank[k] = ok[k] ? atomicAdd(&cnt[d[k]], 1u) : 0;
  }
  __syncthreads();
  wave_exclusive_scan(cnt, start, nbins);
  __syncthreads();
  /* per digit: its count in this tile -> its output position */
  for (int i = threadIdx.x; i < nbins; i += SPLIT_BS) reserve(i, cnt[i]);
#pragma unroll
  for (int k = 0; k < ITEMS; ++k)
    if (ok[k]) buf[start[d[k]] + rank[k]] = r[k];
  __syncthreads();
  /* records in digit order; each digit's run goes to consecutive places */
  const uint32_t total = start[nbins - 1] + cnt[nbins - 1];
  for (uint32_t i = threadIdx.x; i < total; i += SPLIT_BS) {
    const R x = buf[i];
    const uint32_t dx = digit(rec_key(x));
    store(dx, i - start[dx], x);
  }
  __syncthreads();
}

/* Both passes sort 8-byte records in tiles of 8,192 (16 per thread), so a
 * sub-bucket's or region's run of a tile is twice as long as with 4,096;
 * 16-byte records keep 4,096 (the tile fills the same 64 KB of LDS). */
template <typename R> struct SplitShape {
  static constexpr int ITEMS = sizeof(R) == 8 ? 2 * SPLIT_ITEMS : SPLIT_ITEMS;
  static constexpr int TILE = SPLIT_BS * ITEMS;
};

/* the records of tile t of the split segments (ok: a real record) */
template <typename R, int ITEMS>
__device__ __forceinline__ void seg_tile_load(const Split1Args& a, uint64_t t, R (&r)[ITEMS], bool (&ok)[ITEMS],
                                              uint32_t& bucket) {
  int g = 0;
  while (g + 1 < a.n_seg && a.seg[g + 1].first_tile <= t) ++g;
  const SplitSeg& sg = a.seg[g];
  const uint64_t lo = sg.begin + (t - sg.first_tile) * (uint64_t)(SPLIT_BS * ITEMS);
  const R* in = (const R*)a.in;
  bucket = sg.bucket;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint64_t i = lo + (uint64_t)k * SPLIT_BS + threadIdx.x;
    r[k] = i < sg.end ? in[i] : rec_zero<R>();
    ok[k] = i < sg.end && (uint32_t)(rec_key(r[k]) & 31) != REC_CODE_FILL;
  }
}

/* The buckets of one XCD are split by that XCD's workgroups only, so the
 * runs that neighbouring tiles append next to each other (sharing cache
 * lines at their ends) meet in one L2 instead of being written back in
 * parts by two.  The grid is a multiple of SPLIT_XCDS. */
template <typename R>
__global__ __launch_bounds__(SPLIT_BS) __attribute__((amdgpu_waves_per_eu(4))) void k_split1(Split1Args a) {
  constexpr int ITEMS = SplitShape<R>::ITEMS;
  __shared__ R buf[SplitShape<R>::TILE];
  __shared__ uint32_t cnt[1 << SPLIT1_MAX_BITS], start[1 << SPLIT1_MAX_BITS];
  __shared__ unsigned long long pos[1 << SPLIT1_MAX_BITS], lim[1 << SPLIT1_MAX_BITS];
  R* out = (R*)a.out;
  const int b1 = a.b1;
  const int nb = 1 << b1;
  const uint32_t x = blockIdx.x % SPLIT_XCDS, per = gridDim.x / SPLIT_XCDS;
  const uint64_t t0 = a.xcd_tiles[x] + blockIdx.x / SPLIT_XCDS, t1 = a.xcd_tiles[x + 1];
  R rn[ITEMS];
  bool okn[ITEMS];
  uint32_t bn = 0;
  if (t0 < t1) seg_tile_load(a, t0, rn, okn, bn);
  for (uint64_t t = t0; t < t1; t += per) {
    R r[ITEMS];
    bool ok[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      r[k] = rn[k];
      ok[k] = okn[k];
    }
    const uint32_t bucket = bn;
    /* the next tile's loads go out before this tile is sorted and written */
    if (t + per < t1) seg_tile_load(a, t + per, rn, okn, bn);
    unsigned long long* cur = a.cursor + ((uint64_t)bucket << b1);
    const uint64_t* limit = a.limit + ((uint64_t)bucket << b1);
    tile_split<R, ITEMS>(
        r, ok, nb, buf, cnt, start, [&](uint64_t x) -> uint32_t { return b1 ? (uint32_t)(x >> (64 - b1)) : 0u; },
        [&](int i, uint32_t c) {
          pos[i] = c ? atomicAdd(&cur[i], (unsigned long long)c) : 0ULL;
          lim[i] = limit[i];
        },
        [&](uint32_t dx, uint32_t j, const R& x) {
          /* a sub-bucket past its room: reported, and the pass runs again
           * with exact room */
          const uint64_t at = pos[dx] + j;
          if (at < lim[dx]) out[at] = x;
          else atomicOr(a.overflow, 1ULL);
        });
  }
}

/* Exact records per (bucket, top b1 region bits): the room k_split1 takes
 * when the estimate ran short. */
template <typename R>
__global__ __launch_bounds__(SPLIT_BS) void k_split_count(Split1Args a, unsigned long long* counts) {
  __shared__ uint32_t cnt[1 << SPLIT1_MAX_BITS];
  const int b1 = a.b1;
  const int nb = 1 << b1;
  constexpr int ITEMS = SplitShape<R>::ITEMS;
  for (uint64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
    R r[ITEMS];
    bool ok[ITEMS];
    uint32_t bucket = 0;
    seg_tile_load(a, t, r, ok, bucket);
    for (int i = threadIdx.x; i < nb; i += SPLIT_BS) cnt[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      if (ok[k]) atomicAdd(&cnt[b1 ? (uint32_t)(rec_key(r[k]) >> (64 - b1)) : 0u], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += SPLIT_BS)
      if (cnt[i]) atomicAdd(&counts[((uint64_t)bucket << b1) + i], (unsigned long long)cnt[i]);
    __syncthreads();
  }
}

/* region_start of every sub-bucket's first region (its output start), so a
 * region build of the sub-buckets before sb finds its last region's end
 * before the pass over sb has run */
__global__ void k_region_bounds(const uint64_t* out_start, uint64_t n_sb, int b2, uint64_t* region_start) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s <= n_sb) region_start[s << b2] = out_start[s];
}

struct Split2Args {
  const void* in;                 /* records (RecOf<KT>::T) */
  const uint64_t* in_start;       /* [n_sb]: sub-bucket s holds in[in_start[s], + its count) */
  const uint64_t* out_start;      /* [n_sb + 1]: and goes to out[out_start[s], out_start[s + 1]) */
  uint64_t sb0;                   /* this launch: sub-buckets [sb0, sb0 + n_sb) */
  uint64_t n_sb;
  int shift;                      /* digit = (rec >> shift) & (2^b2 - 1) */
  int b2;
  void* out;                      /* grouped by region */
  uint64_t* region_start;         /* [n_sb << b2 (+1)]: first record of every region */
};

template <typename R>
__global__ __launch_bounds__(SPLIT_BS) __attribute__((amdgpu_waves_per_eu(4))) void k_split2(Split2Args a) {
  constexpr int ITEMS = SplitShape<R>::ITEMS, TILE = SplitShape<R>::TILE;
  __shared__ R buf[TILE];
  const R* in = (const R*)a.in;
  R* out = (R*)a.out;
  __shared__ uint32_t cnt[1 << SPLIT2_MAX_BITS], start[1 << SPLIT2_MAX_BITS];
  __shared__ unsigned long long pos[1 << SPLIT2_MAX_BITS], next[1 << SPLIT2_MAX_BITS];
  const int b2 = a.b2, shift = a.shift;
  const uint32_t nb = 1u << b2, m = nb - 1;
  auto digit = [&](uint64_t x) -> uint32_t { return b2 ? (uint32_t)(x >> shift) & m : 0u; };
  for (uint64_t s = a.sb0 + blockIdx.x; s < a.sb0 + a.n_sb; s += gridDim.x) {
    const uint64_t o0 = a.out_start[s], n = a.out_start[s + 1] - o0;
    const uint64_t lo = a.in_start[s], hi = lo + n;
    /* count pass: every digit's records in the sub-bucket (a region's
     * count cannot be estimated: a k-mer repeated thousands of times puts
     * all its instances in one region) */
    for (uint32_t i = threadIdx.x; i < nb; i += SPLIT_BS) cnt[i] = 0;
    __syncthreads();
    for (uint64_t i0 = lo; i0 < hi; i0 += 2 * SPLIT_TILE) {
      uint64_t r[2 * SPLIT_ITEMS];
#pragma unroll
      for (int k = 0; k < 2 * SPLIT_ITEMS; ++k) {
        const uint64_t i = i0 + (uint64_t)k * SPLIT_BS + threadIdx.x;
        r[k] = i < hi ? rec_key(in[i]) : 0;
      }
#pragma unroll
      for (int k = 0; k < 2 * SPLIT_ITEMS; ++k)
        if (i0 + (uint64_t)k * SPLIT_BS + threadIdx.x < hi) atomicAdd(&cnt[digit(r[k])], 1u);
    }
    __syncthreads();
    wave_exclusive_scan(cnt, start, (int)nb);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += SPLIT_BS) {
      next[i] = o0 + start[i];
      a.region_start[(s << b2) + i] = o0 + start[i];
    }
    __syncthreads();
    /* scatter pass: tile by tile, each digit's run after the previous
     * tile's; the next tile loads while one is sorted and written */
    R rn[ITEMS];
    bool okn[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const uint64_t i = lo + (uint64_t)k * SPLIT_BS + threadIdx.x;
      okn[k] = i < hi;
      rn[k] = okn[k] ? in[i] : rec_zero<R>();
    }
    for (uint64_t i0 = lo; i0 < hi; i0 += TILE) {
      R r[ITEMS];
      bool ok[ITEMS];
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        r[k] = rn[k];
        ok[k] = okn[k];
      }
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const uint64_t i = i0 + TILE + (uint64_t)k * SPLIT_BS + threadIdx.x;
        okn[k] = i < hi;
        rn[k] = okn[k] ? in[i] : rec_zero<R>();
      }
      tile_split<R, ITEMS>(
          r, ok, (int)nb, buf, cnt, start, digit,
          [&](int i, uint32_t c) {
            pos[i] = next[i];
            next[i] += c;
          },
          [&](uint32_t dx, uint32_t j, const R& x) { out[pos[dx] + j] = x; });
    }
  }
}

}  // namespace br

#endif
