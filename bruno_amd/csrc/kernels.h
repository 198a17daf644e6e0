/*
 * kernels.h -- HIP kernels of the de Bruijn hot path (gfx950, wave64).
 *
 * Phases and the reference code each one replaces:
 *   k_parse_count   parse reads into canonical (k-mer, biedge) records and
 *                   count them in the HBM table (debruijn_biedge_loader.hpp
 *                   :394-623, debruijn_common.hpp:53-58,
 *                   debruijn_graph_map.hpp:1085-1141 / :175-197)
 *   k_filter        local_insert_by_freqencies per k-mer block
 *                   (debruijn_graph_map.hpp:1258-1623) + IsBranchPoint /
 *                   IsChainNode (debruijn_graph_filters.hpp:57-220)
 *   k_split         branch / chain lists
 *   k_chain_init    to_simple_biedge (debruijn_chain_node.hpp:84-141) with
 *                   neighbours resolved to chain indices
 *   k_termini       setup_chain_termini + terminus_update
 *                   (debruijn_chain_graph.hpp:486-596,
 *                   debruijn_chain_operations.hpp:58-93)
 *   k_lr_send/recv  one synchronous round of list_rank / list_rank_min_update
 *                   (debruijn_chain_graph.hpp:793-994, :1040-1310) with
 *                   chain_update (debruijn_chain_operations.hpp:108-256)
 *   k_separate      separate_cycles / separate_isolated (:670-777)
 *   k_ranked        to_listranked_chain_node (debruijn_chain_operations.hpp
 *                   :261-310)
 *   k_fasta_*       print_chain_as_fasta / print_graph_node_fasta bytes
 */
#ifndef BRUNO_KERNELS_H_
#define BRUNO_KERNELS_H_

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "kmer.h"
#include "table.h"

namespace br {

constexpr int BLOCK = 256;

/* ------------------------------------------------------------------ */
/* chain-node encodings (debruijn_chain_node.hpp:53-80)                 */
/* ------------------------------------------------------------------ */
constexpr uint32_t DIST_MASK = 0x7FFFFFFFu;
constexpr uint32_t TERM_MASK = 0x80000000u;
__device__ __forceinline__ uint32_t cdist(uint32_t x) { return x & DIST_MASK; }
__device__ __forceinline__ bool is_terminal(uint32_t x) { return (x & DIST_MASK) == 0; }
__device__ __forceinline__ bool points_to_self(uint32_t x) { return x == TERM_MASK; }
__device__ __forceinline__ bool points_to_chain(uint32_t x) { return (x ^ TERM_MASK) > TERM_MASK; }
__device__ __forceinline__ bool is_unfinished(uint32_t din, uint32_t dout) {
  return points_to_chain(din) || points_to_chain(dout);
}

/* A chain-end pointer: node index (31 bits) and the end of that node that
 * faces the pointing node (bit 31: 1 = OUT end, 0 = IN end). */
constexpr uint32_t END_OUT = 0x80000000u;
constexpr uint32_t IDX_MASK = 0x7FFFFFFFu;

/* Per-node list-ranking state with its two mailboxes (one per end). */
struct LRNode {
  uint32_t p[2];  /* [0] = in pointer, [1] = out pointer                 */
  uint32_t d[2];  /* [0] = in dist, [1] = out dist (term_mask encoding)  */
  unsigned long long mail[2]; /* (dist << 32) | ptr, 0 = empty          */
};
static_assert(sizeof(LRNode) == 32, "lr node");

struct Counters {
  unsigned long long n_instances, n_dropped, n_distinct, n_k2;
  unsigned long long n_present, n_branch, n_chain, n_virt;
  unsigned long long lr_count, lr_next, lr_cycle, lr_conflict;
  unsigned long long n_cycle, n_isolated, n_out, n_spill;
  unsigned long long n_cand;
  unsigned long long lr_ill;    /* chain ends that break well-formedness (k_chain_init) */
  uint32_t err;
  uint32_t pad;
};

/* ------------------------------------------------------------------ */
/* block-wide reservation of output positions                           */
/* ------------------------------------------------------------------ */
typedef hipcub::BlockScan<uint32_t, BLOCK> BScan;
struct ReserveSmem {
  typename BScan::TempStorage scan;
  unsigned long long base;
};
/* All threads of the block call this; returns this thread's first output
 * position.  One global atomic per block. */
__device__ __forceinline__ unsigned long long block_reserve(uint32_t my, unsigned long long* counter,
                                                            ReserveSmem& sm) {
  uint32_t excl, total;
  BScan(sm.scan).ExclusiveSum(my, excl, total);
  if (threadIdx.x == 0) sm.base = total ? atomicAdd(counter, (unsigned long long)total) : 0ULL;
  __syncthreads();
  unsigned long long b = sm.base;
  __syncthreads();
  return b + excl;
}

__device__ __forceinline__ void block_add(unsigned long long v, unsigned long long* counter) {
  typedef hipcub::BlockReduce<unsigned long long, BLOCK> BR;
  __shared__ typename BR::TempStorage tmp;
  unsigned long long t = BR(tmp).Sum(v);
  if (threadIdx.x == 0 && t) atomicAdd(counter, t);
}

/* ------------------------------------------------------------------ */
/* fill                                                                 */
/* ------------------------------------------------------------------ */
/* n 64-bit words of value v; 16-byte stores, 4 per thread per iteration */
__global__ void k_fill64(uint64_t* p, uint64_t n, uint64_t v) {
  const uint64_t n2 = n / 2;
  ulonglong2* q = (ulonglong2*)p;
  const ulonglong2 w = make_ulonglong2(v, v);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    q[i] = w;
    q[i + stride] = w;
    q[i + 2 * stride] = w;
    q[i + 3 * stride] = w;
  }
  for (; i < n2; i += stride) q[i] = w;
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) p[n - 1] = v;
}

/* ------------------------------------------------------------------ */
/* parse (+ count)                                                      */
/* ------------------------------------------------------------------ */
/* A/a 0, C/c 1, G/g 2, T/t 3, any other byte -1, without branches: the
 * byte is lower-cased (only these eight bytes map onto a, c, g, t), tested
 * against the bit set {a, c, g, t} - 'a', and coded as ((c >> 1) ^ (c >> 2))
 * & 3 (a 0x61 -> 0, c 0x63 -> 1, g 0x67 -> 2, t 0x74 -> 3). */
__device__ __forceinline__ int base_code(uint8_t b) {
  const uint32_t c = (uint32_t)b | 0x20u;
  const uint32_t d = c - 0x61u;
  const bool ok = d < 20u && ((0x80045u >> d) & 1u);
  return ok ? (int)(((c >> 1) ^ (c >> 2)) & 3u) : -1;
}

/* Pack 64 bases (one per lane, code 0..3) into two 32-base words with the
 * first base in the top bits: bit p of w0 is bit (p & 1) of base 31 - p / 2
 * (w1: base 63 - p / 2), so lane p fetches that base's code from its lane
 * and one ballot per word gives the packed word.  The whole wave must be
 * active. */
__device__ __forceinline__ void pack64(int code, uint64_t& w0, uint64_t& w1) {
  const int lane = (int)__lane_id();
  const int bit = lane & 1;
  const int c0 = __shfl(code, 31 - (lane >> 1), 64);
  const int c1 = __shfl(code, 63 - (lane >> 1), 64);
  w0 = __ballot((c0 >> bit) & 1);
  w1 = __ballot((c1 >> bit) & 1);
}

/* word q (0..3) of [a b c d] as a two-level select on the bits of q (a
 * per-lane q then costs selects, not branches) */
__device__ __forceinline__ uint64_t sel4(int q, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  const uint64_t lo = (q & 1) ? b : a, hi = (q & 1) ? d : c;
  return (q & 2) ? hi : lo;
}
/* 64 bits of the 256-bit window [a b c d] starting at base position pos
 * (pos <= 96; at pos 96 the word after d is not read). */
__device__ __forceinline__ uint64_t window64(int pos, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  int q = pos >> 5, r = pos & 31;
  uint64_t hi = sel4(q, a, b, c, d), lo = sel4(q + 1, a, b, c, d);
  return r ? ((hi << (2 * r)) | (lo >> (64 - 2 * r))) : hi;
}
__device__ __forceinline__ int window_base(int pos, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  uint64_t w = sel4(pos >> 5, a, b, c, d);
  return (int)((w >> (62 - 2 * (pos & 31))) & 3);
}

template <typename KT> __device__ __forceinline__ KT window_kmer(int l, int K, uint64_t a, uint64_t b, uint64_t c, uint64_t d);
template <> __device__ __forceinline__ uint64_t window_kmer<uint64_t>(int l, int K, uint64_t a, uint64_t b, uint64_t c,
                                                                      uint64_t d) {
  return window64(l, a, b, c, d) >> (64 - 2 * K);
}
template <> __device__ __forceinline__ u128 window_kmer<u128>(int l, int K, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  uint64_t h = window64(l, a, b, c, d), lo = window64(l + 32, a, b, c, d);
  u128 t = ((u128)h << 64) | lo;
  return t >> (128 - 2 * K);
}

/* Index j of the segment holding position x: the largest j < n with
 * first[j] <= x (first[0] <= x < first[n]; empty segments are skipped). */
__device__ __forceinline__ uint32_t job_of(const uint64_t* first, uint32_t n, uint64_t x) {
  uint32_t lo = 0, hi = n; /* first[lo] <= x < first[hi] */
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct ParseArgs {
  const uint8_t* bases;
  const uint64_t* offsets;
  /* optional: read r is bases[starts[r], ends[r]) instead of
   * bases[offsets[r], offsets[r+1]) (reads left in place in a FASTQ file) */
  const uint64_t* starts;
  const uint64_t* ends;
  uint64_t n_reads;
  int K;
  /* count mode: cap is the probe span (the whole table, or one region
   * when rbits > 0; a batch job's span is its segment) */
  void* table;
  uint64_t cap;
  uint32_t rbits;
  Spill* spill;
  uint64_t spcap;
  /* batches: job j owns reads [job_first[j], job_first[j+1]) of this call
   * and table slots [seg[j], seg[j+1]); n_jobs == 1 uses the whole table */
  uint32_t n_jobs;
  const uint64_t* job_first;
  const uint64_t* seg;
  unsigned long long* job_inst; /* per-job k-mer instances (n_jobs > 1) */
  unsigned long long* job_drop; /* per-job dropped reads (n_jobs > 1) */
  /* sketch mode: SKETCH_M HyperLogLog registers per job */
  uint32_t* sketch;
  /* record mode */
  const uint64_t* rec_base; /* per-read first record index */
  uint64_t* out_kmers;
  uint8_t* out_biedges;
  uint64_t* per_read;       /* per-read record count (count pass) */
  Counters* ctr;
};

/* Workgroup barrier ordering LDS accesses only: global stores issued
 * before it need not have completed. */
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

/* Distinct-k-mer sketch: canonical k-mers whose hash has its top
 * SKETCH_SAMPLE_BITS bits zero (1 in 64) go to a HyperLogLog of SKETCH_M
 * registers (index = the next 12 hash bits, value = 1 + leading zeros of
 * the remaining 46 bits). */
constexpr int SKETCH_SAMPLE_BITS = 6;
constexpr int SKETCH_INDEX_BITS = 12;
constexpr int SKETCH_M = 1 << SKETCH_INDEX_BITS;

/* The canonical (k-mer, biedge) instances of one read of L >= K + 2 bases,
 * parsed by one wave: lane l takes instances p0 + l of a sliding window of
 * four packed words covering bases [p0, p0 + 128).  code_of(q) gives the
 * lane's base code in chunk q (base 64 q + l; 0 past the read).  f(i, km,
 * e) runs for each instance i < L - K + 1. */
template <typename KT, typename G, typename F>
__device__ __forceinline__ void for_each_instance_from(uint64_t L, int K, int lane, const kops<KT>& ops, G&& code_of,
                                                       F&& f) {
  const uint64_t nk = L - K + 1;
  uint64_t A, B, C, D, prevB = 0;
  pack64(code_of(0), A, B);
  pack64(code_of(1), C, D);
  for (uint64_t p0 = 0; p0 < nk; p0 += 64) {
    const uint64_t i = p0 + lane;
    if (i < nk) {
      KT fw = window_kmer<KT>(lane, K, A, B, C, D);
      int inb = (lane > 0) ? window_base(lane - 1, A, B, C, D) : (int)(prevB & 3);
      int outb = window_base(lane + K, A, B, C, D);
      uint8_t left = (i > 0) ? (uint8_t)(1u << inb) : (uint8_t)0;
      uint8_t right = (i + K < L) ? (uint8_t)(1u << outb) : (uint8_t)0;
      uint8_t e = (uint8_t)((left << 4) | right);
      /* lex_less (debruijn_common.hpp:53-58): keep if x < rc(x), else
       * rc the k-mer and the biedge. */
      KT rv = ops.rc(fw);
      KT km = fw;
      if (!(fw < rv)) {
        km = rv;
        e = biedge_rc(e);
      }
      f(i, km, e);
    }
    /* advance the window by 64 bases */
    prevB = B;
    A = C;
    B = D;
    pack64(code_of(p0 / 64 + 2), C, D);
  }
}

/* The same with the bases read from s. */
template <typename KT, typename F>
__device__ __forceinline__ void for_each_instance(const uint8_t* s, uint64_t L, int K, int lane, const kops<KT>& ops,
                                                  F&& f) {
  for_each_instance_from<KT>(L, K, lane, ops, [&](uint64_t q) -> int {
    const uint64_t i = 64 * q + (uint64_t)lane;
    return i < L ? base_code(s[i]) : 0;
  }, f);
}

/* lane l's value of a 64-bit register (l wave-uniform) */
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

/* MODE 0: count records per read; MODE 1: write records; MODE 2: count into
 * the table; MODE 3: distinct-k-mer sketch only.  BS is the workgroup size
 * (the work is per wave, so any multiple of 64 works; it is a tuning knob). */
template <typename KT, int MODE, int BS = BLOCK>
__global__ __launch_bounds__(BS) void k_parse(ParseArgs a) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const int K = a.K;
  kops<KT> ops(K);
  unsigned long long my_inst = 0, my_drop = 0, my_new = 0, my_k2 = 0;
  /* batch: the job of the wave's current read, its read range [j_lo, j_hi)
   * and table segment, and the wave's instance count for it (flushed when
   * the job changes, so the per-job counters see few atomics) */
  uint32_t cj = 0;
  uint64_t j_lo = 0, j_hi = 0, tbase = 0, tcap = a.cap;
  unsigned long long j_inst = 0;
  for (uint64_t r = wave; r < a.n_reads; r += nwaves) {
    if ((MODE == 2 || MODE == 3) && a.n_jobs > 1 && !(r >= j_lo && r < j_hi)) {
      if (lane == 0 && j_inst) atomicAdd(&a.job_inst[cj], j_inst);
      j_inst = 0;
      cj = job_of(a.job_first, a.n_jobs, r);
      j_lo = a.job_first[cj];
      j_hi = a.job_first[cj + 1];
      tbase = a.seg[cj];
      tcap = a.seg[cj + 1] - tbase;
    }
    const uint64_t off = a.starts ? a.starts[r] : a.offsets[r];
    const uint64_t L = (a.starts ? a.ends[r] : a.offsets[r + 1]) - off;
    const uint8_t* s = a.bases + off;
    /* NFilterSequencesIterator: a read with any non-ACGT base is dropped
     * (compact_dbg_freq_construct_clean.cpp:210-213). */
    int bad = 0;
    for (uint64_t i = lane; i < L; i += 64) bad |= (base_code(s[i]) < 0);
    if (__any(bad)) {
      if (lane == 0) ++my_drop;
      if (MODE == 0 && lane == 0) a.per_read[r] = 0;
      if (MODE == 2 && a.n_jobs > 1 && lane == 0) atomicAdd(&a.job_drop[cj], 1ULL);
      continue;
    }
    /* parser window is k+2 (debruijn_biedge_loader.hpp:402, :464-471) */
    if (L < (uint64_t)K + 2) {
      if (MODE == 0 && lane == 0) a.per_read[r] = 0;
      continue;
    }
    const uint64_t nk = L - K + 1;
    if (MODE == 0) {
      if (lane == 0) a.per_read[r] = nk;
      continue;
    }
    uint64_t rbase = (MODE == 1) ? a.rec_base[r] : 0;
    if (MODE == 2 && a.n_jobs > 1) j_inst += nk;
    for_each_instance<KT>(s, L, K, lane, ops, [&](uint64_t i, KT km, uint8_t e) {
      if (MODE == 3) {
        const uint64_t h = sketch_hash(km);
        if ((h >> (64 - SKETCH_SAMPLE_BITS)) == 0) {
          const uint32_t idx = (uint32_t)(h >> (64 - SKETCH_SAMPLE_BITS - SKETCH_INDEX_BITS)) & (SKETCH_M - 1);
          const uint64_t rest = (h << (SKETCH_SAMPLE_BITS + SKETCH_INDEX_BITS)) |
                                (1ULL << (SKETCH_SAMPLE_BITS + SKETCH_INDEX_BITS - 1));
          const uint32_t rank = (uint32_t)__clzll(rest) + 1;
          uint32_t* reg = a.sketch + (uint64_t)cj * SKETCH_M + idx;
          if (rank > __hip_atomic_load(reg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(reg, rank);
        }
      } else if (MODE == 1) {
        a.out_kmers[(rbase + i) * kt_traits<KT>::words] = (uint64_t)km;
        if (kt_traits<KT>::words == 2)
          a.out_kmers[(rbase + i) * kt_traits<KT>::words + 1] = (uint64_t)((u128)km >> 64);
        a.out_biedges[rbase + i] = e;
      } else {
        int claimed = 0;
        uint64_t slot =
            SlotOps<KT>::find_or_insert((Slot<KT>*)a.table + tbase, tcap, a.rbits, km, &a.ctr->err, &claimed);
        if (slot != ~0ULL) {
          my_new += claimed;
          my_k2 += add_k2<KT>((Slot<KT>*)a.table, tbase + slot, e, a.spill, a.spcap, &a.ctr->err);
        }
      }
      ++my_inst;
    });
  }
  if (MODE == 2 && a.n_jobs > 1 && lane == 0 && j_inst) atomicAdd(&a.job_inst[cj], j_inst);
  if (MODE == 2) {
    /* one atomic per wave and counter */
    const unsigned long long si = wave_sum(my_inst), sd = wave_sum(my_drop), sn = wave_sum(my_new),
                             sk = wave_sum(my_k2);
    if (lane == 0) {
      if (si) atomicAdd(&a.ctr->n_instances, si);
      if (sd) atomicAdd(&a.ctr->n_dropped, sd);
      if (sn) atomicAdd(&a.ctr->n_distinct, sn);
      if (sk) atomicAdd(&a.ctr->n_k2, sk);
    }
  }
}

/* ------------------------------------------------------------------ */
/* per-slot k2 gathering                                                */
/* ------------------------------------------------------------------ */
/* All 25 biedge bytes possible for ACGT-only reads (nibbles 0,1,2,4,8). */
__device__ __constant__ uint8_t BIEDGE25[25] = {
    0x00, 0x01, 0x02, 0x04, 0x08, 0x10, 0x11, 0x12, 0x14, 0x18, 0x20, 0x21, 0x22,
    0x24, 0x28, 0x40, 0x41, 0x42, 0x44, 0x48, 0x80, 0x81, 0x82, 0x84, 0x88};

/* k2 entries of one slot.  N is a compile-time bound so that every array
 * index below is a constant after unrolling and the lists stay in registers
 * (N = NSUB for slots without spilled entries, 25 for the rest). */
template <int N>
struct K2L {
  uint8_t e[N];
  uint32_t c[N];
  int n;
};
typedef K2L<25> K2List;

template <typename KT, int N>
__device__ __forceinline__ uint64_t sub_entries(const Slot<KT>* p, K2L<N>& L) {
  static_assert(N >= Slot<KT>::NSUB, "list too small");
  L.n = 0;
  uint64_t kf = 0;
#pragma unroll
  for (int j = 0; j < Slot<KT>::NSUB; ++j) {
    uint64_t v = p->sub[j];
    bool ok = v != EMPTY64;
    /* sub-entries fill in order, so the valid ones are a prefix */
    L.e[j] = ok ? (uint8_t)((v >> 48) - 1) : 0;
    L.c[j] = ok ? (uint32_t)(v & 0xFFFFFFFFFFFFULL) : 0;
    L.n += ok;
    kf += ok ? (v & 0xFFFFFFFFFFFFULL) : 0;
  }
  return kf;
}

template <typename KT>
__device__ __forceinline__ void gather_k2(const Slot<KT>* p, uint64_t s, const Spill* sp, uint64_t spcap, K2List& L) {
  sub_entries<KT, 25>(p, L);
  if (SlotOps<KT>::flags(p) & F_SPILL) {
    spill_each(sp, spcap, s, [&](uint64_t code, uint64_t cnt) {
      if (L.n < 25) {
        L.e[L.n] = (uint8_t)(code - 1);
        L.c[L.n] = (uint32_t)cnt;
        ++L.n;
      }
    });
  }
}

__device__ __forceinline__ int nib_idx(int n) { return n == 1 ? 0 : n == 2 ? 1 : n == 4 ? 2 : n == 8 ? 3 : 8; }
__device__ __forceinline__ uint16_t sat_add16(uint32_t a, uint32_t b) {
  uint32_t c = a + b;
  return (uint16_t)(c > 65535u ? 65535u : c);
}
/* compact_multi_biedge::update(edges, cnt) (debruijn_graph_node.hpp:200-211),
 * written with constant indices only */
__device__ __forceinline__ void node_update(uint16_t (&c)[9], uint8_t e, uint32_t cnt) {
  int o = nib_idx(e & 0xF), i = nib_idx(e >> 4);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q == o) c[q] = sat_add16(c[q], cnt);
    if (q == i) c[4 + q] = sat_add16(c[4 + q], cnt);
  }
  c[8] = sat_add16(c[8], cnt);
}
__device__ __forceinline__ void add_sel4(uint64_t (&a)[4], int idx, uint64_t v) {
#pragma unroll
  for (int q = 0; q < 4; ++q) a[q] += (q == idx) ? v : 0;
}
__device__ __forceinline__ uint64_t get_sel4(const uint64_t (&a)[4], int idx, uint64_t dflt) {
  uint64_t r = dflt;
#pragma unroll
  for (int q = 0; q < 4; ++q) r = (q == idx) ? a[q] : r;
  return r;
}

/* P1: unthresholded counters from the k2 list. */
template <int N>
__device__ __forceinline__ void p1_counts(const K2L<N>& L, uint16_t (&c)[9]) {
#pragma unroll
  for (int t = 0; t < 9; ++t) c[t] = 0;
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (j < L.n) node_update(c, L.e[j], L.c[j] > 65535u ? 65535u : L.c[j]);
}

struct Thresh {
  uint64_t t0lo, t0hi, t1lo, t1hi, t2lo, t2hi;
};

/* local_insert_by_freqencies for one k-mer block (debruijn_graph_map.hpp
 * :1351-1609).  Returns true if any k2-mer is kept. */
template <typename KT, int N>
__device__ bool p2_counts(const kops<KT>& ops, KT k, const K2L<N>& L, const Thresh& t, uint16_t (&c)[9]) {
  const uint64_t cnt_max = 65535;
  bool kpal = ops.is_rc_palindrome(k);
  bool high = ops.km1_high(k), low = ops.km1_low(k);
  int lastb = ops.last(k), firstb = ops.first(k);
  uint64_t kf = 0, k1in[4] = {0, 0, 0, 0}, k1out[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if (j >= L.n) break;
    uint64_t cu = L.c[j];
    kf += cu;
    int ch = L.e[j] >> 4;
    if (ch) {
      int ci = nib_idx(ch);
      /* is_k1_rc_palindrome(l, k, high): last base == comp(l) && high */
      bool split = ((ops.K & 1) && lastb == 3 - ci && high) && !kpal;
      add_sel4(k1in, ci, split ? ((cu + 1) >> 1) : cu);
    }
    ch = L.e[j] & 0xF;
    if (ch) {
      int ci = nib_idx(ch);
      bool split = ((ops.K & 1) && firstb == 3 - ci && low) && !kpal;
      add_sel4(k1out, ci, split ? ((cu + 1) >> 1) : cu);
    }
  }
  kf = kf < cnt_max ? kf : cnt_max;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    k1in[q] = k1in[q] < cnt_max ? k1in[q] : cnt_max;
    k1out[q] = k1out[q] < cnt_max ? k1out[q] : cnt_max;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) c[q] = 0;
  bool have = false;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if (j >= L.n) break;
    uint8_t flag = 0;
    uint64_t cnt = L.c[j] < cnt_max ? L.c[j] : cnt_max;
    if (cnt < t.t2lo || cnt >= t.t2hi) flag |= 4;
    if (kf < t.t0lo || kf >= t.t0hi) flag |= 8;
    int in = L.e[j] >> 4, out = L.e[j] & 0xF;
    uint64_t k1 = get_sel4(k1in, nib_idx(in), t.t1hi);
    if (k1 < t.t1lo || k1 >= t.t1hi) flag |= 2;
    k1 = get_sel4(k1out, nib_idx(out), t.t1hi);
    if (k1 < t.t1lo || k1 >= t.t1hi) flag |= 1;
    if (flag >= 3) continue;
    uint8_t e = L.e[j];
    if (flag == 1) e &= 0xF0;
    if (flag == 2) e &= 0x0F;
    node_update(c, e, (uint32_t)cnt);
    have = true;
  }
  return have;
}

/* What the filter and the classification need of the k-mer itself. */
struct KmerFacts {
  bool kpal;            /* rc palindrome (even k) */
  bool low, high;       /* k-1 prefix / suffix rc palindromes (odd k) */
  uint32_t first, last; /* first and last base */
};
template <typename KT>
__device__ __forceinline__ KmerFacts kmer_facts(const kops<KT>& ops, KT k) {
  return KmerFacts{ops.is_rc_palindrome(k), ops.km1_low(k), ops.km1_high(k), (uint32_t)ops.first(k),
                   (uint32_t)ops.last(k)};
}

/* p2_counts for the few k2 entries a slot holds, without branches: every
 * entry is evaluated (an empty one, count 0 and biedge 0, adds nothing and
 * is never kept).  Sums are taken in 32 bits over terms clamped at 65535,
 * which gives the same saturated values as the 64-bit sums clamped after;
 * thresholds are clamped at 65536, which no clamped value reaches. */
template <typename KT, int N>
__device__ __forceinline__ bool p2_counts_slot(const kops<KT>& ops, const KmerFacts& f, const K2L<N>& L,
                                               const Thresh& t, uint16_t (&c)[9]) {
  const uint32_t M = 65535;
  auto th = [](uint64_t x) -> uint32_t { return x < 65536 ? (uint32_t)x : 65536u; };
  const uint32_t lo0 = th(t.t0lo), hi0 = th(t.t0hi), lo1 = th(t.t1lo), hi1 = th(t.t1hi), lo2 = th(t.t2lo),
                 hi2 = th(t.t2hi);
  const bool odd = ops.K & 1;
  /* the edge whose k+1-mer is an rc palindrome counts half (is_k1_rc_palindrome) */
  const uint32_t split_in = (odd && f.high && !f.kpal) ? (1u << (3 - f.last)) : 0u;
  const uint32_t split_out = (odd && f.low && !f.kpal) ? (1u << (3 - f.first)) : 0u;
  uint32_t kf = 0, kin[4] = {0, 0, 0, 0}, kout[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const uint32_t cu = j < L.n ? L.c[j] : 0u, e = j < L.n ? L.e[j] : 0u;
    const uint32_t ni = e >> 4, no = e & 15, half = (cu >> 1) + (cu & 1);
    const uint32_t vi = min(ni == split_in ? half : cu, M), vo = min(no == split_out ? half : cu, M);
    kf += min(cu, M);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      kin[q] += ni == (1u << q) ? vi : 0u;
      kout[q] += no == (1u << q) ? vo : 0u;
    }
  }
  kf = min(kf, M);
  const bool kf_bad = kf < lo0 || kf >= hi0;
  uint32_t acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool have = false;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const uint32_t cu = j < L.n ? min((uint32_t)L.c[j], M) : 0u, e = j < L.n ? L.e[j] : 0u;
    const uint32_t ni = e >> 4, no = e & 15;
    /* an entry without an edge on a side takes t1hi there: flagged */
    uint32_t k1i = 65536u, k1o = 65536u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      k1i = ni == (1u << q) ? min(kin[q], M) : k1i;
      k1o = no == (1u << q) ? min(kout[q], M) : k1o;
    }
    const bool f2 = k1i < lo1 || k1i >= hi1, f1 = k1o < lo1 || k1o >= hi1;
    const bool keep = j < L.n && !(cu < lo2 || cu >= hi2) && !kf_bad && !(f1 && f2);
    const uint32_t ki = keep && !f2 ? ni : 0u, ko = keep && !f1 ? no : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc[q] += ko == (1u << q) ? cu : 0u;
      acc[4 + q] += ki == (1u << q) ? cu : 0u;
    }
    acc[8] += keep ? cu : 0u;
    have |= keep;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) c[q] = (uint16_t)min(acc[q], M);
  return have;
}

/* make_node with the k-mer's facts at hand, without branches. */
template <typename KT>
__device__ __forceinline__ NodeRec make_node_facts(const kops<KT>& ops, const KmerFacts& f, const uint16_t (&c)[9]) {
  uint32_t om = 0, im = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    om |= c[q] > 0 ? (1u << q) : 0u;
    im |= c[4 + q] > 0 ? (1u << q) : 0u;
  }
  /* IsKplus1Palindrome (debruijn_graph_filters.hpp:57-95), odd k: the k-1
   * prefix an rc palindrome and an out edge whose base is the first base
   * or its complement, or the k-1 suffix one and an in edge whose base is
   * the last base or its complement */
  const bool odd = ops.K & 1;
  const uint32_t fm = (1u << f.first) | (1u << (3 - f.first)), lm = (1u << f.last) | (1u << (3 - f.last));
  const bool pal = f.kpal || (odd && ((f.low && (om & fm)) || (f.high && (im & lm))));
  NodeRec rec;
#pragma unroll
  for (int q = 0; q < 9; ++q) rec.c[q] = c[q];
  rec.cls = (__popc(im) >= 2 || __popc(om) >= 2 || pal) ? CLS_BRANCH : CLS_CHAIN;
  rec.pad = 0;
  rec.idx = 0xFFFFFFFFu;
  return rec;
}

/* The k+1-mer edge-frequency filter of the block-construct driver
 * (compact_debruijn_graph_block_construct.cpp:415-480): an edge is kept when
 * the count of its canonical k+1-mer over the reads is in [t1lo, t1hi)
 * (saturating at 65535, debruijn_biedge_filters.hpp:128-176), a k2 entry
 * when one of its edges is (filter_isolated_nodes, :60-69), with the edges
 * that failed cleared (transform_biedges_by_frequency, :330-418).  The
 * window counts come from the k-mer's own k2 entries: the windows of a.K
 * are its entries with in-edge a, except that a window of an rc-palindromic
 * a.K (odd k, K not a palindrome) gives two entries (one per strand), so
 * the sum is halved (it is even); an rc-palindromic K (even k) is recorded
 * from the reverse strand, so the windows c.K also show up as out-edge
 * comp(c).  Returns true if any entry is kept. */
template <typename KT, int N>
__device__ bool edge_counts(const kops<KT>& ops, KT k, const K2L<N>& L, const Thresh& t, uint16_t (&c)[9]) {
  const uint64_t cnt_max = 65535;
  const bool kpal = ops.is_rc_palindrome(k);
  const bool high = ops.km1_high(k), low = ops.km1_low(k);
  const int lastb = ops.last(k), firstb = ops.first(k);
  uint64_t sin[4] = {0, 0, 0, 0}, sout[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if (j >= L.n) break;
    const uint64_t cu = L.c[j];
    int ch = L.e[j] >> 4;
    if (ch) add_sel4(sin, nib_idx(ch), cu);
    ch = L.e[j] & 0xF;
    if (ch) add_sel4(sout, nib_idx(ch), cu);
  }
  uint64_t win_in[4], win_out[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (kpal) {
      win_in[q] = sin[q] + sout[3 - q];
      win_out[q] = sout[q] + sin[3 - q];
    } else {
      const bool split_in = (ops.K & 1) && lastb == 3 - q && high;
      const bool split_out = (ops.K & 1) && firstb == 3 - q && low;
      win_in[q] = split_in ? sin[q] >> 1 : sin[q];
      win_out[q] = split_out ? sout[q] >> 1 : sout[q];
    }
    win_in[q] = win_in[q] < cnt_max ? win_in[q] : cnt_max;
    win_out[q] = win_out[q] < cnt_max ? win_out[q] : cnt_max;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) c[q] = 0;
  bool have = false;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if (j >= L.n) break;
    const int in = L.e[j] >> 4, out = L.e[j] & 0xF;
    const uint64_t wi = get_sel4(win_in, nib_idx(in), 0), wo = get_sel4(win_out, nib_idx(out), 0);
    const bool in_ok = in && wi >= t.t1lo && wi < t.t1hi;
    const bool out_ok = out && wo >= t.t1lo && wo < t.t1hi;
    if (!in_ok && !out_ok) continue;
    uint8_t e = L.e[j];
    if (!in_ok) e &= 0x0F;
    if (!out_ok) e &= 0xF0;
    node_update(c, e, (uint32_t)(L.c[j] < cnt_max ? L.c[j] : cnt_max));
    have = true;
  }
  return have;
}

/* IsKplus1Palindrome (debruijn_graph_filters.hpp:57-95) with
 * is_k1_k1pair_rc_palindrome (kmer_traits.hpp:110-127). */
template <typename KT>
__device__ __forceinline__ bool kplus1_pal(const kops<KT>& ops, KT k, const uint16_t* c) {
  if (!(ops.K & 1)) return false;
  int firstb = ops.first(k), lastb = ops.last(k);
  bool low = ops.km1_low(k);
  if (low)
    for (int q = 0; q < 4; ++q)
      if (c[q] > 0 && (firstb == 3 - q || firstb == q)) return true;
  bool high = ops.km1_high(k);
  if (high)
    for (int q = 0; q < 4; ++q)
      if (c[4 + q] > 0 && (lastb == 3 - q || lastb == q)) return true;
  return false;
}

__device__ __forceinline__ int n_out_edges(const uint16_t* c) { return (c[0] > 0) + (c[1] > 0) + (c[2] > 0) + (c[3] > 0); }
__device__ __forceinline__ int n_in_edges(const uint16_t* c) { return (c[4] > 0) + (c[5] > 0) + (c[6] > 0) + (c[7] > 0); }

/* Classify a surviving node (IsBranchPoint / IsChainNode,
 * debruijn_graph_filters.hpp:142-220). */
template <typename KT>
__device__ __forceinline__ NodeRec make_node(const kops<KT>& ops, KT k, const uint16_t (&c)[9]) {
  bool pal = ops.is_rc_palindrome(k) || kplus1_pal<KT>(ops, k, c);
  int ni = n_in_edges(c), no = n_out_edges(c);
  NodeRec rec;
#pragma unroll
  for (int q = 0; q < 9; ++q) rec.c[q] = c[q];
  rec.cls = (ni >= 2 || no >= 2 || pal) ? CLS_BRANCH : CLS_CHAIN;
  rec.pad = 0;
  rec.idx = 0xFFFFFFFFu;
  return rec;
}

/* Node lists written by the filter: chain and branch slots (the index in
 * its list goes into the node record) and, in a batch, each node's job. */
struct NodeLists {
  uint64_t* chain_slot;
  uint64_t* branch_slot;
  uint32_t* chain_job;
  uint32_t* branch_job;
  uint64_t cap; /* capacity of each list */
  const uint64_t* seg;
  uint32_t n_jobs;
  /* optional, by chain-list position: the chain node's k-mer (KT) and its
   * edge byte (in-edge base + 1 | (out-edge base + 1) << 4, 0 = none), so
   * the chain extraction reads them in order instead of the node's slot */
  void* chain_kmer;
  uint8_t* chain_edge;
};

/* A chain node's edge byte from its counters (at most one in and one out
 * edge). */
__device__ __forceinline__ uint8_t chain_edge_byte(const uint16_t (&c)[9]) {
  uint32_t ci = 0, co = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ci = c[4 + q] > 0 ? (uint32_t)q + 1 : ci;
    co = c[q] > 0 ? (uint32_t)q + 1 : co;
  }
  return (uint8_t)(ci | (co << 4));
}

/* Write the record in place (sub area of the slot) with its list index,
 * mark the slot present and append it to its list at position pos. */
template <typename KT>
__device__ __forceinline__ void store_node(Slot<KT>* T, uint64_t s, NodeRec rec, unsigned long long pos,
                                           const NodeLists& nl, Counters* ctr, KT k = KT(0)) {
  if (pos >= nl.cap || pos > 0x7FFFFFFFULL) {
    atomicOr(&ctr->err, ERR_LIST_FULL);
    return;
  }
  rec.idx = (uint32_t)pos;
  Slot<KT>* p = T + s;
  *(NodeRec*)&p->sub[0] = rec;
  uint64_t* fw = SlotOps<KT>::flag_word(p);
  *fw = *fw | F_PRESENT;
  const bool chain = rec.cls == CLS_CHAIN;
  (chain ? nl.chain_slot : nl.branch_slot)[pos] = s;
  if (nl.n_jobs > 1) (chain ? nl.chain_job : nl.branch_job)[pos] = job_of(nl.seg, nl.n_jobs, s);
  if (chain && nl.chain_kmer) {
    ((KT*)nl.chain_kmer)[pos] = k;
    nl.chain_edge[pos] = chain_edge_byte(rec.c);
  }
}

}  // namespace br

#endif
