/*
 * bruno_hip.hip -- MI355X (gfx950) implementation of the C ABI in
 * include/bruno_hip.h: job state in HBM, phase orchestration on one HIP
 * stream, and the graph / chain / list-ranking / output kernels.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/bruno_hip.h"
#include "kernels.h"
#include "regions.h"
#include "kmer.h"
#include "pool.h"
#include "synth.h"
#include "table.h"

using namespace br;

namespace {

/* ================================================================== */
/* filter + classify (one thread per table slot)                        */
/* ================================================================== */
constexpr int FILTER_SPT = 4; /* slots per thread per reservation */

/* Flag word and k-mer count (sum of the in-slot k2 counts) of one slot,
 * read with 16-byte streaming loads. */
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void scan_slot(const Slot<uint64_t>* p, uint64_t& fw, uint64_t& kf) {
  const u64x2* q = (const u64x2*)p;
  const u64x2 a = __builtin_nontemporal_load(q), b = __builtin_nontemporal_load(q + 1);
  fw = a.x;
  const uint64_t v[3] = {a.y, b.x, b.y};
  kf = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) kf += v[j] != EMPTY64 ? (v[j] & 0xFFFFFFFFFFFFULL) : 0;
}
__device__ __forceinline__ void scan_slot(const Slot<u128>* p, uint64_t& fw, uint64_t& kf) {
  const u64x2* q = (const u64x2*)p;
  const u64x2 a = __builtin_nontemporal_load(q), b = __builtin_nontemporal_load(q + 1);
  const u64x2 c = __builtin_nontemporal_load(q + 2), d = __builtin_nontemporal_load(q + 3);
  fw = a.x;
  const uint64_t v[5] = {b.y, c.x, c.y, d.x, d.y};
  kf = 0;
#pragma unroll
  for (int j = 0; j < 5; ++j) kf += v[j] != EMPTY64 ? (v[j] & 0xFFFFFFFFFFFFULL) : 0;
}

/* Output indices staged in LDS and appended to a global list with one
 * atomic per flush (the filter scan keeps a few percent of the slots, and a
 * global atomic per tile would serialise on the list counter). */
template <int CAP, typename V = uint64_t>
struct LdsList {
  V v[CAP];
  uint32_t n;
  unsigned long long base;
};
template <int CAP, typename V>
__device__ __forceinline__ void lds_flush(LdsList<CAP, V>& L, V* out, uint64_t out_cap, unsigned long long* counter,
                                          uint32_t* err) {
  __syncthreads();
  const uint32_t n = L.n;
  if (n == 0) return; /* uniform: every thread read L.n after the barrier */
  if (threadIdx.x == 0) L.base = atomicAdd(counter, (unsigned long long)n);
  __syncthreads();
  const unsigned long long b = L.base;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
    if (b + i < out_cap) out[b + i] = L.v[i];
    else atomicOr(err, ERR_LIST_FULL);
  }
  __syncthreads();
  if (threadIdx.x == 0) L.n = 0;
  __syncthreads();
}

constexpr int SCAN_LDS = 2048; /* staged candidates per block */
constexpr int SPILL_LDS = 1536;

/* Pass 1 of the filter: stream the table once and list the slots that can
 * survive.  Every edge count is at most the k-mer count, so a k-mer seen
 * fewer than k1_lo times has both edges of every k2-mer out of range (flag
 * 3, debruijn_graph_map.hpp:1573) and nothing of it survives; in the
 * unthresholded mode every k-mer is a node.  Slots with spilled k2 entries
 * go to their own list (their count needs the side table). */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_filter_scan(const Slot<KT>* T, uint64_t cap, int thresholding,
                                                       uint64_t t1lo, uint64_t* cand, uint64_t cand_cap,
                                                       uint64_t* spilled, uint64_t spilled_cap, Counters* ctr) {
  constexpr uint32_t TILE = BLOCK * FILTER_SPT;
  static_assert(SCAN_LDS > (int)TILE && SPILL_LDS > (int)TILE, "a tile must fit the staging lists");
  __shared__ LdsList<SCAN_LDS> Lc;
  __shared__ LdsList<SPILL_LDS> Ls;
  if (threadIdx.x == 0) {
    Lc.n = 0;
    Ls.n = 0;
  }
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * TILE;
  for (uint64_t base = (uint64_t)blockIdx.x * TILE; base < cap; base += stride) {
    uint64_t fw[FILTER_SPT], kf[FILTER_SPT];
#pragma unroll
    for (int q = 0; q < FILTER_SPT; ++q) {
      const uint64_t s = base + (uint64_t)q * BLOCK + threadIdx.x;
      fw[q] = EMPTY64;
      kf[q] = 0;
      if (s < cap) scan_slot(T + s, fw[q], kf[q]);
    }
    uint32_t keep = 0, spl = 0;
#pragma unroll
    for (int q = 0; q < FILTER_SPT; ++q) {
      if (fw[q] == EMPTY64) continue;
      if (fw[q] & F_SPILL) spl |= 1u << q;
      else if (!thresholding || kf[q] >= t1lo) keep |= 1u << q;
    }
    if (keep) {
      uint32_t p = atomicAdd(&Lc.n, (uint32_t)__popc(keep));
#pragma unroll
      for (int q = 0; q < FILTER_SPT; ++q)
        if (keep & (1u << q)) Lc.v[p++] = base + (uint64_t)q * BLOCK + threadIdx.x;
    }
    if (spl) {
      uint32_t p = atomicAdd(&Ls.n, (uint32_t)__popc(spl));
#pragma unroll
      for (int q = 0; q < FILTER_SPT; ++q)
        if (spl & (1u << q)) Ls.v[p++] = base + (uint64_t)q * BLOCK + threadIdx.x;
    }
    /* every thread reads both counts between two barriers, so all take the
     * same branch; then keep room for one more tile in each list */
    __syncthreads();
    const uint32_t nc = Lc.n, ns = Ls.n;
    __syncthreads();
    if (nc > SCAN_LDS - TILE) lds_flush(Lc, cand, cand_cap, &ctr->n_cand, &ctr->err);
    if (ns > SPILL_LDS - TILE) lds_flush(Ls, spilled, spilled_cap, &ctr->n_spill, &ctr->err);
  }
  lds_flush(Lc, cand, cand_cap, &ctr->n_cand, &ctr->err);
  lds_flush(Ls, spilled, spilled_cap, &ctr->n_spill, &ctr->err);
}

/* The filter pre-pass from the bits a region build left (regions.h):
 * candidate and spilled slots of a region-layout table in slot order,
 * FILTER_BITS_WPT words per thread per reservation. */
constexpr int FILTER_BITS_WPT = 8;
__global__ __launch_bounds__(BLOCK) void k_filter_bits(const uint64_t* cbits, const uint64_t* sbits, uint64_t n_regions,
                                                       uint64_t stride, uint64_t span, uint64_t* cand,
                                                       uint64_t cand_cap, uint64_t* spilled, uint64_t spilled_cap,
                                                       Counters* ctr) {
  __shared__ ReserveSmem sm;
  __shared__ ReserveSmem sm2;
  const uint64_t n_words = n_regions * stride, used = (span + 63) / 64;
  const uint64_t tile = (uint64_t)BLOCK * FILTER_BITS_WPT;
  for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n_words; base += (uint64_t)gridDim.x * tile) {
    uint64_t cw[FILTER_BITS_WPT], sw[FILTER_BITS_WPT];
    uint32_t nc = 0, ns = 0;
#pragma unroll
    for (int q = 0; q < FILTER_BITS_WPT; ++q) {
      const uint64_t w = base + (uint64_t)q * BLOCK + threadIdx.x;
      cw[q] = sw[q] = 0;
      if (w < n_words && w % stride < used) {
        cw[q] = cbits[w];
        sw[q] = sbits[w];
      }
      nc += __popcll(cw[q]);
      ns += __popcll(sw[q]);
    }
    unsigned long long pc = block_reserve(nc, &ctr->n_cand, sm);
    unsigned long long ps = block_reserve(ns, &ctr->n_spill, sm2);
#pragma unroll
    for (int q = 0; q < FILTER_BITS_WPT; ++q) {
      const uint64_t w = base + (uint64_t)q * BLOCK + threadIdx.x;
      const uint64_t s0 = (w / stride) * span + (w % stride) * 64;
      for (uint64_t m = cw[q]; m; m &= m - 1) {
        if (pc < cand_cap) cand[pc] = s0 + __builtin_ctzll(m);
        ++pc;
      }
      for (uint64_t m = sw[q]; m; m &= m - 1) {
        if (ps < spilled_cap) spilled[ps] = s0 + __builtin_ctzll(m);
        ++ps;
      }
    }
  }
}

/* Node lists after a count with the filter fused: positions [0, e) of a
 * list hold nodes and the unused positions of the build's waves' last
 * chunks (holes[2 w + l] = start << 7 | count for list l, marked
 * LIST_HOLE).  The holes are closed by moving the nodes found above the
 * final count into the holes below it; a moved node's record gets its new
 * index.  k_hole_count: holes below e0 / e1, per list. */
__global__ __launch_bounds__(BLOCK) void k_hole_count(const uint64_t* holes, uint64_t n_waves, uint64_t e0,
                                                      uint64_t e1, unsigned long long* out) {
  const uint64_t w = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  unsigned long long c[2] = {0, 0};
  if (w < n_waves) {
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const uint64_t h = holes[2 * w + l], start = h >> 7, cnt = h & 127, e = l ? e1 : e0;
      c[l] = start < e ? (cnt < e - start ? cnt : e - start) : 0;
    }
  }
  const unsigned long long s0 = wave_sum(c[0]), s1 = wave_sum(c[1]);
  if ((threadIdx.x & 63) == 0) {
    if (s0) atomicAdd(out, s0);
    if (s1) atomicAdd(out + 1, s1);
  }
}
/* The nodes at [n, e) of a list, gathered (one atomic per wave). */
__global__ __launch_bounds__(BLOCK) void k_list_tail(const uint64_t* list, uint64_t n, uint64_t e, uint64_t* tail,
                                                     uint64_t tail_cap, unsigned long long* cnt) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t p0 = n + (uint64_t)blockIdx.x * BLOCK; p0 < e; p0 += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t p = p0 + threadIdx.x;
    const bool v = p < e && list[p] != LIST_HOLE;
    const uint64_t m = __ballot(v);
    uint64_t b = 0;
    if (lane == 0 && m) b = atomicAdd(cnt, (unsigned long long)__popcll(m));
    b = readlane64(b, 0);
    const uint64_t at = b + __popcll(m & ((1ULL << lane) - 1));
    if (v && at < tail_cap) tail[at] = p;
  }
}
/* Holes of list l below n filled from the gathered nodes. */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_list_fill(const uint64_t* holes, uint64_t n_waves, int l, uint64_t* list,
                                                     uint64_t n, const uint64_t* tail,
                                                     const unsigned long long* n_tail, unsigned long long* cnt,
                                                     Slot<KT>* T) {
  const uint64_t w = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (w >= n_waves) return;
  const uint64_t h = holes[2 * w + l], start = h >> 7, c = h & 127;
  const unsigned long long nt = *n_tail;
  for (uint64_t p = start; p < start + c && p < n; ++p) {
    const unsigned long long k = atomicAdd(cnt, 1ULL);
    if (k >= nt) continue; /* more holes than nodes above: counted, and reported by the host */
    const uint64_t s = list[tail[k]];
    list[p] = s;
    ((NodeRec*)&T[s].sub[0])->idx = (uint32_t)p;
  }
}

/* Pass 2: local_insert_by_freqencies on the candidate slots (all k2
 * entries in the slot), node classification, record written in place and
 * the node appended to the chain or branch list. */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_filter_cand(Slot<KT>* T, const uint64_t* cand, uint64_t n_cand, int K,
                                                       int thresholding, Thresh th, NodeLists nl, Counters* ctr) {
  __shared__ ReserveSmem sm;
  __shared__ ReserveSmem sm2;
  kops<KT> ops(K);
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n_cand; base += stride) {
    const uint64_t i = base + threadIdx.x;
    bool kept = false;
    uint64_t s = 0;
    KT k = 0;
    NodeRec rec;
    rec.cls = 0;
    if (i < n_cand) {
      s = cand[i];
      Slot<KT>* p = T + s;
      K2L<Slot<KT>::NSUB> L;
      sub_entries<KT, Slot<KT>::NSUB>(p, L);
      k = SlotOps<KT>::kmer(p);
      const KmerFacts f = kmer_facts(ops, k);
      uint16_t c[9];
      if (thresholding == 2) {
        kept = edge_counts<KT, Slot<KT>::NSUB>(ops, k, L, th, c);
      } else if (thresholding) {
        kept = p2_counts_slot<KT, Slot<KT>::NSUB>(ops, f, L, th, c);
      } else {
        p1_counts(L, c);
        kept = true;
      }
      if (kept) rec = make_node_facts<KT>(ops, f, c);
    }
    const bool ch = kept && rec.cls == CLS_CHAIN, br = kept && rec.cls == CLS_BRANCH;
    unsigned long long pc = block_reserve(ch, &ctr->n_chain, sm);
    unsigned long long pb = block_reserve(br, &ctr->n_branch, sm2);
    if (kept) store_node<KT>(T, s, rec, ch ? pc : pb, nl, ctr, k);
  }
}

/* Slots with spilled k2 entries: gather up to 25 entries from the slot and
 * the side table, then the same filter. */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_filter_spilled(Slot<KT>* T, const Spill* sp, uint64_t spcap, int K,
                                                          int thresholding, Thresh th, const uint64_t* spilled,
                                                          uint64_t n_spilled, NodeLists nl, Counters* ctr) {
  __shared__ ReserveSmem sm;
  __shared__ ReserveSmem sm2;
  kops<KT> ops(K);
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n_spilled; base += stride) {
    const uint64_t i = base + threadIdx.x;
    bool kept = false;
    uint64_t s = 0;
    KT k = 0;
    NodeRec rec;
    rec.cls = 0;
    if (i < n_spilled) {
      s = spilled[i];
      Slot<KT>* p = T + s;
      K2List L;
      gather_k2<KT>(p, s, sp, spcap, L);
      k = SlotOps<KT>::kmer(p);
      uint16_t c[9];
      if (thresholding == 2) {
        kept = edge_counts<KT, 25>(ops, k, L, th, c);
      } else if (thresholding) {
        kept = p2_counts<KT, 25>(ops, k, L, th, c);
      } else {
        p1_counts(L, c);
        kept = true;
      }
      if (kept) rec = make_node<KT>(ops, k, c);
    }
    const bool ch = kept && rec.cls == CLS_CHAIN, br = kept && rec.cls == CLS_BRANCH;
    unsigned long long pc = block_reserve(ch, &ctr->n_chain, sm);
    unsigned long long pb = block_reserve(br, &ctr->n_branch, sm2);
    if (kept) store_node<KT>(T, s, rec, ch ? pc : pb, nl, ctr, k);
  }
}

/* ================================================================== */
/* chain extraction                                                     */
/* ================================================================== */
/* Resolve neighbour k-mer n (on the pointing node's strand) to a chain
 * pointer.  facing_if_canonical: which end of n faces the pointing node
 * when n is itself canonical.  Non-chain targets get a virtual index
 * holding the canonical k-mer (updates to them are dropped, as
 * map.update(..., false, ...) drops keys absent from the chain map). */
template <typename KT>
__device__ uint32_t resolve(const Slot<KT>* T, uint64_t tbase, uint64_t tcap, uint32_t rbits, const kops<KT>& ops,
                            KT n, uint32_t facing_if_canonical, uint64_t n_chain, KT* vkmer, uint64_t vcap,
                            Counters* ctr) {
  KT rn = ops.rc(n);
  bool flipped = !(n < rn);
  KT cn = flipped ? rn : n;
  uint32_t end = flipped ? (facing_if_canonical ^ END_OUT) : facing_if_canonical;
  uint64_t s = SlotOps<KT>::find(T + tbase, tcap, rbits, cn, true);
  if (s != ~0ULL) {
    const NodeRec* nr = (const NodeRec*)&T[tbase + s].sub[0];
    if (nr->cls == CLS_CHAIN) return end | nr->idx;
  }
  unsigned long long v = atomicAdd(&ctr->n_virt, 1ULL);
  if (v >= vcap) {
    atomicOr(&ctr->err, ERR_LIST_FULL);
    return end | IDX_MASK;
  }
  vkmer[v] = cn;
  return end | (uint32_t)(n_chain + v);
}

/* resolve() of a chain node's in neighbour (its OUT end faces the node when
 * canonical) and out neighbour (its IN end) at once: the two probe
 * sequences advance in lockstep, so their table reads are in flight
 * together.  Each pointer is written only when its neighbour exists.
 *
 * It also tests, from the neighbours' records it has read, that the chain
 * graph is well formed for list ranking: every edge of the node (k, chain
 * index self) must be answered by the neighbour's edge back, so that the
 * neighbour's facing end points back at k (a chain neighbour) or gets
 * marked terminal by k_termini (a branch neighbour).  Each end that fails
 * (neighbour absent, no edge back, or an end pointing at itself) adds one
 * to ill. */
template <typename KT>
__device__ void resolve_pair(const Slot<KT>* T, uint64_t tbase, uint64_t tcap, uint32_t rbits, const kops<KT>& ops,
                             KT k, uint32_t self, bool has_in, KT nin, bool has_out, KT nout, uint64_t n_chain,
                             KT* vkmer, uint64_t vcap, Counters* ctr, uint32_t& pin, uint32_t& pout,
                             unsigned long long& ill) {
  const KT ri = ops.rc(nin), ro = ops.rc(nout);
  const bool fi = !(nin < ri), fo = !(nout < ro);
  const KT cin = fi ? ri : nin, cout = fo ? ro : nout;
  const uint32_t ein = fi ? 0u : END_OUT, eout = fo ? END_OUT : 0u;
  const Slot<KT>* B = T + tbase;
  Probe qi(kmer_hash(cin), tcap, rbits), qo(kmer_hash(cout), tcap, rbits);
  const uint64_t lim = tcap < MAX_PROBE ? tcap : MAX_PROBE;
  bool di = !has_in, dout = !has_out;
  uint64_t si = ~0ULL, so = ~0ULL;
  for (uint64_t probe = 0; probe < lim && !(di && dout); ++probe) {
    const int a = SlotOps<KT>::probe_state(B, qi.s, cin, true);
    const int b = SlotOps<KT>::probe_state(B, qo.s, cout, true);
    if (!di) {
      if (a == 1) si = qi.s;
      if (a) di = true;
      else qi.next();
    }
    if (!dout) {
      if (b == 1) so = qo.s;
      if (b) dout = true;
      else qo.next();
    }
  }
  const NodeRec* ni = (const NodeRec*)&B[si == ~0ULL ? 0 : si].sub[0];
  const NodeRec* no = (const NodeRec*)&B[so == ~0ULL ? 0 : so].sub[0];
  const uint8_t cls_i = ni->cls, cls_o = no->cls;
  const uint32_t idx_i = ni->idx, idx_o = no->idx;
  /* the edge back: the in neighbour c.k[..K-2] reaches k with base last(k)
   * out of its OUT end (canonical) or comp(last(k)) into its IN end
   * (flipped); the out neighbour k[1..].c reaches it with first(k) into its
   * IN end or comp(first(k)) out of its OUT end */
  const int lb = ops.last(k), fb = ops.first(k);
  if (has_in) {
    const bool back = si != ~0ULL && (fi ? ni->c[4 + 3 - lb] : ni->c[lb]) > 0;
    const bool loop = si != ~0ULL && cls_i == CLS_CHAIN && idx_i == self && ein == 0u;
    ill += (!back || loop) ? 1 : 0;
  }
  if (has_out) {
    const bool back = so != ~0ULL && (fo ? no->c[3 - fb] : no->c[4 + fb]) > 0;
    const bool loop = so != ~0ULL && cls_o == CLS_CHAIN && idx_o == self && eout == END_OUT;
    ill += (!back || loop) ? 1 : 0;
  }
  auto finish = [&](bool found, uint8_t cls, uint32_t idx, KT cn, uint32_t end) -> uint32_t {
    if (found && cls == CLS_CHAIN) return end | idx;
    unsigned long long v = atomicAdd(&ctr->n_virt, 1ULL);
    if (v >= vcap) {
      atomicOr(&ctr->err, ERR_LIST_FULL);
      return end | IDX_MASK;
    }
    vkmer[v] = cn;
    return end | (uint32_t)(n_chain + v);
  };
  if (has_in) pin = finish(si != ~0ULL, cls_i, idx_i, cin, ein);
  if (has_out) pout = finish(so != ~0ULL, cls_o, idx_o, cout, eout);
}

/* to_simple_biedge (debruijn_chain_node.hpp:84-141): one neighbour per side
 * at dist 1, else term_mask. */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_chain_init(const Slot<KT>* T, uint64_t cap, uint32_t rbits, int K,
                                                      const uint64_t* chain_slot,
                                                      const uint32_t* chain_job, const uint64_t* seg, uint32_t n_jobs,
                                                      uint64_t n_chain, KT* ckmer, const uint8_t* cedge, LRNode* lr,
                                                      KT* vkmer, uint64_t vcap, Counters* ctr) {
  kops<KT> ops(K);
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  unsigned long long ill = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_chain; i += stride) {
    uint64_t tbase = 0, tcap = cap;
    if (n_jobs > 1) {
      const uint32_t job = chain_job[i];
      tbase = seg[job];
      tcap = seg[job + 1] - tbase;
    }
    /* a chain node has at most one in and one out edge: from the filter's
     * k-mer and edge byte when it left them, else from the node's slot */
    KT k;
    int ci = -1, co = -1;
    if (cedge) {
      k = ckmer[i];
      const uint32_t eb = cedge[i];
      ci = (int)(eb & 15) - 1;
      co = (int)(eb >> 4) - 1;
    } else {
      const Slot<KT>* p = T + chain_slot[i];
      k = SlotOps<KT>::kmer(p);
      const NodeRec* nr = (const NodeRec*)&p->sub[0];
      for (int c = 0; c < 4; ++c) {
        if (nr->c[4 + c] > 0) ci = c;
        if (nr->c[c] > 0) co = c;
      }
    }
    LRNode st;
    st.p[0] = st.p[1] = IDX_MASK;
    st.d[0] = st.d[1] = TERM_MASK;
    st.mail[0] = st.mail[1] = 0ULL;
    /* in neighbour c.k[0..K-2]: when canonical its OUT end faces k; out
     * neighbour k[1..].c: when canonical its IN end faces k */
    resolve_pair<KT>(T, tbase, tcap, rbits, ops, k, (uint32_t)i, ci >= 0, ops.in_nbr(k, ci < 0 ? 0 : ci), co >= 0,
                     ops.out_nbr(k, co < 0 ? 0 : co), n_chain, vkmer, vcap, ctr, st.p[0], st.p[1], ill);
    if (ci >= 0) st.d[0] = 1;
    if (co >= 0) st.d[1] = 1;
    if (!cedge) ckmer[i] = k;
    lr[i] = st;
  }
  block_add(ill, &ctr->lr_ill);
}

/* setup_chain_termini: each branch node marks the facing end of its chain
 * neighbours as pointing to a branch (dist 0). */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_termini(const Slot<KT>* T, uint64_t cap, uint32_t rbits, int K,
                                                   const uint64_t* branch_slot,
                                                   const uint32_t* branch_job, const uint64_t* seg, uint32_t n_jobs,
                                                   uint64_t n_branch, LRNode* lr, KT* bkmer) {
  kops<KT> ops(K);
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_branch; i += stride) {
    const Slot<KT>* p = T + branch_slot[i];
    uint64_t tbase = 0, tcap = cap;
    if (n_jobs > 1) {
      const uint32_t job = branch_job[i];
      tbase = seg[job];
      tcap = seg[job + 1] - tbase;
    }
    KT k = SlotOps<KT>::kmer(p);
    bkmer[i] = k;
    const NodeRec* nr = (const NodeRec*)&p->sub[0];
    for (int side = 0; side < 2; ++side)
      for (int c = 0; c < 4; ++c) {
        /* out neighbours get (t, IN); in neighbours get (t, OUT); lex_less
         * flips the orientation with the key (debruijn_chain_operations.hpp
         * :491-497). */
        if (!(nr->c[(side ? 4 : 0) + c] > 0)) continue;
        KT n = side ? ops.in_nbr(k, c) : ops.out_nbr(k, c);
        KT rn = ops.rc(n);
        bool flipped = !(n < rn);
        KT cn = flipped ? rn : n;
        uint64_t s = SlotOps<KT>::find(T + tbase, tcap, rbits, cn, true);
        if (s == ~0ULL) continue;
        const NodeRec* q = (const NodeRec*)&T[tbase + s].sub[0];
        if (q->cls != CLS_CHAIN) continue;
        int orient_in = (side == 0);  /* IN for out neighbours */
        if (flipped) orient_in = !orient_in;
        lr[q->idx].d[orient_in ? 0 : 1] = 0;
      }
  }
}

/* ================================================================== */
/* list ranking: one synchronous round = send + recv                    */
/* ================================================================== */
__global__ __launch_bounds__(BLOCK) void k_lr_init_list(const LRNode* lr, uint64_t n_chain, uint32_t* list,
                                                        const uint32_t* cjob, unsigned long long* jctr, Counters* ctr) {
  __shared__ ReserveSmem sm;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n_chain; base += stride) {
    const uint64_t i = base + threadIdx.x;
    bool u = (i < n_chain) && is_unfinished(lr[i].d[0], lr[i].d[1]);
    if (u && jctr) atomicAdd(&jctr[3 * (uint64_t)cjob[i] + 1], 1ULL);
    unsigned long long pos = block_reserve(u, &ctr->lr_next, sm);
    if (u) list[pos] = (uint32_t)i;
  }
}

__device__ __forceinline__ void lr_post(LRNode* lr, uint64_t n_chain, uint32_t target, unsigned long long msg,
                                        Counters* ctr, unsigned long long& conflicts) {
  uint32_t j = target & IDX_MASK;
  if (j >= n_chain) return; /* not in the chain map: dropped */
  int e = (target & END_OUT) ? 1 : 0;
  /* chain_update only acts on an end that still points to a chain node and
   * only with a distance >= the current one; the end's distance never
   * decreases within a round, so such updates are no-ops in every
   * application order and are not posted. */
  const uint32_t dr = lr[j].d[e];
  if (!points_to_chain(dr)) return;
  if (cdist((uint32_t)(msg >> 32)) < cdist(dr)) return;
  /* Several updates to one end in a round (list_rank sends from every
   * non-terminal side) are applied by the reference in hash-table order.
   * Keeping the largest (dist, ptr) word is the result of applying them in
   * decreasing order: the first one applied is accepted, and a marked
   * (terminal) distance or a larger plain distance then blocks or does not
   * change the rest. */
  unsigned long long old = atomicCAS(&lr[j].mail[e], 0ULL, msg);
  if (old != 0ULL) {
    ++conflicts;
    atomicMax(&lr[j].mail[e], msg);
  }
}

/* variant 0: send on every non-terminal side (list_rank :858-885);
 * variant 1: only on sides pointing to a chain node (:1130-1155). */
__global__ __launch_bounds__(BLOCK) void k_lr_send(LRNode* lr, uint64_t n_chain, const uint32_t* list, uint64_t n,
                                                   int variant, const uint32_t* cjob, const uint8_t* active,
                                                   Counters* ctr) {
  unsigned long long conflicts = 0;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    const uint32_t t = list[i];
    /* a finished job of a batch takes no further rounds */
    if (active && !active[cjob[t]]) continue;
    const LRNode st = lr[t];
    const uint32_t din = st.d[0], dout = st.d[1];
    const uint32_t dist = cdist(din) + cdist(dout);
    bool send_in = variant ? points_to_chain(din) : !is_terminal(din);
    bool send_out = variant ? points_to_chain(dout) : !is_terminal(dout);
    if (send_in) {
      /* to the in target's facing end: our out target (or ourselves, IN end) */
      uint32_t ptr = is_terminal(dout) ? t : st.p[1];
      uint32_t dd = points_to_chain(dout) ? dist : (dist | TERM_MASK);
      lr_post(lr, n_chain, st.p[0], ((unsigned long long)dd << 32) | ptr, ctr, conflicts);
    }
    if (send_out) {
      uint32_t ptr = is_terminal(din) ? (t | END_OUT) : st.p[0];
      uint32_t dd = points_to_chain(din) ? dist : (dist | TERM_MASK);
      lr_post(lr, n_chain, st.p[1], ((unsigned long long)dd << 32) | ptr, ctr, conflicts);
    }
  }
  block_add(conflicts, &ctr->lr_conflict);
}

/* Per-job counters of a batch round (jctr[3 * job + {0, 1, 2}]: updates,
 * unfinished nodes, cycle-like nodes); a single job uses ctr. */
__global__ __launch_bounds__(BLOCK) void k_lr_recv(LRNode* lr, const uint32_t* list, uint64_t n, uint32_t* next,
                                                   uint32_t maxd, const uint32_t* cjob, const uint8_t* active,
                                                   unsigned long long* jctr, Counters* ctr) {
  __shared__ ReserveSmem sm;
  unsigned long long count = 0, cyc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n; base += stride) {
    const uint64_t i = base + threadIdx.x;
    bool unf = false;
    uint32_t t = 0;
    if (i < n && !(active && !active[cjob[list[i]]])) {
      t = list[i];
      unsigned long long my_count = 0;
      LRNode st = lr[t];
      bool dirty = false;
      for (int e = 0; e < 2; ++e) {
        unsigned long long m = st.mail[e];
        if (!m) continue;
        dirty = true;
        st.mail[e] = 0ULL;
        uint32_t yd = (uint32_t)(m >> 32), yp = (uint32_t)m;
        uint32_t dist = cdist(yd), distx = cdist(st.d[e]);
        /* chain_update (debruijn_chain_operations.hpp:192-198, :245-251) */
        if (points_to_chain(st.d[e])) {
          if (dist > distx) st.p[e] = yp;
          if (dist >= distx) {
            st.d[e] = yd;
            ++my_count;
          }
        }
      }
      if (dirty) lr[t] = st;
      unf = is_unfinished(st.d[0], st.d[1]);
      const bool my_cyc = unf && cdist(st.d[0]) == maxd && cdist(st.d[1]) == maxd;
      if (jctr) {
        unsigned long long* jc = jctr + 3 * (uint64_t)cjob[t];
        if (my_count) atomicAdd(jc, my_count);
        if (unf) atomicAdd(jc + 1, 1ULL);
        if (my_cyc) atomicAdd(jc + 2, 1ULL);
      }
      count += my_count;
      cyc += my_cyc;
    }
    unsigned long long pos = block_reserve(unf, &ctr->lr_next, sm);
    if (unf) next[pos] = t;
  }
  block_add(count, &ctr->lr_count);
  __syncthreads();
  block_add(cyc, &ctr->lr_cycle);
}

/* ================================================================== */
/* list ranking by ruler contraction (well-formed chain graphs)         */
/* ================================================================== */
/* On a chain graph where every end pointing at a chain node is pointed back
 * at, the chain nodes form disjoint paths and cycles, and the end state of
 * bruno's pointer jumping is fixed: every non-terminal end points at the
 * chain end it leads to, at its exact distance, marked as pointing to a
 * terminal; cycle nodes end with both distances 2^T.  The number of rounds
 * T is the largest ceil(log2 L) over paths of L >= 2 nodes, at least 1 when
 * a cycle exists.  This computes that state without the rounds: rulers
 * (chain ends plus sampled nodes) walk to the next ruler, heads walk the
 * ruler graph, then every node reads its two chain ends off its rulers. */
constexpr uint32_t NONE32 = 0xFFFFFFFFu;
constexpr uint32_t WALK_MAX = 1u << 16;
constexpr uint32_t ERR_NOT_WELL_FORMED = 16;
constexpr uint32_t ERR_WALK = 32;
constexpr uint32_t ERR_NO_NODE = 64; /* an output k-mer without its filtered-graph record */

__device__ __forceinline__ uint32_t end_bit(int e) { return e ? END_OUT : 0u; }

constexpr int RULER_LDS = 2048; /* rulers staged per block before an append */

/* ruler_mask + 1 = sampling period of the rulers (a power of two).  Marks
 * the rulers (chain ends and sampled nodes) and lists them (LDS-staged
 * appends).  (Whether the chain graph is well formed was counted by
 * k_chain_init, which read every neighbour's record anyway.) */
__global__ __launch_bounds__(BLOCK) void k_lr_rulers(const LRNode* lr, uint64_t n_chain, uint64_t* ruler,
                                                     uint64_t ruler_mask, uint32_t* rlist, uint64_t rlist_cap,
                                                     Counters* ctr) {
  __shared__ LdsList<RULER_LDS, uint32_t> Lr;
  if (threadIdx.x == 0) Lr.n = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n_chain; base += stride) {
    const uint64_t t = base + threadIdx.x;
    bool is_ruler = false;
    if (t < n_chain) {
      const LRNode st = lr[t];
      const bool head = is_terminal(st.d[0]) || is_terminal(st.d[1]);
      const bool sampled = (fmix64(t * 0x9e3779b97f4a7c15ULL + 0x1234567ULL) & ruler_mask) == 0;
      is_ruler = head || sampled;
    }
    /* the ruler bitmap, one word per wave (a wave covers 64 aligned nodes) */
    const uint64_t word = __ballot(is_ruler);
    if ((threadIdx.x & 63) == 0 && t < n_chain) ruler[t >> 6] = word;
    if (is_ruler) Lr.v[atomicAdd(&Lr.n, 1u)] = (uint32_t)t;
    /* all threads read the count between two barriers, then keep room for
     * one more tile */
    __syncthreads();
    const uint32_t nr = Lr.n;
    __syncthreads();
    if (nr > RULER_LDS - BLOCK) lds_flush(Lr, rlist, rlist_cap, &ctr->lr_next, &ctr->err);
  }
  lds_flush(Lr, rlist, rlist_cap, &ctr->lr_next, &ctr->err);
}

struct WalkRec {
  uint32_t who;   /* ruler (node walk) or chain end (ruler walk) */
  uint32_t edist; /* bit 31: end, low 31 bits: distance */
};
/* A node walk's record kept in the node's own mailbox word of the end it
 * was entered through (the mailboxes serve only the round-by-round path,
 * which finds them zero): (end | distance) << 32 | ruler number; zero = no
 * walk reached it (a real walk has distance >= 1). */
__device__ __forceinline__ unsigned long long walk_word(uint32_t rid, uint32_t edist) {
  return ((unsigned long long)edist << 32) | rid;
}

/* From every ruler (number rid in the ruler list), along each non-terminal
 * end, to the next ruler.  Each visited node records, in its mailbox of
 * the end it was entered through, the ruler number, the ruler's end the
 * walk left from and the distance: the node's next pointer is read first,
 * so the record goes into a line the walk has just brought in. */
__global__ __launch_bounds__(BLOCK) void k_lr_walk_nodes(LRNode* lr, const uint64_t* ruler, const uint32_t* rlist,
                                                         uint64_t n_rulers, WalkRec* link, Counters* ctr) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < 2 * n_rulers; i += stride) {
    const uint32_t rid = (uint32_t)(i >> 1), r = rlist[rid];
    const int E = (int)(i & 1);
    if (is_terminal(lr[r].d[E])) continue;
    uint32_t ptr = lr[r].p[E];
    uint32_t dist = 0;
    for (;;) {
      const uint32_t j = ptr & IDX_MASK;
      const int e = (ptr & END_OUT) ? 1 : 0;
      ++dist;
      if ((ruler[j >> 6] >> (j & 63)) & 1) {
        link[2 * (uint64_t)r + E] = WalkRec{j, end_bit(e) | dist};
        break;
      }
      ptr = lr[j].p[1 - e];
      lr[j].mail[e] = walk_word(rid, end_bit(E) | dist);
      if (dist >= WALK_MAX) {
        atomicOr(&ctr->err, ERR_WALK);
        break;
      }
    }
  }
}

/* Per ruler number and end Er: what lies past the ruler when a walk left
 * it through Er, i.e. through its other end: the ruler itself when that
 * end is terminal (distance 0 more), else the chain end the ruler walk
 * found that way (NONE32: a cycle).  Dense by ruler number, so the finish
 * pass reads one small table instead of two scattered records. */
__global__ __launch_bounds__(BLOCK) void k_lr_ruler_ends(const LRNode* lr, const uint32_t* rlist, uint64_t n_rulers,
                                                         const WalkRec* info, WalkRec* rend) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= 2 * n_rulers) return;
  const uint32_t r = rlist[i >> 1];
  const int Er = (int)(i & 1);
  rend[i] = is_terminal(lr[r].d[1 - Er]) ? WalkRec{r, end_bit(Er)} : info[2 * (uint64_t)r + (1 - Er)];
}

/* The mailboxes back to empty, for the round-by-round path after a walk
 * that ran past its bound. */
__global__ __launch_bounds__(BLOCK) void k_lr_zero_mail(LRNode* lr, uint64_t n_chain) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t < n_chain; t += stride) {
    lr[t].mail[0] = 0ULL;
    lr[t].mail[1] = 0ULL;
  }
}

__global__ __launch_bounds__(BLOCK) void k_lr_list_heads(const LRNode* lr, const uint32_t* rlist, uint64_t n_rulers,
                                                         uint32_t* heads, Counters* ctr) {
  __shared__ ReserveSmem sm;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < 2 * n_rulers; base += stride) {
    const uint64_t i = base + threadIdx.x;
    bool h = false;
    uint32_t r = 0;
    int E = (int)(i & 1);
    if (i < 2 * n_rulers) {
      r = rlist[i >> 1];
      /* a head end: this end leads into the chain, the other is terminal */
      h = !is_terminal(lr[r].d[E]) && is_terminal(lr[r].d[1 - E]);
    }
    unsigned long long pos = block_reserve(h, &ctr->lr_next, sm);
    if (h) heads[pos] = r | end_bit(E);
  }
}

/* From every head along the ruler graph: each ruler learns, for the end
 * facing the head, which chain end lies that way and how far. */
__global__ __launch_bounds__(BLOCK) void k_lr_walk_rulers(const LRNode* lr, const WalkRec* link, const uint32_t* heads,
                                                          uint64_t n_heads, uint64_t n_rulers, WalkRec* info,
                                                          Counters* ctr) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_heads; i += stride) {
    const uint32_t h = heads[i] & IDX_MASK;
    const int E = (heads[i] & END_OUT) ? 1 : 0;
    uint32_t cur = h;
    int leave = E;
    uint32_t D = 0;
    uint64_t steps = 0;
    for (;;) {
      WalkRec l = link[2 * (uint64_t)cur + leave];
      if (l.who == NONE32) {
        atomicOr(&ctr->err, ERR_WALK);
        break;
      }
      uint32_t r = l.who;
      int e = (l.edist & END_OUT) ? 1 : 0;
      D += l.edist & DIST_MASK;
      /* from r through e lies head h, entered through its end E */
      info[2 * (uint64_t)r + e] = WalkRec{h, end_bit(E) | D};
      if (is_terminal(lr[r].d[1 - e])) {
        info[2 * (uint64_t)h + E] = WalkRec{r, end_bit(e) | D};
        break;
      }
      if (++steps > n_rulers) {
        atomicOr(&ctr->err, ERR_WALK);
        break;
      }
      cur = r;
      leave = 1 - e;
    }
  }
}

__device__ __forceinline__ uint32_t ceil_log2(uint32_t L) { return L <= 1 ? 0 : 32 - __clz(L - 1); }

/* Final state of every chain node; cycle nodes get both distances 2^T
 * (written in a second pass once T is known). */
/* jT[job]: largest ceil(log2 L) over the job's paths; jcyc[job] = 1 when
 * the job has a cycle. */
__global__ __launch_bounds__(BLOCK) void k_lr_finish(LRNode* lr, uint64_t n_chain, const uint64_t* ruler,
                                                     const WalkRec* rend, const WalkRec* info, uint8_t* cyc,
                                                     const uint32_t* cjob, uint32_t n_jobs, uint32_t* jT,
                                                     uint32_t* jcyc, Counters* ctr) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  unsigned long long ncyc = 0, maxr = 0;
  for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t < n_chain; t += stride) {
    LRNode st = lr[t];
    bool is_cycle = false;
    uint32_t np[2], nd[2];
    for (int E = 0; E < 2; ++E) {
      np[E] = st.p[E];
      nd[E] = st.d[E];
      if (is_terminal(st.d[E])) continue;
      uint32_t Z, f, D;
      if ((ruler[t >> 6] >> (t & 63)) & 1) {
        WalkRec w = info[2 * t + E];
        if (w.who == NONE32) {
          is_cycle = true;
          break;
        }
        Z = w.who;
        f = w.edist & END_OUT;
        D = w.edist & DIST_MASK;
      } else {
        /* the walk that reached this end: ruler number, its end, distance */
        const uint32_t rid = (uint32_t)st.mail[E], edist = (uint32_t)(st.mail[E] >> 32);
        const uint32_t d = edist & DIST_MASK;
        if (d == 0) {
          is_cycle = true;
          break;
        }
        /* continue past that ruler, leaving through its other end */
        const WalkRec v = rend[2 * (uint64_t)rid + ((edist & END_OUT) ? 1 : 0)];
        if (v.who == NONE32) {
          is_cycle = true;
          break;
        }
        Z = v.who;
        f = v.edist & END_OUT;
        D = d + (v.edist & DIST_MASK);
      }
      np[E] = Z | f;
      nd[E] = D | TERM_MASK;
    }
    const uint32_t job = n_jobs > 1 ? cjob[t] : 0u;
    if (is_cycle) {
      cyc[t] = 1;
      ++ncyc;
      if (!jcyc[job]) jcyc[job] = 1;
      continue;
    }
    cyc[t] = 0;
    /* path length from a head's point of view: L = 1 + dist to the far end */
    if (is_terminal(st.d[0]) != is_terminal(st.d[1])) {
      uint32_t L = 1 + cdist(is_terminal(st.d[0]) ? nd[1] : nd[0]);
      uint32_t R = ceil_log2(L);
      if (R > maxr) maxr = R;
      if (n_jobs > 1 && R > jT[job]) atomicMax(&jT[job], R);
    }
    st.p[0] = np[0];
    st.p[1] = np[1];
    st.d[0] = nd[0];
    st.d[1] = nd[1];
    lr[t] = st;
  }
  block_add(ncyc, &ctr->lr_cycle);
  typedef hipcub::BlockReduce<unsigned long long, BLOCK> BR;
  __shared__ typename BR::TempStorage tmp;
  __syncthreads();
  unsigned long long m = BR(tmp).Reduce(maxr, hipcub::Max());
  if (n_jobs == 1 && threadIdx.x == 0 && m) atomicMax(&jT[0], (uint32_t)m);
}

/* cycle nodes end with both distances 2^T of their job: jcm[job] */
__global__ __launch_bounds__(BLOCK) void k_lr_cycles(LRNode* lr, uint64_t n_chain, const uint8_t* cyc,
                                                     const uint32_t* cjob, uint32_t n_jobs, const uint32_t* jcm) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t < n_chain; t += stride)
    if (cyc[t]) {
      const uint32_t maxd = jcm[n_jobs > 1 ? cjob[t] : 0u];
      lr[t].d[0] = maxd;
      lr[t].d[1] = maxd;
    }
}

/* separate_cycles (when unfinished nodes remain) and separate_isolated.
 * jcm[job] = 2^T of the job when its cycles are separated, else 0 (2^T >= 2
 * whenever cycles are separated).  jsep[2 * job + {0, 1}] counts cycle and
 * isolated nodes per job. */
__global__ __launch_bounds__(BLOCK) void k_separate(const LRNode* lr, uint64_t n_chain, const uint32_t* cjob,
                                                    uint32_t n_jobs, const uint32_t* jcm, uint8_t* alive,
                                                    unsigned long long* jsep) {
  unsigned long long nc = 0, ni = 0;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_chain; i += stride) {
    uint32_t din = lr[i].d[0], dout = lr[i].d[1];
    const uint32_t job = n_jobs > 1 ? cjob[i] : 0u;
    const uint32_t maxd = jcm[job];
    uint8_t a = 1;
    if (maxd && cdist(din) == maxd && cdist(dout) == maxd) {
      a = 0;
      if (n_jobs > 1) atomicAdd(&jsep[2 * (uint64_t)job], 1ULL);
      else ++nc;
    } else if (points_to_self(din) && points_to_self(dout)) {
      a = 0;
      if (n_jobs > 1) atomicAdd(&jsep[2 * (uint64_t)job + 1], 1ULL);
      else ++ni;
    }
    alive[i] = a;
  }
  if (n_jobs == 1) {
    block_add(nc, &jsep[0]);
    __syncthreads();
    block_add(ni, &jsep[1]);
  }
}

/* ================================================================== */
/* output                                                               */
/* ================================================================== */
/* k-mer on the pointing node's strand from a pointer. */
template <typename KT>
__device__ __forceinline__ KT ptr_kmer(const kops<KT>& ops, uint32_t ptr, const KT* ckmer, uint64_t n_chain,
                                       const KT* vkmer, bool in_side) {
  uint32_t j = ptr & IDX_MASK;
  KT c = (j < n_chain) ? ckmer[j] : vkmer[j - n_chain];
  bool out_end = (ptr & END_OUT) != 0;
  /* in side: the target is on our strand iff its OUT end faces us;
   * out side: iff its IN end faces us. */
  bool same = in_side ? out_end : !out_end;
  return same ? c : ops.rc(c);
}

template <typename KT>
struct RankedRec {
  KT k, rep;
  uint32_t dist, head, len, job;
};

/* to_listranked_chain_node (debruijn_chain_operations.hpp:261-310) for
 * chain node i: the node on the rep's strand, the rep (the smaller end
 * k-mer), the distance from the rep end, plus the chain node at the rep end
 * (NONE32 when that end is not a chain node) and 1 + the distance to the
 * far end (the chain's length when the node is the head). */
template <typename KT>
__device__ __forceinline__ RankedRec<KT> ranked_one(const kops<KT>& ops, const LRNode* lr, const KT* ckmer,
                                                    uint64_t n_chain, const KT* vkmer, uint64_t i) {
  RankedRec<KT> r;
  const LRNode st = lr[i];
  const KT x = ckmer[i];
  const KT L = is_terminal(st.d[0]) ? x : ptr_kmer<KT>(ops, st.p[0], ckmer, n_chain, vkmer, true);
  const KT R = is_terminal(st.d[1]) ? ops.rc(x) : ops.rc(ptr_kmer<KT>(ops, st.p[1], ckmer, n_chain, vkmer, false));
  const bool sense = !(R < L);
  r.k = sense ? x : ops.rc(x);
  r.rep = sense ? L : R;
  const int rs = sense ? 0 : 1;
  r.dist = cdist(st.d[rs]);
  const uint32_t z = st.p[rs] & IDX_MASK;
  r.head = is_terminal(st.d[rs]) ? (uint32_t)i : (z < n_chain ? z : NONE32);
  r.len = cdist(st.d[1 - rs]) + 1;
  return r;
}

/* Every alive chain node, compacted; RANKED_ITEMS nodes per thread share one
 * output reservation (one global atomic per block tile).  out_job,
 * out_head and out_len are optional. */
constexpr int RANKED_ITEMS = 4;
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_ranked(const LRNode* lr, const KT* ckmer, uint64_t n_chain, const KT* vkmer,
                                                  const uint8_t* alive, int K, const uint32_t* cjob, KT* out_k,
                                                  KT* out_rep, uint32_t* out_dist, uint32_t* out_job,
                                                  uint32_t* out_head, uint32_t* out_len, Counters* ctr) {
  __shared__ ReserveSmem sm;
  kops<KT> ops(K);
  const uint64_t tile = (uint64_t)BLOCK * RANKED_ITEMS;
  for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n_chain; base += (uint64_t)gridDim.x * tile) {
    RankedRec<KT> r[RANKED_ITEMS];
    bool a[RANKED_ITEMS];
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < RANKED_ITEMS; ++q) {
      const uint64_t i = base + (uint64_t)q * BLOCK + threadIdx.x;
      a[q] = i < n_chain && alive[i];
      if (a[q]) {
        r[q] = ranked_one<KT>(ops, lr, ckmer, n_chain, vkmer, i);
        r[q].job = cjob ? cjob[i] : 0u;
      }
      m += a[q];
    }
    unsigned long long pos = block_reserve(m, &ctr->n_out, sm);
#pragma unroll
    for (int q = 0; q < RANKED_ITEMS; ++q)
      if (a[q]) {
        out_k[pos] = r[q].k;
        out_rep[pos] = r[q].rep;
        out_dist[pos] = r[q].dist;
        if (out_job) out_job[pos] = r[q].job;
        if (out_head) {
          out_head[pos] = r[q].head;
          out_len[pos] = r[q].len;
        }
        ++pos;
      }
  }
}

/* positions of the chain heads (dist 0) among the listranked nodes;
 * HEAD_ITEMS positions per thread share one reservation */
constexpr int HEAD_ITEMS = 16;
__global__ __launch_bounds__(BLOCK) void k_list_heads_out(const uint32_t* out_dist, uint64_t n, uint32_t* heads,
                                                          Counters* ctr) {
  __shared__ ReserveSmem sm;
  const uint64_t tile = (uint64_t)BLOCK * HEAD_ITEMS;
  for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
    uint32_t hm = 0, m = 0;
#pragma unroll
    for (int q = 0; q < HEAD_ITEMS; ++q) {
      const uint64_t p = base + (uint64_t)q * BLOCK + threadIdx.x;
      const bool h = p < n && out_dist[p] == 0;
      hm |= (uint32_t)h << q;
      m += h;
    }
    unsigned long long at = block_reserve(m, &ctr->n_cand, sm);
#pragma unroll
    for (int q = 0; q < HEAD_ITEMS; ++q)
      if (hm >> q & 1u) heads[at++] = (uint32_t)(base + (uint64_t)q * BLOCK + threadIdx.x);
  }
}

/* per chain, by the node at its rep end */
template <typename KT>
struct HeadMeta {
  KT rep;
  uint32_t off, len, job, pad;
};

/* for the heads in (job, rep) order (hs[r] = position of the r-th head):
 * each chain's range [coff[r], coff[r] + len) and key, stored by head
 * node; two heads with one key count as bad */
template <typename KT>
__global__ void k_head_meta(const uint32_t* hs, const uint64_t* coff, const uint32_t* out_head,
                            const uint32_t* out_len, const KT* out_rep, const uint32_t* out_job, uint64_t nh,
                            uint64_t n_chain, HeadMeta<KT>* meta, unsigned long long* bad) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nh) return;
  const uint32_t p = hs[r];
  const uint32_t node = out_head[p];
  const uint32_t job = out_job ? out_job[p] : 0u;
  const KT rep = out_rep[p];
  bool ok = node < n_chain;
  if (r > 0) {
    const uint32_t q = hs[r - 1];
    ok = ok && !(out_rep[q] == rep && (out_job ? out_job[q] : 0u) == job);
  }
  if (!ok) {
    atomicAdd(bad, 1ULL);
    return;
  }
  HeadMeta<KT> m;
  m.rep = rep;
  m.off = (uint32_t)coff[r];
  m.len = out_len[p];
  m.job = job;
  m.pad = 0;
  meta[node] = m;
}

__global__ void k_take_len(const uint32_t* hs, const uint32_t* out_len, uint64_t nh, uint64_t* len) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nh) len[r] = out_len[hs[r]];
}

/* A listranked node at its place in output order: its k-mer (K on the
 * rep strand), its index among the listranked nodes and its distance. */
template <typename KT>
struct SortedRec {
  KT k;
  uint32_t p, dist;
};

/* every node to its chain's offset + its dist, as one record (its k-mer,
 * index and dist), so the output order's arrays and the FASTA writer read
 * in sequence; bad counts nodes whose head is missing, whose key differs
 * from the head's, or whose dist is past the chain's end (meta.len is 0
 * for a node that heads no chain) */
template <typename KT>
__global__ void k_scatter_ranked(const uint32_t* out_head, const uint32_t* out_dist, const KT* out_rep, const KT* out_k,
                                 const uint32_t* out_job, uint64_t n, uint64_t n_chain, const HeadMeta<KT>* meta,
                                 SortedRec<KT>* sr, unsigned long long* bad) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t h = out_head[p];
  bool ok = h < n_chain;
  if (ok) {
    const HeadMeta<KT> m = meta[h];
    const uint32_t d = out_dist[p];
    const uint64_t pos = (uint64_t)m.off + d;
    ok = d < m.len && pos < n && m.rep == out_rep[p] && m.job == (out_job ? out_job[p] : 0u);
    if (ok) sr[pos] = SortedRec<KT>{out_k[p], (uint32_t)p, d};
  }
  if (!ok) atomicAdd(bad, 1ULL);
}

/* perm and the sorted distances from the placed records; positions no node
 * was placed at (the records are pre-filled with ones) count as bad */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_sorted_split(const SortedRec<KT>* sr, uint64_t n, uint32_t* perm,
                                                        uint32_t* dist_sorted, unsigned long long* bad) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  uint32_t holes = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    const SortedRec<KT> r = sr[i];
    perm[i] = r.p;
    dist_sorted[i] = r.dist;
    holes += r.p == ~0u;
  }
  if (holes) atomicAdd(bad, (unsigned long long)holes);
}

template <typename KT>
__global__ void k_gather_key(const KT* src, const uint32_t* perm, uint64_t* dst, uint64_t n, int word) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    KT v = src[perm[i]];
    dst[i] = word ? (uint64_t)((u128)v >> 64) : (uint64_t)v;
  }
}
template <typename KT>
__global__ void k_gather_kmer(const KT* src, const uint32_t* perm, KT* dst, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
__global__ void k_add_u64(unsigned long long* dst, const unsigned long long* src, uint32_t n) {
  if (threadIdx.x < n) dst[threadIdx.x] += src[threadIdx.x];
}
__global__ void k_gather_u32(const uint32_t* src, const uint32_t* perm, uint32_t* dst, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
/* start[q] = first position of job q in a job-sorted array (start[n_jobs]
 * = n): position i starts every job in (job[i-1], job[i]]. */
__global__ void k_seg_bounds(const uint32_t* job_sorted, uint64_t n, uint32_t n_jobs, uint64_t* start) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int64_t prev = i ? (int64_t)job_sorted[i - 1] : -1;
  const int64_t cur = i < n ? (int64_t)job_sorted[i] : (int64_t)n_jobs;
  for (int64_t q = prev + 1; q <= cur; ++q) start[q] = i;
}
__global__ void k_take64(const uint64_t* src, const uint64_t* idx, uint64_t n, uint64_t* dst) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_iota(uint32_t* p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

template <typename KT>
__device__ __forceinline__ void write_kmer(char* out, KT v, int K) {
  const char A[4] = {'A', 'C', 'G', 'T'};
  for (int i = 0; i < K; ++i) out[i] = A[(int)((v >> (2 * (K - 1 - i))) & 3)];
}

/* bytes each sorted chain node contributes to .chain.fasta */
__global__ void k_fasta_len(const uint32_t* dist_sorted, uint64_t n, int K, uint64_t* len) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) len[i] = dist_sorted[i] == 0 ? (uint64_t)(2 * K + 3) : 1ULL;
}

/* print_chain_as_fasta (debruijn_chain_operations.hpp:352-385): a node at
 * dist 0 starts a record "\n>" canonical(K) "\n" rep; every other node adds
 * the last base of K (its k-mer on the rep strand).  From the records in
 * output order (sr, after placement by chains) or through perm. */
template <typename KT>
__global__ void k_fasta_write_sorted(const SortedRec<KT>* sr, const KT* out_rep, const uint64_t* off, uint64_t n,
                                     int K, char* buf) {
  kops<KT> ops(K);
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SortedRec<KT> r = sr[i];
  char* o = buf + off[i];
  if (r.dist == 0) {
    o[0] = '\n';
    o[1] = '>';
    write_kmer<KT>(o + 2, ops.canon(r.k), K);
    o[2 + K] = '\n';
    write_kmer<KT>(o + 3 + K, out_rep[r.p], K);
  } else {
    const char A[4] = {'A', 'C', 'G', 'T'};
    o[0] = A[(int)(r.k & 3)];
  }
}
template <typename KT>
__global__ void k_fasta_write(const uint32_t* perm, const KT* out_k, const KT* out_rep, const uint32_t* dist_sorted,
                              const uint64_t* off, uint64_t n, int K, char* buf) {
  kops<KT> ops(K);
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t m = perm[i];
  KT Kv = out_k[m];
  char* o = buf + off[i];
  if (dist_sorted[i] == 0) {
    o[0] = '\n';
    o[1] = '>';
    write_kmer<KT>(o + 2, ops.canon(Kv), K);
    o[2 + K] = '\n';
    write_kmer<KT>(o + 3 + K, out_rep[m], K);
  } else {
    const char A[4] = {'A', 'C', 'G', 'T'};
    o[0] = A[(int)(Kv & 3)];
  }
}

/* print_graph_node_fasta (debruijn_graph_operations.hpp:81-90) */
template <typename KT>
__global__ void k_branch_fasta(const KT* sorted, uint64_t n, int K, char* buf) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  char* o = buf + i * (uint64_t)(2 * K + 3);
  o[0] = '>';
  write_kmer<KT>(o + 1, sorted[i], K);
  o[1 + K] = '\n';
  write_kmer<KT>(o + 2 + K, sorted[i], K);
  o[2 + 2 * K] = '\n';
}

/* ================================================================== */
/* getters                                                              */
/* ================================================================== */
template <typename KT>
__device__ __forceinline__ void put_kmer(bruno_node& o, KT k) {
  o.kmer[0] = (uint64_t)k;
  o.kmer[1] = (uint64_t)((u128)k >> 64);
}

/* P1 counters (before filter) of slots [s0, s1); out == nullptr counts only */
/* Moving a table into a larger one (a job sized from a sketch of its first
 * batches whose later batches hold more distinct k-mers): every k-mer is
 * inserted into the new table with its k2 entries and their counts, so its
 * probe window follows the new size.  remap[old slot] = new slot. */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_move_slots(const Slot<KT>* O, uint64_t ocap, Slot<KT>* T, uint64_t span,
                                                     uint32_t rbits, Spill* sp, uint64_t spcap, uint64_t* remap,
                                                     Counters* ctr) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < ocap; i += (uint64_t)gridDim.x * BLOCK) {
    const Slot<KT>* p = O + i;
    if (!SlotOps<KT>::occupied(p)) continue;
    int claimed = 0;
    const uint64_t s = SlotOps<KT>::find_or_insert(T, span, rbits, SlotOps<KT>::kmer(p), &ctr->err, &claimed);
    remap[i] = s;
    if (s == ~0ULL) continue;
#pragma unroll
    for (int q = 0; q < Slot<KT>::NSUB; ++q) {
      const uint64_t v = p->sub[q];
      if (v != EMPTY64)
        add_k2<KT>(T, s, (uint8_t)((v >> 48) - 1), sp, spcap, &ctr->err, v & 0xFFFFFFFFFFFFULL);
    }
  }
}

/* the old side table's entries (keyed by old slot) into the new table */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_move_spill(const Spill* O, uint64_t ocap, const uint64_t* remap,
                                                     Slot<KT>* T, Spill* sp, uint64_t spcap, Counters* ctr) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < ocap; i += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t key = O[i].key;
    if (key == EMPTY64) continue;
    const uint64_t s = remap[key >> 2];
    if (s == ~0ULL) continue;
    for (int q = 0; q < SPILL_SUB; ++q) {
      const uint64_t v = O[i].sub[q];
      if (v != EMPTY64)
        add_k2<KT>(T, s, (uint8_t)((v >> 48) - 1), sp, spcap, &ctr->err, v & 0xFFFFFFFFFFFFULL);
    }
  }
}

template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_get_counts(const Slot<KT>* T, uint64_t s0, uint64_t s1, const Spill* sp,
                                                      uint64_t spcap, bruno_node* out, uint64_t out_cap, Counters* ctr) {
  __shared__ ReserveSmem sm;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = s0 + (uint64_t)blockIdx.x * BLOCK; base < s1; base += stride) {
    const uint64_t s = base + threadIdx.x;
    bool occ = (s < s1) && SlotOps<KT>::occupied(T + s);
    bruno_node nd;
    if (occ && out) {
      K2List L;
      gather_k2<KT>(T + s, s, sp, spcap, L);
      p1_counts(L, nd.counts);
      put_kmer<KT>(nd, SlotOps<KT>::kmer(T + s));
      nd.pad[0] = nd.pad[1] = nd.pad[2] = 0;
    }
    unsigned long long pos = block_reserve(occ, &ctr->n_out, sm);
    if (occ && out) {
      if (pos < out_cap) out[pos] = nd;
      else atomicOr(&ctr->err, ERR_LIST_FULL);
    }
  }
}

/* k2 list (before filter) of slots [s0, s1): one thread per slot, up to 25
 * entries; kmers == nullptr counts only */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_get_k2(const Slot<KT>* T, uint64_t s0, uint64_t s1, const Spill* sp,
                                                  uint64_t spcap, uint64_t* kmers, uint8_t* biedges, uint32_t* counts,
                                                  uint64_t out_cap, Counters* ctr) {
  __shared__ ReserveSmem sm;
  const int W = kt_traits<KT>::words;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = s0 + (uint64_t)blockIdx.x * BLOCK; base < s1; base += stride) {
    const uint64_t s = base + threadIdx.x;
    bool occ = (s < s1) && SlotOps<KT>::occupied(T + s);
    K2List L;
    L.n = 0;
    if (occ) gather_k2<KT>(T + s, s, sp, spcap, L);
    unsigned long long pos = block_reserve((uint32_t)L.n, &ctr->n_out, sm);
    if (occ && kmers) {
      KT k = SlotOps<KT>::kmer(T + s);
      for (int j = 0; j < L.n; ++j) {
        if (pos + j >= out_cap) {
          atomicOr(&ctr->err, ERR_LIST_FULL);
          break;
        }
        kmers[(pos + j) * W] = (uint64_t)k;
        if (W == 2) kmers[(pos + j) * W + 1] = (uint64_t)((u128)k >> 64);
        biedges[pos + j] = L.e[j];
        counts[pos + j] = L.c[j];
      }
    }
  }
}

/* present nodes whose slot lies in [s0, s1) (one job of a batch) */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_get_nodes_range(const Slot<KT>* T, const uint64_t* present, uint64_t n,
                                                           uint64_t s0, uint64_t s1, bruno_node* out,
                                                           uint64_t out_cap, Counters* ctr) {
  __shared__ ReserveSmem sm;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n; base += stride) {
    const uint64_t i = base + threadIdx.x;
    uint64_t s = 0;
    bool in = false;
    if (i < n) {
      s = present[i];
      in = s >= s0 && s < s1;
    }
    unsigned long long pos = block_reserve(in, &ctr->n_out, sm);
    if (in && out) {
      if (pos >= out_cap) {
        atomicOr(&ctr->err, ERR_LIST_FULL);
        continue;
      }
      const Slot<KT>* p = T + s;
      const NodeRec* nr = (const NodeRec*)&p->sub[0];
      bruno_node nd;
      put_kmer<KT>(nd, SlotOps<KT>::kmer(p));
      for (int q = 0; q < 9; ++q) nd.counts[q] = nr->c[q];
      nd.pad[0] = nd.pad[1] = nd.pad[2] = 0;
      out[pos] = nd;
    }
  }
}

template <typename KT>
__global__ void k_get_nodes(const Slot<KT>* T, const uint64_t* present, uint64_t n, bruno_node* out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Slot<KT>* p = T + present[i];
  const NodeRec* nr = (const NodeRec*)&p->sub[0];
  bruno_node nd;
  put_kmer<KT>(nd, SlotOps<KT>::kmer(p));
  for (int q = 0; q < 9; ++q) nd.counts[q] = nr->c[q];
  nd.pad[0] = nd.pad[1] = nd.pad[2] = 0;
  out[i] = nd;
}

template <typename KT>
__global__ void k_get_chain_nodes(const uint32_t* perm, const KT* out_k, const KT* out_rep, const uint32_t* dist_sorted,
                                  uint64_t n, bruno_chain_node* out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t m = perm[i];
  bruno_chain_node c;
  c.k[0] = (uint64_t)out_k[m];
  c.k[1] = (uint64_t)((u128)out_k[m] >> 64);
  c.rep[0] = (uint64_t)out_rep[m];
  c.rep[1] = (uint64_t)((u128)out_rep[m] >> 64);
  c.dist = (int32_t)dist_sorted[i];
  c.pad = 0;
  out[i] = c;
}

/* The filtered-graph record of canonical k-mer k in one job's table
 * segment (null when absent). */
template <typename KT>
__device__ __forceinline__ const NodeRec* node_rec(const Slot<KT>* T, uint64_t tbase, uint64_t tcap, uint32_t rbits,
                                                   KT k) {
  const uint64_t s = SlotOps<KT>::find(T + tbase, tcap, rbits, k, true);
  return s == ~0ULL ? nullptr : (const NodeRec*)&T[tbase + s].sub[0];
}

/* .branch.edges: the sorted branch k-mers with their counts
 * (print_branch_edge_frequencies, compact_dbg_stats.hpp:219-262). */
template <typename KT>
__global__ void k_branch_nodes(const Slot<KT>* T, uint64_t tbase, uint64_t tcap, uint32_t rbits, const KT* sorted,
                               uint64_t n, bruno_node* out, Counters* ctr) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const NodeRec* r = node_rec<KT>(T, tbase, tcap, rbits, sorted[i]);
  bruno_node nd;
  put_kmer<KT>(nd, sorted[i]);
  for (int q = 0; q < 9; ++q) nd.counts[q] = r ? r->c[q] : 0;
  nd.pad[0] = nd.pad[1] = nd.pad[2] = 0;
  if (!r) atomicOr(&ctr->err, ERR_NO_NODE);
  out[i] = nd;
}

/* .chain.edges (compute_freq_map, compact_dbg_stats.hpp:783-816, and
 * print_chain_frequencies, :879-1065), one wave per unitig: the unitig's
 * sorted chain nodes [starts[u], starts[u+1]) run from its rep L (dist 0)
 * to its last k-mer R on the rep strand.  The lanes look up the self counts
 * of the nodes in the table and reduce their sum, min and max; lane 0 reads
 * L's in-edge and R's out-edge counts, complemented in reverse order when
 * the k-mer is the reverse strand of its canonical form. */
template <typename KT>
__global__ __launch_bounds__(BLOCK) void k_chain_edges(const Slot<KT>* T, uint64_t tbase, uint64_t tcap, uint32_t rbits,
                                                       int K, const uint32_t* perm, const KT* out_k,
                                                       const uint32_t* starts, uint64_t n_unitigs,
                                                       bruno_chain_edge* out, Counters* ctr) {
  kops<KT> ops(K);
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const uint64_t n_waves = ((uint64_t)gridDim.x * BLOCK) >> 6;
  for (uint64_t u = wave; u < n_unitigs; u += n_waves) {
    const uint32_t b = starts[u], e = starts[u + 1];
    uint64_t sum = 0;
    uint32_t mn = 65535, mx = 0, missing = 0;
    for (uint32_t p = b + lane; p < e; p += 64) {
      const NodeRec* r = node_rec<KT>(T, tbase, tcap, rbits, ops.canon(out_k[perm[p]]));
      const uint32_t c = r ? r->c[8] : 0;
      missing |= r == nullptr;
      sum += c;
      mn = c < mn ? c : mn;
      mx = c > mx ? c : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sum += __shfl_xor(sum, off);
      const uint32_t a = __shfl_xor(mn, off), z = __shfl_xor(mx, off);
      mn = a < mn ? a : mn;
      mx = z > mx ? z : mx;
      missing |= __shfl_xor(missing, off);
    }
    if (lane == 0) {
      const KT L = out_k[perm[b]], R = out_k[perm[e - 1]];
      const KT cL = ops.canon(L), cR = ops.canon(R);
      const NodeRec* rl = node_rec<KT>(T, tbase, tcap, rbits, cL);
      const NodeRec* rr = node_rec<KT>(T, tbase, tcap, rbits, cR);
      bruno_chain_edge o;
      o.left[0] = (uint64_t)L;
      o.left[1] = (uint64_t)((u128)L >> 64);
      o.right[0] = (uint64_t)R;
      o.right[1] = (uint64_t)((u128)R >> 64);
      o.sum = sum;
      o.n_nodes = e - b;
      for (int c = 0; c < 4; ++c) {
        o.in[c] = rl ? (cL == L ? rl->c[4 + c] : rl->c[3 - c]) : 0;
        o.out[c] = rr ? (cR == R ? rr->c[c] : rr->c[7 - c]) : 0;
      }
      o.min = (uint16_t)mn;
      o.max = (uint16_t)mx;
      if (missing || !rl || !rr) atomicOr(&ctr->err, ERR_NO_NODE);
      out[u] = o;
    }
  }
}

/* 1 where a sorted chain node starts a unitig (dist 0) */
__global__ void k_head_flags(const uint32_t* dist_sorted, uint64_t n, uint8_t* flags) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = dist_sorted[i] == 0;
}

#include "clean.h"

/* the sort keys of the bubble candidates, one array per field */
template <typename KT>
__global__ void k_summary_keys(const ChainSummary<KT>* S, uint64_t n, KT* b5, KT* t5, KT* t3, KT* b3, uint32_t* len) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    b5[i] = S[i].b5;
    t5[i] = S[i].t5;
    t3[i] = S[i].t3;
    b3[i] = S[i].b3;
    len[i] = S[i].len;
  }
}

/* reads of one synthetic job at bases, offsets[r] = obase + r * L */
__global__ void k_synth(br::synth_params p, char* bases, uint64_t* offsets, uint64_t obase) {
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < p.n_reads) {
    make_read(p, r, bases + r * p.L);
    offsets[r] = obase + r * p.L;
  }
  if (r == 0) offsets[p.n_reads] = obase + p.n_reads * p.L;
}

/* ================================================================== */
/* host side                                                            */
/* ================================================================== */
inline unsigned grid_for(uint64_t n, unsigned block = BLOCK, unsigned maxg = 1u << 16) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > maxg) g = maxg;
  return (unsigned)g;
}

/* compute units of the current device (256 on MI355X) */
inline uint32_t device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      n <= 0) {
    (void)hipGetLastError();
    return 256;
  }
  return (uint32_t)n;
}

/* One side stream per device, shared by every job and kept for the
 * process: creating and destroying a stream per job costs milliseconds.
 * Jobs order their work on it only through events, so sharing it can only
 * serialise two jobs' side work, never deadlock. */
inline hipStream_t shared_side_stream() {
  static std::mutex mu;
  static hipStream_t s[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> g(mu);
  if (!s[dev] && hipStreamCreateWithFlags(&s[dev], hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    s[dev] = nullptr;
  }
  return s[dev];
}

/* A second stream per device for the region builds of a count, beside the
 * multisplit pass on the job's stream (ordered through events, as the side
 * stream). */
inline hipStream_t shared_count_stream() {
  static std::mutex mu;
  static hipStream_t s[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> g(mu);
  if (!s[dev] && hipStreamCreateWithFlags(&s[dev], hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    s[dev] = nullptr;
  }
  return s[dev];
}
/* sub-bucket groups of a pipelined region count: at most this many chosen
 * automatically, at most MAX_COUNT_GROUPS by the tuning knob */
constexpr uint64_t COUNT_GROUPS = 16, MAX_COUNT_GROUPS = 64;

enum Stage { ST_CREATED = 0, ST_COUNTED = 1, ST_FILTERED = 2, ST_CHAINS = 3, ST_RANKED = 4, ST_COMPACTED = 5 };

}  // namespace

struct bruno_job {
  int K = 31;
  int words = 1;
  int thresholding = 0;
  Thresh th{};
  double load_factor = 0.5;
  hipStream_t stream = nullptr;
  int stage = ST_CREATED;

  uint64_t cap = 0, spcap = 0;
  void* table = nullptr;
  /* Region layout (plain jobs with k <= 31 and a large table): 2^rbits
   * regions of rspan slots, each small enough for one workgroup's LDS, so
   * a count can build every region on chip from instance records grouped
   * by region.  rbits == 0: one probe window over the whole table (or a
   * batch job's segment). */
  uint32_t rbits = 0;
  uint64_t rspan = 0;
  bool region_ok = false;     /* plain job: the region layout may be chosen */
  bool table_valid = false;   /* every slot holds a key or EMPTY (else it is stale) */
  bool by_regions = false;    /* the last count built the table region by region */
  /* the only count since the reset built every region and left the filter
   * scan's candidate / spilled-slot bits */
  bool bits_fresh = false;
  uint64_t* cbits = nullptr;
  uint64_t* sbits = nullptr;
  /* fuse_filter: a count that builds every region also filters (asked by
   * bruno_job_set_fused_filter); fused: the last count did, leaving the
   * node lists with its waves' unused positions (holes, 2 per wave) */
  bool fuse_filter = false;
  bool fused = false;
  uint64_t* holes = nullptr;
  uint64_t n_hole_waves = 0;
  Spill* spill = nullptr;
  /* capacity 0 at create: the table is sized from a HyperLogLog sketch of
   * every batch counted since the reset (SKETCH_M registers, merged), and
   * moved into a larger table when a later batch needs more slots */
  bool auto_size = false;
  uint32_t* sketch = nullptr;
  uint64_t auto_bases = 0;      /* bases of the batches sketched since the reset */
  Counters* ctr = nullptr;      /* device */
  Counters* hctr = nullptr;     /* pinned host mirror */
  uint64_t device_bytes = 0;
  uint64_t peak_bytes = 0;
  uint64_t n_reads = 0;

  /* A batch of n_jobs independent jobs shares the table: job i owns slots
   * [seg[i], seg[i+1]).  A plain job is the one-segment case. */
  uint32_t n_jobs = 1;
  std::vector<uint64_t> seg;
  uint64_t* d_seg = nullptr;
  unsigned long long* d_jstat = nullptr; /* [0, n): instances, [n, 2n): dropped reads */
  std::vector<uint64_t> job_reads;
  uint32_t* chain_job = nullptr;
  uint32_t* branch_job = nullptr;
  uint32_t* out_job = nullptr;
  std::vector<uint32_t> job_iters;
  std::vector<uint64_t> job_cycle, job_isolated;
  std::vector<uint64_t> chain_start, fasta_off, branch_start; /* n_jobs + 1 each, after compact */

  /* after filter: the nodes are the chain list followed by the branch list */
  uint64_t n_present = 0, n_branch = 0, n_chain = 0, n_virt = 0;
  uint64_t* chain_slot = nullptr;
  uint64_t* branch_slot = nullptr;
  uint64_t list_cap = 0; /* entries each list has room for */

  /* chains (ckmer and cedge: written by the filter, by chain-list
   * position, when it ran as its own pass) */
  uint8_t* cedge = nullptr;
  void* ckmer = nullptr;
  void* vkmer = nullptr;
  void* bkmer = nullptr;
  LRNode* lr = nullptr;
  uint32_t* list[2] = {nullptr, nullptr};
  uint8_t* alive = nullptr;
  uint64_t vcap = 0;
  uint32_t iterations = 0;
  uint64_t n_cycle = 0, n_isolated = 0, conflicts = 0;
  bool lr_contracted = false; /* ranked by the contraction path (well-formed chains) */
  bool out_by_chains = false; /* unitig order placed from chain offsets, not sorted */

  /* output */
  uint64_t n_out = 0, n_unitigs = 0;
  void* out_k = nullptr;
  void* out_rep = nullptr;
  uint32_t* out_dist = nullptr;
  uint32_t* perm = nullptr;
  uint32_t* dist_sorted = nullptr;
  char* fasta = nullptr;
  uint64_t fasta_len = 0;
  void* bsorted = nullptr;
  char* bfasta = nullptr;
  uint64_t bfasta_len = 0;

  hipEvent_t ev[12];
  bool ev_ok = false;

  /* recycle: once the chains are built the table is cleared on a side
   * stream, overlapping list ranking and output; the next reset waits */
  bool recycle = false;
  hipStream_t side = nullptr;
  hipEvent_t ev_chains = nullptr, ev_cleared = nullptr;
  /* the pipelined count's events: a group split (job stream -> build
   * stream; each wait takes the record made just before it), all builds
   * done (back) */
  hipEvent_t cev[2] = {};
  bool clear_pending = false;
  bruno_phase_ms ms{};

  std::vector<void*> allocs;
};

namespace {

}  // namespace

/* fastq.hip: sequence-line ranges with arrays from the caller's allocator
 * (not in the C ABI) */
bruno_status bruno_fastq_ranges(const char* d_fastq, uint64_t bytes, void* (*alloc)(void*, uint64_t), void* ctx,
                                uint64_t** starts, uint64_t** ends, uint64_t* n_reads, uint64_t* n_bases,
                                void* stream);

namespace {

/* Process-wide tuning knobs (bruno_set_tuning). */
std::mutex g_tuning_mu;
/* defaults: the autotuner's picks on the 100 Mbp / 33 M-read workload */
bruno_tuning g_tuning = {256, 65536, 32768, 64, 1, 2560, 0, 0};
bruno_tuning tuning() {
  std::lock_guard<std::mutex> g(g_tuning_mu);
  return g_tuning;
}

/* All device memory comes from the per-device pool. */
hipError_t pool_alloc(void** p, uint64_t bytes) { return DevicePool::current().alloc(p, bytes); }
void pool_free(void* p) { DevicePool::current().release(p); }

bruno_status dev_alloc(bruno_job* j, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (pool_alloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    return BRUNO_OOM;
  }
  j->allocs.push_back(*p);
  j->device_bytes += bytes;
  j->peak_bytes = std::max(j->peak_bytes, j->device_bytes);
  return BRUNO_OK;
}
/* The block goes back to the pool once the job's stream has finished with
 * it (another stream may be handed the block next). */
void dev_free(bruno_job* j, void*& p, size_t bytes) {
  if (!p) return;
  (void)hipStreamSynchronize(j->stream);
  pool_free(p);
  auto it = std::find(j->allocs.begin(), j->allocs.end(), p);
  if (it != j->allocs.end()) j->allocs.erase(it);
  j->device_bytes -= std::min<uint64_t>(j->device_bytes, bytes);
  p = nullptr;
}

#define HIP_CHECK(x)                               \
  do {                                             \
    if ((x) != hipSuccess) {                       \
      (void)hipGetLastError();                     \
      return BRUNO_INTERNAL;                       \
    }                                              \
  } while (0)

#define TRY(x)                                     \
  do {                                             \
    bruno_status _s = (x);                         \
    if (_s != BRUNO_OK) return _s;                 \
  } while (0)

bruno_status sync_counters(bruno_job* j) {
  HIP_CHECK(hipMemcpyAsync(j->hctr, j->ctr, sizeof(Counters), hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  if (j->hctr->err & (ERR_TABLE_FULL | ERR_SPILL_FULL | ERR_LIST_FULL)) return BRUNO_CAPACITY;
  if (j->hctr->err) return BRUNO_INTERNAL;
  return BRUNO_OK;
}

void fill64(bruno_job* j, void* p, uint64_t n_words, uint64_t v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill64, dim3(grid_for(n_words, BLOCK, 1u << 14)), dim3(BLOCK), 0, st, (uint64_t*)p, n_words,
                     v);
}
void fill64(bruno_job* j, void* p, uint64_t n_words, uint64_t v) { fill64(j, p, n_words, v, j->stream); }

size_t slot_bytes(const bruno_job* j) { return j->words == 1 ? sizeof(Slot<uint64_t>) : sizeof(Slot<u128>); }

/* table, side table and per-job counters back to empty on stream st.  A
 * region-layout table is only marked stale: the next count writes every
 * slot (or ensure_table_valid fills it). */
void clear_table(bruno_job* j, hipStream_t st) {
  if (j->rbits) {
    j->table_valid = false;
  } else {
    fill64(j, j->table, j->cap * slot_bytes(j) / 8, EMPTY64, st);
    j->table_valid = true;
  }
  fill64(j, j->spill, j->spcap * sizeof(Spill) / 8, EMPTY64, st);
  fill64(j, j->d_jstat, 2 * j->n_jobs, 0, st);
}

void ev_record(bruno_job* j, int i) {
  if (j->ev_ok) (void)hipEventRecord(j->ev[i], j->stream);
}
double ev_ms(bruno_job* j, int a, int b) {
  float ms = 0;
  if (!j->ev_ok) return 0;
  if (hipEventElapsedTime(&ms, j->ev[a], j->ev[b]) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return ms;
}


/* Tables of at least this many slots get the region layout (count_regions
 * 1 or 2).  Measured on jobs of 2^22 to 2^26 slots counted alone, the
 * region count and the filter after it save 1.1 to 6.7 ms per job over
 * the atomic count (below 2^22 the layout is not offered). */
constexpr uint64_t REGION_MIN_CAP = 1ULL << 22;
constexpr uint64_t REGION_MIN_CAP_ANY = 1ULL << 22;

/* One table for all jobs: job i gets caps[i] slots from seg[i].  A plain
 * k <= 31 job with a large table gets 2^rbits regions of at most
 * region_slots slots (the capacity rounds up to a whole number of
 * regions); its slots are written by the first count, not filled here. */
bruno_status alloc_table(bruno_job* j, const uint64_t* caps) {
  const uint32_t nj = j->n_jobs;
  j->seg.assign(nj + 1, 0);
  for (uint32_t i = 0; i < nj; ++i) j->seg[i + 1] = j->seg[i] + caps[i];
  j->rbits = 0;
  j->rspan = 0;
  const uint32_t cr = tuning().count_regions;
  if (nj == 1 && j->region_ok && cr && j->seg[1] >= (cr >= 2 ? REGION_MIN_CAP_ANY : REGION_MIN_CAP)) {
    uint32_t b = REC_BUCKET_BITS;
    /* the knob counts 32-byte slots: a 64-byte-slot region holds half */
    const uint64_t most = j->words == 1 ? tuning().region_slots : tuning().region_slots / 2;
    while ((j->seg[1] + (1ULL << b) - 1) >> b > most) ++b;
    j->rbits = b;
    j->rspan = (j->seg[1] + (1ULL << b) - 1) >> b;
    j->seg[1] = j->rspan << b;
  }
  const uint64_t cap = j->seg[nj];
  j->cap = cap;
  TRY(dev_alloc(j, &j->table, cap * slot_bytes(j)));
  /* side-table records: one per 32 slots (the same bytes as a quarter of
   * the slots' sub-entries) */
  j->spcap = std::max<uint64_t>(1024, cap / 32);
  TRY(dev_alloc(j, (void**)&j->spill, j->spcap * sizeof(Spill)));
  TRY(dev_alloc(j, (void**)&j->d_seg, (nj + 1) * sizeof(uint64_t)));
  TRY(dev_alloc(j, (void**)&j->d_jstat, 2 * nj * sizeof(unsigned long long)));
  HIP_CHECK(hipMemcpyAsync(j->d_seg, j->seg.data(), (nj + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, j->stream));
  j->table_valid = false;
  if (!j->rbits) {
    fill64(j, j->table, cap * slot_bytes(j) / 8, EMPTY64);
    j->table_valid = true;
  }
  fill64(j, j->spill, j->spcap * sizeof(Spill) / 8, EMPTY64);
  fill64(j, j->d_jstat, 2 * nj, 0);
  return BRUNO_OK;
}

/* probe span of a plain job: one region, or the whole table */
uint64_t probe_span(const bruno_job* j) { return j->rbits ? j->rspan : j->cap; }

/* The filter pre-pass keeps a k-mer when its count can reach the lower
 * bound.  An edge count of the k2 filter is at most the k-mer count, and so
 * is a canonical k+1-mer window count of a k-mer that is not its own
 * reverse complement; for one that is (even k only) the window count is at
 * most twice it (the window c.K shows up once as in-edge c and once as
 * out-edge comp(c), both in one read "c K comp(c)"). */
uint64_t scan_lo(const bruno_job* j) {
  return (j->thresholding == 2 && !(j->K & 1)) ? (j->th.t1lo + 1) / 2 : j->th.t1lo;
}

/* every slot EMPTY unless a count has written the table since the reset */
void ensure_table_valid(bruno_job* j) {
  if (j->table && !j->table_valid) {
    fill64(j, j->table, j->cap * slot_bytes(j) / 8, EMPTY64);
    j->table_valid = true;
  }
}

bool persistent(const bruno_job* j, const void* p) {
  return p == j->table || p == j->spill || p == j->d_seg || p == j->d_jstat || (p && p == j->sketch);
}
uint64_t persistent_bytes(const bruno_job* j) {
  const uint64_t sk = j->sketch ? SKETCH_M * 4 : 0;
  if (!j->table) return sk;
  return j->cap * slot_bytes(j) + j->spcap * sizeof(Spill) + (j->n_jobs + 1) * sizeof(uint64_t) +
         2 * j->n_jobs * sizeof(unsigned long long) + sk;
}

/* HyperLogLog estimate from SKETCH_M registers of a 1-in-64 hash sample,
 * scaled back to all k-mers; small cardinalities by linear counting. */
double hll_estimate(const uint32_t* M) {
  const double m = SKETCH_M;
  const double alpha = 0.7213 / (1.0 + 1.079 / m);
  double sum = 0;
  int zeros = 0;
  for (int q = 0; q < SKETCH_M; ++q) {
    sum += std::ldexp(1.0, -(int)M[q]);
    zeros += M[q] == 0;
  }
  double e = alpha * m * m / sum;
  if (e <= 2.5 * m && zeros > 0) e = m * std::log(m / zeros);
  return e * (double)(1u << SKETCH_SAMPLE_BITS);
}

/* Hash slots for an estimated distinct count: a margin of four standard
 * errors (register and sampling noise), at most one slot per instance. */
uint64_t capacity_from_estimate(double est, uint64_t n_instances, double load_factor) {
  const double sampled = std::max(1.0, est / (double)(1u << SKETCH_SAMPLE_BITS));
  const double rel = 1.04 / std::sqrt((double)SKETCH_M) + 1.0 / std::sqrt(sampled);
  const double d = std::min((double)n_instances, est * (1.0 + 4.0 * rel));
  return (uint64_t)(d / load_factor) + 4096;
}

/* sketch of the canonical k-mers of reads (device pointers) */
template <typename KT>
bruno_status sketch_t(int K, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads, uint32_t* regs,
                      Counters* ctr, hipStream_t st, const uint64_t* d_starts, const uint64_t* d_ends) {
  ParseArgs a{};
  a.bases = (const uint8_t*)d_bases;
  a.offsets = d_offsets;
  a.starts = d_starts;
  a.ends = d_ends;
  a.n_reads = n_reads;
  a.K = K;
  a.n_jobs = 1;
  a.sketch = regs;
  a.ctr = ctr;
  const uint64_t waves = std::min<uint64_t>(n_reads, tuning().parse_grid);
  const unsigned g = (unsigned)std::max<uint64_t>(1, (waves + 3) / 4);
  hipLaunchKernelGGL((k_parse<KT, 3, 256>), dim3(g), dim3(256), 0, st, a);
  return BRUNO_OK;
}

bruno_status estimate_distinct(int K, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads, double* est,
                               hipStream_t st, const uint64_t* d_starts = nullptr, const uint64_t* d_ends = nullptr) {
  uint32_t* regs = nullptr;
  if (pool_alloc((void**)&regs, SKETCH_M * 4) != hipSuccess) return BRUNO_OOM;
  std::vector<uint32_t> h(SKETCH_M);
  bruno_status s = BRUNO_OK;
  if (hipMemsetAsync(regs, 0, SKETCH_M * 4, st) != hipSuccess) s = BRUNO_INTERNAL;
  if (s == BRUNO_OK) {
    if (K <= 31) s = sketch_t<uint64_t>(K, d_bases, d_offsets, n_reads, regs, nullptr, st, d_starts, d_ends);
    else s = sketch_t<u128>(K, d_bases, d_offsets, n_reads, regs, nullptr, st, d_starts, d_ends);
  }
  if (s == BRUNO_OK && (hipMemcpyAsync(h.data(), regs, SKETCH_M * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess)) {
    (void)hipGetLastError();
    s = BRUNO_INTERNAL;
  }
  (void)hipStreamSynchronize(st);
  pool_free(regs);
  if (s == BRUNO_OK) *est = hll_estimate(h.data());
  return s;
}

/* Move a plain job's counted table into one of `want` slots (more than it
 * has): a new table (region layout if it is large enough), every k-mer and
 * k2 entry inserted again with its count, then the old table released.
 * Only between counts: the filter has not run. */
bruno_status grow_table(bruno_job* j, uint64_t want) {
  HIP_CHECK(hipStreamSynchronize(j->stream));
  void* otab = j->table;
  void* osp = j->spill;
  void* oseg = j->d_seg;
  void* ojs = j->d_jstat;
  const uint64_t ocap = j->cap, ospcap = j->spcap;
  const uint32_t orbits = j->rbits;
  const uint64_t orspan = j->rspan;
  const std::vector<uint64_t> oseg_h = j->seg;
  j->table = j->spill = nullptr;
  j->d_seg = nullptr;
  j->d_jstat = nullptr;
  bruno_status st = alloc_table(j, &want);
  uint64_t* remap = nullptr;
  if (st == BRUNO_OK) st = dev_alloc(j, (void**)&remap, ocap * 8);
  if (st != BRUNO_OK) {
    /* keep the old table */
    void* p;
    p = j->table; dev_free(j, p, j->cap * slot_bytes(j));
    p = j->spill; dev_free(j, p, j->spcap * sizeof(Spill));
    p = j->d_seg; dev_free(j, p, (j->n_jobs + 1) * 8);
    p = j->d_jstat; dev_free(j, p, 2 * j->n_jobs * 8);
    j->table = otab;
    j->spill = (Spill*)osp;
    j->d_seg = (uint64_t*)oseg;
    j->d_jstat = (unsigned long long*)ojs;
    j->cap = ocap;
    j->spcap = ospcap;
    j->rbits = orbits;
    j->rspan = orspan;
    j->seg = oseg_h;
    j->table_valid = true;
    return st;
  }
  ensure_table_valid(j);
  const unsigned g = grid_for(ocap, BLOCK, 1u << 15);
  if (j->words == 1) {
    hipLaunchKernelGGL((k_move_slots<uint64_t>), dim3(g), dim3(BLOCK), 0, j->stream, (const Slot<uint64_t>*)otab,
                       ocap, (Slot<uint64_t>*)j->table, probe_span(j), j->rbits, j->spill, j->spcap, remap, j->ctr);
    hipLaunchKernelGGL((k_move_spill<uint64_t>), dim3(grid_for(ospcap, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                       (const Spill*)osp, ospcap, remap, (Slot<uint64_t>*)j->table, j->spill, j->spcap, j->ctr);
  } else {
    hipLaunchKernelGGL((k_move_slots<u128>), dim3(g), dim3(BLOCK), 0, j->stream, (const Slot<u128>*)otab, ocap,
                       (Slot<u128>*)j->table, probe_span(j), j->rbits, j->spill, j->spcap, remap, j->ctr);
    hipLaunchKernelGGL((k_move_spill<u128>), dim3(grid_for(ospcap, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                       (const Spill*)osp, ospcap, remap, (Slot<u128>*)j->table, j->spill, j->spcap, j->ctr);
  }
  HIP_CHECK(hipMemcpyAsync(j->d_jstat, ojs, 2 * j->n_jobs * 8, hipMemcpyDeviceToDevice, j->stream));
  bruno_status s2 = sync_counters(j);
  void* p = remap;
  dev_free(j, p, ocap * 8);
  dev_free(j, otab, ocap * slot_bytes(j));
  dev_free(j, osp, ospcap * sizeof(Spill));
  dev_free(j, oseg, (j->n_jobs + 1) * 8);
  dev_free(j, ojs, 2 * j->n_jobs * 8);
  j->bits_fresh = false;
  j->by_regions = false;
  return s2;
}

/* Size (capacity-0 jobs) or grow the table for one more batch: its
 * canonical k-mers are sketched into the job's registers, merged with
 * every batch since the reset, and the union's estimate gives the slots
 * needed (instances bounded by the bases so far).  A table that is short
 * grows to at least half again its size, so a run of batches moves it few
 * times.  The reference's hash maps grow as they fill instead. */
bruno_status autosize(bruno_job* j, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads,
                      uint64_t n_bases, const uint64_t* d_starts, const uint64_t* d_ends) {
  if (!j->sketch) {
    TRY(dev_alloc(j, (void**)&j->sketch, SKETCH_M * 4));
    HIP_CHECK(hipMemsetAsync(j->sketch, 0, SKETCH_M * 4, j->stream));
  }
  j->auto_size = true;
  if (j->K <= 31) TRY(sketch_t<uint64_t>(j->K, d_bases, d_offsets, n_reads, j->sketch, nullptr, j->stream, d_starts, d_ends));
  else TRY(sketch_t<u128>(j->K, d_bases, d_offsets, n_reads, j->sketch, nullptr, j->stream, d_starts, d_ends));
  std::vector<uint32_t> h(SKETCH_M);
  HIP_CHECK(hipMemcpyAsync(h.data(), j->sketch, SKETCH_M * 4, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  j->auto_bases += n_bases;
  const uint64_t want = capacity_from_estimate(hll_estimate(h.data()), j->auto_bases, j->load_factor);
  if (j->cap == 0) return alloc_table(j, &want);
  if (want > j->cap) return grow_table(j, std::max<uint64_t>(want, j->cap + j->cap / 2));
  return BRUNO_OK;
}

/* Count one batch of reads into a region-layout table (regions.h): bucketed
 * instance records, two multisplit passes that group each bucket's records
 * by region (partition.h), one on-chip build per region.  *done = false
 * (and nothing counted) when the record buffers do not fit, a bucket
 * overflowed its share, or the table has more regions than two passes
 * split; the caller then counts with atomics. */
/* record staging per bucket in the parser's LDS: 48 KB of records */
/* the record parser's shape: 16-byte records (k > 31) stage half as many
 * per bucket in the same LDS, so the workgroup appends four times as often
 * (a quarter of the records per append, under the halved threshold) */
template <typename KT> struct RecShape {
  static constexpr bool wide = sizeof(typename RecOf<KT>::T) == 16;
  static constexpr int BS = REC_BS, PF = wide ? REC_PF / 4 : REC_PF;
  static constexpr int STAGE = wide ? REC_STAGE / 2 : REC_STAGE, FLUSH = wide ? REC_FLUSH / 2 : REC_FLUSH;
};

template <typename KT>
bruno_status count_by_regions(bruno_job* j, const ParseArgs& a, bool* done) {
  typedef typename RecOf<KT>::T R;
  constexpr int STAGE = RecShape<KT>::STAGE;
  constexpr uint64_t RB = sizeof(R);
  *done = false;
  const uint64_t n_reads = a.n_reads;
  const int rb = (int)j->rbits - REC_BUCKET_BITS; /* region bits inside a bucket */
  const int b1 = std::min(std::max(rb - 8, 0), SPLIT1_MAX_BITS), b2 = rb - b1;
  if (rb < 0 || b2 > SPLIT2_MAX_BITS) return BRUNO_OK;
  const uint64_t n_sb = (uint64_t)REC_BUCKETS << b1, nreg = 1ULL << j->rbits;
  /* scratch: front / back per bucket, instances and dropped reads, the
   * instance bound, the first pass's overflow flag, its cursors and exact
   * counts (when the first try overflows) */
  const uint64_t NS = 2 * REC_BUCKETS + 4 + 2 * n_sb;
  unsigned long long* scratch = nullptr;
  TRY(dev_alloc(j, (void**)&scratch, NS * 8));
  unsigned long long *front = scratch, *back = scratch + REC_BUCKETS, *stat = scratch + 2 * REC_BUCKETS,
                     *bound = stat + 2, *overflow = bound + 1, *cursor = overflow + 1, *counts = cursor + n_sb;
  void* sp = scratch;
  HIP_CHECK(hipMemsetAsync(scratch, 0, NS * 8, j->stream));
  hipLaunchKernelGGL(k_count_instances, dim3(grid_for(n_reads)), dim3(BLOCK), 0, j->stream, a, bound);
  unsigned long long n_bound = 0;
  HIP_CHECK(hipMemcpyAsync(&n_bound, bound, 8, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  typedef RecShape<KT> RS;
  auto parse_kernel = k_parse_records<KT, RS::BS, RS::PF, RS::STAGE, RS::FLUSH>;
  /* persistent: exactly the workgroups that fit on the device at once */
  static int per_cu = 0;
  if (per_cu == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, parse_kernel, RS::BS, 0) != hipSuccess || per_cu <= 0) {
      (void)hipGetLastError();
      per_cu = 1;
    }
  }
  const uint64_t groups =
      std::max<uint64_t>(1, std::min<uint64_t>((n_reads + RS::BS - 1) / RS::BS, (uint64_t)device_cus() * per_cu));
  /* chunks of up to 2048 records, smaller for small inputs so padding
   * stays under ~1/8 of the records; a flush appends at most REC_STAGE
   * records, so a chunk holds at least that many */
  uint32_t chunk = 2048;
  while (chunk / 2 >= (uint32_t)STAGE && (uint64_t)chunk * groups * REC_BUCKETS * 8 > n_bound) chunk >>= 1;
  const uint64_t bcap = n_bound / REC_BUCKETS + n_bound / (4 * REC_BUCKETS) + 65536 + groups * chunk;
  /* each sub-bucket's room in the first pass's output: its share of the
   * bucket plus 3% and a constant (a sub-bucket's count strays from its
   * share by about its square root) */
  auto room = [&](uint64_t n_b) { return (n_b >> b1) + (n_b >> b1) / 32 + 256; };
  const uint64_t out_cap = n_sb * room(bcap);
  R* recs = nullptr;
  /* the parser's buffer also takes the second pass's output (<= the first's) */
  const uint64_t recs_n = std::max<uint64_t>(REC_BUCKETS * bcap, out_cap);
  if (dev_alloc(j, (void**)&recs, recs_n * RB) != BRUNO_OK) {
    dev_free(j, sp, NS * 8);
    return BRUNO_OK;
  }
  void* rp = recs;
  RecArgs ra{recs, bcap, front, back, stat, chunk};
  hipLaunchKernelGGL(parse_kernel, dim3((unsigned)groups), dim3(RS::BS), 0,
                     j->stream, a, ra);
  std::vector<unsigned long long> hcur(2 * REC_BUCKETS);
  HIP_CHECK(hipMemcpyAsync(hcur.data(), front, 2 * REC_BUCKETS * 8, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  bool over = false;
  for (int b = 0; b < REC_BUCKETS; ++b) over |= hcur[b] + hcur[REC_BUCKETS + b] > bcap;
  /* sub-bucket rooms from the bucket sizes (records plus chunk padding) */
  std::vector<uint64_t> base(n_sb + 1, 0);
  for (uint64_t q = 0; q < n_sb; ++q) base[q + 1] = base[q] + room(hcur[q >> b1] + hcur[REC_BUCKETS + (q >> b1)]);
  R* recs2 = nullptr;
  uint64_t *d_base = nullptr, *d_out = nullptr, *rstart = nullptr;
  const uint64_t C = base[n_sb];
  auto give_up = [&]() {
    void* p = rstart;
    dev_free(j, p, (nreg + 1) * 8);
    p = d_out;
    dev_free(j, p, (n_sb + 1) * 8);
    p = d_base;
    dev_free(j, p, (n_sb + 1) * 8);
    p = recs2;
    dev_free(j, p, std::max<uint64_t>(1, C) * RB);
    dev_free(j, rp, recs_n * RB);
    dev_free(j, sp, NS * 8);
  };
  if (over || dev_alloc(j, (void**)&recs2, std::max<uint64_t>(1, C) * RB) != BRUNO_OK ||
      dev_alloc(j, (void**)&d_base, (n_sb + 1) * 8) != BRUNO_OK ||
      dev_alloc(j, (void**)&d_out, (n_sb + 1) * 8) != BRUNO_OK ||
      dev_alloc(j, (void**)&rstart, (nreg + 1) * 8) != BRUNO_OK) {
    give_up();
    return BRUNO_OK;
  }
  /* pass 1: the chunked front of every bucket and the records appended
   * from its back, by the top b1 region bits */
  Split1Args s1{};
  s1.in = recs;
  s1.b1 = b1;
  s1.cursor = cursor;
  s1.limit = d_base + 1;
  s1.out = recs2;
  s1.overflow = overflow;
  /* segments (tiles numbered through them) grouped by the XCD that splits
   * their bucket, bucket % SPLIT_XCDS */
  uint64_t tiles = 0;
  for (int x = 0; x < SPLIT_XCDS; ++x) {
    s1.xcd_tiles[x] = tiles;
    for (int b = x; b < REC_BUCKETS; b += SPLIT_XCDS) {
      const uint64_t ranges[2][2] = {{(uint64_t)b * bcap, (uint64_t)b * bcap + hcur[b]},
                                     {(uint64_t)(b + 1) * bcap - hcur[REC_BUCKETS + b], (uint64_t)(b + 1) * bcap}};
      for (auto& r : ranges) {
        if (r[1] <= r[0]) continue;
        s1.seg[s1.n_seg++] = SplitSeg{r[0], r[1], tiles, (uint32_t)b, 0};
        tiles += (r[1] - r[0] + SplitShape<R>::TILE - 1) / SplitShape<R>::TILE;
      }
    }
  }
  s1.xcd_tiles[SPLIT_XCDS] = tiles;
  s1.n_tiles = tiles;
  const unsigned split_grid = device_cus() * 4;
  /* k_split1: a multiple of SPLIT_XCDS workgroups */
  const dim3 g1((unsigned)(SPLIT_XCDS * std::max<uint64_t>(1, std::min<uint64_t>(tiles, split_grid) / SPLIT_XCDS)));
  std::vector<unsigned long long> cur(n_sb + 1);
  for (int attempt = 0; attempt < 2; ++attempt) {
    HIP_CHECK(hipMemcpyAsync(d_base, base.data(), (n_sb + 1) * 8, hipMemcpyHostToDevice, j->stream));
    HIP_CHECK(hipMemcpyAsync(cursor, base.data(), n_sb * 8, hipMemcpyHostToDevice, j->stream));
    if (tiles) hipLaunchKernelGGL(k_split1<R>, g1, dim3(SPLIT_BS), 0, j->stream, s1);
    HIP_CHECK(hipMemcpyAsync(cur.data(), cursor, n_sb * 8, hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipMemcpyAsync(&cur[n_sb], overflow, 8, hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipStreamSynchronize(j->stream));
    if (!cur[n_sb]) break;
    if (attempt == 1) {
      give_up();
      return BRUNO_INTERNAL; /* exact room cannot overflow */
    }
    /* a sub-bucket ran out of room: exact counts, then the pass again */
    HIP_CHECK(hipMemsetAsync(counts, 0, n_sb * 8, j->stream));
    HIP_CHECK(hipMemsetAsync(overflow, 0, 8, j->stream));
    hipLaunchKernelGGL(k_split_count<R>, g1, dim3(SPLIT_BS), 0, j->stream, s1, counts);
    std::vector<unsigned long long> hc(n_sb);
    HIP_CHECK(hipMemcpyAsync(hc.data(), counts, n_sb * 8, hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipStreamSynchronize(j->stream));
    for (uint64_t q = 0; q < n_sb; ++q) base[q + 1] = base[q] + hc[q];
    if (base[n_sb] > C) {
      give_up();
      return BRUNO_INTERNAL;
    }
  }
  /* pass 2: each sub-bucket by the remaining b2 region bits, into the
   * parser's buffer without gaps, with every region's first record */
  std::vector<uint64_t> out(n_sb + 1, 0);
  for (uint64_t q = 0; q < n_sb; ++q) out[q + 1] = out[q] + (cur[q] - base[q]);
  HIP_CHECK(hipMemcpyAsync(d_out, out.data(), (n_sb + 1) * 8, hipMemcpyHostToDevice, j->stream));
  HIP_CHECK(hipMemcpyAsync(d_base, base.data(), n_sb * 8, hipMemcpyHostToDevice, j->stream));
  HIP_CHECK(hipMemcpyAsync(rstart + nreg, &out[n_sb], 8, hipMemcpyHostToDevice, j->stream));
  hipLaunchKernelGGL(k_region_bounds, dim3(grid_for(n_sb + 1, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, d_out, n_sb,
                     b2, rstart);
  Split2Args s2{recs2, d_base, d_out, 0, n_sb, 64 - b1 - b2, b2, recs, rstart};
  const int valid = j->table_valid ? 1 : 0;
  /* a first count builds every region: it also leaves the filter scan's
   * bits (stale once another count follows) */
  const uint64_t stride = (j->rspan + 63) / 64, nbits = nreg * stride;
  uint64_t *cbits = nullptr, *sbits = nullptr;
  if (!valid && dev_alloc(j, (void**)&cbits, nbits * 8) == BRUNO_OK &&
      dev_alloc(j, (void**)&sbits, nbits * 8) != BRUNO_OK) {
    void* cp = cbits;
    dev_free(j, cp, nbits * 8);
    cbits = nullptr;
  }
  if (!sbits) cbits = nullptr;
  const dim3 g4((unsigned)std::min<uint64_t>(nreg, 4 * REGION_BUILD_CU_SLOTS)),
      g2((unsigned)std::min<uint64_t>(nreg, 2 * REGION_BUILD_CU_SLOTS));
  const bool small = RB == 8 ? j->rspan <= REGION_SMALL_SLOTS : j->rspan <= REGION_WIDE_SMALL_SLOTS;
  const uint64_t n_waves = small ? (uint64_t)g4.x * 4 : (uint64_t)g2.x * 8;
  /* side-table queues: a wave's share of the instances / 128 (a k-mer
   * spills only past its slot's sub-entries), at least 1024 entries; a
   * full queue adds inline */
  uint32_t spq_cap = (uint32_t)std::min<uint64_t>(65536, std::max<uint64_t>(1024, n_bound / 128 / n_waves));
  uint64_t* spq = nullptr;
  uint32_t* spq_n = nullptr;
  TRY(dev_alloc(j, (void**)&spq_n, n_waves * 4));
  if (dev_alloc(j, (void**)&spq, n_waves * spq_cap * 8) != BRUNO_OK) spq_cap = 0;
  /* the filter fused into the build (a first count, by request): node
   * lists with room for the nodes the table can hold (at most half its
   * slots, and at most the instances / the count threshold) plus every
   * wave's open chunks; a node that finds no room stays a candidate for
   * the filter pass */
  bool fuse = j->fuse_filter && cbits && j->n_jobs == 1;
  NodeLists nl{};
  uint64_t* holes = nullptr;
  uint64_t lcap = 0;
  if (fuse) {
    const uint64_t lo = j->thresholding ? std::max<uint64_t>(1, scan_lo(j)) : 1;
    lcap = std::min<uint64_t>(j->cap / 2, n_bound / lo) + 2 * FUSE_CHUNK * n_waves + 1024;
    lcap = std::min<uint64_t>(lcap, 0x7FFFFFFFULL);
    if (tuning().fuse_list_cap) lcap = std::min<uint64_t>(lcap, tuning().fuse_list_cap);
    uint64_t *cs = nullptr, *bs = nullptr;
    if (dev_alloc(j, (void**)&cs, lcap * 8) == BRUNO_OK && dev_alloc(j, (void**)&bs, lcap * 8) == BRUNO_OK &&
        dev_alloc(j, (void**)&holes, 2 * n_waves * 8) == BRUNO_OK) {
      j->chain_slot = cs;
      j->branch_slot = bs;
      j->list_cap = lcap;
      nl = NodeLists{cs, bs, nullptr, nullptr, lcap, j->d_seg, 1};
      HIP_CHECK(hipMemsetAsync(&j->ctr->n_present, 0, 3 * sizeof(unsigned long long), j->stream));
    } else {
      void* p = cs;
      dev_free(j, p, lcap * 8);
      p = bs;
      dev_free(j, p, lcap * 8);
      p = holes;
      dev_free(j, p, 2 * n_waves * 8);
      holes = nullptr;
      fuse = false;
    }
  }
  RegionArgs g{j->table, j->rspan, j->rbits, nreg, 0, nreg, recs, rstart, valid, j->spill, j->spcap, spq, spq_n,
               spq_cap, j->ctr, cbits, sbits, stride, j->thresholding, scan_lo(j), fuse ? 1 : 0, j->K, j->th, nl,
               holes};
  auto build = [&](auto fuse_tag, hipStream_t st, dim3 b4, dim3 b2_) {
    constexpr bool F = decltype(fuse_tag)::value;
    if (RB == 8) {
      if (j->rspan <= REGION_SMALL_SLOTS)
        hipLaunchKernelGGL((k_region_build<uint64_t, REGION_SMALL_SLOTS, 256, F>), b4, dim3(256), 0, st, g);
      else
        hipLaunchKernelGGL((k_region_build<uint64_t, REGION_MAX_SLOTS, 512, F>), b2_, dim3(512), 0, st, g);
    } else {
      if (j->rspan <= REGION_WIDE_SMALL_SLOTS)
        hipLaunchKernelGGL((k_region_build<u128, REGION_WIDE_SMALL_SLOTS, 256, F>), b4, dim3(256), 0, st, g);
      else
        hipLaunchKernelGGL((k_region_build<u128, REGION_WIDE_SLOTS, 512, F>), b2_, dim3(512), 0, st, g);
    }
  };
  /* Pass 2 and the builds, pipelined by groups of sub-buckets: the build
   * of a group (bound by instruction issue and LDS) runs on a second stream
   * beside the next group's multisplit (bound by HBM).  A pipelined pass 2
   * runs one workgroup per CU, which leaves every CU's LDS room for a
   * region build beside it.  By default only large counts are pipelined,
   * from 8 sub-buckets per CU, in groups of at least half a sub-bucket per
   * CU: on fewer sub-buckets the split's narrower launches and the builds'
   * shared CUs cost more than the overlap gives (measured: 512 sub-buckets
   * in two groups, 4.5 -> 5.6 ms).  One group when the filter is fused (its
   * list chunks and holes belong to one launch). */
  const uint64_t cus = device_cus();
  uint64_t want = tuning().count_groups;
  if (want == 0) want = n_sb >= 8 * cus ? std::min<uint64_t>(COUNT_GROUPS, 2 * n_sb / cus) : 1;
  want = std::min<uint64_t>(std::min<uint64_t>(want, MAX_COUNT_GROUPS), n_sb);
  hipStream_t bst = (fuse || want < 2) ? nullptr : shared_count_stream();
  if (bst) {
    for (hipEvent_t& e : j->cev)
      if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        e = nullptr;
        bst = nullptr;
      }
  }
  const uint64_t ng = bst ? want : 1;
  if (!bst) bst = j->stream;
  for (uint64_t gi = 0; gi < ng; ++gi) {
    const uint64_t sb0 = n_sb * gi / ng, sb1 = n_sb * (gi + 1) / ng;
    if (sb1 == sb0) continue;
    s2.sb0 = sb0;
    s2.n_sb = sb1 - sb0;
    hipLaunchKernelGGL(k_split2<R>, dim3((unsigned)std::min<uint64_t>(sb1 - sb0, ng > 1 ? device_cus() : split_grid)),
                       dim3(SPLIT_BS), 0, j->stream, s2);
    if (bst != j->stream) {
      HIP_CHECK(hipEventRecord(j->cev[0], j->stream));
      HIP_CHECK(hipStreamWaitEvent(bst, j->cev[0], 0));
    }
    g.q0 = sb0 << b2;
    g.q1 = sb1 << b2;
    const uint64_t nq = g.q1 - g.q0;
    const dim3 b4((unsigned)std::min<uint64_t>(nq, 4 * REGION_BUILD_CU_SLOTS)),
        b2d((unsigned)std::min<uint64_t>(nq, 2 * REGION_BUILD_CU_SLOTS));
    if (fuse) build(std::true_type{}, bst, b4, b2d);
    else build(std::false_type{}, bst, b4, b2d);
    /* this launch's waves' side-table queues */
    if (spq_cap)
      hipLaunchKernelGGL(k_spill_flush, dim3(small ? b4.x * 4 : b2d.x * 8), dim3(BLOCK), 0, bst, spq, spq_n, spq_cap,
                         j->spill, j->spcap, j->ctr);
  }
  if (bst != j->stream) {
    HIP_CHECK(hipEventRecord(j->cev[1], bst));
    HIP_CHECK(hipStreamWaitEvent(j->stream, j->cev[1], 0));
  }
  /* instances and dropped reads into the job's counters */
  hipLaunchKernelGGL(k_add_u64, dim3(1), dim3(64), 0, j->stream, (unsigned long long*)&j->ctr->n_instances, stat, 1);
  hipLaunchKernelGGL(k_add_u64, dim3(1), dim3(64), 0, j->stream, (unsigned long long*)&j->ctr->n_dropped, stat + 1,
                     1);
  HIP_CHECK(hipStreamSynchronize(j->stream));
  void* qp = spq;
  dev_free(j, qp, n_waves * spq_cap * 8);
  qp = spq_n;
  dev_free(j, qp, n_waves * 4);
  give_up(); /* every buffer of the count back to the pool */
  j->table_valid = true;
  j->cbits = cbits;
  j->sbits = sbits;
  j->bits_fresh = cbits != nullptr;
  j->fused = fuse;
  j->holes = holes;
  j->n_hole_waves = fuse ? n_waves : 0;
  *done = true;
  return BRUNO_OK;
}

template <typename KT>
bruno_status parse_count_t(bruno_job* j, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads,
                           const uint64_t* d_first, const uint64_t* d_starts = nullptr,
                           const uint64_t* d_ends = nullptr) {
  ParseArgs a{};
  a.starts = d_starts;
  a.ends = d_ends;
  a.n_jobs = j->n_jobs;
  a.job_first = d_first;
  a.seg = j->d_seg;
  a.job_inst = j->d_jstat;
  a.job_drop = j->d_jstat + j->n_jobs;
  a.bases = (const uint8_t*)d_bases;
  a.offsets = d_offsets;
  a.n_reads = n_reads;
  a.K = j->K;
  a.table = j->table;
  a.cap = probe_span(j);
  a.rbits = j->rbits;
  a.spill = j->spill;
  a.spcap = j->spcap;
  a.ctr = j->ctr;
  const bruno_tuning t = tuning();
  j->by_regions = false;
  j->bits_fresh = false;
  if (j->rbits && t.count_regions) {
    bool done = false;
    TRY(count_by_regions<KT>(j, a, &done));
    if (done) {
      j->by_regions = true;
      return BRUNO_OK;
    }
  }
  ensure_table_valid(j);
  const unsigned wpb = t.parse_block / 64;
  const uint64_t waves = std::min<uint64_t>(n_reads, t.parse_grid);
  const unsigned g = (unsigned)std::max<uint64_t>(1, (waves + wpb - 1) / wpb);
  switch (t.parse_block) {
    case 128: hipLaunchKernelGGL((k_parse<KT, 2, 128>), dim3(g), dim3(128), 0, j->stream, a); break;
    case 512: hipLaunchKernelGGL((k_parse<KT, 2, 512>), dim3(g), dim3(512), 0, j->stream, a); break;
    default: hipLaunchKernelGGL((k_parse<KT, 2, 256>), dim3(g), dim3(256), 0, j->stream, a); break;
  }
  return BRUNO_OK;
}

/* After a count with the filter fused: the node lists without holes, the
 * list counters at the nodes placed (returned in placed[2]). */
template <typename KT>
bruno_status close_list_holes(bruno_job* j, uint64_t* placed) {
  TRY(sync_counters(j));
  const uint64_t nw = j->n_hole_waves, cap = j->list_cap;
  const uint64_t e[2] = {std::min<uint64_t>(j->hctr->n_chain, cap), std::min<uint64_t>(j->hctr->n_branch, cap)};
  unsigned long long* tmp = nullptr;
  uint64_t* tail = nullptr;
  TRY(dev_alloc(j, (void**)&tmp, 4 * 8));
  TRY(dev_alloc(j, (void**)&tail, std::max<uint64_t>(1, nw * FUSE_CHUNK) * 8));
  HIP_CHECK(hipMemsetAsync(tmp, 0, 4 * 8, j->stream));
  if (nw) hipLaunchKernelGGL(k_hole_count, dim3(grid_for(nw)), dim3(BLOCK), 0, j->stream, j->holes, nw, e[0], e[1], tmp);
  unsigned long long h[4];
  HIP_CHECK(hipMemcpyAsync(h, tmp, 2 * 8, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  uint64_t* lists[2] = {j->chain_slot, j->branch_slot};
  for (int l = 0; l < 2; ++l) {
    if (h[l] > e[l]) return BRUNO_INTERNAL;
    const uint64_t n = e[l] - h[l];
    placed[l] = n;
    if (n == e[l]) continue;
    HIP_CHECK(hipMemsetAsync(tmp + 2, 0, 2 * 8, j->stream));
    hipLaunchKernelGGL(k_list_tail, dim3(grid_for(e[l] - n)), dim3(BLOCK), 0, j->stream, lists[l], n, e[l], tail,
                       std::max<uint64_t>(1, nw * FUSE_CHUNK), tmp + 2);
    hipLaunchKernelGGL(k_list_fill<KT>, dim3(grid_for(nw)), dim3(BLOCK), 0, j->stream, j->holes, nw, l, lists[l], n,
                       tail, tmp + 2, tmp + 3, (Slot<KT>*)j->table);
    HIP_CHECK(hipMemcpyAsync(h + 2, tmp + 2, 2 * 8, hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipStreamSynchronize(j->stream));
    /* the nodes above n and the holes below it pair up exactly */
    if (h[2] != h[3] || h[2] > nw * FUSE_CHUNK) return BRUNO_INTERNAL;
  }
  unsigned long long nc[2] = {placed[1], placed[0]}; /* n_branch, n_chain */
  HIP_CHECK(hipMemcpyAsync(&j->ctr->n_branch, nc, 2 * 8, hipMemcpyHostToDevice, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  void* a = tmp;
  dev_free(j, a, 4 * 8);
  a = tail;
  dev_free(j, a, std::max<uint64_t>(1, nw * FUSE_CHUNK) * 8);
  a = j->holes;
  dev_free(j, a, 2 * nw * 8);
  j->holes = nullptr;
  return BRUNO_OK;
}

/* The node lists with room for cap entries each, keeping the first nc /
 * nb entries. */
static bruno_status grow_lists(bruno_job* j, uint64_t cap, uint64_t nc, uint64_t nb) {
  uint64_t *cs = nullptr, *bs = nullptr;
  TRY(dev_alloc(j, (void**)&cs, cap * 8));
  TRY(dev_alloc(j, (void**)&bs, cap * 8));
  if (nc) HIP_CHECK(hipMemcpyAsync(cs, j->chain_slot, nc * 8, hipMemcpyDeviceToDevice, j->stream));
  if (nb) HIP_CHECK(hipMemcpyAsync(bs, j->branch_slot, nb * 8, hipMemcpyDeviceToDevice, j->stream));
  void* a = j->chain_slot;
  dev_free(j, a, j->list_cap * 8);
  a = j->branch_slot;
  dev_free(j, a, j->list_cap * 8);
  j->chain_slot = cs;
  j->branch_slot = bs;
  j->list_cap = cap;
  return BRUNO_OK;
}

template <typename KT>
bruno_status filter_t(bruno_job* j) {
  ensure_table_valid(j);
  /* nodes the count placed (filter fused into it) */
  uint64_t placed[2] = {0, 0};
  if (j->fused) TRY(close_list_holes<KT>(j, placed));
  else HIP_CHECK(hipMemsetAsync(&j->ctr->n_present, 0, 3 * sizeof(unsigned long long), j->stream));
  HIP_CHECK(hipMemsetAsync(&j->ctr->n_spill, 0, sizeof(unsigned long long), j->stream));
  HIP_CHECK(hipMemsetAsync(&j->ctr->n_cand, 0, sizeof(unsigned long long), j->stream));
  /* candidates: at most the distinct k-mers; slots with spilled entries: at
   * most the side-table entries */
  const uint64_t n_distinct = std::max<uint64_t>(1, j->hctr->n_distinct);
  const uint64_t n_spill_cap = std::max<uint64_t>(1, std::min<uint64_t>(j->hctr->n_k2, j->spcap));
  uint64_t *cand = nullptr, *spilled = nullptr;
  TRY(dev_alloc(j, (void**)&cand, n_distinct * sizeof(uint64_t)));
  TRY(dev_alloc(j, (void**)&spilled, n_spill_cap * sizeof(uint64_t)));
  if (j->bits_fresh) {
    /* the last count built every region and left the scan's choice */
    const uint64_t nreg = 1ULL << j->rbits, stride = (j->rspan + 63) / 64;
    hipLaunchKernelGGL(k_filter_bits, dim3(grid_for((nreg * stride + FILTER_BITS_WPT - 1) / FILTER_BITS_WPT)),
                       dim3(BLOCK), 0, j->stream, j->cbits, j->sbits, nreg, stride, j->rspan, cand, n_distinct,
                       spilled, n_spill_cap, j->ctr);
  } else {
    hipLaunchKernelGGL((k_filter_scan<KT>),
                       dim3(grid_for((j->cap + FILTER_SPT - 1) / FILTER_SPT, BLOCK, tuning().scan_grid)),
                       dim3(BLOCK), 0, j->stream, (const Slot<KT>*)j->table, j->cap, j->thresholding, scan_lo(j),
                       cand, n_distinct, spilled, n_spill_cap, j->ctr);
  }
  TRY(sync_counters(j));
  const uint64_t n_cand = j->hctr->n_cand, n_spilled = j->hctr->n_spill;
  if (n_cand > n_distinct || n_spilled > n_spill_cap) return BRUNO_CAPACITY;
  /* nodes: at most the candidates plus the spilled slots, in either list,
   * after the nodes already placed */
  const uint64_t lcap = std::max<uint64_t>(1, std::max(placed[0], placed[1]) + n_cand + n_spilled);
  if (!j->chain_slot) {
    TRY(dev_alloc(j, (void**)&j->chain_slot, lcap * sizeof(uint64_t)));
    TRY(dev_alloc(j, (void**)&j->branch_slot, lcap * sizeof(uint64_t)));
    j->list_cap = lcap;
  } else if (lcap > j->list_cap) {
    TRY(grow_lists(j, lcap, placed[0], placed[1]));
  }
  if (j->n_jobs > 1) {
    TRY(dev_alloc(j, (void**)&j->chain_job, j->list_cap * sizeof(uint32_t)));
    TRY(dev_alloc(j, (void**)&j->branch_job, j->list_cap * sizeof(uint32_t)));
  }
  /* the chain nodes' k-mers and edge bytes for the chain extraction (not
   * after a fused count, whose build placed most nodes without them) */
  if (!j->fused) {
    TRY(dev_alloc(j, &j->ckmer, j->list_cap * sizeof(KT)));
    TRY(dev_alloc(j, (void**)&j->cedge, j->list_cap));
  }
  NodeLists nl{j->chain_slot, j->branch_slot, j->chain_job, j->branch_job, j->list_cap, j->d_seg, j->n_jobs,
               j->fused ? nullptr : j->ckmer, j->fused ? nullptr : j->cedge};
  if (n_cand)
    hipLaunchKernelGGL((k_filter_cand<KT>), dim3(grid_for(n_cand)), dim3(BLOCK), 0, j->stream, (Slot<KT>*)j->table,
                       cand, n_cand, j->K, j->thresholding, j->th, nl, j->ctr);
  if (n_spilled) {
    /* slot order (the bits give it; the scan's lists are put in it), so
     * neighbouring threads read neighbouring slots */
    uint64_t* sorted = spilled;
    void* t = nullptr;
    size_t tmp = 0;
    if (!j->bits_fresh) {
      TRY(dev_alloc(j, (void**)&sorted, n_spilled * sizeof(uint64_t)));
      int bits = 1;
      while (bits < 64 && (1ULL << bits) < j->cap) ++bits;
      HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, spilled, sorted, (int)n_spilled, 0, bits, j->stream));
      TRY(dev_alloc(j, &t, tmp));
      HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(t, tmp, spilled, sorted, (int)n_spilled, 0, bits, j->stream));
    }
    hipLaunchKernelGGL((k_filter_spilled<KT>), dim3(grid_for(n_spilled)), dim3(BLOCK), 0, j->stream,
                       (Slot<KT>*)j->table, (const Spill*)j->spill, j->spcap, j->K, j->thresholding, j->th, sorted,
                       n_spilled, nl, j->ctr);
    if (sorted != spilled) {
      dev_free(j, t, tmp);
      void* a = sorted;
      dev_free(j, a, n_spilled * sizeof(uint64_t));
    }
  }
  TRY(sync_counters(j));
  void* fa = cand;
  dev_free(j, fa, n_distinct * sizeof(uint64_t));
  fa = spilled;
  dev_free(j, fa, n_spill_cap * sizeof(uint64_t));
  j->n_branch = j->hctr->n_branch;
  j->n_chain = j->hctr->n_chain;
  j->n_present = j->n_branch + j->n_chain;
  return BRUNO_OK;
}

template <typename KT>
bruno_status chains_t(bruno_job* j) {
  uint64_t nc = j->n_chain;
  if (!j->cedge) TRY(dev_alloc(j, &j->ckmer, std::max<uint64_t>(1, nc) * sizeof(KT)));
  j->vcap = 2 * nc + 16;
  TRY(dev_alloc(j, &j->vkmer, j->vcap * sizeof(KT)));
  TRY(dev_alloc(j, &j->bkmer, std::max<uint64_t>(1, j->n_branch) * sizeof(KT)));
  TRY(dev_alloc(j, (void**)&j->lr, std::max<uint64_t>(1, nc) * sizeof(LRNode)));
  HIP_CHECK(hipMemsetAsync(&j->ctr->lr_ill, 0, sizeof(unsigned long long), j->stream));
  hipLaunchKernelGGL((k_chain_init<KT>), dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, (const Slot<KT>*)j->table,
                     probe_span(j), j->rbits, j->K, j->chain_slot, j->chain_job, j->d_seg, j->n_jobs, nc,
                     (KT*)j->ckmer, j->cedge, j->lr, (KT*)j->vkmer, j->vcap, j->ctr);
  hipLaunchKernelGGL((k_termini<KT>), dim3(grid_for(j->n_branch)), dim3(BLOCK), 0, j->stream,
                     (const Slot<KT>*)j->table, probe_span(j), j->rbits, j->K, j->branch_slot, j->branch_job,
                     j->d_seg, j->n_jobs, j->n_branch, j->lr, (KT*)j->bkmer);
  TRY(sync_counters(j));
  j->n_virt = j->hctr->n_virt;
  if (nc + j->n_virt >= (1ULL << 31)) return BRUNO_CAPACITY;
  return BRUNO_OK;
}

/* separate_cycles / separate_isolated with per-job cycle distances jcm
 * (device, 0 = no cycle separation for that job); fills the per-job and
 * total cycle / isolated counts. */
bruno_status separate(bruno_job* j, const uint32_t* jcm) {
  const uint32_t nj = j->n_jobs;
  const uint64_t nc = j->n_chain;
  unsigned long long* jsep = nullptr;
  TRY(dev_alloc(j, (void**)&jsep, 2 * nj * sizeof(unsigned long long)));
  HIP_CHECK(hipMemsetAsync(jsep, 0, 2 * nj * sizeof(unsigned long long), j->stream));
  hipLaunchKernelGGL(k_separate, dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, j->lr, nc, j->chain_job, nj, jcm,
                     j->alive, jsep);
  std::vector<unsigned long long> h(2 * nj);
  HIP_CHECK(hipMemcpyAsync(h.data(), jsep, 2 * nj * sizeof(unsigned long long), hipMemcpyDeviceToHost, j->stream));
  TRY(sync_counters(j));
  j->job_cycle.assign(nj, 0);
  j->job_isolated.assign(nj, 0);
  j->n_cycle = j->n_isolated = 0;
  for (uint32_t q = 0; q < nj; ++q) {
    j->job_cycle[q] = h[2 * q];
    j->job_isolated[q] = h[2 * q + 1];
    j->n_cycle += h[2 * q];
    j->n_isolated += h[2 * q + 1];
  }
  void* a = jsep;
  dev_free(j, a, 2 * nj * sizeof(unsigned long long));
  return BRUNO_OK;
}

/* Returns BRUNO_OK with *used = true when the contraction path produced the
 * final state; *used = false means the graph is not well-formed (or a walk
 * ran out) and the caller runs the round-by-round emulation instead. */
bruno_status list_rank_fast(bruno_job* j, bool* used) {
  *used = false;
  uint64_t nc = j->n_chain;
  if (nc == 0) return BRUNO_OK;
  /* ruler flags as a bitmap (n_chain / 8 bytes: small enough to stay in
   * the device caches for the walks' random tests) */
  uint64_t* ruler = nullptr;
  const uint64_t ruler_bytes = (nc + 63) / 64 * 8;
  uint32_t* rlist = nullptr;
  uint32_t* heads = nullptr;
  WalkRec *rend = nullptr, *link = nullptr, *info = nullptr;
  /* k_chain_init counted the ill-formed ends */
  const bool well_formed = j->hctr->lr_ill == 0;
  if (!well_formed) return BRUNO_OK;
  TRY(dev_alloc(j, (void**)&ruler, ruler_bytes));
  j->hctr->lr_conflict = 0;
  j->hctr->lr_next = 0;
  j->hctr->lr_count = 0;
  j->hctr->lr_cycle = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->lr_count, &j->hctr->lr_count, 4 * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, j->stream));
  TRY(dev_alloc(j, (void**)&rlist, nc * 4));
  hipLaunchKernelGGL(k_lr_rulers, dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, j->lr, nc, ruler,
                     (uint64_t)tuning().ruler_sample - 1, rlist, nc, j->ctr);
  TRY(sync_counters(j));
  uint64_t nr = j->hctr->lr_next;
  TRY(dev_alloc(j, (void**)&rend, std::max<uint64_t>(1, nr) * 2 * sizeof(WalkRec)));
  TRY(dev_alloc(j, (void**)&link, nc * 2 * sizeof(WalkRec)));
  TRY(dev_alloc(j, (void**)&info, nc * 2 * sizeof(WalkRec)));
  fill64(j, link, nc * 2, ~0ULL);
  fill64(j, info, nc * 2, ~0ULL);
  if (nr)
    hipLaunchKernelGGL(k_lr_walk_nodes, dim3(grid_for(2 * nr, BLOCK, 1u << 20)), dim3(BLOCK), 0, j->stream, j->lr,
                       ruler, rlist, nr, link, j->ctr);
  TRY(dev_alloc(j, (void**)&heads, std::max<uint64_t>(1, 2 * nr) * 4));
  j->hctr->lr_next = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->lr_next, &j->hctr->lr_next, sizeof(unsigned long long), hipMemcpyHostToDevice,
                           j->stream));
  if (nr)
    hipLaunchKernelGGL(k_lr_list_heads, dim3(grid_for(2 * nr, BLOCK, 1u << 20)), dim3(BLOCK), 0, j->stream, j->lr,
                       rlist, nr, heads, j->ctr);
  /* a walk that ran past its bound leaves the chain state untouched: clear
   * the flag and let the round-by-round path run */
  auto walk_failed = [&](bruno_status st, bool* fell_back) -> bruno_status {
    *fell_back = false;
    if (j->hctr->err == ERR_WALK) {
      j->hctr->err = 0;
      HIP_CHECK(hipMemsetAsync(&j->ctr->err, 0, 4, j->stream));
      hipLaunchKernelGGL(k_lr_zero_mail, dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, j->lr, nc);
      HIP_CHECK(hipStreamSynchronize(j->stream));
      *fell_back = true;
      return BRUNO_OK;
    }
    return st;
  };
  bool fell = false;
  bruno_status s = sync_counters(j);
  if (s != BRUNO_OK) {
    s = walk_failed(s, &fell);
    return s;
  }
  uint64_t nh = j->hctr->lr_next;
  if (nh)
    hipLaunchKernelGGL(k_lr_walk_rulers, dim3(grid_for(nh)), dim3(BLOCK), 0, j->stream, j->lr, link, heads, nh, nr,
                       info, j->ctr);
  s = sync_counters(j);
  if (s != BRUNO_OK) {
    s = walk_failed(s, &fell);
    return s;
  }
  const uint32_t nj = j->n_jobs;
  uint8_t* cyc = nullptr;
  uint32_t* jbuf = nullptr; /* jT[nj], jcyc[nj], jcm[nj] */
  TRY(dev_alloc(j, (void**)&cyc, nc));
  TRY(dev_alloc(j, (void**)&jbuf, 3 * nj * sizeof(uint32_t)));
  HIP_CHECK(hipMemsetAsync(jbuf, 0, 2 * nj * sizeof(uint32_t), j->stream));
  if (nr)
    hipLaunchKernelGGL(k_lr_ruler_ends, dim3(grid_for(2 * nr, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, j->lr,
                       rlist, nr, info, rend);
  hipLaunchKernelGGL(k_lr_finish, dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, j->lr, nc, ruler, rend, info, cyc,
                     j->chain_job, nj, jbuf, jbuf + nj, j->ctr);
  std::vector<uint32_t> hj(3 * nj, 0);
  HIP_CHECK(hipMemcpyAsync(hj.data(), jbuf, 2 * nj * sizeof(uint32_t), hipMemcpyDeviceToHost, j->stream));
  TRY(sync_counters(j));
  const uint64_t ncyc = j->hctr->lr_cycle;
  /* per job: T = largest ceil(log2 L), at least 1 with a cycle; cycle nodes
   * end at distance 2^T and are separated */
  j->job_iters.assign(nj, 0);
  uint32_t Tmax = 0;
  for (uint32_t q = 0; q < nj; ++q) {
    uint32_t T = hj[q];
    if (hj[nj + q] && T < 1) T = 1;
    if (T >= 30) return BRUNO_INTERNAL; /* the reference stops at 30 rounds; not reachable below 2^29 nodes */
    j->job_iters[q] = T;
    hj[2 * nj + q] = hj[nj + q] ? (1u << T) : 0u;
    Tmax = std::max(Tmax, T);
  }
  uint32_t* jcm = jbuf + 2 * nj;
  HIP_CHECK(hipMemcpyAsync(jcm, hj.data() + 2 * nj, nj * sizeof(uint32_t), hipMemcpyHostToDevice, j->stream));
  if (ncyc)
    hipLaunchKernelGGL(k_lr_cycles, dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, j->lr, nc, cyc, j->chain_job, nj,
                       jcm);
  j->iterations = Tmax;
  j->conflicts = 0;
  TRY(dev_alloc(j, (void**)&j->alive, nc));
  TRY(separate(j, jcm));
  *used = true;
  return BRUNO_OK;
}

bruno_status list_rank_rounds(bruno_job* j, int variant) {
  const uint64_t nc = j->n_chain;
  const uint32_t nj = j->n_jobs;
  const bool batch = nj > 1;
  TRY(dev_alloc(j, (void**)&j->list[0], std::max<uint64_t>(1, nc) * sizeof(uint32_t)));
  TRY(dev_alloc(j, (void**)&j->list[1], std::max<uint64_t>(1, nc) * sizeof(uint32_t)));
  TRY(dev_alloc(j, (void**)&j->alive, std::max<uint64_t>(1, nc)));
  /* batch: per-job round counters and active flags */
  unsigned long long* jctr = nullptr;
  uint8_t* active = nullptr;
  uint32_t* jcm = nullptr;
  std::vector<unsigned long long> hc(3 * nj, 0);
  std::vector<uint8_t> act(nj, 0);
  TRY(dev_alloc(j, (void**)&jcm, nj * sizeof(uint32_t)));
  if (batch) {
    TRY(dev_alloc(j, (void**)&jctr, 3 * nj * sizeof(unsigned long long)));
    TRY(dev_alloc(j, (void**)&active, nj));
    HIP_CHECK(hipMemsetAsync(jctr, 0, 3 * nj * sizeof(unsigned long long), j->stream));
  }
  j->hctr->lr_next = 0;
  j->hctr->lr_count = 0;
  j->hctr->lr_cycle = 0;
  j->hctr->lr_conflict = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->lr_count, &j->hctr->lr_count, 4 * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, j->stream));
  hipLaunchKernelGGL(k_lr_init_list, dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, j->lr, nc, j->list[0],
                     j->chain_job, jctr, j->ctr);
  if (batch) HIP_CHECK(hipMemcpyAsync(hc.data(), jctr, 3 * nj * 8, hipMemcpyDeviceToHost, j->stream));
  TRY(sync_counters(j));
  uint64_t n = j->hctr->lr_next;
  /* unfinished nodes per job; a job without any takes no round */
  std::vector<uint64_t> unf(nj, 0);
  if (batch) {
    for (uint32_t q = 0; q < nj; ++q) unf[q] = hc[3 * q + 1];
  } else {
    unf[0] = n;
  }
  uint32_t n_active = 0;
  for (uint32_t q = 0; q < nj; ++q) {
    act[q] = unf[q] > 0;
    n_active += act[q];
  }
  if (batch) HIP_CHECK(hipMemcpyAsync(active, act.data(), nj, hipMemcpyHostToDevice, j->stream));
  j->job_iters.assign(nj, 0);
  uint32_t iters = 0;
  int cur = 0;
  uint64_t conflicts = 0;
  while (n_active > 0) {
    j->hctr->lr_count = 0;
    j->hctr->lr_next = 0;
    j->hctr->lr_cycle = 0;
    j->hctr->lr_conflict = 0;
    HIP_CHECK(hipMemcpyAsync(&j->ctr->lr_count, &j->hctr->lr_count, 4 * sizeof(unsigned long long),
                             hipMemcpyHostToDevice, j->stream));
    if (batch) HIP_CHECK(hipMemsetAsync(jctr, 0, 3 * nj * sizeof(unsigned long long), j->stream));
    hipLaunchKernelGGL(k_lr_send, dim3(grid_for(n)), dim3(BLOCK), 0, j->stream, j->lr, nc, j->list[cur], n, variant,
                       j->chain_job, active, j->ctr);
    /* after this round the iteration count is iters+1 and IsCycleNode
     * compares against 1 << (iters+1) (debruijn_chain_filters.hpp:189-213);
     * every active job has taken every round so far */
    uint32_t maxd = 1u << (iters + 1);
    hipLaunchKernelGGL(k_lr_recv, dim3(grid_for(n)), dim3(BLOCK), 0, j->stream, j->lr, j->list[cur], n,
                       j->list[cur ^ 1], maxd, j->chain_job, active, jctr, j->ctr);
    if (batch) HIP_CHECK(hipMemcpyAsync(hc.data(), jctr, 3 * nj * 8, hipMemcpyDeviceToHost, j->stream));
    TRY(sync_counters(j));
    n = j->hctr->lr_next;
    conflicts += j->hctr->lr_conflict;
    cur ^= 1;
    ++iters;
    if (!batch) {
      hc[0] = j->hctr->lr_count;
      hc[1] = n;
      hc[2] = j->hctr->lr_cycle;
    }
    bool changed = false;
    for (uint32_t q = 0; q < nj; ++q) {
      if (!act[q]) continue;
      ++j->job_iters[q];
      const uint64_t count = hc[3 * q], u = hc[3 * q + 1], cyc = hc[3 * q + 2];
      unf[q] = u;
      /* stop when nothing changed or only cycle-like nodes remain, or at
       * 30 rounds (debruijn_chain_graph.hpp:1238-1277) */
      if (j->job_iters[q] >= 30 || count == 0 || cyc == u) {
        act[q] = 0;
        --n_active;
        changed = true;
      }
    }
    if (batch && changed) HIP_CHECK(hipMemcpyAsync(active, act.data(), nj, hipMemcpyHostToDevice, j->stream));
  }
  j->iterations = iters;
  j->conflicts = conflicts;
  /* separate_cycles only when unfinished nodes remain (:1284-1291) */
  std::vector<uint32_t> hm(nj, 0);
  for (uint32_t q = 0; q < nj; ++q) hm[q] = unf[q] > 0 ? (1u << j->job_iters[q]) : 0u;
  HIP_CHECK(hipMemcpyAsync(jcm, hm.data(), nj * sizeof(uint32_t), hipMemcpyHostToDevice, j->stream));
  return separate(j, jcm);
}

template <typename KT>
bruno_status radix_pairs(bruno_job* j, uint64_t* keys_in, uint64_t* keys_out, uint32_t* vals_in, uint32_t* vals_out,
                         uint64_t n, int end_bit) {
  size_t tmp = 0;
  HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys_in, keys_out, vals_in, vals_out, (int)n, 0, end_bit,
                                               j->stream));
  void* t = nullptr;
  TRY(dev_alloc(j, &t, tmp));
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(t, tmp, keys_in, keys_out, vals_in, vals_out, (int)n, 0, end_bit,
                                                    j->stream);
  HIP_CHECK(hipStreamSynchronize(j->stream));
  dev_free(j, t, tmp);
  HIP_CHECK(e);
  return BRUNO_OK;
}

/* stable LSD sort of (key words) producing a permutation of [0, n) */
template <typename KT>
bruno_status sort_perm_kmer(bruno_job* j, const KT* keys, uint32_t* perm, uint64_t n, int K, bool have_perm) {
  if (n == 0) return BRUNO_OK;
  uint64_t *kin = nullptr, *kout = nullptr;
  uint32_t* pout = nullptr;
  TRY(dev_alloc(j, (void**)&kin, n * 8));
  TRY(dev_alloc(j, (void**)&kout, n * 8));
  TRY(dev_alloc(j, (void**)&pout, n * 4));
  if (!have_perm) hipLaunchKernelGGL(k_iota, dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, perm, n);
  int W = kt_traits<KT>::words;
  for (int w = 0; w < W; ++w) {
    int bits = std::min(64, 2 * K - 64 * w);
    if (bits <= 0) break;
    hipLaunchKernelGGL((k_gather_key<KT>), dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, keys, perm,
                       kin, n, w);
    TRY(radix_pairs<KT>(j, kin, kout, perm, pout, n, bits));
    HIP_CHECK(hipMemcpyAsync(perm, pout, n * 4, hipMemcpyDeviceToDevice, j->stream));
  }
  void* a = kin;
  dev_free(j, a, n * 8);
  a = kout;
  dev_free(j, a, n * 8);
  a = pout;
  dev_free(j, a, n * 4);
  return BRUNO_OK;
}

/* Stable pass over a 32-bit key: perm is reordered by keys[perm[i]]'s low
 * `bits` bits, keeping the current order within equal keys. */
bruno_status sort_perm_u32(bruno_job* j, const uint32_t* keys, uint32_t* perm, uint64_t n, int bits) {
  if (n == 0 || bits <= 0) return BRUNO_OK;
  uint32_t *kin = nullptr, *kout = nullptr, *pout = nullptr;
  TRY(dev_alloc(j, (void**)&kin, n * 4));
  TRY(dev_alloc(j, (void**)&kout, n * 4));
  TRY(dev_alloc(j, (void**)&pout, n * 4));
  hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, keys, perm, kin, n);
  size_t tmp = 0;
  HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, perm, pout, (int)n, 0, bits, j->stream));
  void* t = nullptr;
  TRY(dev_alloc(j, &t, tmp));
  HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(t, tmp, kin, kout, perm, pout, (int)n, 0, bits, j->stream));
  HIP_CHECK(hipMemcpyAsync(perm, pout, n * 4, hipMemcpyDeviceToDevice, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  dev_free(j, t, tmp);
  void* a = kin;
  dev_free(j, a, n * 4);
  a = kout;
  dev_free(j, a, n * 4);
  a = pout;
  dev_free(j, a, n * 4);
  return BRUNO_OK;
}

int job_bits(uint32_t n_jobs) {
  int b = 0;
  while (b < 32 && (1ull << b) < n_jobs) ++b;
  return b;
}

/* Per-job start positions in a job-sorted order: out[0..n_jobs] on the
 * host.  job_of_item == nullptr (one job) gives {0, n}. */
bruno_status job_bounds(bruno_job* j, const uint32_t* job_of_item, const uint32_t* perm, uint64_t n,
                        std::vector<uint64_t>& out) {
  const uint32_t nj = j->n_jobs;
  out.assign(nj + 1, 0);
  out[nj] = n;
  if (nj == 1) return BRUNO_OK;
  uint32_t* js = nullptr;
  uint64_t* st = nullptr;
  TRY(dev_alloc(j, (void**)&js, std::max<uint64_t>(1, n) * 4));
  TRY(dev_alloc(j, (void**)&st, (nj + 1) * 8));
  if (n)
    hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, job_of_item, perm,
                       js, n);
  hipLaunchKernelGGL(k_seg_bounds, dim3(grid_for(n + 1, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, js, n, nj, st);
  HIP_CHECK(hipMemcpyAsync(out.data(), st, (nj + 1) * 8, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  void* a = js;
  dev_free(j, a, std::max<uint64_t>(1, n) * 4);
  a = st;
  dev_free(j, a, (nj + 1) * 8);
  return BRUNO_OK;
}

/* chain_rep_less order without sorting every node: after the contraction
 * ranking each chain's nodes carry distances 0..m-1 from its rep end, so
 * only the chain heads are sorted (by job, then rep) and every node goes to
 * its chain's offset + its distance.  The result is the sorted order when
 * no node is rejected and no position stays empty: n nodes then fill n
 * places one each, the chains' keys strictly increase and each chain's
 * nodes sit in dist order.  Otherwise *placed = false and the caller
 * sorts. */
template <typename KT>
bruno_status place_by_chains(bruno_job* j, uint64_t n, const uint32_t* out_head, const uint32_t* out_len, bool* placed,
                             SortedRec<KT>** sr_out) {
  *placed = false;
  *sr_out = nullptr;
  const uint64_t nc = j->n_chain;
  const uint32_t nj = j->n_jobs;
  j->hctr->n_cand = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->n_cand, &j->hctr->n_cand, 8, hipMemcpyHostToDevice, j->stream));
  uint32_t* heads = nullptr;
  TRY(dev_alloc(j, (void**)&heads, n * 4));
  hipLaunchKernelGGL(k_list_heads_out, dim3(grid_for((n + HEAD_ITEMS - 1) / HEAD_ITEMS)), dim3(BLOCK), 0, j->stream,
                     j->out_dist, n, heads, j->ctr);
  TRY(sync_counters(j));
  const uint64_t nh = j->hctr->n_cand;
  void* a = nullptr;
  if (nh == 0 || nh > n) {
    a = heads;
    dev_free(j, a, n * 4);
    return BRUNO_OK;
  }
  const dim3 gh(grid_for(nh, BLOCK, 1u << 30)), gn(grid_for(n, BLOCK, 1u << 30));
  /* heads in (job, rep) order */
  KT* hkeys = nullptr;
  uint32_t *hperm = nullptr, *hs = nullptr, *hjob = nullptr;
  TRY(dev_alloc(j, (void**)&hkeys, nh * sizeof(KT)));
  TRY(dev_alloc(j, (void**)&hperm, nh * 4));
  TRY(dev_alloc(j, (void**)&hs, nh * 4));
  hipLaunchKernelGGL((k_gather_kmer<KT>), gh, dim3(BLOCK), 0, j->stream, (const KT*)j->out_rep, heads, hkeys, nh);
  TRY(sort_perm_kmer<KT>(j, hkeys, hperm, nh, j->K, false));
  if (nj > 1) {
    TRY(dev_alloc(j, (void**)&hjob, nh * 4));
    hipLaunchKernelGGL(k_gather_u32, gh, dim3(BLOCK), 0, j->stream, j->out_job, heads, hjob, nh);
    TRY(sort_perm_u32(j, hjob, hperm, nh, job_bits(nj)));
  }
  hipLaunchKernelGGL(k_gather_u32, gh, dim3(BLOCK), 0, j->stream, heads, hperm, hs, nh);
  /* chain offsets in that order */
  uint64_t *hl = nullptr, *coff = nullptr;
  HeadMeta<KT>* meta = nullptr;
  unsigned long long* bad = nullptr;
  TRY(dev_alloc(j, (void**)&hl, nh * 8));
  TRY(dev_alloc(j, (void**)&coff, nh * 8));
  hipLaunchKernelGGL(k_take_len, gh, dim3(BLOCK), 0, j->stream, hs, out_len, nh, hl);
  size_t tmp = 0;
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, hl, coff, (int)nh, j->stream));
  void* t = nullptr;
  TRY(dev_alloc(j, &t, tmp));
  HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t, tmp, hl, coff, (int)nh, j->stream));
  TRY(dev_alloc(j, (void**)&meta, nc * sizeof(HeadMeta<KT>)));
  TRY(dev_alloc(j, (void**)&bad, 8));
  /* len 0 marks a node that heads no chain */
  SortedRec<KT>* sr = nullptr;
  TRY(dev_alloc(j, (void**)&sr, n * sizeof(SortedRec<KT>)));
  HIP_CHECK(hipMemsetAsync(meta, 0, nc * sizeof(HeadMeta<KT>), j->stream));
  HIP_CHECK(hipMemsetAsync(sr, 0xFF, n * sizeof(SortedRec<KT>), j->stream));
  HIP_CHECK(hipMemsetAsync(bad, 0, 8, j->stream));
  hipLaunchKernelGGL((k_head_meta<KT>), gh, dim3(BLOCK), 0, j->stream, hs, coff, out_head, out_len,
                     (const KT*)j->out_rep, j->out_job, nh, nc, meta, bad);
  hipLaunchKernelGGL((k_scatter_ranked<KT>), gn, dim3(BLOCK), 0, j->stream, out_head, j->out_dist,
                     (const KT*)j->out_rep, (const KT*)j->out_k, j->out_job, n, nc, meta, sr, bad);
  hipLaunchKernelGGL((k_sorted_split<KT>), dim3(grid_for(n)), dim3(BLOCK), 0, j->stream, sr, n, j->perm,
                     j->dist_sorted, bad);
  unsigned long long hbad = 0;
  HIP_CHECK(hipMemcpyAsync(&hbad, bad, 8, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  *placed = hbad == 0;
  if (*placed) {
    *sr_out = sr;
  } else {
    a = sr;
    dev_free(j, a, n * sizeof(SortedRec<KT>));
  }
  a = heads;
  dev_free(j, a, n * 4);
  a = hkeys;
  dev_free(j, a, nh * sizeof(KT));
  a = hperm;
  dev_free(j, a, nh * 4);
  a = hs;
  dev_free(j, a, nh * 4);
  if (hjob) {
    a = hjob;
    dev_free(j, a, nh * 4);
  }
  a = hl;
  dev_free(j, a, nh * 8);
  a = coff;
  dev_free(j, a, nh * 8);
  dev_free(j, t, tmp);
  a = meta;
  dev_free(j, a, nc * sizeof(HeadMeta<KT>));
  a = bad;
  dev_free(j, a, 8);
  return BRUNO_OK;
}

template <typename KT>
bruno_status compact_t(bruno_job* j) {
  uint64_t nc = j->n_chain;
  int K = j->K;
  TRY(dev_alloc(j, &j->out_k, std::max<uint64_t>(1, nc) * sizeof(KT)));
  TRY(dev_alloc(j, &j->out_rep, std::max<uint64_t>(1, nc) * sizeof(KT)));
  TRY(dev_alloc(j, (void**)&j->out_dist, std::max<uint64_t>(1, nc) * 4));
  TRY(dev_alloc(j, (void**)&j->perm, std::max<uint64_t>(1, nc) * 4));
  TRY(dev_alloc(j, (void**)&j->dist_sorted, std::max<uint64_t>(1, nc) * 4));
  const uint32_t nj = j->n_jobs;
  if (nj > 1) TRY(dev_alloc(j, (void**)&j->out_job, std::max<uint64_t>(1, nc) * 4));
  j->hctr->n_out = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->n_out, &j->hctr->n_out, sizeof(unsigned long long), hipMemcpyHostToDevice,
                           j->stream));
  uint32_t *out_head = nullptr, *out_len = nullptr;
  if (j->lr_contracted) {
    TRY(dev_alloc(j, (void**)&out_head, std::max<uint64_t>(1, nc) * 4));
    TRY(dev_alloc(j, (void**)&out_len, std::max<uint64_t>(1, nc) * 4));
  }
  hipLaunchKernelGGL((k_ranked<KT>), dim3(grid_for((nc + RANKED_ITEMS - 1) / RANKED_ITEMS)), dim3(BLOCK), 0,
                     j->stream, j->lr, (const KT*)j->ckmer, nc, (const KT*)j->vkmer, j->alive, K, j->chain_job,
                     (KT*)j->out_k, (KT*)j->out_rep, j->out_dist, j->out_job, out_head, out_len, j->ctr);
  TRY(sync_counters(j));
  uint64_t n = j->hctr->n_out;
  j->n_out = n;
  /* chain_rep_less: by rep, then |dist| -> LSD: dist first, then rep words;
   * a batch adds a last pass on the job, so each job's nodes are contiguous */
  j->fasta_off.assign(nj + 1, 0);
  bool placed = false;
  SortedRec<KT>* sr = nullptr; /* the nodes in output order, when placed by chains */
  if (n > 0 && out_head) TRY(place_by_chains<KT>(j, n, out_head, out_len, &placed, &sr));
  j->out_by_chains = placed;
  if (n > 0 && !placed) {
    uint32_t* pin = nullptr;
    TRY(dev_alloc(j, (void**)&pin, n * 4));
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, pin, n);
    uint32_t* dout = nullptr;
    TRY(dev_alloc(j, (void**)&dout, n * 4));
    size_t tmp = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, j->out_dist, dout, pin, j->perm, (int)n, 0, 32,
                                                 j->stream));
    void* t = nullptr;
    TRY(dev_alloc(j, &t, tmp));
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(t, tmp, j->out_dist, dout, pin, j->perm, (int)n, 0, 32, j->stream));
    HIP_CHECK(hipStreamSynchronize(j->stream));
    dev_free(j, t, tmp);
    void* a = dout;
    dev_free(j, a, n * 4);
    a = pin;
    dev_free(j, a, n * 4);
    TRY(sort_perm_kmer<KT>(j, (const KT*)j->out_rep, j->perm, n, K, true));
    if (nj > 1) TRY(sort_perm_u32(j, j->out_job, j->perm, n, job_bits(nj)));
  }
  void* a = nullptr;
  if (out_head) {
    a = out_head;
    dev_free(j, a, std::max<uint64_t>(1, nc) * 4);
    a = out_len;
    dev_free(j, a, std::max<uint64_t>(1, nc) * 4);
  }
  if (n > 0) {
    void* t = nullptr;
    size_t tmp = 0;
    TRY(job_bounds(j, j->out_job, j->perm, n, j->chain_start));
    if (!sr)
      hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, j->out_dist,
                         j->perm, j->dist_sorted, n);
    /* .chain.fasta */
    uint64_t *len = nullptr, *off = nullptr;
    TRY(dev_alloc(j, (void**)&len, n * 8));
    TRY(dev_alloc(j, (void**)&off, (n + 1) * 8));
    hipLaunchKernelGGL(k_fasta_len, dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, j->dist_sorted, n,
                       K, len);
    tmp = 0;
    HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, len, off + 1, (int)n, j->stream));
    TRY(dev_alloc(j, &t, tmp));
    HIP_CHECK(hipMemsetAsync(off, 0, 8, j->stream));
    HIP_CHECK(hipcub::DeviceScan::InclusiveSum(t, tmp, len, off + 1, (int)n, j->stream));
    uint64_t total = 0;
    HIP_CHECK(hipMemcpyAsync(&total, off + n, 8, hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipStreamSynchronize(j->stream));
    dev_free(j, t, tmp);
    TRY(dev_alloc(j, (void**)&j->fasta, total + 1));
    if (sr)
      hipLaunchKernelGGL((k_fasta_write_sorted<KT>), dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                         sr, (const KT*)j->out_rep, off, n, K, j->fasta);
    else
      hipLaunchKernelGGL((k_fasta_write<KT>), dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, j->perm,
                         (const KT*)j->out_k, (const KT*)j->out_rep, j->dist_sorted, off, n, K, j->fasta);
    j->fasta_len = total;
    /* byte range of each job's records */
    j->fasta_off[nj] = total;
    if (nj > 1) {
      uint64_t *cs = nullptr, *fo = nullptr;
      TRY(dev_alloc(j, (void**)&cs, (nj + 1) * 8));
      TRY(dev_alloc(j, (void**)&fo, (nj + 1) * 8));
      HIP_CHECK(hipMemcpyAsync(cs, j->chain_start.data(), (nj + 1) * 8, hipMemcpyHostToDevice, j->stream));
      hipLaunchKernelGGL(k_take64, dim3(grid_for(nj + 1)), dim3(BLOCK), 0, j->stream, off, cs, nj + 1, fo);
      HIP_CHECK(hipMemcpyAsync(j->fasta_off.data(), fo, (nj + 1) * 8, hipMemcpyDeviceToHost, j->stream));
      HIP_CHECK(hipStreamSynchronize(j->stream));
      a = cs;
      dev_free(j, a, (nj + 1) * 8);
      a = fo;
      dev_free(j, a, (nj + 1) * 8);
    }
    a = len;
    dev_free(j, a, n * 8);
    a = off;
    dev_free(j, a, (n + 1) * 8);
    a = sr;
    dev_free(j, a, n * sizeof(SortedRec<KT>));
  } else {
    j->fasta_len = 0;
    j->chain_start.assign(nj + 1, 0);
    TRY(dev_alloc(j, (void**)&j->fasta, 16));
  }
  /* .branch.fasta: sorted branch k-mers */
  uint64_t nb = j->n_branch;
  j->branch_start.assign(nj + 1, 0);
  TRY(dev_alloc(j, &j->bsorted, std::max<uint64_t>(1, nb) * sizeof(KT)));
  j->bfasta_len = nb * (uint64_t)(2 * K + 3);
  TRY(dev_alloc(j, (void**)&j->bfasta, j->bfasta_len + 1));
  if (nb > 0) {
    uint32_t* bp = nullptr;
    TRY(dev_alloc(j, (void**)&bp, nb * 4));
    TRY(sort_perm_kmer<KT>(j, (const KT*)j->bkmer, bp, nb, K, false));
    if (nj > 1) TRY(sort_perm_u32(j, j->branch_job, bp, nb, job_bits(nj)));
    TRY(job_bounds(j, j->branch_job, bp, nb, j->branch_start));
    hipLaunchKernelGGL((k_gather_kmer<KT>), dim3(grid_for(nb, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                       (const KT*)j->bkmer, bp, (KT*)j->bsorted, nb);
    hipLaunchKernelGGL((k_branch_fasta<KT>), dim3(grid_for(nb, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                       (const KT*)j->bsorted, nb, K, j->bfasta);
    void* a = bp;
    dev_free(j, a, nb * 4);
  }
  HIP_CHECK(hipStreamSynchronize(j->stream));
  /* the file drops the first character (compact_dbg_io.hpp:351-355) */
  return BRUNO_OK;
}

}  // namespace


/* ================================================================== */
/* C ABI                                                                */
/* ================================================================== */
namespace {
bool valid_job(const bruno_job* j) { return j != nullptr; }
}  // namespace

extern "C" {

const char* bruno_status_string(bruno_status s) {
  switch (s) {
    case BRUNO_OK: return "ok";
    case BRUNO_BAD_INPUT: return "bad input";
    case BRUNO_OOM: return "device out of memory";
    case BRUNO_CAPACITY: return "capacity exceeded (hash table, side table or list)";
    case BRUNO_INTERNAL: return "internal device error";
    case BRUNO_STATE: return "call out of order";
    case BRUNO_NO_DEVICE: return "no HIP device";
  }
  return "unknown";
}

static bruno_status create_common(const bruno_job_desc* d, uint32_t n_jobs, const uint64_t* caps, bruno_job** out,
                                  bool plain = false) {
  if (!d || !out) return BRUNO_BAD_INPUT;
  *out = nullptr;
  if (d->k < 3 || d->k > 63 || n_jobs == 0 || d->thresholding < 0 || d->thresholding > 2) return BRUNO_BAD_INPUT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return BRUNO_NO_DEVICE;
  }
  bruno_job* j = new (std::nothrow) bruno_job();
  if (!j) return BRUNO_OOM;
  j->K = d->k;
  j->words = d->k <= 31 ? 1 : 2;
  j->thresholding = d->thresholding;
  j->th = Thresh{d->thresholds[0], d->thresholds[1], d->thresholds[2],
                 d->thresholds[3], d->thresholds[4], d->thresholds[5]};
  j->load_factor = (d->load_factor > 0 && d->load_factor < 1) ? d->load_factor : 0.5;
  j->stream = (hipStream_t)d->stream;
  j->n_jobs = n_jobs;
  j->region_ok = plain;
  j->job_reads.assign(n_jobs, 0);
  if (pool_alloc((void**)&j->ctr, sizeof(Counters)) != hipSuccess ||
      hipHostMalloc((void**)&j->hctr, sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    pool_free(j->ctr);
    delete j;
    return BRUNO_OOM;
  }
  memset(j->hctr, 0, sizeof(Counters));
  if (hipMemsetAsync(j->ctr, 0, sizeof(Counters), j->stream) != hipSuccess) {
    (void)hipGetLastError();
  }
  j->ev_ok = true;
  for (int i = 0; i < 12; ++i)
    if (hipEventCreate(&j->ev[i]) != hipSuccess) {
      (void)hipGetLastError();
      j->ev_ok = false;
    }
  if (caps) {
    bruno_status s = alloc_table(j, caps);
    if (s == BRUNO_OK && hipStreamSynchronize(j->stream) != hipSuccess) {
      (void)hipGetLastError();
      s = BRUNO_INTERNAL;
    }
    if (s != BRUNO_OK) {
      bruno_job_destroy(j);
      return s;
    }
  }
  *out = j;
  return BRUNO_OK;
}

bruno_status bruno_job_create(const bruno_job_desc* d, bruno_job** out) {
  if (!d) return BRUNO_BAD_INPUT;
  const uint64_t cap = d->capacity;
  return create_common(d, 1, cap ? &cap : nullptr, out, true);
}

bruno_status bruno_batch_create(const bruno_job_desc* d, uint32_t n_jobs, const uint64_t* capacities,
                                bruno_job** out) {
  if (!capacities || n_jobs == 0 || n_jobs > BRUNO_MAX_BATCH_JOBS) return BRUNO_BAD_INPUT;
  for (uint32_t i = 0; i < n_jobs; ++i)
    if (capacities[i] == 0) return BRUNO_BAD_INPUT;
  /* one job: a plain job (the region layout when its table is large) */
  return create_common(d, n_jobs, capacities, out, n_jobs == 1);
}

/* Slots from which a one-job table counts by regions (0: never, the
 * region count switched off); for the request server, not exported. */
__attribute__((visibility("hidden"))) uint64_t br_region_count_min_slots() {
  const uint32_t cr = tuning().count_regions;
  return cr == 0 ? 0 : cr >= 2 ? REGION_MIN_CAP_ANY : REGION_MIN_CAP;
}

bruno_status bruno_job_destroy(bruno_job* j) {
  if (!j) return BRUNO_BAD_INPUT;
  (void)hipStreamSynchronize(j->stream);
  /* the side stream is shared: wait for this job's last work on it only
   * (the count joins it back to the job's stream; the table clear records
   * ev_cleared) */
  if (j->clear_pending && j->ev_cleared) (void)hipEventSynchronize(j->ev_cleared);
  if (j->ev_chains) (void)hipEventDestroy(j->ev_chains);
  for (hipEvent_t e : j->cev)
    if (e) (void)hipEventDestroy(e);
  if (j->ev_cleared) (void)hipEventDestroy(j->ev_cleared);
  for (void* p : j->allocs) pool_free(p);
  j->allocs.clear();
  pool_free(j->ctr);
  if (j->hctr) (void)hipHostFree(j->hctr);
  if (j->ev_ok)
    for (int i = 0; i < 12; ++i) (void)hipEventDestroy(j->ev[i]);
  delete j;
  return BRUNO_OK;
}

bruno_status bruno_job_reset(bruno_job* j) {
  if (!valid_job(j)) return BRUNO_BAD_INPUT;
  HIP_CHECK(hipStreamSynchronize(j->stream));
  std::vector<void*> keep;
  for (void* p : j->allocs) {
    if (persistent(j, p)) keep.push_back(p);
    else pool_free(p);
  }
  j->allocs = keep;
  j->device_bytes = persistent_bytes(j);
  j->peak_bytes = j->device_bytes;
  j->chain_slot = j->branch_slot = nullptr;
  j->ckmer = j->vkmer = j->bkmer = j->out_k = j->out_rep = j->bsorted = nullptr;
  j->cedge = nullptr;
  j->lr = nullptr;
  j->list[0] = j->list[1] = nullptr;
  j->alive = nullptr;
  j->out_dist = j->perm = j->dist_sorted = nullptr;
  j->chain_job = j->branch_job = j->out_job = nullptr;
  j->fasta = j->bfasta = nullptr;
  j->n_present = j->n_branch = j->n_chain = j->n_virt = 0;
  j->iterations = 0;
  j->n_cycle = j->n_isolated = j->conflicts = 0;
  j->lr_contracted = j->out_by_chains = false;
  j->bits_fresh = false;
  j->fused = false;
  j->holes = nullptr;
  j->n_hole_waves = 0;
  j->cbits = j->sbits = nullptr;
  j->n_out = j->n_unitigs = j->fasta_len = j->bfasta_len = 0;
  j->n_reads = 0;
  j->job_reads.assign(j->n_jobs, 0);
  j->job_iters.clear();
  j->job_cycle.clear();
  j->job_isolated.clear();
  j->chain_start.clear();
  j->fasta_off.clear();
  j->branch_start.clear();
  j->ms = bruno_phase_ms{};
  memset(j->hctr, 0, sizeof(Counters));
  HIP_CHECK(hipMemsetAsync(j->ctr, 0, sizeof(Counters), j->stream));
  j->auto_bases = 0;
  if (j->sketch) HIP_CHECK(hipMemsetAsync(j->sketch, 0, SKETCH_M * 4, j->stream));
  if (j->table) {
    ev_record(j, 10);
    if (j->clear_pending) {
      /* cleared in the background since the chains were built */
      HIP_CHECK(hipStreamWaitEvent(j->stream, j->ev_cleared, 0));
      j->clear_pending = false;
    } else {
      clear_table(j, j->stream);
    }
    ev_record(j, 11);
  }
  HIP_CHECK(hipStreamSynchronize(j->stream));
  j->stage = ST_CREATED;
  return BRUNO_OK;
}

static bruno_status parse_count_common(bruno_job* j, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads,
                                       const uint64_t* d_first, const uint64_t* d_starts = nullptr,
                                       const uint64_t* d_ends = nullptr) {
  if (j->fused) return BRUNO_STATE; /* the filter already ran on the table */
  ev_record(j, 0);
  if (j->words == 1) TRY(parse_count_t<uint64_t>(j, d_bases, d_offsets, n_reads, d_first, d_starts, d_ends));
  else TRY(parse_count_t<u128>(j, d_bases, d_offsets, n_reads, d_first, d_starts, d_ends));
  ev_record(j, 1);
  TRY(sync_counters(j));
  j->ms.parse_count_ms += ev_ms(j, 0, 1);
  j->n_reads += n_reads;
  /* No limit on a job's instances: sub-entries count to 2^48, and every
   * reader takes a k2 count modulo 2^32, as the reference's uint32 k2
   * counter (K2FreqType, debruijn_graph_map.hpp:160) holds it. */
  j->stage = ST_COUNTED;
  return BRUNO_OK;
}

bruno_status bruno_parse_count(bruno_job* j, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads) {
  if (!valid_job(j)) return BRUNO_BAD_INPUT;
  if (j->n_jobs != 1) return BRUNO_BAD_INPUT; /* a batch names each read's job */
  if (j->stage > ST_COUNTED || j->fused) return BRUNO_STATE;
  if (n_reads == 0) {
    j->stage = ST_COUNTED;
    return BRUNO_OK;
  }
  if (!d_bases || !d_offsets) return BRUNO_BAD_INPUT;
  if (j->cap == 0 || j->auto_size) {
    /* size (or grow) the table from a sketch of the distinct k-mers of
     * every batch so far (instances <= bases bounds them from above) */
    uint64_t ends[2];
    HIP_CHECK(hipMemcpyAsync(&ends[0], d_offsets, 8, hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipMemcpyAsync(&ends[1], d_offsets + n_reads, 8, hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipStreamSynchronize(j->stream));
    TRY(autosize(j, d_bases, d_offsets, n_reads, ends[1] - ends[0], nullptr, nullptr));
  }
  j->job_reads[0] += n_reads;
  return parse_count_common(j, d_bases, d_offsets, n_reads, nullptr);
}

bruno_status bruno_batch_parse_count(bruno_job* j, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads,
                                     const uint64_t* job_first_read) {
  if (!valid_job(j) || !job_first_read) return BRUNO_BAD_INPUT;
  if (j->stage > ST_COUNTED) return BRUNO_STATE;
  const uint32_t nj = j->n_jobs;
  if (job_first_read[0] != 0 || job_first_read[nj] != n_reads) return BRUNO_BAD_INPUT;
  for (uint32_t i = 0; i < nj; ++i)
    if (job_first_read[i + 1] < job_first_read[i]) return BRUNO_BAD_INPUT;
  if (n_reads == 0) {
    j->stage = ST_COUNTED;
    return BRUNO_OK;
  }
  if (!d_bases || !d_offsets || j->cap == 0) return BRUNO_BAD_INPUT;
  uint64_t* d_first = nullptr;
  if (nj > 1) {
    TRY(dev_alloc(j, (void**)&d_first, (nj + 1) * 8));
    HIP_CHECK(hipMemcpyAsync(d_first, job_first_read, (nj + 1) * 8, hipMemcpyHostToDevice, j->stream));
  }
  bruno_status s = parse_count_common(j, d_bases, d_offsets, n_reads, d_first);
  if (d_first) {
    (void)hipStreamSynchronize(j->stream);
    void* a = d_first;
    dev_free(j, a, (nj + 1) * 8);
  }
  if (s == BRUNO_OK)
    for (uint32_t i = 0; i < nj; ++i) j->job_reads[i] += job_first_read[i + 1] - job_first_read[i];
  return s;
}

bruno_status bruno_parse_count_host(bruno_job* j, const char* bases, const uint64_t* offsets, uint64_t n_reads) {
  if (!valid_job(j) || (n_reads && (!bases || !offsets))) return BRUNO_BAD_INPUT;
  if (n_reads == 0) return bruno_parse_count(j, nullptr, nullptr, 0);
  uint64_t nb = offsets[n_reads] - offsets[0];
  char* db = nullptr;
  uint64_t* doff = nullptr;
  if (pool_alloc((void**)&db, nb + 1) != hipSuccess || pool_alloc((void**)&doff, (n_reads + 1) * 8) != hipSuccess) {
    pool_free(db);
    return BRUNO_OOM;
  }
  std::vector<uint64_t> off(offsets, offsets + n_reads + 1);
  for (auto& o : off) o -= offsets[0];
  bruno_status s = BRUNO_OK;
  if (hipMemcpyAsync(db, bases + offsets[0], nb, hipMemcpyHostToDevice, j->stream) != hipSuccess ||
      hipMemcpyAsync(doff, off.data(), (n_reads + 1) * 8, hipMemcpyHostToDevice, j->stream) != hipSuccess ||
      hipStreamSynchronize(j->stream) != hipSuccess) {
    (void)hipGetLastError();
    s = BRUNO_INTERNAL;
  }
  if (s == BRUNO_OK) s = bruno_parse_count(j, db, doff, n_reads);
  (void)hipStreamSynchronize(j->stream);
  pool_free(db);
  pool_free(doff);
  return s;
}

static void* job_alloc(void* ctx, uint64_t bytes) {
  void* p = nullptr;
  return dev_alloc((bruno_job*)ctx, &p, bytes) == BRUNO_OK ? p : nullptr;
}

bruno_status bruno_parse_count_fastq(bruno_job* j, const char* d_fastq, uint64_t bytes) {
  if (!valid_job(j) || (bytes && !d_fastq)) return BRUNO_BAD_INPUT;
  if (j->n_jobs != 1) return BRUNO_BAD_INPUT;
  if (j->stage > ST_COUNTED || j->fused) return BRUNO_STATE;
  uint64_t nr = 0, nb = 0;
  uint64_t *st = nullptr, *en = nullptr;
  /* the sequence lines are parsed where they are in the file: only their
   * byte ranges are materialised */
  bruno_status s = bruno_fastq_ranges(d_fastq, bytes, job_alloc, j, &st, &en, &nr, &nb, j->stream);
  if (s == BRUNO_OK && nr) {
    if (j->cap == 0 || j->auto_size) s = autosize(j, d_fastq, nullptr, nr, nb, st, en);
    if (s == BRUNO_OK) {
      j->job_reads[0] += nr;
      s = parse_count_common(j, d_fastq, nullptr, nr, nullptr, st, en);
    }
  } else if (s == BRUNO_OK) {
    j->stage = ST_COUNTED;
  }
  void* a = st;
  dev_free(j, a, std::max<uint64_t>(1, nr) * 8);
  a = en;
  dev_free(j, a, std::max<uint64_t>(1, nr) * 8);
  return s;
}

bruno_status bruno_parse_count_fastq_host(bruno_job* j, const char* fastq, uint64_t bytes) {
  if (!valid_job(j) || (bytes && !fastq)) return BRUNO_BAD_INPUT;
  char* d = nullptr;
  TRY(dev_alloc(j, (void**)&d, bytes + 1));
  bruno_status s = BRUNO_OK;
  if (bytes && (hipMemcpyAsync(d, fastq, bytes, hipMemcpyHostToDevice, j->stream) != hipSuccess ||
                hipStreamSynchronize(j->stream) != hipSuccess)) {
    (void)hipGetLastError();
    s = BRUNO_INTERNAL;
  }
  if (s == BRUNO_OK) s = bruno_parse_count_fastq(j, d, bytes);
  void* a = d;
  dev_free(j, a, bytes + 1);
  return s;
}

bruno_status bruno_filter(bruno_job* j) {
  if (!valid_job(j)) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COUNTED) return BRUNO_STATE;
  ev_record(j, 2);
  if (j->words == 1) TRY(filter_t<uint64_t>(j));
  else TRY(filter_t<u128>(j));
  ev_record(j, 3);
  HIP_CHECK(hipStreamSynchronize(j->stream));
  j->ms.filter_ms = ev_ms(j, 2, 3);
  j->stage = ST_FILTERED;
  return BRUNO_OK;
}

bruno_status bruno_build_chains(bruno_job* j) {
  if (!valid_job(j)) return BRUNO_BAD_INPUT;
  if (j->stage != ST_FILTERED) return BRUNO_STATE;
  ev_record(j, 4);
  if (j->words == 1) TRY(chains_t<uint64_t>(j));
  else TRY(chains_t<u128>(j));
  ev_record(j, 5);
  if (j->recycle && j->table) {
    /* the table is not read after this point: clear it now, beside the
     * list ranking and the output */
    HIP_CHECK(hipEventRecord(j->ev_chains, j->stream));
    HIP_CHECK(hipStreamWaitEvent(j->side, j->ev_chains, 0));
    clear_table(j, j->side);
    HIP_CHECK(hipEventRecord(j->ev_cleared, j->side));
    j->clear_pending = true;
  }
  HIP_CHECK(hipStreamSynchronize(j->stream));
  j->ms.chains_ms = ev_ms(j, 4, 5);
  j->stage = ST_CHAINS;
  return BRUNO_OK;
}

bruno_status bruno_list_rank(bruno_job* j, int variant, uint32_t* rounds) {
  if (!valid_job(j) || (variant != 0 && variant != 1)) return BRUNO_BAD_INPUT;
  if (j->stage != ST_CHAINS) return BRUNO_STATE;
  ev_record(j, 6);
  bool used = false;
  const char* force = getenv("BRUNO_LR_ROUNDS");
  if (!(force && force[0] == '1')) TRY(list_rank_fast(j, &used));
  if (!used) TRY(list_rank_rounds(j, variant));
  j->lr_contracted = used;
  ev_record(j, 7);
  HIP_CHECK(hipStreamSynchronize(j->stream));
  j->ms.list_rank_ms = ev_ms(j, 6, 7);
  if (rounds) *rounds = j->iterations;
  j->stage = ST_RANKED;
  return BRUNO_OK;
}

bruno_status bruno_compact(bruno_job* j) {
  if (!valid_job(j)) return BRUNO_BAD_INPUT;
  if (j->stage != ST_RANKED) return BRUNO_STATE;
  ev_record(j, 8);
  if (j->words == 1) TRY(compact_t<uint64_t>(j));
  else TRY(compact_t<u128>(j));
  ev_record(j, 9);
  HIP_CHECK(hipStreamSynchronize(j->stream));
  j->ms.output_ms = ev_ms(j, 8, 9);
  /* number of unitigs = number of dist-0 nodes = count of records starting
   * a FASTA record; derived on the host from the byte length */
  uint64_t n_body = j->n_out;
  uint64_t per_head = 2 * (uint64_t)j->K + 3;
  /* fasta_len = heads*per_head + (n - heads)  =>  heads = (len - n)/(per_head-1) */
  j->n_unitigs = (j->fasta_len >= n_body && per_head > 1) ? (j->fasta_len - n_body) / (per_head - 1) : 0;
  j->stage = ST_COMPACTED;
  return BRUNO_OK;
}

/* ================================================================== */
/* graph cleaning (host side)                                           */
/* ================================================================== */
extern "C++" {
/* Everything allocated after the node lists (chains, ranking, output) goes
 * back to the pool; the table and the node lists stay. */
static void free_after_lists(bruno_job* j) {
  (void)hipStreamSynchronize(j->stream);
  std::vector<void*> keep;
  for (void* p : j->allocs) {
    if (persistent(j, p) || p == j->chain_slot || p == j->branch_slot) keep.push_back(p);
    else pool_free(p);
  }
  j->allocs = keep;
  j->device_bytes = persistent_bytes(j) + 2 * j->list_cap * sizeof(uint64_t);
  j->ckmer = j->vkmer = j->bkmer = j->out_k = j->out_rep = j->bsorted = nullptr;
  j->cedge = nullptr;
  j->lr = nullptr;
  j->list[0] = j->list[1] = nullptr;
  j->alive = nullptr;
  j->out_dist = j->perm = j->dist_sorted = nullptr;
  j->out_job = nullptr;
  j->fasta = j->bfasta = nullptr;
  j->cbits = j->sbits = nullptr;
  j->bits_fresh = false;
  j->lr_contracted = j->out_by_chains = false;
  j->n_out = j->n_unitigs = j->fasta_len = j->bfasta_len = 0;
  j->chain_start.clear();
  j->fasta_off.clear();
  j->branch_start.clear();
}

/* new chain / branch lists from the nodes still in the graph, classified
 * by their current edges */
template <typename KT>
static bruno_status relist_t(bruno_job* j) {
  const uint64_t nc = j->n_chain, nb = j->n_branch, lcap = std::max<uint64_t>(1, nc + nb);
  uint64_t *cs = nullptr, *bs = nullptr;
  TRY(dev_alloc(j, (void**)&cs, lcap * sizeof(uint64_t)));
  TRY(dev_alloc(j, (void**)&bs, lcap * sizeof(uint64_t)));
  HIP_CHECK(hipMemsetAsync(&j->ctr->n_branch, 0, 2 * sizeof(unsigned long long), j->stream));
  NodeLists nl{cs, bs, nullptr, nullptr, lcap, j->d_seg, 1};
  if (nc + nb)
    hipLaunchKernelGGL((k_relist<KT>), dim3(grid_for(nc + nb)), dim3(BLOCK), 0, j->stream, (Slot<KT>*)j->table, j->K,
                       j->chain_slot, nc, j->branch_slot, nb, nl, j->ctr);
  TRY(sync_counters(j));
  void* a = j->chain_slot;
  dev_free(j, a, j->list_cap * sizeof(uint64_t));
  a = j->branch_slot;
  dev_free(j, a, j->list_cap * sizeof(uint64_t));
  j->chain_slot = cs;
  j->branch_slot = bs;
  j->list_cap = lcap;
  j->n_chain = j->hctr->n_chain;
  j->n_branch = j->hctr->n_branch;
  j->n_present = j->n_chain + j->n_branch;
  return BRUNO_OK;
}

/* erase_nodes of the separated chain nodes (cycles and edgeless ones) */
template <typename KT>
static bruno_status drop_separated_t(bruno_job* j) {
  const uint64_t nc = j->n_chain;
  if (!nc || !j->alive) return BRUNO_OK;
  const uint64_t tcap = probe_span(j); /* plain jobs: one probe window per k-mer over the table */
  uint8_t* mark = nullptr;
  TRY(dev_alloc(j, (void**)&mark, nc));
  hipLaunchKernelGGL((k_drop_separated<KT>), dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream, (Slot<KT>*)j->table, tcap,
                     j->rbits, j->K, j->chain_slot, j->alive, nc, mark);
  hipLaunchKernelGGL((k_drop_separated_edges<KT>), dim3(grid_for(nc)), dim3(BLOCK), 0, j->stream,
                     (Slot<KT>*)j->table, tcap, j->rbits, j->K, j->chain_slot, mark, nc);
  void* a = mark;
  dev_free(j, a, nc);
  return BRUNO_OK;
}

/* erase_if(IsIsolated) over both node lists */
template <typename KT>
static bruno_status drop_edgeless_t(bruno_job* j) {
  if (j->n_chain)
    hipLaunchKernelGGL((k_drop_edgeless<KT>), dim3(grid_for(j->n_chain)), dim3(BLOCK), 0, j->stream,
                       (Slot<KT>*)j->table, j->chain_slot, j->n_chain);
  if (j->n_branch)
    hipLaunchKernelGGL((k_drop_edgeless<KT>), dim3(grid_for(j->n_branch)), dim3(BLOCK), 0, j->stream,
                       (Slot<KT>*)j->table, j->branch_slot, j->n_branch);
  HIP_CHECK(hipGetLastError());
  return BRUNO_OK;
}

/* the compacted graph again, from the node lists on */
template <typename KT>
static bruno_status recompact_t(bruno_job* j, int variant) {
  TRY(relist_t<KT>(j));
  free_after_lists(j);
  HIP_CHECK(hipMemsetAsync(&j->ctr->n_virt, 0, sizeof(unsigned long long), j->stream));
  j->stage = ST_FILTERED;
  TRY(bruno_build_chains(j));
  TRY(bruno_list_rank(j, variant, nullptr));
  return bruno_compact(j);
}

/* Deadend and bubble edges of the compacted graph: dq one query per deadend,
 * bq two per bubble chain (device arrays the caller frees). */
template <typename KT>
static bruno_status pick_t(bruno_job* j, uint64_t min_freq, EdgeQuery<KT>** dq, uint64_t* ndq, EdgeQuery<KT>** bq,
                           uint64_t* nbq, uint64_t* dq_bytes, uint64_t* bq_bytes) {
  *dq = *bq = nullptr;
  *ndq = *nbq = *dq_bytes = *bq_bytes = 0;
  const uint64_t n = j->n_out, nu = j->n_unitigs;
  if (nu == 0) return BRUNO_OK;
  const uint64_t tcap = probe_span(j); /* plain jobs: one probe window per k-mer over the table */
  uint8_t* flags = nullptr;
  uint32_t* starts = nullptr;
  ChainSummary<KT>*S = nullptr, *N = nullptr;
  unsigned long long* cnt = nullptr;
  void* t = nullptr;
  size_t tmp = 0;
  hipcub::CountingInputIterator<uint32_t> pos(0);
  TRY(dev_alloc(j, (void**)&flags, n));
  TRY(dev_alloc(j, (void**)&starts, (n + 1) * 4)); /* room for every node: checked against nu below */
  TRY(dev_alloc(j, (void**)&cnt, 4 * sizeof(unsigned long long)));
  TRY(dev_alloc(j, (void**)&S, nu * sizeof(ChainSummary<KT>)));
  TRY(dev_alloc(j, (void**)&N, nu * sizeof(ChainSummary<KT>)));
  TRY(dev_alloc(j, (void**)dq, nu * sizeof(EdgeQuery<KT>)));
  *dq_bytes = nu * sizeof(EdgeQuery<KT>);
  HIP_CHECK(hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned long long), j->stream));
  hipLaunchKernelGGL(k_head_flags, dim3(grid_for(n, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, j->dist_sorted, n,
                     flags);
  HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, tmp, pos, flags, starts, (uint32_t*)(cnt + 3), (int)n, j->stream));
  TRY(dev_alloc(j, &t, std::max<size_t>(tmp, 1)));
  HIP_CHECK(hipcub::DeviceSelect::Flagged(t, tmp, pos, flags, starts, (uint32_t*)(cnt + 3), (int)n, j->stream));
  uint32_t heads = 0;
  HIP_CHECK(hipMemcpyAsync(&heads, cnt + 3, 4, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  if (heads != nu) return BRUNO_INTERNAL; /* one dist-0 node per unitig */
  const uint32_t n32 = (uint32_t)n;
  HIP_CHECK(hipMemcpyAsync(starts + nu, &n32, 4, hipMemcpyHostToDevice, j->stream));
  hipLaunchKernelGGL((k_chain_summaries<KT>), dim3(grid_for(nu)), dim3(BLOCK), 0, j->stream,
                     (const Slot<KT>*)j->table, tcap, j->rbits, j->K, j->perm, (const KT*)j->out_k, j->dist_sorted,
                     starts, nu, S, j->ctr);
  hipLaunchKernelGGL((k_pick_chains<KT>), dim3(grid_for(nu)), dim3(BLOCK), 0, j->stream, S, nu, j->K, min_freq, *dq,
                     cnt, nu, N, cnt + 1, j->ctr);
  unsigned long long h[4];
  HIP_CHECK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, j->stream));
  TRY(sync_counters(j));
  *ndq = h[0];
  const uint64_t nn = h[1];
  if (nn > 0) {
    KT *b5 = nullptr, *t5 = nullptr, *t3 = nullptr, *b3 = nullptr;
    uint32_t *len = nullptr, *order = nullptr;
    TRY(dev_alloc(j, (void**)&b5, nn * sizeof(KT)));
    TRY(dev_alloc(j, (void**)&t5, nn * sizeof(KT)));
    TRY(dev_alloc(j, (void**)&t3, nn * sizeof(KT)));
    TRY(dev_alloc(j, (void**)&b3, nn * sizeof(KT)));
    TRY(dev_alloc(j, (void**)&len, nn * 4));
    TRY(dev_alloc(j, (void**)&order, nn * 4));
    TRY(dev_alloc(j, (void**)bq, 2 * nn * sizeof(EdgeQuery<KT>)));
    *bq_bytes = 2 * nn * sizeof(EdgeQuery<KT>);
    hipLaunchKernelGGL((k_summary_keys<KT>), dim3(grid_for(nn, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream, N, nn,
                       b5, t5, t3, b3, len);
    /* by (5' branch, 3' branch, length, 5' end, 3' end): stable passes from
     * the last key to the first */
    TRY(sort_perm_kmer<KT>(j, t3, order, nn, j->K, false));
    TRY(sort_perm_kmer<KT>(j, t5, order, nn, j->K, true));
    TRY(sort_perm_u32(j, len, order, nn, 32));
    TRY(sort_perm_kmer<KT>(j, b3, order, nn, j->K, true));
    TRY(sort_perm_kmer<KT>(j, b5, order, nn, j->K, true));
    hipLaunchKernelGGL((k_pick_bubbles<KT>), dim3(grid_for(nn)), dim3(BLOCK), 0, j->stream, N, order, nn, j->K,
                       min_freq, *bq, cnt + 2, 2 * nn, j->ctr);
    HIP_CHECK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, j->stream));
    TRY(sync_counters(j));
    *nbq = h[2];
    void* a = b5;
    dev_free(j, a, nn * sizeof(KT));
    a = t5;
    dev_free(j, a, nn * sizeof(KT));
    a = t3;
    dev_free(j, a, nn * sizeof(KT));
    a = b3;
    dev_free(j, a, nn * sizeof(KT));
    a = len;
    dev_free(j, a, nn * 4);
    a = order;
    dev_free(j, a, nn * 4);
  }
  if (j->hctr->err) return BRUNO_INTERNAL;
  void* a = flags;
  dev_free(j, a, n);
  a = starts;
  dev_free(j, a, (n + 1) * 4);
  a = cnt;
  dev_free(j, a, 4 * sizeof(unsigned long long));
  a = S;
  dev_free(j, a, nu * sizeof(ChainSummary<KT>));
  a = N;
  dev_free(j, a, nu * sizeof(ChainSummary<KT>));
  dev_free(j, t, std::max<size_t>(tmp, 1));
  return BRUNO_OK;
}

/* the queried edges below min_freq, erased on both nodes */
template <typename KT>
static bruno_status erase_t(bruno_job* j, const EdgeQuery<KT>* q, uint64_t n, uint64_t min_freq, uint64_t* erased) {
  *erased = 0;
  if (n == 0) return BRUNO_OK;
  const uint64_t tcap = probe_span(j); /* plain jobs: one probe window per k-mer over the table */
  uint8_t* hit = nullptr;
  unsigned long long* cnt = nullptr;
  TRY(dev_alloc(j, (void**)&hit, n));
  TRY(dev_alloc(j, (void**)&cnt, sizeof(unsigned long long)));
  HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), j->stream));
  hipLaunchKernelGGL((k_edge_hits<KT>), dim3(grid_for(n)), dim3(BLOCK), 0, j->stream, (const Slot<KT>*)j->table, tcap,
                     j->rbits, j->K, q, n, min_freq, hit);
  hipLaunchKernelGGL((k_erase_edges<KT>), dim3(grid_for(n)), dim3(BLOCK), 0, j->stream, (Slot<KT>*)j->table, tcap,
                     j->rbits, j->K, q, n, hit, cnt);
  unsigned long long h = 0;
  HIP_CHECK(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  *erased = h;
  void* a = hit;
  dev_free(j, a, n);
  a = cnt;
  dev_free(j, a, sizeof(unsigned long long));
  return BRUNO_OK;
}

template <typename KT>
static bruno_status clean_t(bruno_job* j, int variant, uint64_t k1_lo, bruno_clean_stats* st) {
  const uint64_t min_freq = k1_lo + 1; /* edge_freq_filter(threshes[2] + 1) */
  /* do_work drops cycles and edgeless nodes before cleaning */
  TRY(drop_separated_t<KT>(j));
  TRY(drop_edgeless_t<KT>(j));
  for (;;) {
    EdgeQuery<KT>*dq = nullptr, *bq = nullptr;
    uint64_t ndq = 0, nbq = 0, e1 = 0, e2 = 0, dqb = 0, bqb = 0;
    bruno_status s = pick_t<KT>(j, min_freq, &dq, &ndq, &bq, &nbq, &dqb, &bqb);
    if (s == BRUNO_OK) s = erase_t<KT>(j, dq, ndq, min_freq, &e1);
    if (s == BRUNO_OK) s = drop_edgeless_t<KT>(j);
    if (s == BRUNO_OK) s = erase_t<KT>(j, bq, nbq, min_freq, &e2);
    if (s == BRUNO_OK) s = drop_edgeless_t<KT>(j);
    void* a = dq;
    dev_free(j, a, dqb);
    a = bq;
    dev_free(j, a, bqb);
    if (s != BRUNO_OK) return s;
    st->deadends += ndq;
    st->bubble_chains += nbq / 2;
    st->edges_erased += e1 + e2;
    /* nothing found, or nothing weak enough to erase: the same chains
     * would be found again */
    if (e1 + e2 == 0) break;
    TRY(recompact_t<KT>(j, variant));
    TRY(drop_separated_t<KT>(j));
    TRY(drop_edgeless_t<KT>(j));
    ++st->rounds;
  }
  /* the outputs of the cleaned graph (its branch points and unitigs) */
  return recompact_t<KT>(j, variant);
}
}  // extern "C++"

bruno_status bruno_clean(bruno_job* j, int variant, uint64_t k1_lo, bruno_clean_stats* out) {
  if (!valid_job(j) || (variant != 0 && variant != 1)) return BRUNO_BAD_INPUT;
  if (j->n_jobs != 1) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COMPACTED || j->recycle || j->clear_pending) return BRUNO_STATE;
  bruno_clean_stats st{};
  bruno_status s = j->words == 1 ? clean_t<uint64_t>(j, variant, k1_lo, &st) : clean_t<u128>(j, variant, k1_lo, &st);
  if (s != BRUNO_OK) return s;
  if (out) *out = st;
  return BRUNO_OK;
}

bruno_status bruno_run_pipeline(bruno_job* j, int variant) {
  TRY(bruno_filter(j));
  TRY(bruno_build_chains(j));
  TRY(bruno_list_rank(j, variant, nullptr));
  return bruno_compact(j);
}

bruno_status bruno_get_stats(bruno_job* j, bruno_stats* o) {
  if (!valid_job(j) || !o) return BRUNO_BAD_INPUT;
  TRY(sync_counters(j));
  memset(o, 0, sizeof(*o));
  o->n_reads = j->n_reads;
  o->n_reads_dropped = j->hctr->n_dropped;
  o->n_instances = j->hctr->n_instances;
  o->n_distinct = j->hctr->n_distinct;
  o->n_k2 = j->hctr->n_k2;
  o->n_overflow = j->hctr->n_spill;
  o->n_nodes = j->n_present;
  o->n_branches = j->n_branch;
  o->n_chain_nodes = j->n_chain;
  o->n_cycle_nodes = j->n_cycle;
  o->n_isolated = j->n_isolated;
  o->n_unitig_nodes = j->n_out;
  o->n_unitigs = j->n_unitigs;
  o->iterations = j->iterations;
  o->n_conflicts = j->conflicts;
  o->capacity = j->cap;
  o->device_bytes = j->device_bytes;
  o->peak_device_bytes = j->peak_bytes;
  o->output_by_chains = j->out_by_chains;
  o->count_by_regions = j->by_regions;
  o->filter_fused = j->fused;
  return BRUNO_OK;
}

bruno_status bruno_get_phase_ms(bruno_job* j, bruno_phase_ms* o) {
  if (!valid_job(j) || !o) return BRUNO_BAD_INPUT;
  *o = j->ms;
  return BRUNO_OK;
}

/* ---- per-job views ------------------------------------------------- */
/* device -> host on the job's stream (never the null stream, which would
 * wait for every other job's work) */
static hipError_t copy_out(bruno_job* j, void* dst, const void* src, uint64_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, j->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(j->stream);
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}
static bool valid_idx(const bruno_job* j, uint32_t idx) { return valid_job(j) && idx < j->n_jobs; }

static uint64_t unitigs_of(const bruno_job* j, uint64_t bytes, uint64_t nodes) {
  /* bytes = heads * (2K + 3) + (nodes - heads) */
  const uint64_t per_head = 2 * (uint64_t)j->K + 3;
  return bytes >= nodes ? (bytes - nodes) / (per_head - 1) : 0;
}

bruno_status bruno_batch_get_stats(bruno_job* j, uint32_t idx, bruno_stats* o) {
  if (!valid_idx(j, idx) || !o) return BRUNO_BAD_INPUT;
  if (j->n_jobs == 1) return bruno_get_stats(j, o);
  TRY(sync_counters(j));
  memset(o, 0, sizeof(*o));
  const uint32_t nj = j->n_jobs;
  std::vector<unsigned long long> js(2 * nj);
  HIP_CHECK(hipMemcpyAsync(js.data(), j->d_jstat, 2 * nj * 8, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  o->n_reads = j->job_reads[idx];
  o->n_instances = js[idx];
  o->n_reads_dropped = js[nj + idx];
  o->capacity = j->seg[idx + 1] - j->seg[idx];
  if (j->stage >= ST_RANKED) {
    o->iterations = j->job_iters[idx];
    o->n_cycle_nodes = j->job_cycle[idx];
    o->n_isolated = j->job_isolated[idx];
  }
  if (j->stage >= ST_COMPACTED) {
    o->n_branches = j->branch_start[idx + 1] - j->branch_start[idx];
    o->n_unitig_nodes = j->chain_start[idx + 1] - j->chain_start[idx];
    o->n_chain_nodes = o->n_unitig_nodes + o->n_cycle_nodes + o->n_isolated;
    o->n_nodes = o->n_branches + o->n_chain_nodes;
    o->n_unitigs = unitigs_of(j, j->fasta_off[idx + 1] - j->fasta_off[idx], o->n_unitig_nodes);
    o->output_by_chains = j->out_by_chains;
  }
  return BRUNO_OK;
}

bruno_status bruno_batch_get_counts(bruno_job* j, uint32_t idx, bruno_node* buf, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COUNTED || j->fused) return BRUNO_STATE; /* fused: the nodes replaced the k2 entries */
  ensure_table_valid(j);
  const uint64_t s0 = j->seg.empty() ? 0 : j->seg[idx], s1 = j->seg.empty() ? 0 : j->seg[idx + 1];
  const bool one = j->n_jobs == 1;
  if (!buf) {
    if (one) {
      *n = j->hctr->n_distinct;
      return BRUNO_OK;
    }
    j->hctr->n_out = 0;
    HIP_CHECK(hipMemcpyAsync(&j->ctr->n_out, &j->hctr->n_out, 8, hipMemcpyHostToDevice, j->stream));
    if (j->words == 1)
      hipLaunchKernelGGL((k_get_counts<uint64_t>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                         (const Slot<uint64_t>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, nullptr, 0,
                         j->ctr);
    else
      hipLaunchKernelGGL((k_get_counts<u128>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                         (const Slot<u128>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, nullptr, 0, j->ctr);
    TRY(sync_counters(j));
    *n = j->hctr->n_out;
    return BRUNO_OK;
  }
  const uint64_t cap_out = *n;
  bruno_node* d = nullptr;
  TRY(dev_alloc(j, (void**)&d, std::max<uint64_t>(1, cap_out) * sizeof(bruno_node)));
  j->hctr->n_out = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->n_out, &j->hctr->n_out, 8, hipMemcpyHostToDevice, j->stream));
  if (j->words == 1)
    hipLaunchKernelGGL((k_get_counts<uint64_t>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                       (const Slot<uint64_t>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, d, cap_out, j->ctr);
  else
    hipLaunchKernelGGL((k_get_counts<u128>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                       (const Slot<u128>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, d, cap_out, j->ctr);
  bruno_status s = sync_counters(j);
  const uint64_t got = j->hctr->n_out;
  if (s == BRUNO_CAPACITY) s = BRUNO_BAD_INPUT; /* buffer smaller than the count */
  if (s == BRUNO_OK && got) {
    if (copy_out(j, buf, d, got * sizeof(bruno_node)) != hipSuccess) {
      (void)hipGetLastError();
      s = BRUNO_INTERNAL;
    }
  }
  if (s == BRUNO_BAD_INPUT) {
    /* clear the list-full flag so the job stays usable */
    j->hctr->err = 0;
    (void)hipMemsetAsync(&j->ctr->err, 0, 4, j->stream);
    (void)hipStreamSynchronize(j->stream);
  }
  void* a = d;
  dev_free(j, a, std::max<uint64_t>(1, cap_out) * sizeof(bruno_node));
  *n = got;
  return s;
}

bruno_status bruno_batch_get_k2(bruno_job* j, uint32_t idx, uint64_t* kmers, uint8_t* biedges, uint32_t* counts,
                                uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COUNTED || j->fused) return BRUNO_STATE; /* fused: the nodes replaced the k2 entries */
  ensure_table_valid(j);
  const uint64_t s0 = j->seg.empty() ? 0 : j->seg[idx], s1 = j->seg.empty() ? 0 : j->seg[idx + 1];
  const bool one = j->n_jobs == 1;
  if (!kmers) {
    if (one) {
      *n = j->hctr->n_k2;
      return BRUNO_OK;
    }
    j->hctr->n_out = 0;
    HIP_CHECK(hipMemcpyAsync(&j->ctr->n_out, &j->hctr->n_out, 8, hipMemcpyHostToDevice, j->stream));
    if (j->words == 1)
      hipLaunchKernelGGL((k_get_k2<uint64_t>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                         (const Slot<uint64_t>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, nullptr, nullptr,
                         nullptr, 0, j->ctr);
    else
      hipLaunchKernelGGL((k_get_k2<u128>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                         (const Slot<u128>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, nullptr, nullptr,
                         nullptr, 0, j->ctr);
    TRY(sync_counters(j));
    *n = j->hctr->n_out;
    return BRUNO_OK;
  }
  if (!biedges || !counts) return BRUNO_BAD_INPUT;
  const uint64_t m = std::max<uint64_t>(1, *n);
  uint64_t* dk = nullptr;
  uint8_t* de = nullptr;
  uint32_t* dc = nullptr;
  TRY(dev_alloc(j, (void**)&dk, m * 8 * j->words));
  TRY(dev_alloc(j, (void**)&de, m));
  TRY(dev_alloc(j, (void**)&dc, m * 4));
  j->hctr->n_out = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->n_out, &j->hctr->n_out, 8, hipMemcpyHostToDevice, j->stream));
  if (j->words == 1)
    hipLaunchKernelGGL((k_get_k2<uint64_t>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                       (const Slot<uint64_t>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, dk, de, dc, *n,
                       j->ctr);
  else
    hipLaunchKernelGGL((k_get_k2<u128>), dim3(grid_for(s1 - s0, BLOCK, 1u << 15)), dim3(BLOCK), 0, j->stream,
                       (const Slot<u128>*)j->table, s0, s1, (const Spill*)j->spill, j->spcap, dk, de, dc, *n, j->ctr);
  bruno_status s = sync_counters(j);
  const uint64_t got = j->hctr->n_out;
  if (s == BRUNO_CAPACITY) s = BRUNO_BAD_INPUT;
  if (s == BRUNO_OK && got) {
    if (copy_out(j, kmers, dk, got * 8 * j->words) != hipSuccess || copy_out(j, biedges, de, got) != hipSuccess ||
        copy_out(j, counts, dc, got * 4) != hipSuccess) {
      (void)hipGetLastError();
      s = BRUNO_INTERNAL;
    }
  }
  if (s == BRUNO_BAD_INPUT) {
    j->hctr->err = 0;
    (void)hipMemsetAsync(&j->ctr->err, 0, 4, j->stream);
    (void)hipStreamSynchronize(j->stream);
  }
  void* a = dk;
  dev_free(j, a, m * 8 * j->words);
  a = de;
  dev_free(j, a, m);
  a = dc;
  dev_free(j, a, m * 4);
  *n = got;
  return s;
}

bruno_status bruno_batch_get_nodes(bruno_job* j, uint32_t idx, bruno_node* buf, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage < ST_FILTERED || j->clear_pending) return BRUNO_STATE;
  const uint64_t np = j->n_present;
  const uint64_t* lists[2] = {j->chain_slot, j->branch_slot};
  const uint64_t lens[2] = {j->n_chain, j->n_branch};
  if (j->n_jobs == 1) {
    if (!buf) {
      *n = np;
      return BRUNO_OK;
    }
    if (*n < np) return BRUNO_BAD_INPUT;
    bruno_node* d = nullptr;
    TRY(dev_alloc(j, (void**)&d, std::max<uint64_t>(1, np) * sizeof(bruno_node)));
    uint64_t at = 0;
    for (int l = 0; l < 2; ++l) {
      if (!lens[l]) continue;
      if (j->words == 1)
        hipLaunchKernelGGL((k_get_nodes<uint64_t>), dim3(grid_for(lens[l], BLOCK, 1u << 30)), dim3(BLOCK), 0,
                           j->stream, (const Slot<uint64_t>*)j->table, lists[l], lens[l], d + at);
      else
        hipLaunchKernelGGL((k_get_nodes<u128>), dim3(grid_for(lens[l], BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                           (const Slot<u128>*)j->table, lists[l], lens[l], d + at);
      at += lens[l];
    }
    hipError_t e = np ? copy_out(j, buf, d, np * sizeof(bruno_node)) : hipSuccess;
    void* a = d;
    dev_free(j, a, std::max<uint64_t>(1, np) * sizeof(bruno_node));
    if (e != hipSuccess) return BRUNO_INTERNAL;
    *n = np;
    return BRUNO_OK;
  }
  const uint64_t s0 = j->seg[idx], s1 = j->seg[idx + 1];
  const uint64_t cap_out = buf ? *n : 0;
  bruno_node* d = nullptr;
  if (buf) TRY(dev_alloc(j, (void**)&d, std::max<uint64_t>(1, cap_out) * sizeof(bruno_node)));
  j->hctr->n_out = 0;
  HIP_CHECK(hipMemcpyAsync(&j->ctr->n_out, &j->hctr->n_out, 8, hipMemcpyHostToDevice, j->stream));
  for (int l = 0; l < 2; ++l) {
    if (!lens[l]) continue;
    if (j->words == 1)
      hipLaunchKernelGGL((k_get_nodes_range<uint64_t>), dim3(grid_for(lens[l])), dim3(BLOCK), 0, j->stream,
                         (const Slot<uint64_t>*)j->table, lists[l], lens[l], s0, s1, d, cap_out, j->ctr);
    else
      hipLaunchKernelGGL((k_get_nodes_range<u128>), dim3(grid_for(lens[l])), dim3(BLOCK), 0, j->stream,
                         (const Slot<u128>*)j->table, lists[l], lens[l], s0, s1, d, cap_out, j->ctr);
  }
  bruno_status s = sync_counters(j);
  const uint64_t got = j->hctr->n_out;
  if (s == BRUNO_CAPACITY) s = BRUNO_BAD_INPUT;
  if (s == BRUNO_OK && buf && got) {
    if (copy_out(j, buf, d, got * sizeof(bruno_node)) != hipSuccess) s = BRUNO_INTERNAL;
  }
  if (s == BRUNO_BAD_INPUT) {
    j->hctr->err = 0;
    (void)hipMemsetAsync(&j->ctr->err, 0, 4, j->stream);
    (void)hipStreamSynchronize(j->stream);
  }
  if (d) {
    void* a = d;
    dev_free(j, a, std::max<uint64_t>(1, cap_out) * sizeof(bruno_node));
  }
  *n = got;
  return s;
}

bruno_status bruno_batch_get_branches(bruno_job* j, uint32_t idx, uint64_t* kmers, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COMPACTED) return BRUNO_STATE;
  const uint64_t b0 = j->branch_start[idx], nb = j->branch_start[idx + 1] - b0;
  if (!kmers) {
    *n = nb;
    return BRUNO_OK;
  }
  if (*n < nb) return BRUNO_BAD_INPUT;
  const uint64_t w = 8 * (uint64_t)j->words;
  if (nb) HIP_CHECK(hipMemcpyAsync(kmers, (const char*)j->bsorted + b0 * w, nb * w, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  *n = nb;
  return BRUNO_OK;
}

bruno_status bruno_batch_get_chain_nodes(bruno_job* j, uint32_t idx, bruno_chain_node* buf, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COMPACTED) return BRUNO_STATE;
  const uint64_t c0 = j->chain_start[idx], m = j->chain_start[idx + 1] - c0;
  if (!buf) {
    *n = m;
    return BRUNO_OK;
  }
  if (*n < m) return BRUNO_BAD_INPUT;
  if (m) {
    bruno_chain_node* d = nullptr;
    TRY(dev_alloc(j, (void**)&d, m * sizeof(bruno_chain_node)));
    if (j->words == 1)
      hipLaunchKernelGGL((k_get_chain_nodes<uint64_t>), dim3(grid_for(m, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                         j->perm + c0, (const uint64_t*)j->out_k, (const uint64_t*)j->out_rep, j->dist_sorted + c0, m,
                         d);
    else
      hipLaunchKernelGGL((k_get_chain_nodes<u128>), dim3(grid_for(m, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                         j->perm + c0, (const u128*)j->out_k, (const u128*)j->out_rep, j->dist_sorted + c0, m, d);
    HIP_CHECK(hipMemcpyAsync(buf, d, m * sizeof(bruno_chain_node), hipMemcpyDeviceToHost, j->stream));
    HIP_CHECK(hipStreamSynchronize(j->stream));
    void* a = d;
    dev_free(j, a, m * sizeof(bruno_chain_node));
  }
  *n = m;
  return BRUNO_OK;
}

static bruno_status get_text(bruno_job* j, const char* dptr, uint64_t len, char* buf, uint64_t* n) {
  if (!buf) {
    *n = len;
    return BRUNO_OK;
  }
  if (*n < len) return BRUNO_BAD_INPUT;
  if (len) HIP_CHECK(hipMemcpyAsync(buf, dptr, len, hipMemcpyDeviceToHost, j->stream));
  HIP_CHECK(hipStreamSynchronize(j->stream));
  *n = len;
  return BRUNO_OK;
}

bruno_status bruno_batch_get_unitigs(bruno_job* j, uint32_t idx, char* buf, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COMPACTED) return BRUNO_STATE;
  /* each job's file drops the leading newline of its first record
   * (compact_dbg_io.hpp:351-355) */
  const uint64_t b0 = j->fasta_off[idx], len = j->fasta_off[idx + 1] - b0;
  return get_text(j, j->fasta + b0 + (len ? 1 : 0), len ? len - 1 : 0, buf, n);
}

bruno_status bruno_batch_get_branch_fasta(bruno_job* j, uint32_t idx, char* buf, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COMPACTED) return BRUNO_STATE;
  const uint64_t rec = 2 * (uint64_t)j->K + 3;
  const uint64_t b0 = j->branch_start[idx] * rec, len = j->branch_start[idx + 1] * rec - b0;
  return get_text(j, j->bfasta + b0, len, buf, n);
}

/* table segment of job idx: first slot and probe span */
static void job_segment(const bruno_job* j, uint32_t idx, uint64_t& tbase, uint64_t& tcap) {
  if (j->n_jobs == 1) {
    tbase = 0;
    tcap = probe_span(j);
  } else {
    tbase = j->seg[idx];
    tcap = j->seg[idx + 1] - j->seg[idx];
  }
}

extern "C++" {
template <typename KT>
static bruno_status branch_nodes_t(bruno_job* j, uint32_t idx, uint64_t b0, uint64_t nb, bruno_node* buf) {
  bruno_node* d = nullptr;
  TRY(dev_alloc(j, (void**)&d, nb * sizeof(bruno_node)));
  uint64_t tbase, tcap;
  job_segment(j, idx, tbase, tcap);
  hipLaunchKernelGGL((k_branch_nodes<KT>), dim3(grid_for(nb, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                     (const Slot<KT>*)j->table, tbase, tcap, j->rbits, (const KT*)j->bsorted + b0, nb, d, j->ctr);
  bruno_status st = sync_counters(j);
  if (st == BRUNO_OK && copy_out(j, buf, d, nb * sizeof(bruno_node)) != hipSuccess) st = BRUNO_INTERNAL;
  void* a = d;
  dev_free(j, a, nb * sizeof(bruno_node));
  return st;
}

}  // extern "C++"

bruno_status bruno_batch_get_branch_nodes(bruno_job* j, uint32_t idx, bruno_node* buf, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COMPACTED || j->clear_pending) return BRUNO_STATE;
  const uint64_t b0 = j->branch_start[idx], nb = j->branch_start[idx + 1] - b0;
  if (!buf) {
    *n = nb;
    return BRUNO_OK;
  }
  if (*n < nb) return BRUNO_BAD_INPUT;
  if (nb) TRY(j->words == 1 ? branch_nodes_t<uint64_t>(j, idx, b0, nb, buf) : branch_nodes_t<u128>(j, idx, b0, nb, buf));
  *n = nb;
  return BRUNO_OK;
}

extern "C++" {
template <typename KT>
static bruno_status chain_edges_t(bruno_job* j, uint32_t idx, uint64_t c0, uint64_t m, uint64_t nu,
                                  bruno_chain_edge* buf) {
  /* unitig starts: positions of the dist-0 nodes among the job's sorted
   * chain nodes, in order, then m (room for m + 1: the selection is checked
   * against the unitig count before anything reads it) */
  uint8_t* flags = nullptr;
  uint32_t *starts = nullptr, *n_sel = nullptr;
  bruno_chain_edge* d = nullptr;
  void* t = nullptr;
  size_t tmp = 0;
  bruno_status st = BRUNO_OK;
  uint32_t got = 0;
  hipcub::CountingInputIterator<uint32_t> pos(0);
  TRY(dev_alloc(j, (void**)&flags, m));
  if ((st = dev_alloc(j, (void**)&starts, (m + 1) * 4)) != BRUNO_OK) goto done;
  if ((st = dev_alloc(j, (void**)&n_sel, 4)) != BRUNO_OK) goto done;
  if ((st = dev_alloc(j, (void**)&d, nu * sizeof(bruno_chain_edge))) != BRUNO_OK) goto done;
  hipLaunchKernelGGL(k_head_flags, dim3(grid_for(m, BLOCK, 1u << 30)), dim3(BLOCK), 0, j->stream,
                     j->dist_sorted + c0, m, flags);
  if (hipcub::DeviceSelect::Flagged(nullptr, tmp, pos, flags, starts, n_sel, (int)m, j->stream) != hipSuccess) {
    st = BRUNO_INTERNAL;
    goto done;
  }
  if ((st = dev_alloc(j, &t, std::max<size_t>(tmp, 1))) != BRUNO_OK) goto done;
  if (hipcub::DeviceSelect::Flagged(t, tmp, pos, flags, starts, n_sel, (int)m, j->stream) != hipSuccess ||
      copy_out(j, &got, n_sel, 4) != hipSuccess) {
    st = BRUNO_INTERNAL;
    goto done;
  }
  if (got != nu) {
    st = BRUNO_INTERNAL;
    goto done;
  }
  {
    const uint32_t mm = (uint32_t)m;
    if (hipMemcpyAsync(starts + nu, &mm, 4, hipMemcpyHostToDevice, j->stream) != hipSuccess) {
      st = BRUNO_INTERNAL;
      goto done;
    }
    uint64_t tbase, tcap;
    job_segment(j, idx, tbase, tcap);
    hipLaunchKernelGGL((k_chain_edges<KT>), dim3(grid_for(nu * 64, BLOCK, 1u << 16)), dim3(BLOCK), 0, j->stream,
                       (const Slot<KT>*)j->table, tbase, tcap, j->rbits, j->K, j->perm + c0, (const KT*)j->out_k,
                       starts, nu, d, j->ctr);
    st = sync_counters(j);
    if (st == BRUNO_OK && copy_out(j, buf, d, nu * sizeof(bruno_chain_edge)) != hipSuccess) st = BRUNO_INTERNAL;
  }
done:
  void* a = flags;
  dev_free(j, a, m);
  if (starts) {
    a = starts;
    dev_free(j, a, (m + 1) * 4);
  }
  if (n_sel) {
    a = n_sel;
    dev_free(j, a, 4);
  }
  if (d) {
    a = d;
    dev_free(j, a, nu * sizeof(bruno_chain_edge));
  }
  if (t) dev_free(j, t, std::max<size_t>(tmp, 1));
  return st;
}

}  // extern "C++"

bruno_status bruno_batch_get_chain_edges(bruno_job* j, uint32_t idx, bruno_chain_edge* buf, uint64_t* n) {
  if (!valid_idx(j, idx) || !n) return BRUNO_BAD_INPUT;
  if (j->stage != ST_COMPACTED || j->clear_pending) return BRUNO_STATE;
  const uint64_t c0 = j->chain_start[idx], m = j->chain_start[idx + 1] - c0;
  const uint64_t nu = unitigs_of(j, j->fasta_off[idx + 1] - j->fasta_off[idx], m);
  if (!buf) {
    *n = nu;
    return BRUNO_OK;
  }
  if (*n < nu) return BRUNO_BAD_INPUT;
  if (nu) TRY(j->words == 1 ? chain_edges_t<uint64_t>(j, idx, c0, m, nu, buf) : chain_edges_t<u128>(j, idx, c0, m, nu, buf));
  *n = nu;
  return BRUNO_OK;
}

/* whole-job getters: the one-job case of the per-job views */
#define ONE_JOB(j) \
  if (!valid_job(j)) return BRUNO_BAD_INPUT; \
  if ((j)->n_jobs != 1) return BRUNO_BAD_INPUT

bruno_status bruno_get_counts(bruno_job* j, bruno_node* buf, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_counts(j, 0, buf, n);
}
bruno_status bruno_get_k2(bruno_job* j, uint64_t* kmers, uint8_t* biedges, uint32_t* counts, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_k2(j, 0, kmers, biedges, counts, n);
}
bruno_status bruno_get_nodes(bruno_job* j, bruno_node* buf, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_nodes(j, 0, buf, n);
}
bruno_status bruno_get_branches(bruno_job* j, uint64_t* kmers, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_branches(j, 0, kmers, n);
}
bruno_status bruno_get_chain_nodes(bruno_job* j, bruno_chain_node* buf, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_chain_nodes(j, 0, buf, n);
}
bruno_status bruno_get_unitigs(bruno_job* j, char* buf, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_unitigs(j, 0, buf, n);
}
bruno_status bruno_get_branch_fasta(bruno_job* j, char* buf, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_branch_fasta(j, 0, buf, n);
}
bruno_status bruno_get_branch_nodes(bruno_job* j, bruno_node* buf, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_branch_nodes(j, 0, buf, n);
}
bruno_status bruno_get_chain_edges(bruno_job* j, bruno_chain_edge* buf, uint64_t* n) {
  ONE_JOB(j);
  return bruno_batch_get_chain_edges(j, 0, buf, n);
}

bruno_status bruno_parse_records(int k, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads,
                                 uint64_t* d_out_kmers, uint8_t* d_out_biedges, uint64_t* n_out, void* stream) {
  if (k < 3 || k > 63 || !n_out || (n_reads && (!d_bases || !d_offsets))) return BRUNO_BAD_INPUT;
  hipStream_t st = (hipStream_t)stream;
  if (n_reads == 0) {
    *n_out = 0;
    return BRUNO_OK;
  }
  uint64_t *per = nullptr, *base = nullptr;
  if (pool_alloc((void**)&per, n_reads * 8) != hipSuccess || pool_alloc((void**)&base, (n_reads + 1) * 8) != hipSuccess) {
    pool_free(per);
    return BRUNO_OOM;
  }
  ParseArgs a{};
  a.bases = (const uint8_t*)d_bases;
  a.offsets = d_offsets;
  a.n_reads = n_reads;
  a.K = k;
  a.per_read = per;
  a.rec_base = base;
  a.out_kmers = d_out_kmers;
  a.out_biedges = d_out_biedges;
  unsigned g = grid_for((n_reads + 3) / 4, 1, 8192);
  bruno_status s = BRUNO_OK;
  if (k <= 31) hipLaunchKernelGGL((k_parse<uint64_t, 0>), dim3(g), dim3(BLOCK), 0, st, a);
  else hipLaunchKernelGGL((k_parse<u128, 0>), dim3(g), dim3(BLOCK), 0, st, a);
  size_t tmp = 0;
  void* t = nullptr;
  uint64_t total = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, per, base, (int)n_reads, st) != hipSuccess ||
      pool_alloc(&t, tmp + 16) != hipSuccess ||
      hipcub::DeviceScan::ExclusiveSum(t, tmp, per, base, (int)n_reads, st) != hipSuccess) {
    (void)hipGetLastError();
    s = BRUNO_INTERNAL;
  }
  uint64_t lastc = 0, lastb = 0;
  if (s == BRUNO_OK && (hipMemcpyAsync(&lastc, per + n_reads - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipMemcpyAsync(&lastb, base + n_reads - 1, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess)) {
    (void)hipGetLastError();
    s = BRUNO_INTERNAL;
  }
  total = lastc + lastb;
  if (s == BRUNO_OK && d_out_kmers) {
    if (k <= 31) hipLaunchKernelGGL((k_parse<uint64_t, 1>), dim3(g), dim3(BLOCK), 0, st, a);
    else hipLaunchKernelGGL((k_parse<u128, 1>), dim3(g), dim3(BLOCK), 0, st, a);
    if (hipStreamSynchronize(st) != hipSuccess) {
      (void)hipGetLastError();
      s = BRUNO_INTERNAL;
    }
  }
  (void)hipStreamSynchronize(st);
  pool_free(t);
  pool_free(per);
  pool_free(base);
  *n_out = total;
  return s;
}

bruno_status bruno_dev_alloc(uint64_t bytes, void** out) {
  if (!out) return BRUNO_BAD_INPUT;
  *out = nullptr;
  if (pool_alloc(out, bytes ? bytes : 16) != hipSuccess) {
    *out = nullptr;
    return BRUNO_OOM;
  }
  return BRUNO_OK;
}

bruno_status bruno_dev_free(void* p) {
  if (!p) return BRUNO_OK;
  /* the caller may have queued work on any stream that uses the block */
  if (hipDeviceSynchronize() != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  pool_free(p);
  return BRUNO_OK;
}

bruno_status bruno_job_set_fused_filter(bruno_job* j, int on) {
  if (!valid_job(j)) return BRUNO_BAD_INPUT;
  j->fuse_filter = on != 0;
  return BRUNO_OK;
}

bruno_status bruno_job_set_recycle(bruno_job* j, int on) {
  if (!valid_job(j)) return BRUNO_BAD_INPUT;
  if (on) {
    if (!j->side) j->side = shared_side_stream();
    if (!j->side ||
        (!j->ev_chains && hipEventCreateWithFlags(&j->ev_chains, hipEventDisableTiming) != hipSuccess) ||
        (!j->ev_cleared && hipEventCreateWithFlags(&j->ev_cleared, hipEventDisableTiming) != hipSuccess)) {
      (void)hipGetLastError();
      return BRUNO_INTERNAL;
    }
  }
  j->recycle = on != 0;
  return BRUNO_OK;
}

bruno_status bruno_estimate_distinct(int k, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads,
                                     double* estimate, uint64_t* capacity, double load_factor, void* stream) {
  if (k < 3 || k > 63 || !estimate || (n_reads && (!d_bases || !d_offsets))) return BRUNO_BAD_INPUT;
  *estimate = 0;
  if (n_reads) TRY(estimate_distinct(k, d_bases, d_offsets, n_reads, estimate, (hipStream_t)stream));
  if (capacity) {
    uint64_t ends[2] = {0, 0};
    if (n_reads &&
        (hipMemcpyAsync(&ends[0], d_offsets, 8, hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
         hipMemcpyAsync(&ends[1], d_offsets + n_reads, 8, hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
         hipStreamSynchronize((hipStream_t)stream) != hipSuccess)) {
      (void)hipGetLastError();
      return BRUNO_INTERNAL;
    }
    const double lf = (load_factor > 0 && load_factor < 1) ? load_factor : 0.5;
    *capacity = capacity_from_estimate(*estimate, ends[1] - ends[0], lf);
  }
  return BRUNO_OK;
}

bruno_status bruno_set_tuning(const bruno_tuning* t) {
  if (!t) return BRUNO_BAD_INPUT;
  const bool pow2 = t->ruler_sample >= 8 && t->ruler_sample <= 4096 && !(t->ruler_sample & (t->ruler_sample - 1));
  if ((t->parse_block != 128 && t->parse_block != 256 && t->parse_block != 512) || t->parse_grid < 64 ||
      t->scan_grid < 1 || !pow2 || t->count_regions > 2 || t->count_groups > MAX_COUNT_GROUPS ||
      (t->region_slots != REGION_MAX_SLOTS && t->region_slots != REGION_SMALL_SLOTS))
    return BRUNO_BAD_INPUT;
  std::lock_guard<std::mutex> g(g_tuning_mu);
  g_tuning = *t;
  return BRUNO_OK;
}

bruno_status bruno_get_tuning(bruno_tuning* t) {
  if (!t) return BRUNO_BAD_INPUT;
  *t = tuning();
  return BRUNO_OK;
}

bruno_status bruno_pool_reserve(uint64_t bytes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return BRUNO_NO_DEVICE;
  }
  return DevicePool::current().reserve(bytes) == hipSuccess ? BRUNO_OK : BRUNO_OOM;
}

bruno_status bruno_pool_trim(void) {
  DevicePool::current().trim();
  return BRUNO_OK;
}

bruno_status bruno_pool_stats(uint64_t* reserved, uint64_t* in_use, uint64_t* largest_free) {
  DevicePool::current().stats(reserved, in_use, largest_free);
  return BRUNO_OK;
}

bruno_status bruno_dev_copy(void* dst, const void* src, uint64_t bytes, int kind) {
  if ((!dst || !src) && bytes) return BRUNO_BAD_INPUT;
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (kind < 0 || kind > 2) return BRUNO_BAD_INPUT;
  if (bytes && hipMemcpy(dst, src, bytes, k) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  return BRUNO_OK;
}

bruno_status bruno_dev_sync(void) {
  if (hipDeviceSynchronize() != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  return BRUNO_OK;
}

bruno_status bruno_dev_mem_info(uint64_t* free_bytes, uint64_t* total_bytes) {
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_NO_DEVICE;
  }
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return BRUNO_OK;
}

static bool make_params(const bruno_synth_desc* d, br::synth_params& p) {
  if (!d || d->genome_len < 2 || d->read_len == 0 || d->read_len > d->genome_len) return false;
  memset(&p, 0, sizeof(p));
  p.seed = d->seed;
  p.G = d->genome_len;
  p.n_reads = d->n_reads;
  p.L = d->read_len;
  auto thr = [](double x) -> uint64_t {
    if (x <= 0) return 0;
    if (x >= 1) return ~0ULL;
    return (uint64_t)(x * 18446744073709551616.0);
  };
  p.sub_thresh = thr(d->sub_rate);
  p.indel_thresh = thr(d->indel_rate);
  p.div_thresh = thr(d->family_divergence);
  if (d->n_families > 0 && d->repeat_fraction > 0) {
    p.n_families = d->n_families;
    p.fmin = d->family_min_len;
    p.fmax = std::max(d->family_max_len, d->family_min_len);
    double mean = 0.5 * (p.fmin + p.fmax);
    uint64_t copies = (uint64_t)(d->repeat_fraction * (double)p.G / mean + 0.5);
    if (copies < 1) copies = 1;
    uint64_t spacing = p.G / copies;
    if (spacing < p.fmax) {
      copies = p.G / p.fmax;
      spacing = copies ? p.G / copies : 0;
    }
    p.n_copies = copies;
    p.spacing = spacing;
    if (!copies || !spacing) p.n_copies = 0;
  }
  return true;
}

bruno_status bruno_synth_reads_device(const bruno_synth_desc* d, char* d_bases, uint64_t* d_offsets, void* stream) {
  br::synth_params p;
  if (!make_params(d, p) || !d_bases || !d_offsets) return BRUNO_BAD_INPUT;
  uint64_t g = (p.n_reads + 255) / 256;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_synth, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, p, d_bases, d_offsets, 0ULL);
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  return BRUNO_OK;
}

bruno_status bruno_synth_batch_device(const bruno_synth_desc* descs, uint32_t n_jobs, char* d_bases,
                                      uint64_t* d_offsets, void* stream) {
  if (!descs || !d_bases || !d_offsets) return BRUNO_BAD_INPUT;
  std::vector<br::synth_params> ps(n_jobs);
  for (uint32_t i = 0; i < n_jobs; ++i)
    if (!make_params(&descs[i], ps[i])) return BRUNO_BAD_INPUT;
  uint64_t rsum = 0, bsum = 0;
  for (uint32_t i = 0; i < n_jobs; ++i) {
    const br::synth_params& p = ps[i];
    uint64_t g = std::max<uint64_t>(1, (p.n_reads + 255) / 256);
    hipLaunchKernelGGL(k_synth, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, p, d_bases + bsum,
                       d_offsets + rsum, bsum);
    rsum += p.n_reads;
    bsum += p.n_reads * p.L;
  }
  if (n_jobs == 0 && hipMemsetAsync(d_offsets, 0, 8, (hipStream_t)stream) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    (void)hipGetLastError();
    return BRUNO_INTERNAL;
  }
  return BRUNO_OK;
}

bruno_status bruno_synth_reads_host(const bruno_synth_desc* d, char* bases, uint64_t* offsets) {
  return bruno_synth_batch_host(d, 1, bases, offsets);
}

bruno_status bruno_synth_batch_host(const bruno_synth_desc* descs, uint32_t n_jobs, char* bases, uint64_t* offsets) {
  if (!descs || !bases || !offsets) return BRUNO_BAD_INPUT;
  std::vector<br::synth_params> ps(n_jobs);
  for (uint32_t i = 0; i < n_jobs; ++i)
    if (!make_params(&descs[i], ps[i])) return BRUNO_BAD_INPUT;
  uint64_t rsum = 0, bsum = 0;
  offsets[0] = 0;
  for (uint32_t i = 0; i < n_jobs; ++i) {
    const br::synth_params& p = ps[i];
    for (uint64_t r = 0; r < p.n_reads; ++r) {
      br::make_read(p, r, bases + bsum + r * p.L);
      offsets[rsum + r] = bsum + r * p.L;
    }
    rsum += p.n_reads;
    bsum += p.n_reads * p.L;
    offsets[rsum] = bsum;
  }
  return BRUNO_OK;
}

}  // extern "C"
