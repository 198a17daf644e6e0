/*
 * clean.h -- device side of graph cleaning (deadends and bubbles).
 *
 * Included by bruno_hip.hip inside its kernel section (it uses that file's
 * node records, lookups and list writers).  bruno's cleaning loop
 * (clean_deadend_and_bubbles_naive, test/common/compact_dbg_clean.hpp
 * :560-850) works on the compacted graph: every chain is summarised by its
 * two ends (summarized_chain, debruijn_chain_node.hpp:165-174; built by
 * to_summarized_chains, debruijn_chain_graph.hpp:2044-2180), short weak
 * deadends and weak bubbles are picked (find_deadends / find_bubbles,
 * debruijn_topo_operations.hpp:152-409), the edges that tie them to their
 * branch points are erased when weaker than the edge threshold
 * (erase_edges with edge_freq_filter, debruijn_graph_map.hpp:737-773), nodes
 * left without edges are dropped (IsIsolated), and the graph is compacted
 * again.  Here a chain's summary comes straight from the compacted output
 * (each unitig's first and last k-mer on the representative's strand and
 * the node records of those two k-mers), which is what the reference's
 * terminus halves merge into.
 */
#ifndef BRUNO_CLEAN_H_
#define BRUNO_CLEAN_H_

/* summarized_chain: branch before the 5' end (0 when none), 5' k-mer (the
 * chain representative), 3' k-mer, branch after the 3' end (0 when none),
 * distance between the ends, total count of the edges into the 5' end and
 * out of the 3' end (saturating, as the 16-bit counters add) */
template <typename KT>
struct ChainSummary {
  KT b5, t5, t3, b3;
  uint32_t len;
  uint16_t f5, f6;
};

/* An edge to erase: oriented k-mer, side (1 = out, 0 = in) and base. */
template <typename KT>
struct EdgeQuery {
  KT x;
  uint32_t out;
  uint32_t c;
};

/* counter index of oriented k-mer u's edge (side out, base c) in the node
 * record of canon(u) */
template <typename KT>
__device__ __forceinline__ int edge_index(bool canonical, bool out, int c) {
  if (canonical) return out ? c : 4 + c;
  return out ? 4 + (3 - c) : 3 - c;
}

/* One thread per unitig u: starts[u] and starts[u + 1] - 1 are the sorted
 * positions of its first and last node. */
template <typename KT>
__global__ void k_chain_summaries(const Slot<KT>* T, uint64_t tcap, uint32_t rbits, int K, const uint32_t* perm,
                                  const KT* out_k, const uint32_t* dist_sorted, const uint32_t* starts, uint64_t nu,
                                  ChainSummary<KT>* out, Counters* ctr) {
  kops<KT> ops(K);
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nu; u += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t h = starts[u], t = starts[u + 1] - 1;
    ChainSummary<KT> s;
    s.t5 = out_k[perm[h]];
    s.t3 = out_k[perm[t]];
    s.len = dist_sorted[t];
    s.b5 = 0;
    s.b3 = 0;
    const KT c5 = ops.canon(s.t5), c3 = ops.canon(s.t3);
    const NodeRec* r5 = node_rec<KT>(T, 0, tcap, rbits, c5);
    const NodeRec* r3 = node_rec<KT>(T, 0, tcap, rbits, c3);
    uint32_t f5 = 0, f6 = 0;
    if (!r5 || !r3) {
      atomicOr(&ctr->err, ERR_NO_NODE);
    } else {
      /* get_in_edge_total_frequency of the 5' end / out of the 3' end on
       * the representative's strand; a chain end has at most one such edge,
       * whose far k-mer is the branch point */
      for (int c = 3; c >= 0; --c) {
        const uint16_t a = r5->c[edge_index<KT>(c5 == s.t5, false, c)];
        if (a) s.b5 = ops.in_nbr(s.t5, c);
        f5 = min(65535u, f5 + a);
        const uint16_t b = r3->c[edge_index<KT>(c3 == s.t3, true, c)];
        if (b) s.b3 = ops.out_nbr(s.t3, c);
        f6 = min(65535u, f6 + b);
      }
    }
    s.f5 = (uint16_t)f5;
    s.f6 = (uint16_t)f6;
    out[u] = s;
  }
}

/* Deadends (find_deadends + deadend_filter, compact_dbg_clean.hpp:61-76):
 * one open end, shorter than 2k - 1, both end counts below min_freq; each
 * closed end's edge goes to the query list.  Chains with both ends on
 * branch points (find_normal_chains, debruijn_topo_operations.hpp:256-283)
 * go to the bubble candidates, on the strand whose 5' branch is the
 * smaller (:319-325). */
template <typename KT>
__global__ void k_pick_chains(const ChainSummary<KT>* S, uint64_t nu, int K, uint64_t min_freq, EdgeQuery<KT>* dq,
                              unsigned long long* n_dq, uint64_t dcap, ChainSummary<KT>* normal,
                              unsigned long long* n_normal, Counters* ctr) {
  __shared__ ReserveSmem sm1, sm2;
  kops<KT> ops(K);
  const uint32_t max_len = 2 * (uint32_t)K - 1;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < nu; base += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t i = base + threadIdx.x;
    ChainSummary<KT> s{};
    bool dead = false, norm = false;
    if (i < nu) {
      s = S[i];
      dead = s.len < max_len && s.f5 < min_freq && s.f6 < min_freq && ((s.f5 == 0) != (s.f6 == 0));
      norm = s.len > 0 && s.f5 > 0 && s.f6 > 0;
    }
    const uint32_t nq = dead ? 1u : 0u; /* exactly one end is closed */
    const unsigned long long q = block_reserve(nq, n_dq, sm1);
    const unsigned long long p = block_reserve(norm ? 1u : 0u, n_normal, sm2);
    if (dead) {
      if (q >= dcap) {
        atomicOr(&ctr->err, ERR_LIST_FULL);
      } else {
        /* the branch's out edge to the 5' end (that end's last base) or its
         * in edge from the 3' end (that end's first base) */
        dq[q] = s.f5 > 0 ? EdgeQuery<KT>{s.b5, 1u, (uint32_t)ops.last(s.t5)}
                         : EdgeQuery<KT>{s.b3, 0u, (uint32_t)ops.first(s.t3)};
      }
    }
    if (norm) {
      if (s.b5 > ops.rc(s.b3)) {
        const ChainSummary<KT> r{ops.rc(s.b3), ops.rc(s.t3), ops.rc(s.t5), ops.rc(s.b5), s.len, s.f6, s.f5};
        s = r;
      }
      normal[p] = s;
    }
  }
}

/* Bubbles over the candidates sorted by (5' branch, 3' branch, length):
 * neighbours i - 1, i are linked when they join the same two branch points,
 * differ in length by at most one and one of their four end counts is below
 * min_freq (bubble_filter, compact_dbg_clean.hpp:79-101); a run of two or
 * more linked chains is a bubble (the scan at :350-397), and every chain in
 * it queries both of its end edges. */
template <typename KT>
__device__ __forceinline__ bool bubble_link(const ChainSummary<KT>& a, const ChainSummary<KT>& b, uint64_t min_freq) {
  const uint32_t d = a.len > b.len ? a.len - b.len : b.len - a.len;
  return a.b5 == b.b5 && a.b3 == b.b3 && d <= 1 &&
         (a.f5 < min_freq || a.f6 < min_freq || b.f5 < min_freq || b.f6 < min_freq);
}
template <typename KT>
__global__ void k_pick_bubbles(const ChainSummary<KT>* N, const uint32_t* order, uint64_t n, int K, uint64_t min_freq,
                               EdgeQuery<KT>* bq, unsigned long long* n_bq, uint64_t bcap, Counters* ctr) {
  __shared__ ReserveSmem sm;
  kops<KT> ops(K);
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n; base += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t i = base + threadIdx.x;
    bool bub = false;
    ChainSummary<KT> s{};
    if (i < n) {
      s = N[order[i]];
      bub = (i > 0 && bubble_link(N[order[i - 1]], s, min_freq)) ||
            (i + 1 < n && bubble_link(s, N[order[i + 1]], min_freq));
    }
    const unsigned long long q = block_reserve(bub ? 2u : 0u, n_bq, sm);
    if (bub) {
      if (q + 1 >= bcap) {
        atomicOr(&ctr->err, ERR_LIST_FULL);
      } else {
        bq[q] = EdgeQuery<KT>{s.b5, 1u, (uint32_t)ops.last(s.t5)};
        bq[q + 1] = EdgeQuery<KT>{s.b3, 0u, (uint32_t)ops.first(s.t3)};
      }
    }
  }
}

/* find_edges with the edge-count predicate (debruijn_graph_map.hpp
 * :616-666): hit[q] = the queried edge exists with a count below min_freq.
 * All queries of a list are judged before any of them is erased. */
template <typename KT>
__global__ void k_edge_hits(const Slot<KT>* T, uint64_t tcap, uint32_t rbits, int K, const EdgeQuery<KT>* Q, uint64_t n,
                            uint64_t min_freq, uint8_t* hit) {
  kops<KT> ops(K);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const EdgeQuery<KT> e = Q[i];
    const KT c = ops.canon(e.x);
    const NodeRec* r = node_rec<KT>(T, 0, tcap, rbits, c);
    uint16_t v = 0;
    if (r) v = r->c[edge_index<KT>(c == e.x, e.out != 0, (int)e.c)];
    hit[i] = v > 0 && v < min_freq;
  }
}

/* zero the counter of oriented k-mer u's edge (side out, base c), if u is
 * a node */
template <typename KT>
__device__ __forceinline__ void clear_edge(Slot<KT>* T, uint64_t tcap, uint32_t rbits, const kops<KT>& ops, KT u,
                                           bool out, int c) {
  const KT cu = ops.canon(u);
  const uint64_t s = SlotOps<KT>::find(T, tcap, rbits, cu, true);
  if (s == ~0ULL) return;
  NodeRec* r = (NodeRec*)&T[s].sub[0];
  r->c[edge_index<KT>(cu == u, out, c)] = 0;
}

/* erase_edges_internal (debruijn_graph_map.hpp:737-747): the edge on the
 * queried node and the reverse edge on its neighbour (get_remote_edges,
 * :466-503); self counts stay. */
template <typename KT>
__global__ void k_erase_edges(Slot<KT>* T, uint64_t tcap, uint32_t rbits, int K, const EdgeQuery<KT>* Q, uint64_t n,
                              const uint8_t* hit, unsigned long long* n_erased) {
  kops<KT> ops(K);
  unsigned long long mine = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!hit[i]) continue;
    const EdgeQuery<KT> e = Q[i];
    const bool out = e.out != 0;
    clear_edge<KT>(T, tcap, rbits, ops, e.x, out, (int)e.c);
    if (out) clear_edge<KT>(T, tcap, rbits, ops, ops.out_nbr(e.x, (int)e.c), false, ops.first(e.x));
    else clear_edge<KT>(T, tcap, rbits, ops, ops.in_nbr(e.x, (int)e.c), true, ops.last(e.x));
    ++mine;
  }
  block_add(mine, n_erased);
}

/* erase_if(IsIsolated) (debruijn_graph_filters.hpp:163-172) over the node
 * lists: a node left without edges leaves the graph. */
template <typename KT>
__global__ void k_drop_edgeless(Slot<KT>* T, const uint64_t* slots, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    Slot<KT>* p = T + slots[i];
    uint64_t* fw = SlotOps<KT>::flag_word(p);
    if (!(*fw & F_PRESENT)) continue;
    const NodeRec* r = (const NodeRec*)&p->sub[0];
    bool any = false;
#pragma unroll
    for (int q = 0; q < 8; ++q) any |= r->c[q] > 0;
    if (!any) *fw = *fw & ~F_PRESENT;
  }
}

/* erase_nodes of the chain nodes list ranking separated (cycles and
 * isolated nodes; compact_dbg_freq_construct_clean.cpp:606-610,
 * debruijn_graph_map.hpp:822-856): the node leaves the graph, and every
 * neighbour that is still a node loses its edge back.  Called after every
 * node to drop has been marked, so neighbours in the same cycle find each
 * other gone. */
template <typename KT>
__global__ void k_drop_separated(Slot<KT>* T, uint64_t tcap, uint32_t rbits, int K, const uint64_t* chain_slot,
                                 const uint8_t* alive, uint64_t nc, uint8_t* mark) {
  kops<KT> ops(K);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (uint64_t)gridDim.x * blockDim.x) {
    mark[i] = 0;
    if (alive[i]) continue;
    Slot<KT>* p = T + chain_slot[i];
    uint64_t* fw = SlotOps<KT>::flag_word(p);
    if (!(*fw & F_PRESENT)) continue;
    *fw = *fw & ~F_PRESENT;
    mark[i] = 1;
  }
}
template <typename KT>
__global__ void k_drop_separated_edges(Slot<KT>* T, uint64_t tcap, uint32_t rbits, int K, const uint64_t* chain_slot,
                                       const uint8_t* mark, uint64_t nc) {
  kops<KT> ops(K);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!mark[i]) continue;
    const Slot<KT>* p = T + chain_slot[i];
    const KT x = SlotOps<KT>::kmer(p);
    const NodeRec* r = (const NodeRec*)&p->sub[0];
    for (int c = 0; c < 4; ++c) {
      if (r->c[c] > 0) clear_edge<KT>(T, tcap, rbits, ops, ops.out_nbr(x, c), false, ops.first(x));
      if (r->c[4 + c] > 0) clear_edge<KT>(T, tcap, rbits, ops, ops.in_nbr(x, c), true, ops.last(x));
    }
  }
}

/* The node lists again after edges changed: every node still in the graph
 * is classified from its counters (IsBranchPoint / IsChainNode) and
 * appended to the chain or the branch list (one thread per old list entry,
 * chain list first). */
template <typename KT>
__global__ void k_relist(Slot<KT>* T, int K, const uint64_t* old_chain, uint64_t nc, const uint64_t* old_branch,
                         uint64_t nb, NodeLists nl, Counters* ctr) {
  __shared__ ReserveSmem sm, sm2;
  kops<KT> ops(K);
  const uint64_t n = nc + nb;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < n; base += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t i = base + threadIdx.x;
    bool kept = false;
    uint64_t s = 0;
    NodeRec rec;
    rec.cls = 0;
    if (i < n) {
      s = i < nc ? old_chain[i] : old_branch[i - nc];
      const Slot<KT>* p = T + s;
      if (SlotOps<KT>::flags(p) & F_PRESENT) {
        const NodeRec* r = (const NodeRec*)&p->sub[0];
        uint16_t c[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) c[q] = r->c[q];
        rec = make_node<KT>(ops, SlotOps<KT>::kmer(p), c);
        kept = true;
      }
    }
    const bool ch = kept && rec.cls == CLS_CHAIN, br = kept && rec.cls == CLS_BRANCH;
    const unsigned long long pc = block_reserve(ch, &ctr->n_chain, sm);
    const unsigned long long pb = block_reserve(br, &ctr->n_branch, sm2);
    if (kept) store_node<KT>(T, s, rec, ch ? pc : pb, nl, ctr);
  }
}

#endif
