/*
 * bruno_oracle.cpp -- CPU restatement of bruno's de Bruijn hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle; see bruno_oracle.h).
 *
 * K-mers are held as unsigned __int128 values v = sum_i c_i * 4^(K-1-i) with
 * the DNA codes A0 C1 G2 T3, i.e. the first base is the most significant
 * digit.  This is the value order bruno's Kmer<K,DNA,uint64_t> compares by
 * (one word for K <= 31, two words for K <= 63), so numeric order is the
 * lexicographic order lex_less relies on (debruijn_common.hpp:44-67).
 *
 * Edge bytes use DNA16 nibbles A1 C2 G4 T8, 0 = no base, in<<4 | out
 * (debruijn_biedge_loader.hpp:47-50, :356-372).
 */
#include "bruno_oracle.h"

#include <algorithm>
#include <map>
#include <array>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* Kmer arithmetic (restated from the kmerind Kmer semantics that bruno's
 * call sites rely on; see kmer_traits.hpp:50-124 and
 * debruijn_graph_node.hpp:315-381 for the consumers).                  */
/* ------------------------------------------------------------------ */

struct KmerOps {
  int K;
  u128 mask;
  explicit KmerOps(int k) : K(k) {
    mask = (K == 64) ? ~(u128)0 : (((u128)1 << (2 * K)) - 1);
  }
  /* base at sequence position p (0 = first base). */
  inline int base(u128 v, int p) const { return (int)((v >> (2 * (K - 1 - p))) & 3); }
  /* getCharsAtPos(pos, 1): pos 0 is the LAST base (debruijn_graph_map.hpp:442). */
  inline int chars_at_pos(u128 v, int pos) const { return (int)((v >> (2 * pos)) & 3); }
  /* reverse complement: complement every 2-bit code (3 - c == ~c & 3) and
   * reverse the order of the codes (swap pairs, nibbles, then bytes). */
  static uint64_t rc64(uint64_t x) {
    x = ~x;
    x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
    return __builtin_bswap64(x);
  }
  u128 rc(u128 v) const {
    u128 r = ((u128)rc64((uint64_t)v) << 64) | (u128)rc64((uint64_t)(v >> 64));
    return r >> (128 - 2 * K);
  }
  /* nextFromChar(c): shift in c at the end (out neighbour). */
  inline u128 next_from_char(u128 v, int c) const { return ((v << 2) | (u128)c) & mask; }
  /* nextReverseFromChar(c): shift in c at the front (in neighbour). */
  inline u128 next_reverse_from_char(u128 v, int c) const {
    return (v >> 2) | ((u128)c << (2 * (K - 1)));
  }
  std::string to_string(u128 v) const {
    static const char A[4] = {'A', 'C', 'G', 'T'};
    std::string s(K, 'A');
    for (int i = 0; i < K; ++i) s[i] = A[base(v, i)];
    return s;
  }
  /* rc-palindrome test of bases [p, p+len) of v. */
  bool sub_is_rc_pal(u128 v, int p, int len) const {
    for (int i = 0; i < len; ++i)
      if (base(v, p + i) != 3 - base(v, p + len - 1 - i)) return false;
    return true;
  }
  /* kmer_traits::is_rc_palindrome (kmer_traits.hpp:70-73): odd K never. */
  bool is_rc_palindrome(u128 v) const {
    if (K & 1) return false;
    return v == rc(v);
  }
  /* is_kminus1_rc_palindrome_high: the first K-1 bases form an rc-palindrome
   * (kmer_traits.hpp:58-62).  Even K: false. */
  bool km1_high(u128 v) const {
    if ((K & 1) == 0) return false;
    return sub_is_rc_pal(v, 0, K - 1);
  }
  /* is_kminus1_rc_palindrome_low: the last K-1 bases (kmer_traits.hpp:63-67). */
  bool km1_low(u128 v) const {
    if ((K & 1) == 0) return false;
    return sub_is_rc_pal(v, 1, K - 1);
  }
  /* is_k1_rc_palindrome(k, r, low): k.r is an rc-palindrome
   * (kmer_traits.hpp:77-82): first base == comp(r) and low. */
  bool k1_pal_out(u128 v, int r, bool low) const {
    if ((K & 1) == 0) return false;
    if (chars_at_pos(v, K - 1) != 3 - r) return false;
    return low;
  }
  /* is_k1_rc_palindrome(l, k, high): l.k (kmer_traits.hpp:84-89). */
  bool k1_pal_in(int l, u128 v, bool high) const {
    if ((K & 1) == 0) return false;
    if (chars_at_pos(v, 0) != 3 - l) return false;
    return high;
  }
  /* is_k1_k1pair_rc_palindrome(k, r, low) (kmer_traits.hpp:110-118). */
  int k1_k1pair_out(u128 v, int r, bool low) const {
    if ((K & 1) == 0) return 0;
    if (!low) return 0;
    int f = chars_at_pos(v, K - 1);
    return (f == 3 - r ? 1 : 0) | (f == r ? 2 : 0);
  }
  /* is_k1_k1pair_rc_palindrome(l, k, high) (kmer_traits.hpp:119-127). */
  int k1_k1pair_in(int l, u128 v, bool high) const {
    if ((K & 1) == 0) return 0;
    if (!high) return 0;
    int f = chars_at_pos(v, 0);
    return (f == 3 - l ? 1 : 0) | (f == l ? 2 : 0);
  }
};

/* DNA16 nibble of an ASCII base; 0 for anything outside ACGT (the read is
 * then dropped before this is used). */
inline int dna_code(char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
  }
}

/* DNA16 reverse complement of the biedge byte: Kmer<2,DNA16>::reverse_complement
 * swaps the two nibbles and complements each; the DNA16 complement is the
 * nibble bit reversal (debruijn_biedge_loader.hpp:705-713). */
inline uint8_t dna16_comp(uint8_t n) {
  return (uint8_t)(((n & 1) << 3) | ((n & 2) << 1) | ((n & 4) >> 1) | ((n & 8) >> 3));
}
inline uint8_t biedge_rc(uint8_t e) {
  return (uint8_t)((dna16_comp(e & 0xF) << 4) | dna16_comp(e >> 4));
}

/* FROM_DNA16 of compact_multi_biedge<DNA> (debruijn_graph_node.hpp:118-136):
 * A,C,G,T nibbles map to 0..3, anything else (incl. 0) to 8 (ignored). */
inline int from_dna16(int n) {
  switch (n) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    default: return 8;
  }
}

/* ------------------------------------------------------------------ */
/* compact_multi_biedge<DNA, uint16_t> (debruijn_graph_node.hpp:99-392) */
/* ------------------------------------------------------------------ */
typedef std::array<uint16_t, 9> Counts;

inline uint16_t sat_add16(uint16_t a, uint16_t b) {  // :81-85
  uint16_t c = (uint16_t)(a + b);
  return (c < a) ? (uint16_t)0xFFFF : c;
}
inline void sat_incr16(uint16_t& t) { t = (uint16_t)(t + (t < 0xFFFF)); }  // :88-91

inline void node_update(Counts& c, uint8_t e) {  // update(edges) :187-198
  int out = from_dna16(e & 0xF);
  if (out < 4) sat_incr16(c[out]);
  int in = from_dna16(e >> 4);
  if (in < 4) sat_incr16(c[in + 4]);
  sat_incr16(c[8]);
}
inline void node_update_cnt(Counts& c, uint8_t e, uint16_t cnt) {  // :200-211
  int out = from_dna16(e & 0xF);
  if (out < 4) c[out] = sat_add16(c[out], cnt);
  int in = from_dna16(e >> 4);
  if (in < 4) c[in + 4] = sat_add16(c[in + 4], cnt);
  c[8] = sat_add16(c[8], cnt);
}
inline int out_edge_count(const Counts& c) {  // :283-289
  int n = 0;
  for (int i = 0; i < 4; ++i) n += c[i] > 0;
  return n;
}
inline int in_edge_count(const Counts& c) {  // :291-297
  int n = 0;
  for (int i = 4; i < 8; ++i) n += c[i] > 0;
  return n;
}

/* ------------------------------------------------------------------ */
/* Records and hashing                                                  */
/* ------------------------------------------------------------------ */
struct Rec {
  u128 kmer;
  uint8_t e;
};
struct K2 {
  u128 kmer;
  uint8_t e;
  uint32_t cnt;
};
struct Node {
  u128 kmer;
  Counts c;
};

inline uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}
inline uint64_t hash128(u128 v) { return mix64((uint64_t)v ^ mix64((uint64_t)(v >> 64) + 0x9e3779b97f4a7c15ULL)); }

/* ------------------------------------------------------------------ */
/* Chain node record (debruijn_chain_node.hpp:49-80)                    */
/* ------------------------------------------------------------------ */
const uint32_t dist_mask = 0x7FFFFFFFu;
const uint32_t term_mask = 0x80000000u;
const uint32_t branch_neighbor = 0u;
inline uint32_t get_chain_dist(uint32_t x) { return x & dist_mask; }
inline bool is_chain_terminal(uint32_t x) { return (x & dist_mask) == 0; }
inline bool points_to_self(uint32_t x) { return x == term_mask; }
inline bool points_to_chain_node(uint32_t x) { return (x ^ term_mask) > term_mask; }
inline uint32_t mark_dist_as_point_to_terminal(uint32_t x) { return x | term_mask; }

struct Biedge {  // simple_biedge<K> = tuple<K in, K out, uint in_dist, uint out_dist>
  u128 in, out;
  uint32_t din, dout;
};

struct ChainMap {
  std::vector<u128> keys;      // sorted canonical k-mers
  std::vector<Biedge> vals;
  std::vector<uint8_t> alive;  // erased entries are kept but marked dead
  long find(u128 k) const {
    auto it = std::lower_bound(keys.begin(), keys.end(), k);
    if (it == keys.end() || *it != k) return -1;
    long i = (long)(it - keys.begin());
    return alive[i] ? i : -1;
  }
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

/* ================================================================== */
struct oracle_job {
  int K;
  KmerOps ops;
  bool thresholding;
  bool edge_filter = false;  /* k+1-mer edge-frequency filter (block-construct driver) */
  uint64_t th[6];
  int nthreads;
  uint64_t order_seed = 0;  /* 0: generation order; else shuffle updates */

  std::vector<std::string> reads;  // kept reads (ACGT only)
  std::vector<Rec> records;        // P0 (canonical)
  std::vector<K2> k2;              // distinct canonical k2-mers (sorted)
  std::vector<Node> unfiltered;    // P1
  std::vector<Node> nodes;         // P2 (sorted by k-mer)
  std::vector<u128> branches;      // P3 (sorted)
  ChainMap chain;
  uint64_t n_cycle = 0, n_isolated = 0, iterations = 0;
  std::vector<u128> cycle_kmers;         // the chain nodes separate_cycles removed
  std::vector<oracle_chain_node> ranked;  // P4 (sorted by chain_rep_less)
  std::string branch_fasta, chain_fasta, chain_components;
  std::string branch_edges, chain_edges;  /* non-benchmark outputs */
  uint64_t chain_edge_misses = 0;         /* chain ends without a frequency summary */
  double phase[6] = {0, 0, 0, 0, 0, 0};

  oracle_job(int k, const uint64_t* t, int nt) : K(k), ops(k), thresholding(t != nullptr), nthreads(nt < 1 ? 1 : nt) {
    for (int i = 0; i < 6; ++i) th[i] = t ? t[i] : 65536;
  }

  long node_find(u128 k) const {
    Node key;
    key.kmer = k;
    auto it = std::lower_bound(nodes.begin(), nodes.end(), key,
                               [](const Node& a, const Node& b) { return a.kmer < b.kmer; });
    if (it == nodes.end() || it->kmer != k) return -1;
    return (long)(it - nodes.begin());
  }

  /* ---------------- parse (debruijn_biedge_loader.hpp:394-623) -------- */
  /* For read s of length L: records i = 0..L-K with the k-mer s[i..i+K) and
   * the biedge built by biedge_generating_iterator::operator* (:356-372):
   * in nibble = DNA16(s[i-1]) unless i == 0, out nibble = DNA16(s[i+K])
   * unless i+K == L.  Reads containing a non-ACGT base are dropped whole
   * (NFilterSequencesIterator, compact_dbg_freq_construct_clean.cpp:210-213).
   * Reads shorter than the parser window K+2 (:402, :464-471) produce no
   * records; this gate lives in kmerind's get_valid_iterator_range, which is
   * absent from the mount, so the exact boundary is an assumption. */
  void parse_read(const std::string& s, std::vector<Rec>& out) const {
    size_t L = s.size();
    if (L < (size_t)K + 2) return;
    u128 v = 0;
    for (int i = 0; i < K; ++i) v = (v << 2) | (u128)dna_code(s[i]);
    for (size_t i = 0; i + K <= L; ++i) {
      if (i > 0) v = ops.next_from_char(v, dna_code(s[i + K - 1]));
      uint8_t left = (i > 0) ? (uint8_t)(1u << dna_code(s[i - 1])) : 0;
      uint8_t right = (i + K < L) ? (uint8_t)(1u << dna_code(s[i + K])) : 0;
      Rec r;
      r.kmer = v;
      r.e = (uint8_t)((left << 4) | right);
      /* lex_less on (k-mer, value): keep if x < rc(x), else rc both
       * (debruijn_common.hpp:53-58); applied by transform_input
       * (debruijn_graph_map.hpp:1099). */
      u128 y = ops.rc(r.kmer);
      if (!(r.kmer < y)) {
        r.kmer = y;
        r.e = biedge_rc(r.e);
      }
      out.push_back(r);
    }
  }

  void parse_all() {
    records.clear();
    if (nthreads <= 1) {
      for (auto& s : reads) parse_read(s, records);
      return;
    }
    std::vector<std::vector<Rec>> parts(nthreads);
    std::vector<std::thread> ts;
    size_t n = reads.size();
    for (int t = 0; t < nthreads; ++t) {
      ts.emplace_back([&, t]() {
        size_t b = n * t / nthreads, e = n * (t + 1) / nthreads;
        for (size_t i = b; i < e; ++i) parse_read(reads[i], parts[t]);
      });
    }
    for (auto& th_ : ts) th_.join();
    size_t tot = 0;
    for (auto& p : parts) tot += p.size();
    records.reserve(tot);
    for (auto& p : parts) records.insert(records.end(), p.begin(), p.end());
  }

  /* ---------------- k2 counting (debruijn_graph_map.hpp:1085-1141) ---- */
  /* Distinct canonical (k-mer, biedge) pairs with uint32 counts, no
   * saturation (:1122-1132).  Restated as sort + run-length; the result is
   * the same multiset as the hash-map insert. */
  static void count_k2(std::vector<Rec>& recs, std::vector<K2>& out) {
    std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) {
      return a.kmer < b.kmer || (a.kmer == b.kmer && a.e < b.e);
    });
    out.clear();
    for (size_t i = 0; i < recs.size();) {
      size_t j = i;
      while (j < recs.size() && recs[j].kmer == recs[i].kmer && recs[j].e == recs[i].e) ++j;
      out.push_back(K2{recs[i].kmer, recs[i].e, (uint32_t)(j - i)});
      i = j;
    }
  }

  /* ---------------- local_insert_by_freqencies (:1258-1623) ----------- */
  /* k2 is sorted by k-mer (:1275).  Per k-mer block: k_f = sum cnt; k1in/
   * k1out per edge char, halved ((cnt+1)>>1) when the k+1-mer is an
   * rc-palindrome and K is not (:1468-1507); capped at 65535 (:1402-1407).
   * Flags: 4 = k2 out of [t2lo,t2hi), 8 = k_f out of [t0lo,t0hi), 2 = in k1
   * out of [t1lo,t1hi) or no in edge, 1 = likewise out (:1416-1446).  Keep
   * flag < 3; flag 1 clears the out nibble, flag 2 the in nibble
   * (:1560-1597); then local_insert with update(edge, min(cnt, 65535))
   * (:264-285). */
  void filter_block(const K2* b, const K2* e, std::vector<Node>& out) const {
    const uint64_t cnt_max = 65535;
    const uint64_t t0_lo = th[0], t0_hi = th[1], t1_lo = th[2], t1_hi = th[3], t2_lo = th[4], t2_hi = th[5];
    u128 k = b->kmer;
    bool k_pal = ops.is_rc_palindrome(k);
    bool high = ops.km1_high(k);
    bool low = ops.km1_low(k);
    uint64_t k_f = 0, k1in[4] = {0, 0, 0, 0}, k1out[4] = {0, 0, 0, 0};
    for (const K2* it = b; it != e; ++it) {
      uint64_t cnt_unsplit = it->cnt;
      k_f += cnt_unsplit;
      int ch = it->e >> 4;
      if (ch > 0) {
        int c = from_dna16(ch);  // power of two guaranteed for ACGT input
        bool split = ops.k1_pal_in(c, k, high) && !k_pal;
        k1in[c] += split ? ((cnt_unsplit + 1) >> 1) : cnt_unsplit;
      }
      ch = it->e & 0xF;
      if (ch > 0) {
        int c = from_dna16(ch);
        bool split = ops.k1_pal_out(k, c, low) && !k_pal;
        k1out[c] += split ? ((cnt_unsplit + 1) >> 1) : cnt_unsplit;
      }
    }
    k_f = std::min(cnt_max, k_f);
    for (int c = 0; c < 4; ++c) {
      k1in[c] = std::min(cnt_max, k1in[c]);
      k1out[c] = std::min(cnt_max, k1out[c]);
    }
    bool have = false;
    Node n;
    n.kmer = k;
    n.c.fill(0);
    for (const K2* it = b; it != e; ++it) {
      uint8_t flag = 0;
      uint64_t cnt = std::min(cnt_max, (uint64_t)it->cnt);
      if (cnt < t2_lo || cnt >= t2_hi) flag |= 4;
      if (k_f < t0_lo || k_f >= t0_hi) flag |= 8;
      uint64_t k1cnt;
      int in = it->e >> 4, outn = it->e & 0xF;
      k1cnt = (in == 1 || in == 2 || in == 4 || in == 8) ? k1in[from_dna16(in)] : t1_hi;
      if (k1cnt < t1_lo || k1cnt >= t1_hi) flag |= 2;
      k1cnt = (outn == 1 || outn == 2 || outn == 4 || outn == 8) ? k1out[from_dna16(outn)] : t1_hi;
      if (k1cnt < t1_lo || k1cnt >= t1_hi) flag |= 1;
      if (flag >= 3) continue;
      uint8_t edge = it->e;
      if (flag == 1) edge &= 0xF0;
      if (flag == 2) edge &= 0x0F;
      /* local_insert frequency version (:264-285): cnt > 0 always here. */
      node_update_cnt(n.c, edge, (uint16_t)cnt);
      have = true;
    }
    if (have) out.push_back(n);
  }

  /* ---------------- k+1-mer edge-frequency filter ------------------------
   * The block-construct driver (compact_debruijn_graph_block_construct.cpp
   * :415-480): compute_edge_frequency counts canonical k+1-mers over the
   * reads and keeps those with lower <= count < upper (saturating uint16
   * counts, :151; debruijn_biedge_filters.hpp:128-176);
   * transform_biedges_by_frequency (:330-418) clears a record's in / out
   * edge whose canonical k+1-mer is not kept; filter_isolated_nodes (:60-69)
   * drops records left without edges; the rest are inserted.  Restated per
   * k-mer block from the k2 list: the windows of canonical k+1-mer a.K are
   * the records of K with in-edge a, except that a window of an
   * rc-palindromic a.K (K not a palindrome) yields two such records, so its
   * window count is half their sum (the sum is even); an rc-palindromic K
   * (even k) is recorded from the reverse strand, so its in-edge c also
   * collects the records with out-edge comp(c).  A kept record adds
   * its surviving edges to the node (the reference's index stores edge
   * existence; the counts here are those of the kept records).  Thresholds:
   * lower = th[2], upper = th[3]. */
  void filter_block_edges(const K2* b, const K2* e, std::vector<Node>& out) const {
    const uint64_t cnt_max = 65535;
    const uint64_t lo = th[2], hi = th[3];
    u128 k = b->kmer;
    bool k_pal = ops.is_rc_palindrome(k);
    bool high = ops.km1_high(k);
    bool low = ops.km1_low(k);
    uint64_t win_in[4] = {0, 0, 0, 0}, win_out[4] = {0, 0, 0, 0};
    bool split_in[4], split_out[4];
    for (int c = 0; c < 4; ++c) {
      split_in[c] = ops.k1_pal_in(c, k, high) && !k_pal;
      split_out[c] = ops.k1_pal_out(k, c, low) && !k_pal;
    }
    uint64_t s_in[4] = {0, 0, 0, 0}, s_out[4] = {0, 0, 0, 0};
    for (const K2* it = b; it != e; ++it) {
      int ch = it->e >> 4;
      if (ch > 0) s_in[from_dna16(ch)] += it->cnt;
      ch = it->e & 0xF;
      if (ch > 0) s_out[from_dna16(ch)] += it->cnt;
    }
    for (int c = 0; c < 4; ++c) {
      if (k_pal) {
        /* every instance of an rc-palindromic k-mer is recorded from the
         * reverse strand (lex_less flips ties): the windows written c.K
         * show up as out-edge comp(c), those written K.comp(c) as in-edge
         * c; both are the canonical k+1-mer of in-edge c (even k, so no
         * k+1-mer is a palindrome) */
        win_in[c] = s_in[c] + s_out[3 - c];
        win_out[c] = s_out[c] + s_in[3 - c];
      } else {
        win_in[c] = split_in[c] ? s_in[c] >> 1 : s_in[c];
        win_out[c] = split_out[c] ? s_out[c] >> 1 : s_out[c];
      }
      win_in[c] = std::min(cnt_max, win_in[c]);
      win_out[c] = std::min(cnt_max, win_out[c]);
    }
    bool have = false;
    Node n;
    n.kmer = k;
    n.c.fill(0);
    for (const K2* it = b; it != e; ++it) {
      int in = it->e >> 4, outn = it->e & 0xF;
      bool in_ok = in > 0 && win_in[from_dna16(in)] >= lo && win_in[from_dna16(in)] < hi;
      bool out_ok = outn > 0 && win_out[from_dna16(outn)] >= lo && win_out[from_dna16(outn)] < hi;
      if (!in_ok && !out_ok) continue;
      uint8_t edge = it->e;
      if (!in_ok) edge &= 0x0F;
      if (!out_ok) edge &= 0xF0;
      node_update_cnt(n.c, edge, (uint16_t)std::min(cnt_max, (uint64_t)it->cnt));
      have = true;
    }
    if (have) out.push_back(n);
  }

  void filter_range(const std::vector<K2>& v, std::vector<Node>& out) const {
    for (size_t i = 0; i < v.size();) {
      size_t j = i;
      while (j < v.size() && v[j].kmer == v[i].kmer) ++j;
      if (edge_filter) filter_block_edges(&v[i], &v[0] + j, out);
      else filter_block(&v[i], &v[0] + j, out);
      i = j;
    }
  }

  /* unthresholded: local_insert of each record with update(edge)
   * (debruijn_graph_map.hpp:175-197). */
  static void unfiltered_from_k2(const std::vector<K2>& v, std::vector<Node>& out) {
    for (size_t i = 0; i < v.size();) {
      Node n;
      n.kmer = v[i].kmer;
      n.c.fill(0);
      size_t j = i;
      for (; j < v.size() && v[j].kmer == v[i].kmer; ++j)
        for (uint32_t r = 0; r < v[j].cnt; ++r) node_update(n.c, v[j].e);
      out.push_back(n);
      i = j;
    }
  }

  void count_and_filter() {
    double t0 = now_s();
    k2.clear();
    unfiltered.clear();
    nodes.clear();
    if (nthreads <= 1) {
      count_k2(records, k2);
      phase[1] = now_s() - t0;
      t0 = now_s();
      unfiltered_from_k2(k2, unfiltered);
      if (thresholding) filter_range(k2, nodes);
      else nodes = unfiltered;
      phase[2] = now_s() - t0;
      return;
    }
    /* Shards mirror MPI ranks: distribute by key hash
     * (debruijn_graph_map.hpp:1105-1111), then count and filter locally. */
    int T = nthreads;
    std::vector<std::vector<Rec>> shard(T);
    {
      std::vector<std::vector<std::vector<Rec>>> parts(T, std::vector<std::vector<Rec>>(T));
      std::vector<std::thread> ts;
      size_t n = records.size();
      for (int t = 0; t < T; ++t)
        ts.emplace_back([&, t]() {
          size_t b = n * t / T, e = n * (t + 1) / T;
          for (size_t i = b; i < e; ++i) parts[t][hash128(records[i].kmer) % T].push_back(records[i]);
        });
      for (auto& x : ts) x.join();
      for (int s = 0; s < T; ++s) {
        size_t tot = 0;
        for (int t = 0; t < T; ++t) tot += parts[t][s].size();
        shard[s].reserve(tot);
        for (int t = 0; t < T; ++t) {
          shard[s].insert(shard[s].end(), parts[t][s].begin(), parts[t][s].end());
          std::vector<Rec>().swap(parts[t][s]);
        }
      }
    }
    std::vector<std::vector<K2>> sk2(T);
    std::vector<std::vector<Node>> sun(T), sno(T);
    {
      std::vector<std::thread> ts;
      for (int t = 0; t < T; ++t) ts.emplace_back([&, t]() { count_k2(shard[t], sk2[t]); });
      for (auto& x : ts) x.join();
    }
    phase[1] = now_s() - t0;
    t0 = now_s();
    {
      std::vector<std::thread> ts;
      for (int t = 0; t < T; ++t)
        ts.emplace_back([&, t]() {
          unfiltered_from_k2(sk2[t], sun[t]);
          if (thresholding) filter_range(sk2[t], sno[t]);
        });
      for (auto& x : ts) x.join();
    }
    auto by_kmer_n = [](const Node& a, const Node& b) { return a.kmer < b.kmer; };
    for (int t = 0; t < T; ++t) {
      k2.insert(k2.end(), sk2[t].begin(), sk2[t].end());
      unfiltered.insert(unfiltered.end(), sun[t].begin(), sun[t].end());
      nodes.insert(nodes.end(), sno[t].begin(), sno[t].end());
    }
    std::sort(k2.begin(), k2.end(), [](const K2& a, const K2& b) {
      return a.kmer < b.kmer || (a.kmer == b.kmer && a.e < b.e);
    });
    std::sort(unfiltered.begin(), unfiltered.end(), by_kmer_n);
    if (thresholding) std::sort(nodes.begin(), nodes.end(), by_kmer_n);
    else nodes = unfiltered;
    phase[2] = now_s() - t0;
  }

  /* ---------------- node predicates (debruijn_graph_filters.hpp) ------ */
  bool is_kplus1_palindrome(const Node& n) const {  // :57-95
    if ((K & 1) == 0) return false;
    bool low = ops.km1_low(n.kmer);
    for (int c = 0; c < 4; ++c)
      if (n.c[c] > 0 && ops.k1_k1pair_out(n.kmer, c, low)) return true;
    bool high = ops.km1_high(n.kmer);
    for (int c = 0; c < 4; ++c)
      if (n.c[c + 4] > 0 && ops.k1_k1pair_in(c, n.kmer, high)) return true;
    return false;
  }
  bool is_branch(const Node& n) const {  // IsBranchPoint :142-160
    return in_edge_count(n.c) >= 2 || out_edge_count(n.c) >= 2 || ops.is_rc_palindrome(n.kmer) ||
           is_kplus1_palindrome(n);
  }
  bool is_chain(const Node& n) const {  // IsChainNode :206-220
    return in_edge_count(n.c) < 2 && out_edge_count(n.c) < 2 && !ops.is_rc_palindrome(n.kmer) &&
           !is_kplus1_palindrome(n);
  }

  /* ---------------- chain extraction (debruijn_chain_graph.hpp:605-667) */
  void extract_chains() {
    chain.keys.clear();
    chain.vals.clear();
    chain.alive.clear();
    branches.clear();
    for (auto& n : nodes) {
      if (is_branch(n)) branches.push_back(n.kmer);
      if (!is_chain(n)) continue;
      /* to_simple_biedge (debruijn_chain_node.hpp:84-141): neighbour k-mers
       * on the key's strand, dist 1, else KMER() and term_mask. */
      Biedge b{0, 0, term_mask, term_mask};
      for (int c = 0; c < 4; ++c)
        if (n.c[c + 4] > 0) {
          b.in = ops.next_reverse_from_char(n.kmer, c);
          b.din = 1;
        }
      for (int c = 0; c < 4; ++c)
        if (n.c[c] > 0) {
          b.out = ops.next_from_char(n.kmer, c);
          b.dout = 1;
        }
      chain.keys.push_back(n.kmer);
      chain.vals.push_back(b);
      chain.alive.push_back(1);
    }
    setup_chain_termini();
  }

  /* canonicalise an update key with lex_less; returns true if flipped. */
  bool canon_key(u128& key) const {
    u128 y = ops.rc(key);
    if (key < y) return false;
    key = y;
    return true;
  }

  /* setup_chain_termini (:486-596) with terminus_update
   * (debruijn_chain_operations.hpp:58-93). */
  void setup_chain_termini() {
    for (auto& n : nodes) {
      if (!is_branch(n)) continue;
      // out neighbours receive (t, IN); in neighbours receive (t, OUT)
      for (int c = 0; c < 4; ++c)
        if (n.c[c] > 0) apply_terminus(ops.next_from_char(n.kmer, c), -1);
      for (int c = 0; c < 4; ++c)
        if (n.c[c + 4] > 0) apply_terminus(ops.next_reverse_from_char(n.kmer, c), +1);
    }
  }
  void apply_terminus(u128 key, int orient) {
    /* reverse_complement(terminus_update_md) negates the orientation
     * (debruijn_chain_operations.hpp:491-497). */
    if (canon_key(key)) orient = -orient;
    long i = chain.find(key);
    if (i < 0) return;  // update without insert
    if (orient < 0) chain.vals[i].din = branch_neighbor;
    else chain.vals[i].dout = branch_neighbor;
  }

  /* chain_update (debruijn_chain_operations.hpp:108-256). */
  static size_t chain_update(Biedge& x, u128 yk, uint32_t yd, int orient) {
    uint32_t dist = get_chain_dist(yd);
    if (orient < 0) {
      uint32_t distx = get_chain_dist(x.din);
      if (points_to_chain_node(x.din)) {
        if (dist > distx) x.in = yk;
        if (dist >= distx) {
          x.din = yd;
          return 1;
        }
      }
    } else {
      uint32_t distx = get_chain_dist(x.dout);
      if (points_to_chain_node(x.dout)) {
        if (dist > distx) x.out = yk;
        if (dist >= distx) {
          x.dout = yd;
          return 1;
        }
      }
    }
    return 0;
  }

  struct Upd {
    u128 key, k;
    uint32_t d;
    int orient;
  };

  /* Applies the round's updates in generation order.  With nthreads > 1 the
   * canonicalisation and lookups run in parallel and each thread applies, in
   * generation order, the updates whose target falls in its index range, so
   * the result is the same as the sequential loop. */
  size_t apply_updates(std::vector<Upd>& ups) {
    std::vector<long> tgt(ups.size());
    auto prep = [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) {
        Upd& u = ups[q];
        /* lex_less on (key, chain_update_md): rc key and md k-mer, negate the
         * orientation (debruijn_chain_operations.hpp:500-505). */
        if (canon_key(u.key)) {
          u.k = ops.rc(u.k);
          u.orient = -u.orient;
        }
        tgt[q] = chain.find(u.key);
      }
    };
    int T = (nthreads > 1 && ups.size() > 100000) ? nthreads : 1;
    if (T == 1) {
      prep(0, ups.size());
      size_t count = 0;
      for (size_t q = 0; q < ups.size(); ++q)
        if (tgt[q] >= 0) count += chain_update(chain.vals[tgt[q]], ups[q].k, ups[q].d, ups[q].orient);
      return count;
    }
    std::vector<std::thread> ts;
    for (int t = 0; t < T; ++t)
      ts.emplace_back([&, t]() { prep(ups.size() * t / T, ups.size() * (t + 1) / T); });
    for (auto& x : ts) x.join();
    ts.clear();
    std::vector<size_t> cnt(T, 0);
    long n = (long)chain.keys.size();
    for (int t = 0; t < T; ++t)
      ts.emplace_back([&, t]() {
        long lo = n * t / T, hi = n * (t + 1) / T;
        for (size_t q = 0; q < ups.size(); ++q)
          if (tgt[q] >= lo && tgt[q] < hi) cnt[t] += chain_update(chain.vals[tgt[q]], ups[q].k, ups[q].d, ups[q].orient);
      });
    for (auto& x : ts) x.join();
    size_t count = 0;
    for (auto c : cnt) count += c;
    return count;
  }

  static bool is_unfinished(const Biedge& b) {  // PointsToInternalNode
    return points_to_chain_node(b.din) || points_to_chain_node(b.dout);
  }
  static bool is_cycle(const Biedge& b, uint32_t maxd) {  // IsCycleNode
    return get_chain_dist(b.din) == maxd && get_chain_dist(b.dout) == maxd;
  }

  /* list_rank (:793-994, variant 0) and list_rank_min_update (:1040-1310,
   * variant 1).  Updates are generated from the state at the start of each
   * round and then applied in generation order; the unfinished list keeps
   * its relative order (the reference uses std::partition, whose order is
   * unspecified; order only matters for conflicting updates, which
   * consistent graphs do not produce). */
  void list_rank(int variant) {
    std::vector<long> unfinished;
    for (size_t i = 0; i < chain.keys.size(); ++i)
      if (chain.alive[i] && is_unfinished(chain.vals[i])) unfinished.push_back((long)i);
    bool all_compacted = unfinished.empty();
    size_t iters = 0;
    std::vector<Upd> ups;
    while (!all_compacted) {
      ups.clear();
      for (long t : unfinished) {
        const Biedge md = chain.vals[t];
        u128 km = chain.keys[t];
        uint32_t dist = get_chain_dist(md.din) + get_chain_dist(md.dout);
        bool send_in, send_out;
        if (variant == 0) {
          send_in = !is_chain_terminal(md.din);
          send_out = !is_chain_terminal(md.dout);
        } else {
          send_in = points_to_chain_node(md.din);
          send_out = points_to_chain_node(md.dout);
        }
        if (send_in)
          ups.push_back(Upd{md.in, is_chain_terminal(md.dout) ? km : md.out,
                            points_to_chain_node(md.dout) ? dist : mark_dist_as_point_to_terminal(dist), +1});
        if (send_out)
          ups.push_back(Upd{md.out, is_chain_terminal(md.din) ? km : md.in,
                            points_to_chain_node(md.din) ? dist : mark_dist_as_point_to_terminal(dist), -1});
      }
      if (order_seed) {
        /* test knob: apply this round's updates in a shuffled order */
        uint64_t x = order_seed + iters * 0x9e3779b97f4a7c15ULL;
        for (size_t i = ups.size(); i > 1; --i) {
          x ^= x << 13; x ^= x >> 7; x ^= x << 17;
          std::swap(ups[i - 1], ups[x % i]);
        }
      }
      size_t count = apply_updates(ups);
      std::stable_partition(unfinished.begin(), unfinished.end(),
                            [&](long i) { return is_unfinished(chain.vals[i]); });
      while (!unfinished.empty() && !is_unfinished(chain.vals[unfinished.back()])) unfinished.pop_back();
      ++iters;
      uint32_t maxd = 1u << iters;
      size_t cyc = 0;
      for (long t : unfinished) cyc += is_cycle(chain.vals[t], maxd);
      if (iters >= 30) all_compacted = true;
      else all_compacted = (count == 0) || (cyc == unfinished.size());
    }
    iterations = iters;
    /* separate_cycles(iterations) when unfinished nodes remain (:1284-1291,
     * :670-705), then separate_isolated (:744-777). */
    bool any_unf = false;
    for (size_t i = 0; i < chain.keys.size(); ++i)
      if (chain.alive[i] && is_unfinished(chain.vals[i])) any_unf = true;
    n_cycle = 0;
    cycle_kmers.clear();
    if (any_unf) {
      uint32_t maxd = 1u << iters;
      for (size_t i = 0; i < chain.keys.size(); ++i)
        if (chain.alive[i] && is_cycle(chain.vals[i], maxd)) {
          chain.alive[i] = 0;
          cycle_kmers.push_back(chain.keys[i]);
          ++n_cycle;
        }
    }
    n_isolated = 0;
    for (size_t i = 0; i < chain.keys.size(); ++i)
      if (chain.alive[i] && points_to_self(chain.vals[i].din) && points_to_self(chain.vals[i].dout)) {
        chain.alive[i] = 0;
        ++n_isolated;
      }
  }

  /* ---------------- output ------------------------------------------- */
  /* to_listranked_chain_node (debruijn_chain_operations.hpp:261-310). */
  oracle_chain_node to_listranked(u128 key, const Biedge& b) const {
    u128 L = is_chain_terminal(b.din) ? key : b.in;
    u128 R = is_chain_terminal(b.dout) ? ops.rc(key) : ops.rc(b.out);
    bool sense = L <= R;
    u128 Kv = sense ? key : ops.rc(key);
    int32_t dist = (int32_t)(sense ? get_chain_dist(b.din) : get_chain_dist(b.dout));
    u128 rep = sense ? L : R;
    oracle_chain_node o;
    o.k_lo = (uint64_t)Kv;
    o.k_hi = (uint64_t)(Kv >> 64);
    o.rep_lo = (uint64_t)rep;
    o.rep_hi = (uint64_t)(rep >> 64);
    o.dist = dist;
    return o;
  }
  static u128 lo_hi(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

  void make_output() {
    /* print_branch_fasta (compact_dbg_io.hpp:254-299): sorted by k-mer,
     * print_graph_node_fasta (debruijn_graph_operations.hpp:81-90). */
    branch_fasta.clear();
    for (u128 b : branches) {
      std::string s = ops.to_string(b);
      branch_fasta += ">" + s + "\n" + s + "\n";
    }
    ranked.clear();
    for (size_t i = 0; i < chain.keys.size(); ++i)
      if (chain.alive[i]) ranked.push_back(to_listranked(chain.keys[i], chain.vals[i]));
    /* chain_rep_less (:333-340): by rep, then |dist|. */
    std::sort(ranked.begin(), ranked.end(), [](const oracle_chain_node& a, const oracle_chain_node& b) {
      u128 ra = lo_hi(a.rep_lo, a.rep_hi), rb = lo_hi(b.rep_lo, b.rep_hi);
      if (ra != rb) return ra < rb;
      return std::abs(a.dist) < std::abs(b.dist);
    });
    /* print_chain_as_fasta (:352-385) + print_chain_string
     * (compact_dbg_io.hpp:323-363): the leading newline is dropped. */
    static const char A[4] = {'A', 'C', 'G', 'T'};
    std::string s;
    for (auto& x : ranked) {
      u128 Kv = lo_hi(x.k_lo, x.k_hi);
      if (x.dist == 0) {
        u128 y = ops.rc(Kv);
        u128 canon = (Kv < y) ? Kv : y;
        s += "\n>" + ops.to_string(canon) + "\n" + ops.to_string(lo_hi(x.rep_lo, x.rep_hi));
      } else {
        s += A[(int)(Kv & 3)];
      }
    }
    chain_fasta = s.empty() ? s : s.substr(1);
    /* print_chain_nodes (compact_dbg_io.hpp:463-498): sorted by K,
     * print_chain_node (debruijn_chain_operations.hpp:427-442). */
    std::vector<oracle_chain_node> byk(ranked);
    std::sort(byk.begin(), byk.end(), [](const oracle_chain_node& a, const oracle_chain_node& b) {
      return lo_hi(a.k_lo, a.k_hi) < lo_hi(b.k_lo, b.k_hi);
    });
    chain_components.clear();
    for (auto& x : byk) {
      u128 Kv = lo_hi(x.k_lo, x.k_hi);
      u128 y = ops.rc(Kv);
      u128 canon = (Kv < y) ? Kv : y;
      chain_components += ops.to_string(canon) + "\t" + ops.to_string(lo_hi(x.rep_lo, x.rep_hi)) + "\t" +
                          std::to_string(x.dist) + "\t" + (Kv == canon ? "1" : "0") + "\n";
    }
  }

  u128 canon(u128 v) const {
    u128 r = ops.rc(v);
    return v < r ? v : r;
  }
  const Node* node_of(u128 k) const {
    auto it = std::lower_bound(nodes.begin(), nodes.end(), k, [](const Node& n, u128 v) { return n.kmer < v; });
    return (it != nodes.end() && it->kmer == k) ? &*it : nullptr;
  }

  /* The non-benchmark outputs of do_work
   * (compact_dbg_freq_construct_clean.cpp:528-543, :759-829).
   *
   * .branch.edges: print_branch_edge_frequencies (compact_dbg_stats.hpp
   * :219-262) prints the branch points sorted by k-mer with
   * print_graph_node (debruijn_graph_operations.hpp:58-78): k-mer, in-edge
   * counts A C G T, out-edge counts A C G T, self count, tab separated.
   *
   * .chain.edges: compute_freq_map (compact_dbg_stats.hpp:783-816) reduces
   * the self counts of every listranked chain node by chain rep (the map
   * canonicalises its keys, FreqMapParams = CanonicalHashMapParams,
   * compact_dbg_freq_construct_clean.cpp:240-253) into (nodes, sum, min,
   * max) with freq_summary (debruijn_chain_operations.hpp:447-459);
   * print_chain_frequencies (compact_dbg_stats.hpp:879-1065) takes the chain
   * map entries that pass IsCanonicalTerminusOrIsolated
   * (debruijn_chain_filters.hpp:125-147), orients L and R on one strand,
   * reads L's in-edge and R's out-edge counts (complemented and reversed
   * on the other strand), the mean as float(sum) / float(nodes) converted to
   * the 16-bit count type, sorts by L and prints L, R, 4 + 4 counts, mean,
   * min, max, each followed by a tab.  The counts come from the filtered
   * graph: the nodes erased before the print (cycles, edgeless nodes,
   * chain interiors, branch points; :606-610, :806-810) only lose edges
   * that point into a chain's interior or into a cycle, never the L in /
   * R out edges printed. */
  void make_edge_outputs() {
    branch_edges.clear();
    for (u128 b : branches) {
      const Node* n = node_of(b);  /* every branch is a node */
      std::string line = ops.to_string(b);
      for (int q = 4; q < 8; ++q) line += "\t" + std::to_string(n->c[q]);
      for (int q = 0; q < 4; ++q) line += "\t" + std::to_string(n->c[q]);
      line += "\t" + std::to_string(n->c[8]) + "\n";
      branch_edges += line;
    }
    struct Freq {
      uint64_t n = 0, sum = 0;
      uint16_t mn = 65535, mx = 0;
    };
    std::map<u128, Freq> freq;
    for (auto& x : ranked) {
      const Node* nk = node_of(canon(lo_hi(x.k_lo, x.k_hi)));
      const uint16_t self = nk ? nk->c[8] : 0;
      Freq& f = freq[canon(lo_hi(x.rep_lo, x.rep_hi))];
      f.n += 1;
      f.sum += self;
      f.mn = std::min(f.mn, self);
      f.mx = std::max(f.mx, self);
    }
    struct Line {
      u128 L, R;
      uint16_t v[11];
    };
    std::vector<Line> lines;
    chain_edge_misses = 0;
    for (size_t i = 0; i < chain.keys.size(); ++i) {
      if (!chain.alive[i]) continue;
      const u128 key = chain.keys[i];
      const Biedge& b = chain.vals[i];
      const bool lt = is_chain_terminal(b.din), rt = is_chain_terminal(b.dout);
      bool sel;
      if (lt && rt) sel = true;
      else if (lt) sel = key <= ops.rc(b.out);
      else if (rt) sel = ops.rc(key) < b.in;
      else sel = false;
      if (!sel) continue;
      Line l;
      if (lt && rt) {
        l.L = key;
        l.R = key;
      } else if (lt) {
        l.L = key;
        l.R = b.out;
      } else {
        l.L = ops.rc(key);
        l.R = ops.rc(b.in);
      }
      const u128 cL = canon(l.L), cR = canon(l.R);
      const Node* nl = node_of(cL);
      const Node* nr = node_of(cR);
      /* an end that is not in the graph (k-mer-only thresholds keep edges
       * into removed k-mers, and the ranking can carry such a k-mer as a
       * chain end): the reference asserts; the counts here read 0 */
      if (!nl || !nr) ++chain_edge_misses;
      for (int c = 0; c < 4; ++c) {
        l.v[c] = !nl ? 0 : cL == l.L ? nl->c[4 + c] : nl->c[3 - c];
        l.v[4 + c] = !nr ? 0 : cR == l.R ? nr->c[c] : nr->c[4 + 3 - c];
      }
      auto f = freq.find(cL);
      if (f == freq.end()) {
        ++chain_edge_misses;
        l.v[8] = l.v[9] = l.v[10] = 0;
      } else {
        const float mean = (float)f->second.sum / (float)f->second.n;
        l.v[8] = (uint16_t)(int32_t)mean;
        l.v[9] = f->second.mn;
        l.v[10] = f->second.mx;
      }
      lines.push_back(l);
    }
    std::sort(lines.begin(), lines.end(), [](const Line& a, const Line& b) { return a.L < b.L; });
    chain_edges.clear();
    for (auto& l : lines) {
      chain_edges += ops.to_string(l.L) + "\t" + ops.to_string(l.R) + "\t";
      for (int q = 0; q < 11; ++q) chain_edges += std::to_string(l.v[q]) + "\t";
      chain_edges += "\n";
    }
  }

  /* ---------------- graph cleaning ------------------------------------
   * clean_deadend_and_bubbles_naive (test/common/compact_dbg_clean.hpp
   * :560-850), run by do_work when cleaning is on
   * (compact_dbg_freq_construct_clean.cpp:633-701) after the cycle and
   * isolated removal (:606-610).  Each round summarises every chain from its
   * termini, picks deadends and bubbles, erases the edges that tie them to
   * their branch points when those edges are weaker than k1_lo + 1, drops
   * nodes left without edges, and compacts the graph again from scratch. */

  /* summarized_chain (debruijn_chain_node.hpp:165-174): branch before the
   * 5' end, 5' terminus (the chain representative's strand), 3' terminus,
   * branch after the 3' end, distance between the termini, and the total
   * frequency of the edges entering the 5' end / leaving the 3' end. */
  struct Summary {
    u128 b5, t5, t3, b3;
    uint32_t len;
    uint16_t f5, f6;
  };
  Summary rc_summary(const Summary& x) const {  // transform::reverse_complement (:259-268)
    return Summary{ops.rc(x.b3), ops.rc(x.t3), ops.rc(x.t5), ops.rc(x.b5), x.len, x.f6, x.f5};
  }
  Biedge rc_biedge(const Biedge& b) const { return Biedge{ops.rc(b.out), ops.rc(b.in), b.dout, b.din}; }

  static uint16_t in_total(const Counts& c) {  // get_in_edge_total_frequency (debruijn_graph_node.hpp:307-313)
    uint16_t t = 0;
    for (int i = 4; i < 8; ++i) t = sat_add16(t, c[i]);
    return t;
  }
  static uint16_t out_total(const Counts& c) {  // get_out_edge_total_frequency (:299-305)
    uint16_t t = 0;
    for (int i = 0; i < 4; ++i) t = sat_add16(t, c[i]);
    return t;
  }
  /* make_summarized_chain (debruijn_chain_node.hpp:201-240) */
  Summary summary_both(u128 ff, const Biedge& fs, u128 sf, const Biedge& ss) const {
    return Summary{fs.in, ff, sf, ss.out, get_chain_dist(fs.din) + get_chain_dist(fs.dout),
                   (uint16_t)(points_to_self(fs.din) ? 0 : 1), (uint16_t)(points_to_self(ss.dout) ? 0 : 1)};
  }
  Summary summary_5p(u128 ff, const Biedge& fs) const {
    return Summary{fs.in, ff, fs.out, 0, get_chain_dist(fs.din) + get_chain_dist(fs.dout),
                   (uint16_t)(points_to_self(fs.din) ? 0 : 1), 0};
  }
  Summary summary_3p(u128 sf, const Biedge& ss) const {
    return Summary{0, ss.in, sf, ss.out, get_chain_dist(ss.din) + get_chain_dist(ss.dout), 0,
                   (uint16_t)(points_to_self(ss.dout) ? 0 : 1)};
  }

  /* to_summarized_chains(dbg) (debruijn_chain_graph.hpp:2044-2180): one
   * half summary per chain terminus (IsTerminusOrIsolated, :418-420) with
   * the terminus node's edge frequency, then the halves of a chain (same
   * 5' terminus) merged by merge_summarized_chains (debruijn_chain_node.hpp
   * :177-198). */
  std::vector<Summary> summarize() const {
    std::vector<Summary> half;
    for (size_t i = 0; i < chain.keys.size(); ++i) {
      if (!chain.alive[i]) continue;
      const u128 k = chain.keys[i];
      const Biedge& b = chain.vals[i];
      const bool ti = is_chain_terminal(b.din), to = is_chain_terminal(b.dout);
      if (!ti && !to) continue;
      const Node* n = node_of(k);
      const uint16_t fin = n ? in_total(n->c) : 0, fout = n ? out_total(n->c) : 0;
      Summary s;
      if (ti && to) {
        s = summary_both(k, b, k, b);
        s.f5 = fin;
        s.f6 = fout;
      } else if (ti) {
        const u128 r = ops.rc(b.out);
        if (k < r) {
          s = summary_5p(k, b);
          s.f5 = fin;
        } else if (k > r) {
          s = summary_3p(ops.rc(k), rc_biedge(b));
          s.f6 = fin;
        } else {
          continue;  // the reference throws: a chain cannot end in its own reverse complement
        }
      } else {
        const u128 r = ops.rc(k);
        if (b.in < r) {
          s = summary_3p(k, b);
          s.f6 = fout;
        } else if (b.in > r) {
          s = summary_5p(r, rc_biedge(b));
          s.f5 = fout;
        } else {
          continue;
        }
      }
      half.push_back(s);
    }
    /* sorted by 5' terminus, then by the 3' branch; a total order here (the
     * reference's distributed sort leaves ties unordered, and merging is
     * symmetric in them) */
    std::sort(half.begin(), half.end(), [](const Summary& a, const Summary& b) {
      if (a.t5 != b.t5) return a.t5 < b.t5;
      if (a.b3 != b.b3) return a.b3 < b.b3;
      if (a.b5 != b.b5) return a.b5 < b.b5;
      if (a.f5 != b.f5) return a.f5 < b.f5;
      return a.f6 < b.f6;
    });
    std::vector<Summary> out;
    Summary last{};
    bool have_last = false;
    for (const Summary& x : half) {
      if (x.len == 0) {
        out.push_back(x);  // unit chain: both termini in one record
      } else if (have_last && x.t5 == last.t5) {
        Summary m = x;  // merge_summarized_chains(x, last)
        if (m.f5 < last.f5) {
          m.b5 = last.b5;
          m.f5 = last.f5;
        }
        if (m.f6 < last.f6) {
          m.b3 = last.b3;
          m.f6 = last.f6;
        }
        out.push_back(m);
      }
      last = x;
      have_last = true;
    }
    return out;
  }

  /* find_deadends (debruijn_topo_operations.hpp:152-177) with
   * deadend_filter (compact_dbg_clean.hpp:61-76): exactly one open end,
   * shorter than 2k - 1, and both end frequencies below min_freq. */
  std::vector<Summary> find_deadends(const std::vector<Summary>& all, uint64_t min_freq) const {
    std::vector<Summary> out;
    const uint32_t max_len = 2 * (uint32_t)K - 1;
    for (const Summary& x : all)
      if (x.len < max_len && x.f5 < min_freq && x.f6 < min_freq && ((x.f5 == 0) ^ (x.f6 == 0))) out.push_back(x);
    return out;
  }

  /* find_bubbles (debruijn_topo_operations.hpp:290-409) over
   * find_normal_chains (:256-283): chains with both ends on branch points,
   * put on the strand whose 5' branch is the smaller, sorted by (5' branch,
   * 3' branch, length); runs of chains between the same two branch points
   * where each neighbour pair differs in length by at most one and has an
   * end frequency below min_freq (bubble_filter, compact_dbg_clean.hpp
   * :79-101) are bubbles when the run holds two or more chains. */
  std::vector<Summary> find_bubbles(const std::vector<Summary>& all, uint64_t min_freq) const {
    std::vector<Summary> v;
    for (const Summary& x : all)
      if (x.len > 0 && x.f5 > 0 && x.f6 > 0) v.push_back(x.b5 > ops.rc(x.b3) ? rc_summary(x) : x);
    std::sort(v.begin(), v.end(), [](const Summary& a, const Summary& b) {
      if (a.b5 != b.b5) return a.b5 < b.b5;
      if (a.b3 != b.b3) return a.b3 < b.b3;
      if (a.len != b.len) return a.len < b.len;
      if (a.t5 != b.t5) return a.t5 < b.t5;
      return a.t3 < b.t3;
    });
    auto pair_ok = [&](const Summary& l, const Summary& r) {
      const uint32_t d = l.len > r.len ? l.len - r.len : r.len - l.len;
      return d <= 1 && (l.f5 < min_freq || l.f6 < min_freq || r.f5 < min_freq || r.f6 < min_freq);
    };
    std::vector<Summary> out;
    size_t first = 0;
    for (size_t i = 1; i <= v.size(); ++i) {
      if (i < v.size() && v[i - 1].b5 == v[i].b5 && v[i - 1].b3 == v[i].b3 && pair_ok(v[i - 1], v[i])) continue;
      if (i - first > 1)
        for (size_t q = first; q < i; ++q) out.push_back(v[q]);
      first = i;
    }
    return out;
  }

  /* the node record of k-mer x (any strand) and the counter of its edge on
   * side `out` (else in) with base c, on x's strand */
  uint16_t* edge_counter(u128 x, bool out, int c) {
    const u128 y = canon(x);
    auto it = std::lower_bound(nodes.begin(), nodes.end(), y, [](const Node& n, u128 v) { return n.kmer < v; });
    if (it == nodes.end() || it->kmer != y) return nullptr;
    if (y != x) {  // the edge on the other strand: sides swap, bases complement
      out = !out;
      c = 3 - c;
    }
    return &it->c[(out ? 0 : 4) + c];
  }
  /* erase_edges(query, edge_freq_filter) (debruijn_graph_map.hpp:769-773):
   * find_edges with the predicate (:616-666) keeps the queried edges whose
   * count is below min_freq (and not 0), and erase_edges_internal
   * (:737-747) clears them on both nodes (get_remote_edges :466-503); node
   * self counts are not touched. */
  size_t erase_weak_edges(const std::vector<std::pair<u128, std::pair<bool, int>>>& q, uint64_t min_freq) {
    std::vector<std::pair<u128, std::pair<bool, int>>> hit;
    for (auto& e : q) {
      const uint16_t* p = edge_counter(e.first, e.second.first, e.second.second);
      if (p && *p > 0 && *p < min_freq) hit.push_back(e);
    }
    for (auto& e : hit) {
      const u128 x = e.first;
      const bool out = e.second.first;
      const int c = e.second.second;
      if (uint16_t* p = edge_counter(x, out, c)) *p = 0;
      const u128 y = out ? ops.next_from_char(x, c) : ops.next_reverse_from_char(x, c);
      /* y's edge back to x: an in edge with x's first base, or an out edge
       * with x's last base */
      if (uint16_t* p = edge_counter(y, !out, out ? ops.base(x, 0) : ops.base(x, K - 1))) *p = 0;
    }
    return hit.size();
  }
  /* erase_if(IsIsolated) (debruijn_graph_filters.hpp:163-172) */
  void erase_isolated_nodes() {
    nodes.erase(std::remove_if(nodes.begin(), nodes.end(),
                               [](const Node& n) { return in_edge_count(n.c) == 0 && out_edge_count(n.c) == 0; }),
                nodes.end());
  }
  /* erase_nodes (debruijn_graph_map.hpp:822-856): the nodes, and every
   * neighbour's edge back to them */
  void erase_nodes(const std::vector<u128>& ks) {
    for (u128 k : ks) {
      const Node* n = node_of(k);
      if (!n) continue;
      const Counts c = n->c;
      for (int b = 0; b < 4; ++b) {
        if (c[b] > 0)
          if (uint16_t* p = edge_counter(ops.next_from_char(k, b), false, ops.base(k, 0))) *p = 0;
        if (c[4 + b] > 0)
          if (uint16_t* p = edge_counter(ops.next_reverse_from_char(k, b), true, ops.base(k, K - 1))) *p = 0;
      }
    }
    std::vector<u128> sorted(ks);
    std::sort(sorted.begin(), sorted.end());
    nodes.erase(std::remove_if(nodes.begin(), nodes.end(),
                               [&](const Node& n) { return std::binary_search(sorted.begin(), sorted.end(), n.kmer); }),
                nodes.end());
  }

  uint64_t clean_rounds = 0;
  uint64_t clean_deadends = 0, clean_bubbles = 0;

  void clean(int variant) {
    /* do_work: cycles and edgeless nodes leave the graph first */
    erase_nodes(cycle_kmers);
    erase_isolated_nodes();
    const uint64_t min_freq = th[2] + 1;  // edge_freq_filter(threshes[2] + 1)
    std::vector<Summary> all = summarize();
    std::vector<Summary> dead = find_deadends(all, min_freq), bub = find_bubbles(all, min_freq);
    clean_rounds = clean_deadends = clean_bubbles = 0;
    while (!dead.empty() || !bub.empty()) {
      clean_deadends += dead.size();
      clean_bubbles += bub.size();
      /* a deadend's edge to its branch point: the branch's out edge to the
       * 5' terminus (that terminus's last base) or its in edge from the 3'
       * terminus (that terminus's first base) */
      std::vector<std::pair<u128, std::pair<bool, int>>> q;
      for (const Summary& x : dead) {
        if (x.f5 > 0) q.push_back({x.b5, {true, ops.base(x.t5, K - 1)}});
        if (x.f6 > 0) q.push_back({x.b3, {false, ops.base(x.t3, 0)}});
      }
      size_t erased = erase_weak_edges(q, min_freq);
      erase_isolated_nodes();
      q.clear();
      for (const Summary& x : bub) {
        q.push_back({x.b5, {true, ops.base(x.t5, K - 1)}});
        q.push_back({x.b3, {false, ops.base(x.t3, 0)}});
      }
      erased += erase_weak_edges(q, min_freq);
      erase_isolated_nodes();
      /* a round that erased nothing would find the same chains again (the
       * reference loops for ever there: a branch-side edge count at or
       * above min_freq whose terminus-side count is below it) */
      if (erased == 0) break;
      extract_chains();
      list_rank(variant);
      erase_nodes(cycle_kmers);
      erase_isolated_nodes();
      all = summarize();
      dead = find_deadends(all, min_freq);
      bub = find_bubbles(all, min_freq);
      ++clean_rounds;
    }
    /* the branch points printed after cleaning are those of the cleaned
     * graph (print_branch_fasta(idx), :680-686) */
    branches.clear();
    for (auto& n : nodes)
      if (is_branch(n)) branches.push_back(n.kmer);
  }

  std::string clean_branch_fasta, clean_chain_fasta, clean_chain_components, clean_branch_edges, clean_chain_edges;
  std::vector<oracle_chain_node> clean_ranked;

  /* cleaning after run(): the .clean.* outputs of do_work; the first
   * compaction's outputs are kept */
  int run_clean(int variant) {
    const std::string bf = branch_fasta, cf = chain_fasta, cc = chain_components, be = branch_edges,
                      ce = chain_edges;
    const uint64_t it0 = iterations, cy0 = n_cycle, is0 = n_isolated;
    const std::vector<oracle_chain_node> rk0 = ranked;
    clean(variant);
    make_output();
    make_edge_outputs();
    clean_ranked.swap(ranked);
    ranked = rk0;
    iterations = it0;
    n_cycle = cy0;
    n_isolated = is0;
    clean_branch_fasta.swap(branch_fasta);
    clean_chain_fasta.swap(chain_fasta);
    clean_chain_components.swap(chain_components);
    clean_branch_edges.swap(branch_edges);
    clean_chain_edges.swap(chain_edges);
    branch_fasta = bf;
    chain_fasta = cf;
    chain_components = cc;
    branch_edges = be;
    chain_edges = ce;
    return 0;
  }

  int run(int variant) {
    double t0 = now_s();
    parse_all();
    phase[0] = now_s() - t0;
    count_and_filter();
    t0 = now_s();
    extract_chains();
    phase[3] = now_s() - t0;
    t0 = now_s();
    list_rank(variant);
    phase[4] = now_s() - t0;
    t0 = now_s();
    make_output();
    make_edge_outputs();
    phase[5] = now_s() - t0;
    return 0;
  }
};

/* ================================================================== */
extern "C" {

oracle_job* oracle_create(int k, const uint64_t* thresholds, int nthreads) {
  if (k < 1 || k > 63) return nullptr;
  return new oracle_job(k, thresholds, nthreads);
}
void oracle_destroy(oracle_job* j) { delete j; }
void oracle_set_edge_filter(oracle_job* j, int on) { j->edge_filter = on != 0; }

int oracle_add_reads(oracle_job* j, const char* bases, const uint64_t* offsets, uint64_t n_reads) {
  for (uint64_t r = 0; r < n_reads; ++r) {
    const char* s = bases + offsets[r];
    size_t L = (size_t)(offsets[r + 1] - offsets[r]);
    bool ok = true;
    for (size_t i = 0; i < L; ++i)
      if (dna_code(s[i]) < 0) {
        ok = false;
        break;
      }
    if (ok) j->reads.emplace_back(s, L);
  }
  return 0;
}

int oracle_run(oracle_job* j, int lr_variant) { return j->run(lr_variant); }
void oracle_set_order_seed(oracle_job* j, uint64_t seed) { j->order_seed = seed; }

uint64_t oracle_n_instances(oracle_job* j) { return j->records.size(); }
uint64_t oracle_n_k2(oracle_job* j) { return j->k2.size(); }
uint64_t oracle_n_nodes(oracle_job* j) { return j->nodes.size(); }
uint64_t oracle_n_branches(oracle_job* j) { return j->branches.size(); }
uint64_t oracle_n_chain_nodes(oracle_job* j) { return j->ranked.size(); }
uint64_t oracle_n_cycle_nodes(oracle_job* j) { return j->n_cycle; }
uint64_t oracle_n_isolated(oracle_job* j) { return j->n_isolated; }
uint64_t oracle_iterations(oracle_job* j) { return j->iterations; }

uint64_t oracle_get_records(oracle_job* j, uint64_t* lo, uint64_t* hi, uint8_t* e, uint64_t cap) {
  uint64_t n = j->records.size();
  if (!lo) return n;
  std::vector<Rec> v(j->records);
  std::sort(v.begin(), v.end(), [](const Rec& a, const Rec& b) {
    return a.kmer < b.kmer || (a.kmer == b.kmer && a.e < b.e);
  });
  for (uint64_t i = 0; i < n && i < cap; ++i) {
    lo[i] = (uint64_t)v[i].kmer;
    hi[i] = (uint64_t)(v[i].kmer >> 64);
    e[i] = v[i].e;
  }
  return n;
}
uint64_t oracle_get_k2(oracle_job* j, uint64_t* lo, uint64_t* hi, uint8_t* e, uint32_t* cnt, uint64_t cap) {
  uint64_t n = j->k2.size();
  if (!lo) return n;
  for (uint64_t i = 0; i < n && i < cap; ++i) {
    lo[i] = (uint64_t)j->k2[i].kmer;
    hi[i] = (uint64_t)(j->k2[i].kmer >> 64);
    e[i] = j->k2[i].e;
    cnt[i] = j->k2[i].cnt;
  }
  return n;
}
static uint64_t copy_nodes(const std::vector<Node>& v, oracle_node* out, uint64_t cap) {
  if (!out) return v.size();
  for (uint64_t i = 0; i < v.size() && i < cap; ++i) {
    out[i].kmer_lo = (uint64_t)v[i].kmer;
    out[i].kmer_hi = (uint64_t)(v[i].kmer >> 64);
    for (int c = 0; c < 9; ++c) out[i].counts[c] = v[i].c[c];
  }
  return v.size();
}
uint64_t oracle_get_unfiltered(oracle_job* j, oracle_node* out, uint64_t cap) { return copy_nodes(j->unfiltered, out, cap); }
uint64_t oracle_get_nodes(oracle_job* j, oracle_node* out, uint64_t cap) { return copy_nodes(j->nodes, out, cap); }
static uint64_t copy_str(const std::string& s, char* out, uint64_t cap) {
  if (!out) return s.size();
  memcpy(out, s.data(), std::min<uint64_t>(cap, s.size()));
  return s.size();
}
uint64_t oracle_get_branch_fasta(oracle_job* j, char* out, uint64_t cap) { return copy_str(j->branch_fasta, out, cap); }
uint64_t oracle_get_chain_fasta(oracle_job* j, char* out, uint64_t cap) { return copy_str(j->chain_fasta, out, cap); }
uint64_t oracle_get_branch_edges(oracle_job* j, char* out, uint64_t cap) { return copy_str(j->branch_edges, out, cap); }
uint64_t oracle_get_chain_edges(oracle_job* j, char* out, uint64_t cap) { return copy_str(j->chain_edges, out, cap); }
uint64_t oracle_chain_edge_misses(oracle_job* j) { return j->chain_edge_misses; }
uint64_t oracle_get_chain_components(oracle_job* j, char* out, uint64_t cap) {
  return copy_str(j->chain_components, out, cap);
}
uint64_t oracle_get_chain_nodes(oracle_job* j, oracle_chain_node* out, uint64_t cap) {
  if (!out) return j->ranked.size();
  for (uint64_t i = 0; i < j->ranked.size() && i < cap; ++i) out[i] = j->ranked[i];
  return j->ranked.size();
}
int oracle_run_clean(oracle_job* j, int lr_variant) { return j->run_clean(lr_variant); }
uint64_t oracle_clean_rounds(oracle_job* j) { return j->clean_rounds; }
uint64_t oracle_clean_deadends(oracle_job* j) { return j->clean_deadends; }
uint64_t oracle_clean_bubbles(oracle_job* j) { return j->clean_bubbles; }
uint64_t oracle_get_clean_output(oracle_job* j, int which, char* out, uint64_t cap) {
  const std::string* s[5] = {&j->clean_chain_fasta, &j->clean_branch_fasta, &j->clean_chain_components,
                             &j->clean_branch_edges, &j->clean_chain_edges};
  if (which < 0 || which > 4) return 0;
  return copy_str(*s[which], out, cap);
}
uint64_t oracle_n_clean_chain_nodes(oracle_job* j) { return j->clean_ranked.size(); }
uint64_t oracle_get_clean_chain_nodes(oracle_job* j, oracle_chain_node* out, uint64_t cap) {
  const uint64_t n = j->clean_ranked.size();
  if (out)
    for (uint64_t i = 0; i < n && i < cap; ++i) out[i] = j->clean_ranked[i];
  return n;
}
uint64_t oracle_get_clean_nodes(oracle_job* j, oracle_node* out, uint64_t cap) { return copy_nodes(j->nodes, out, cap); }

void oracle_phase_seconds(oracle_job* j, double* out6) {
  for (int i = 0; i < 6; ++i) out6[i] = j->phase[i];
}

}  // extern "C"
