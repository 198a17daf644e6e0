/*
 * debruijn_graph.hpp -- C++ host façade over the C ABI (include/bruno_hip.h).
 *
 * Keeps the class and method names of bruno's graph map, graph and chain graph
 * (src/debruijn/debruijn_graph_map.hpp, debruijn_graph.hpp,
 * debruijn_chain_graph.hpp) so that a driver written against bruno reads the
 * same, while every method body is one call into libbruno_hip.so.  Errors
 * surface as bruno_amd::error (the reference throws std::logic_error).
 */
#ifndef BRUNO_AMD_DEBRUIJN_GRAPH_HPP_
#define BRUNO_AMD_DEBRUIJN_GRAPH_HPP_

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bruno_hip.h"

namespace bruno_amd {

struct error : std::runtime_error {
  bruno_status status;
  error(bruno_status s, const std::string& what)
      : std::runtime_error(what + ": " + bruno_status_string(s)), status(s) {}
};

inline void check(bruno_status s, const char* what) {
  if (s != BRUNO_OK) throw error(s, what);
}

/* Reads in host memory: concatenated bases and n+1 offsets. */
struct read_batch {
  std::string bases;
  std::vector<uint64_t> offsets{0};
  void add(const std::string& s) {
    bases += s;
    offsets.push_back(bases.size());
  }
  uint64_t size() const { return offsets.size() - 1; }
};

/* A k-mer value as the ABI returns it (words least significant first). */
struct kmer_value {
  uint64_t w[2] = {0, 0};
  bool operator<(const kmer_value& o) const { return w[1] != o.w[1] ? w[1] < o.w[1] : w[0] < o.w[0]; }
  bool operator==(const kmer_value& o) const { return w[0] == o.w[0] && w[1] == o.w[1]; }
  std::string to_string(int k) const {
    static const char A[4] = {'A', 'C', 'G', 'T'};
    std::string s(k, 'A');
    for (int i = 0; i < k; ++i) {
      int sh = 2 * (k - 1 - i);
      uint64_t code = sh >= 64 ? (w[1] >> (sh - 64)) : (w[0] >> sh);
      s[i] = A[code & 3];
    }
    return s;
  }
};

/* One print_graph_node line (debruijn_graph_operations.hpp:58-78): the
 * k-mer, then in A C G T, out A C G T and self, tab separated. */
inline std::string node_line(const bruno_node& n, int k) {
  kmer_value km;
  km.w[0] = n.kmer[0];
  km.w[1] = n.kmer[1];
  std::string out = km.to_string(k);
  static const int order[9] = {4, 5, 6, 7, 0, 1, 2, 3, 8};
  for (int q : order) out += "\t" + std::to_string(n.counts[q]);
  out += "\n";
  return out;
}

/* Shared device job (one per graph + chain graph pair). */
class job {
 public:
  /* edge_filter: the block-construct driver's filter (canonical k+1-mer
   * counts in [threshes[2], threshes[3]), compact_debruijn_graph_block_construct.cpp
   * :415-480) in place of local_insert_by_freqencies */
  job(int k, const std::vector<size_t>* threshes, uint64_t capacity = 0, void* stream = nullptr,
      bool edge_filter = false)
      : k_(k) {
    bruno_job_desc d{};
    d.k = k;
    d.thresholding = threshes ? (edge_filter ? 2 : 1) : 0;
    for (int i = 0; i < 6; ++i) d.thresholds[i] = threshes ? (*threshes)[i] : 65536;
    d.capacity = capacity;
    d.load_factor = 0.5;
    d.stream = stream;
    bruno_job* j = nullptr;
    check(bruno_job_create(&d, &j), "bruno_job_create");
    h_.reset(j, [](bruno_job* p) { bruno_job_destroy(p); });
  }
  bruno_job* get() const { return h_.get(); }
  int k() const { return k_; }

 private:
  std::shared_ptr<bruno_job> h_;
  int k_;
};

/* debruijn_graph_map (debruijn_graph_map.hpp:52-1626) */
class debruijn_graph_map {
 public:
  explicit debruijn_graph_map(job& j) : j_(j) {}
  /* compute_biedge_freqencies (:1085-1141): parse + count one batch */
  size_t compute_biedge_freqencies(const read_batch& reads) {
    check(bruno_parse_count_host(j_.get(), reads.bases.data(), reads.offsets.data(), reads.size()),
          "compute_biedge_freqencies");
    return k2_size();
  }
  /* the same from a whole FASTQ file's bytes, split into reads on the device
   * (parse_file_data's reader, compact_dbg_io.hpp:165-217) */
  size_t compute_biedge_freqencies_fastq(const std::string& fastq_bytes) {
    check(bruno_parse_count_fastq_host(j_.get(), fastq_bytes.data(), fastq_bytes.size()),
          "compute_biedge_freqencies (FASTQ)");
    return k2_size();
  }
  /* insert (:871-921), the unthresholded build_index path */
  size_t insert(const read_batch& reads) { return compute_biedge_freqencies(reads); }
  /* the next batch is the last before local_insert_by_freqencies: when it
   * is the first count of a region table, the filter runs inside it
   * (bruno_job_set_fused_filter); no counterpart in the reference */
  void last_batch_next(bool on = true) { check(bruno_job_set_fused_filter(j_.get(), on ? 1 : 0), "fused filter"); }
  /* local_insert_by_freqencies (:1258-1623); thresholds come from the job */
  size_t local_insert_by_freqencies() {
    check(bruno_filter(j_.get()), "local_insert_by_freqencies");
    return local_size();
  }
  size_t k2_size() const { return stats().n_k2; }
  size_t local_size() const { return stats().n_nodes; }
  bruno_stats stats() const {
    bruno_stats s;
    check(bruno_get_stats(j_.get(), &s), "bruno_get_stats");
    return s;
  }

 private:
  job& j_;
};

/* debruijn_graph (debruijn_graph.hpp:46-295) */
class debruijn_graph {
 public:
  explicit debruijn_graph(job& j) : j_(j), map_(j) {}
  debruijn_graph_map& get_map() { return map_; }
  size_t local_size() const { return map_.local_size(); }
  size_t size() const { return local_size(); }
  /* nodes with their 9 counters (print_graph_node order) */
  std::vector<bruno_node> nodes() const {
    uint64_t n = 0;
    check(bruno_get_nodes(j_.get(), nullptr, &n), "bruno_get_nodes");
    std::vector<bruno_node> v(n);
    if (n) check(bruno_get_nodes(j_.get(), v.data(), &n), "bruno_get_nodes");
    return v;
  }
  job& get_job() { return j_; }
  /* the .graph.all bytes of print_graph_edge_frequencies -> print_graph_nodes
   * (compact_dbg_stats.hpp:266-303, :348-370): every node of the filtered
   * graph sorted by k-mer, print_graph_node (debruijn_graph_operations.hpp
   * :58-78) lines: k-mer, in A C G T, out A C G T, self */
  std::string graph_all() const {
    std::vector<bruno_node> v = nodes();
    std::sort(v.begin(), v.end(), [](const bruno_node& a, const bruno_node& b) {
      return a.kmer[1] != b.kmer[1] ? a.kmer[1] < b.kmer[1] : a.kmer[0] < b.kmer[0];
    });
    std::string out;
    for (auto& n : v) out += node_line(n, j_.k());
    return out;
  }

 private:
  job& j_;
  debruijn_graph_map map_;
};

/* listranked_chain_node (debruijn_chain_node.hpp:156) */
struct listranked_chain_node {
  kmer_value kmer, rep;
  int dist;
};

/* debruijn_chain_graph (debruijn_chain_graph.hpp:126-2284) */
class debruijn_chain_graph {
 public:
  explicit debruijn_chain_graph(job& j) : j_(j) {}
  /* extract_chains (:605-667) incl. setup_chain_termini (:486-596) */
  void extract_chains(debruijn_graph&) { check(bruno_build_chains(j_.get()), "extract_chains"); }
  /* list_rank_min_update (:1040-1310) */
  size_t list_rank_min_update() { return rank(1); }
  /* list_rank (:793-994) */
  size_t list_rank() { return rank(0); }
  /* to_ranked_chain_nodes (:1760-1773), sorted by chain_rep_less */
  std::vector<listranked_chain_node> to_ranked_chain_nodes() {
    compact();
    uint64_t n = 0;
    check(bruno_get_chain_nodes(j_.get(), nullptr, &n), "to_ranked_chain_nodes");
    std::vector<bruno_chain_node> v(n);
    if (n) check(bruno_get_chain_nodes(j_.get(), v.data(), &n), "to_ranked_chain_nodes");
    std::vector<listranked_chain_node> out(n);
    for (uint64_t i = 0; i < n; ++i) {
      out[i].kmer.w[0] = v[i].k[0];
      out[i].kmer.w[1] = v[i].k[1];
      out[i].rep.w[0] = v[i].rep[0];
      out[i].rep.w[1] = v[i].rep[1];
      out[i].dist = v[i].dist;
    }
    return out;
  }
  /* the .chain.fasta bytes of print_chain_string (compact_dbg_io.hpp:323-363) */
  std::string chain_fasta() {
    compact();
    return text(bruno_get_unitigs);
  }
  /* the .branch.fasta bytes of print_branch_fasta (compact_dbg_io.hpp:254-299) */
  std::string branch_fasta() {
    compact();
    return text(bruno_get_branch_fasta);
  }
  /* the .branch.edges bytes of print_branch_edge_frequencies
   * (compact_dbg_stats.hpp:219-262): branch points sorted by k-mer,
   * print_graph_node (debruijn_graph_operations.hpp:58-78) lines */
  std::string branch_edges() {
    compact();
    uint64_t n = 0;
    check(bruno_get_branch_nodes(j_.get(), nullptr, &n), "branch_edges");
    std::vector<bruno_node> v(n);
    if (n) check(bruno_get_branch_nodes(j_.get(), v.data(), &n), "branch_edges");
    std::string out;
    for (auto& b : v) out += node_line(b, j_.k());
    return out;
  }
  /* the .chain.edges bytes of compute_freq_map + print_chain_frequencies
   * (compact_dbg_stats.hpp:783-816, :879-1065): per unitig L, R, L's in-edge
   * counts, R's out-edge counts, mean, min, max, each followed by a tab */
  std::string chain_edges() {
    compact();
    uint64_t n = 0;
    check(bruno_get_chain_edges(j_.get(), nullptr, &n), "chain_edges");
    std::vector<bruno_chain_edge> v(n);
    if (n) check(bruno_get_chain_edges(j_.get(), v.data(), &n), "chain_edges");
    std::string out;
    for (auto& e : v) {
      kmer_value L, R;
      L.w[0] = e.left[0];
      L.w[1] = e.left[1];
      R.w[0] = e.right[0];
      R.w[1] = e.right[1];
      /* float division, then the conversion to the 16-bit count type */
      const float mean = (float)e.sum / (float)e.n_nodes;
      const uint16_t vals[11] = {e.in[0], e.in[1], e.in[2], e.in[3], e.out[0], e.out[1], e.out[2], e.out[3],
                                 (uint16_t)(int32_t)mean, e.min, e.max};
      out += L.to_string(j_.k()) + "\t" + R.to_string(j_.k()) + "\t";
      for (uint16_t x : vals) out += std::to_string(x) + "\t";
      out += "\n";
    }
    return out;
  }

 private:
  size_t rank(int variant) {
    uint32_t r = 0;
    check(bruno_list_rank(j_.get(), variant, &r), "list_rank");
    return r;
  }
  void compact() {
    if (!compacted_) check(bruno_compact(j_.get()), "bruno_compact");
    compacted_ = true;
  }
  std::string text(bruno_status (*fn)(bruno_job*, char*, uint64_t*)) {
    uint64_t n = 0;
    check(fn(j_.get(), nullptr, &n), "getter");
    std::string s(n, '\0');
    if (n) check(fn(j_.get(), &s[0], &n), "getter");
    return s;
  }
  job& j_;
  bool compacted_ = false;
  friend bruno_clean_stats clean_deadend_and_bubbles(debruijn_graph&, debruijn_chain_graph&,
                                                     const std::vector<size_t>&, bool);
};

/* clean_deadend_and_bubbles_naive (test/common/compact_dbg_clean.hpp
 * :560-850) with do_work's cycle / edgeless-node erasure before it
 * (compact_dbg_freq_construct_clean.cpp:606-610): afterwards idx and
 * chainmap hold the cleaned graph, whose outputs are the .clean.* files. */
inline bruno_clean_stats clean_deadend_and_bubbles(debruijn_graph&, debruijn_chain_graph& chainmap,
                                                   const std::vector<size_t>& threshes, bool lr_optimized) {
  chainmap.compact();
  bruno_clean_stats st{};
  check(bruno_clean(chainmap.j_.get(), lr_optimized ? 1 : 0, threshes[2], &st), "clean_deadend_and_bubbles");
  return st;
}

}  // namespace bruno_amd

#endif
