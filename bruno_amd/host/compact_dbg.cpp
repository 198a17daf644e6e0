/*
 * compact_dbg.cpp -- host driver with the reference driver's call order and
 * flags (test/test/compact_dbg_freq_construct_clean.cpp:431-840, :848-1030):
 *
 *   compact_dbg [-T] [-E] [-L t | -L k k1 k2] [-U ...] [-R] [-B] [--clean] [-K k] -O prefix reads.fastq...
 *
 * It mirrors the benchmarked binary, which is built with cleaning compiled
 * out (-DNCLEAN; its -N switch then does nothing and is accepted here the
 * same way).  --clean runs the cleaning that binary leaves out
 * (compact_dbg_freq_construct_clean.cpp:633-701): after the files above it
 * writes <prefix>.clean.branch.fasta and <prefix>.clean.chain.fasta, and
 * unless -B <prefix>.clean.chain.components, .clean.branch.edges and
 * .clean.chain.edges, from the graph cleaned of short weak deadends and
 * weak bubbles.
 *
 * -E filters as the block-construct driver does
 * (compact_debruijn_graph_block_construct.cpp:415-480, :686-712): with -T
 * -L lower -U upper an edge is kept when its canonical k+1-mer occurs
 * lower <= count < upper times over the reads.
 *
 * Writes <prefix>.branch.fasta and <prefix>.chain.fasta, and unless -B also
 * <prefix>.graph.all (print_graph_edge_frequencies, compact_dbg_stats.hpp
 * :348-370; the reference writes it in debug builds, :509-526),
 * <prefix>.chain.components (print_chain_nodes, compact_dbg_io.hpp:463-498),
 * <prefix>.branch.edges (print_branch_edge_frequencies,
 * compact_dbg_stats.hpp:219-262) and <prefix>.chain.edges (compute_freq_map +
 * print_chain_frequencies, :783-1065).
 * Each file's bytes go to the device, where the FASTQ records are split into
 * reads; reads with a non-ACGT base are dropped (NFilterSequencesIterator).
 */
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <iostream>
#include <string>
#include <vector>

#include "debruijn_graph.hpp"

using namespace bruno_amd;

static bool read_file(const std::string& fn, std::string& out) {
  std::ifstream f(fn, std::ios::binary);
  if (!f) return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

static void write_file(const std::string& fn, const std::string& data) {
  std::ofstream f(fn, std::ios::binary);
  f.write(data.data(), (std::streamsize)data.size());
}

static std::string canon_string(const std::string& s) {
  std::string r(s.rbegin(), s.rend());
  for (auto& c : r) c = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
  return std::min(s, r);
}

int main(int argc, char** argv) {
  std::vector<size_t> threshes(6, 65536);
  bool thresholding = false, benchmark = false, lr_opt = false, edge_filter = false, clean = false;
  int k = 31;
  std::string prefix = "./output";
  std::vector<std::string> files;
  std::vector<size_t> lower, upper;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto vals = [&](std::vector<size_t>& v) {
      while (i + 1 < argc && argv[i + 1][0] != '-') v.push_back(strtoull(argv[++i], nullptr, 10));
    };
    if (a == "-T") thresholding = true;
    else if (a == "-B") benchmark = true;
    else if (a == "-E") edge_filter = true;
    else if (a == "-R") lr_opt = true;
    else if (a == "--clean") clean = true;
    else if (a == "-N") continue;
    else if (a == "-L") vals(lower);
    else if (a == "-U") vals(upper);
    else if (a == "-K" && i + 1 < argc) k = atoi(argv[++i]);
    else if (a == "-O" && i + 1 < argc) prefix = argv[++i];
    else if (a == "-h" || a == "--help") {
      std::cout << "usage: compact_dbg [-T] [-E] [-L t|-L k k1 k2] [-U ...] [-R] [-B] [--clean] [-K k] -O prefix "
                   "reads.fastq...\n";
      return 0;
    } else files.push_back(a);
  }
  /* the threshes[6] mapping of the reference (:940-968) */
  if (lower.size() == 1) {
    threshes[2] = std::min(lower[0], threshes[2]);
    threshes[0] = threshes[4] = 0;
  } else if (lower.size() == 3) {
    threshes[0] = std::min(lower[0], threshes[0]);
    threshes[2] = std::min(lower[1], threshes[2]);
    threshes[4] = std::min(lower[2], threshes[4]);
  }
  if (upper.size() == 1) {
    threshes[3] = std::min(upper[0], threshes[3]);
  } else if (upper.size() == 3) {
    threshes[1] = std::min(upper[0], threshes[1]);
    threshes[3] = std::min(upper[1], threshes[3]);
    threshes[5] = std::min(upper[2], threshes[5]);
  }
  if (files.empty()) {
    std::cerr << "no input files\n";
    return 1;
  }
  try {
    job j(k, thresholding ? &threshes : nullptr, 0, nullptr, edge_filter);
    debruijn_graph idx(j);
    debruijn_chain_graph chainmap(j);
    /* build_index_thresholded / build_index: one batch per file (:313-330) */
    for (auto& fn : files) {
      std::string bytes;
      if (!read_file(fn, bytes)) {
        std::cerr << "cannot read " << fn << "\n";
        return 1;
      }
      /* the FASTQ records are split into reads on the device */
      idx.get_map().compute_biedge_freqencies_fastq(bytes);
    }
    idx.get_map().local_insert_by_freqencies();
    /* the filtered graph, before chain extraction (:509-526) */
    if (!benchmark) write_file(prefix + ".graph.all", idx.graph_all());
    chainmap.extract_chains(idx);
    size_t iters = lr_opt ? chainmap.list_rank_min_update() : chainmap.list_rank();
    auto outputs = [&](const std::string& p) {
      write_file(p + ".branch.fasta", chainmap.branch_fasta());
      write_file(p + ".chain.fasta", chainmap.chain_fasta());
      if (!benchmark) {
        /* print_chain_nodes: sorted by K, "canon(K)\trep\tdist\t(K==canon)" */
        auto nodes = chainmap.to_ranked_chain_nodes();
        std::sort(nodes.begin(), nodes.end(),
                  [](const listranked_chain_node& a, const listranked_chain_node& b) { return a.kmer < b.kmer; });
        std::string out;
        for (auto& n : nodes) {
          std::string ks = n.kmer.to_string(k), c = canon_string(ks);
          out += c + "\t" + n.rep.to_string(k) + "\t" + std::to_string(n.dist) + "\t" + (ks == c ? "1" : "0") + "\n";
        }
        write_file(p + ".chain.components", out);
        write_file(p + ".branch.edges", chainmap.branch_edges());
        write_file(p + ".chain.edges", chainmap.chain_edges());
      }
    };
    outputs(prefix);
    if (clean) {
      const bruno_clean_stats cs = clean_deadend_and_bubbles(idx, chainmap, threshes, lr_opt);
      std::printf("cleaning: %u rounds, %llu deadends, %llu bubble chains, %llu edges erased\n", cs.rounds,
                  (unsigned long long)cs.deadends, (unsigned long long)cs.bubble_chains,
                  (unsigned long long)cs.edges_erased);
      outputs(prefix + ".clean");
    }
    bruno_stats st;
    check(bruno_get_stats(j.get(), &st), "stats");
    std::printf("k-mer instances %llu, graph nodes %llu, branches %llu, chain nodes %llu, unitigs %llu, rounds %zu\n",
                (unsigned long long)st.n_instances, (unsigned long long)st.n_nodes,
                (unsigned long long)st.n_branches, (unsigned long long)st.n_unitig_nodes,
                (unsigned long long)st.n_unitigs, iters);
  } catch (const error& e) {
    std::cerr << e.what() << "\n";
    return 2;
  }
  return 0;
}
