/*
 * bruno_oracle.h -- CPU restatement of bruno's de Bruijn hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: tests/, the smoke
 * check and bench.py's cpu_baseline leg call it as the checker.  No product
 * code path may link, load or call it.
 *
 * The restatement follows the reference source file by file (citations are
 * relative to the bruno source tree):
 *   src/debruijn/debruijn_biedge_loader.hpp     parse (k-mer, biedge) records
 *   src/debruijn/debruijn_common.hpp            lex_less canonicalisation
 *   src/debruijn/kmer_traits.hpp                palindrome predicates
 *   src/debruijn/debruijn_graph_node.hpp        compact_multi_biedge counters
 *   src/debruijn/debruijn_graph_map.hpp         k2 counting + threshold filter
 *   src/debruijn/debruijn_graph_filters.hpp     branch / chain predicates
 *   src/debruijn/debruijn_chain_node.hpp        chain node record, dist masks
 *   src/debruijn/debruijn_chain_operations.hpp  terminus/chain updates, output
 *   src/debruijn/debruijn_chain_filters.hpp     unfinished / cycle / isolated
 *   src/debruijn/debruijn_chain_graph.hpp       extract_chains + list ranking
 *   test/common/compact_dbg_io.hpp              writers (.branch/.chain.fasta)
 *   test/test/compact_dbg_freq_construct_clean.cpp  do_work call order
 *
 * All functions are plain C ABI so ctypes can drive them.  K-mers cross the
 * ABI as two uint64 words (lo, hi) of the 2-bit packed value with the first
 * base most significant; hi is 0 for k <= 31.
 */
#ifndef BRUNO_ORACLE_H_
#define BRUNO_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_job oracle_job;

/* One graph node as printed by print_graph_node: counts[0..3] = out A,C,G,T,
 * counts[4..7] = in A,C,G,T, counts[8] = self (debruijn_graph_node.hpp:156). */
typedef struct {
  uint64_t kmer_lo, kmer_hi;
  uint16_t counts[9];
} oracle_node;

/* One listranked chain node: (K on rep strand, rep, dist)
 * (debruijn_chain_node.hpp:156). */
typedef struct {
  uint64_t k_lo, k_hi;
  uint64_t rep_lo, rep_hi;
  int32_t dist;
} oracle_chain_node;

/* k in [1, 63].  thresholds: {k_lo,k_hi,k1_lo,k1_hi,k2_lo,k2_hi}; NULL
 * selects the unthresholded build_index path (compact_dbg_build.hpp:35-62).
 * nthreads: 1 = sequential restatement; >1 = hash-partitioned shards that
 * mirror MPI ranks for the parse/count/filter phases (same output). */
oracle_job* oracle_create(int k, const uint64_t* thresholds, int nthreads);
/* on: the thresholded filter is the k+1-mer edge-frequency filter of the
 * block-construct driver (lower = thresholds[2], upper = thresholds[3]) */
void oracle_set_edge_filter(oracle_job* j, int on);
void oracle_destroy(oracle_job* j);

/* Feed reads: concatenated bases and n_reads+1 offsets (ASCII). */
int oracle_add_reads(oracle_job* j, const char* bases, const uint64_t* offsets,
                     uint64_t n_reads);

/* Run the pipeline: count/filter, branches, chains, list ranking (variant
 * 0 = list_rank, 1 = list_rank_min_update), output.  Returns 0 on success. */
int oracle_run(oracle_job* j, int lr_variant);

/* Test knob: apply each list-ranking round's updates in a shuffled order
 * (seed != 0).  The reference applies them in hash-table / rank order, which
 * is unspecified; outputs must not depend on it. */
void oracle_set_order_seed(oracle_job* j, uint64_t seed);

/* Statistics. */
uint64_t oracle_n_instances(oracle_job* j);  /* parsed (k-mer, biedge) records */
uint64_t oracle_n_k2(oracle_job* j);         /* distinct canonical k2-mers     */
uint64_t oracle_n_nodes(oracle_job* j);      /* nodes in the count graph (P2)  */
uint64_t oracle_n_branches(oracle_job* j);
uint64_t oracle_n_chain_nodes(oracle_job* j); /* after cycle/isolated removal */
uint64_t oracle_n_cycle_nodes(oracle_job* j);
uint64_t oracle_n_isolated(oracle_job* j);
uint64_t oracle_iterations(oracle_job* j);

/* Two-call getters: pass buf=NULL to get the size in elements / bytes. */
uint64_t oracle_get_records(oracle_job* j, uint64_t* lo, uint64_t* hi, uint8_t* biedge, uint64_t cap);
uint64_t oracle_get_k2(oracle_job* j, uint64_t* lo, uint64_t* hi, uint8_t* biedge, uint32_t* cnt, uint64_t cap);
uint64_t oracle_get_unfiltered(oracle_job* j, oracle_node* out, uint64_t cap);
uint64_t oracle_get_nodes(oracle_job* j, oracle_node* out, uint64_t cap);
uint64_t oracle_get_branch_fasta(oracle_job* j, char* out, uint64_t cap);
uint64_t oracle_get_chain_fasta(oracle_job* j, char* out, uint64_t cap);
uint64_t oracle_get_chain_components(oracle_job* j, char* out, uint64_t cap);
/* .branch.edges and .chain.edges text (print_branch_edge_frequencies,
 * print_chain_frequencies); chain ends whose chain has no frequency summary
 * (none expected) */
uint64_t oracle_get_branch_edges(oracle_job* j, char* out, uint64_t cap);
uint64_t oracle_get_chain_edges(oracle_job* j, char* out, uint64_t cap);
uint64_t oracle_chain_edge_misses(oracle_job* j);
uint64_t oracle_get_chain_nodes(oracle_job* j, oracle_chain_node* out, uint64_t cap);

/* Graph cleaning after oracle_run (clean_deadend_and_bubbles_naive,
 * test/common/compact_dbg_clean.hpp:560-850, do_work with cleaning on,
 * compact_dbg_freq_construct_clean.cpp:606-701): erases the weak edges
 * that tie short deadends and bubbles to their branch points, round by
 * round, recompacting after each.  The first compaction's outputs stay as
 * they are; the cleaned graph's outputs (.clean.*) come from
 * oracle_get_clean_output: which = 0 chain FASTA, 1 branch FASTA, 2 chain
 * components, 3 branch edges, 4 chain edges.  The graph nodes afterwards
 * are the cleaned graph's (oracle_get_clean_nodes). */
int oracle_run_clean(oracle_job* j, int lr_variant);
uint64_t oracle_clean_rounds(oracle_job* j);
uint64_t oracle_clean_deadends(oracle_job* j);
uint64_t oracle_clean_bubbles(oracle_job* j);
uint64_t oracle_get_clean_output(oracle_job* j, int which, char* out, uint64_t cap);
uint64_t oracle_n_clean_chain_nodes(oracle_job* j);
uint64_t oracle_get_clean_chain_nodes(oracle_job* j, oracle_chain_node* out, uint64_t cap);
uint64_t oracle_get_clean_nodes(oracle_job* j, oracle_node* out, uint64_t cap);

/* Wall-clock seconds of the last oracle_run's phases:
 * [0]=parse [1]=count [2]=filter [3]=chains [4]=list rank [5]=output. */
void oracle_phase_seconds(oracle_job* j, double* out6);

#ifdef __cplusplus
}
#endif
#endif
