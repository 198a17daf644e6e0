/*
 * bruno_hip.h -- C ABI of the MI355X de Bruijn graph pipeline.
 *
 * Drop-in boundary for the hot path of bruno's compact_debruijn_graph_*_freq
 * driver (test/test/compact_dbg_freq_construct_clean.cpp:431-840).  Each entry
 * point replaces the reference call named beside it; host façades (C++
 * bruno_amd/host/debruijn_graph.hpp, Python bruno_amd) keep bruno's class and
 * method names and call these.
 *
 * Conventions
 *  - All entry points return a bruno_status; no exception crosses the ABI.
 *  - One job handle is not thread-safe: calls on one handle are serialised by
 *    the caller.  Distinct handles are independent and may use different
 *    streams from different host threads.
 *  - A job owns all of its device (HBM) state until bruno_job_destroy.  The
 *    state stays resident across calls, so a serving loop can keep many jobs
 *    on the device at once.
 *  - K-mers cross the ABI as `words_per_kmer` uint64 words per k-mer (1 for
 *    k <= 31, 2 for k <= 63), least significant word first.  The value is the
 *    2-bit packed k-mer (A0 C1 G2 T3) with the first base most significant,
 *    i.e. the value bruno's Kmer<K, DNA, uint64_t> holds, so numeric order is
 *    lexicographic order.
 *  - Two-call getters: pass buf == NULL to read the element (or byte) count
 *    into *n, then call again with a buffer of that size.
 */
#ifndef BRUNO_HIP_H_
#define BRUNO_HIP_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  BRUNO_OK = 0,
  BRUNO_BAD_INPUT = 1,  /* invalid argument (k, pointer, thresholds)        */
  BRUNO_OOM = 2,        /* device allocation failed                          */
  BRUNO_CAPACITY = 3,   /* hash table / side table full: recreate larger     */
  BRUNO_INTERNAL = 4,   /* device error or a bounded wait ran out            */
  BRUNO_STATE = 5,      /* call out of order for this job                    */
  BRUNO_NO_DEVICE = 6   /* no HIP device / HIP runtime error at set-up       */
} bruno_status;

typedef struct bruno_job bruno_job;

/* Job description (replaces the template parameters and TCLAP flags of the
 * reference driver, compact_dbg_freq_construct_clean.cpp:99-258, :873-968). */
typedef struct {
  int k;                    /* 1..63 (k <= 31 uses one 64-bit word)                 */
  int thresholding;         /* 0: build_index (compact_dbg_build.hpp:35-62)
                               1: build_index_thresholded (..._clean.cpp:297-348)
                               2: the block-construct driver's build_index_thresholded
                                  (compact_debruijn_graph_block_construct.cpp:415-480):
                                  an edge is kept when its canonical k+1-mer occurs
                                  thresholds[2] <= count < thresholds[3] times;
                                  other values: BRUNO_BAD_INPUT                    */
  uint64_t thresholds[6];   /* {k_lo,k_hi,k1_lo,k1_hi,k2_lo,k2_hi}, reference order
                               (debruijn_graph_map.hpp:1261); -L t = {0,65536,t,65536,0,65536} */
  uint64_t capacity;        /* hash slots (distinct k-mers / load factor);
                               0 = size from a sketch of the first batch's
                               distinct k-mers (bruno_estimate_distinct)           */
  double load_factor;       /* used when capacity == 0; 0 = default 0.5          */
  void* stream;             /* hipStream_t for all of this job's work; NULL = 0   */
} bruno_job_desc;

/* Phase timings (ms, device events on the job's stream). */
typedef struct {
  double parse_count_ms;    /* parse + k2 count                                  */
  double filter_ms;         /* local_insert_by_freqencies + node classification  */
  double chains_ms;         /* extract_chains + setup_chain_termini              */
  double list_rank_ms;      /* list ranking incl. cycle / isolated removal       */
  double output_ms;         /* to_ranked_chain_nodes + sort + unitig / branch text */
} bruno_phase_ms;

typedef struct {
  uint64_t n_reads;         /* reads fed                                         */
  uint64_t n_reads_dropped; /* reads with a non-ACGT base                        */
  uint64_t n_instances;     /* parsed (k-mer, biedge) records                    */
  uint64_t n_distinct;      /* distinct canonical k-mers in the count table      */
  uint64_t n_k2;            /* distinct canonical (k-mer, biedge) pairs          */
  uint64_t n_overflow;      /* table slots with k2 entries in the side table     */
  uint64_t n_nodes;         /* graph nodes after the filter                      */
  uint64_t n_branches;
  uint64_t n_chain_nodes;   /* chain nodes extracted                             */
  uint64_t n_cycle_nodes;
  uint64_t n_isolated;
  uint64_t n_unitig_nodes;  /* chain nodes printed as unitigs                    */
  uint64_t n_unitigs;
  uint64_t iterations;      /* list-ranking rounds                               */
  uint64_t n_conflicts;     /* same-round updates to one chain end in the
                               round-by-round list ranking (outputs do not depend
                               on their order)                                   */
  uint64_t capacity;        /* hash slots                                        */
  uint64_t device_bytes;    /* HBM held by the job now                           */
  uint64_t peak_device_bytes; /* most HBM held at once since create / reset      */
  uint64_t output_by_chains;  /* 1: unitig order placed from chain offsets after the
                                 contraction ranking; 0: full sort (same bytes)  */
  uint64_t count_by_regions;  /* 1: the last count built the table region by region
                                 from instance records; 0: atomic counting in HBM */
  uint64_t filter_fused;      /* 1: that count also ran the filter (bruno_job_set_fused_filter) */
} bruno_stats;

/* Listranked chain node (debruijn_chain_node.hpp:156): K on the rep strand,
 * rep, dist.  Words are least significant first; word [1] is 0 for k <= 31. */
typedef struct {
  uint64_t k[2];
  uint64_t rep[2];
  int32_t dist;
  int32_t pad;
} bruno_chain_node;

/* Graph node (print_graph_node, debruijn_graph_operations.hpp:58-78):
 * counts[0..3] out A,C,G,T, counts[4..7] in A,C,G,T, counts[8] self. */
typedef struct {
  uint64_t kmer[2];
  uint16_t counts[9];
  uint16_t pad[3];
} bruno_node;

/* One line of .chain.edges (print_chain_frequencies,
 * compact_dbg_stats.hpp:879-1065), one per unitig, in unitig order (sorted
 * by L).  The line prints L, R, in[0..3], out[0..3], the mean
 * (uint16_t)((float)sum / (float)n_nodes), min, max. */
typedef struct {
  uint64_t left[2];         /* L: the unitig's first k-mer on the strand it is printed on (its rep) */
  uint64_t right[2];        /* R: its last k-mer on that strand                                */
  uint64_t sum;             /* sum of the self counts of its k-mers (compute_freq_map, :783-816) */
  uint32_t n_nodes;         /* its k-mers                                                      */
  uint16_t in[4];           /* L's in-edge counts A,C,G,T on L's strand                        */
  uint16_t out[4];          /* R's out-edge counts A,C,G,T on R's strand                       */
  uint16_t min, max;        /* smallest / largest self count                                   */
} bruno_chain_edge;

const char* bruno_status_string(bruno_status s);

/* Replaces the construction of CountDBGType idx / ChainGraphType chainmap
 * (compact_dbg_freq_construct_clean.cpp:471-472). */
bruno_status bruno_job_create(const bruno_job_desc* desc, bruno_job** out);
bruno_status bruno_job_destroy(bruno_job* job);

/* Returns the job to its freshly created state while keeping the hash-table
 * allocation: the table is cleared on the job's stream and every phase
 * buffer is released.  Replaces clearing the maps between runs
 * (k2_counter.reset(), debruijn_graph_map.hpp:1270; chainmap.clear(),
 * compact_dbg_freq_construct_clean.cpp:826) so a serving loop reuses HBM. */
bruno_status bruno_job_reset(bruno_job* job);

/* Recycle mode for jobs run again and again (a serving loop, a benchmark):
 * as soon as bruno_build_chains has built the chains, the table is cleared
 * on a side stream, overlapping list ranking and output, and the next
 * bruno_job_reset waits for that instead of clearing.  In this mode the
 * filtered graph is gone after bruno_build_chains: bruno_get_nodes returns
 * BRUNO_STATE.  No counterpart in the reference. */
bruno_status bruno_job_set_recycle(bruno_job* job, int on);

/* Filter fused into the count.  With on != 0, a count that builds every
 * region of the job's table (its first bruno_parse_count, counted by
 * regions) also runs the filter: local_insert_by_freqencies
 * (debruijn_graph_map.hpp:1351-1609; the k+1-mer edge filter or none, as
 * the job's thresholding asks) and the branch / chain classification
 * (debruijn_graph_filters.hpp:142-220) for every slot it writes back, in
 * on-chip memory before the write-back, appending the nodes to the node
 * lists.  bruno_filter then only finishes what remains: slots with more
 * k2-mers than a slot holds, and any node the lists had no room for.
 * Results equal the unfused path's.  After such a count the job takes no
 * further batch (bruno_parse_count* returns BRUNO_STATE until
 * bruno_job_reset) and the k2 / count getters return BRUNO_STATE, since
 * node records replace the k2 entries.  Off by default; a count that falls
 * back to atomics (small tables, batches) is not fused.  No counterpart in
 * the reference, whose filter is a pass of its own over the count map. */
bruno_status bruno_job_set_fused_filter(bruno_job* job, int on);

/* Replaces parse_file_data<debruijn_kmer_simple_biedge_parser> +
 * compute_biedge_freqencies (compact_dbg_freq_construct_clean.cpp:318-328,
 * debruijn_graph_map.hpp:1085-1141), or the unthresholded idx.insert
 * (compact_dbg_build.hpp:45-54).  d_bases / d_offsets are DEVICE pointers:
 * concatenated ASCII reads and n_reads+1 byte offsets.  May be called
 * repeatedly to add batches (the incremental per-file loop, :313-330). */
bruno_status bruno_parse_count(bruno_job* job, const char* d_bases, const uint64_t* d_offsets,
                               uint64_t n_reads);
/* Same from host buffers (copies over PCIe first). */
bruno_status bruno_parse_count_host(bruno_job* job, const char* bases, const uint64_t* offsets,
                                    uint64_t n_reads);

/* Estimated number of distinct canonical k-mers in reads (device pointers),
 * from a HyperLogLog sketch of a 1-in-64 hash sample (one pass over the
 * reads, no table); with capacity != NULL also the hash slots a job needs
 * at load_factor (0 = 0.5), with a margin of four standard errors.  A job
 * created with capacity 0 sizes its table this way from its first batch.
 * No counterpart in the reference (its hash maps grow). */
bruno_status bruno_estimate_distinct(int k, const char* d_bases, const uint64_t* d_offsets, uint64_t n_reads,
                                     double* estimate, uint64_t* capacity, double load_factor, void* stream);

/* FASTQ ingest on the device.  Replaces the reader under parse_file_data
 * (compact_dbg_io.hpp:165-217; kmerind's FASTQ parser, one sequence per
 * record): d_fastq holds a FASTQ file's bytes (4-line records, '\n' or
 * "\r\n" line ends, last newline optional, blank lines at the end
 * ignored).  Writes the sequence lines back to back to d_out_bases and
 * n_reads + 1 offsets to d_out_offsets.  On entry *n_reads / *n_bases are
 * the capacities of the two outputs (CAPACITY if exceeded); pass both
 * outputs NULL to learn the sizes.  A malformed file is BAD_INPUT.  Reads
 * with a non-ACGT base stay in the output; bruno_parse_count drops them. */
bruno_status bruno_fastq_split(const char* d_fastq, uint64_t bytes, char* d_out_bases, uint64_t* d_out_offsets,
                               uint64_t* n_reads, uint64_t* n_bases, void* stream);
/* bruno_fastq_split without any device allocation: same arguments, results
 * and errors, but its scratch (per-chunk line counts, sequence-line
 * bounds, prefix-sum storage) is one fixed 80 MiB device array that the
 * library holds on every GPU it runs on until the process ends, instead of
 * blocks from the memory pool.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Limits: at most 1 GiB of FASTQ bytes and 2^21
 * records per call; larger inputs return BRUNO_CAPACITY.  Every call on a
 * GPU shares that GPU's scratch: calls that overlap on one GPU (from
 * different host threads or streams) race, so the caller serialises them.
 * The call returns after its work on `stream` has finished.  No
 * counterpart in the reference. */
bruno_status bruno_fastq_split_noalloc(const char* d_fastq, uint64_t bytes, char* d_out_bases,
                                       uint64_t* d_out_offsets, uint64_t* n_reads, uint64_t* n_bases, void* stream);
/* parse_file_data + compute_biedge_freqencies on a whole FASTQ file:
 * device bytes, or host bytes copied first. */
bruno_status bruno_parse_count_fastq(bruno_job* job, const char* d_fastq, uint64_t bytes);
bruno_status bruno_parse_count_fastq_host(bruno_job* job, const char* fastq, uint64_t bytes);

/* Replaces local_insert_by_freqencies(k2_counter, threshes)
 * (debruijn_graph_map.hpp:1258-1623); in the unthresholded mode it
 * materialises the count graph (local_insert :175-197).  Also classifies
 * nodes as branch (IsBranchPoint) or chain (IsChainNode)
 * (debruijn_graph_filters.hpp:142-220). */
bruno_status bruno_filter(bruno_job* job);

/* Replaces chainmap.extract_chains(idx) (debruijn_chain_graph.hpp:605-667),
 * including setup_chain_termini (:486-596). */
bruno_status bruno_build_chains(bruno_job* job);

/* Replaces list_rank() (variant 0, debruijn_chain_graph.hpp:793-994) and
 * list_rank_min_update() (variant 1, :1040-1310), including separate_cycles
 * and separate_isolated.  *rounds receives the iteration count. */
bruno_status bruno_list_rank(bruno_job* job, int variant, uint32_t* rounds);

/* Replaces to_ranked_chain_nodes (:1760-1773) + print_chain_string's sort
 * and print_chain_as_fasta (compact_dbg_io.hpp:323-363), and the sort of
 * print_branch_fasta (:254-299).  Leaves the .chain.fasta and .branch.fasta
 * bytes in device memory for the getters. */
bruno_status bruno_compact(bruno_job* job);

/* Convenience: filter + build_chains + list_rank + compact. */
bruno_status bruno_run_pipeline(bruno_job* job, int variant);

/* Graph cleaning after bruno_compact: replaces the cycle / edgeless-node
 * erasure of do_work (compact_dbg_freq_construct_clean.cpp:606-610) and
 * clean_deadend_and_bubbles_naive (test/common/compact_dbg_clean.hpp
 * :560-850) with find_deadends / find_bubbles / erase_edges
 * (debruijn_topo_operations.hpp:152-409, debruijn_graph_map.hpp:737-773).
 * Round by round it erases, on both nodes, the edges that tie short weak
 * deadends and weak bubbles to their branch points when their count is
 * below k1_lo + 1 (k1_lo: the driver's threshes[2], 65536 without -L),
 * drops nodes left without edges and compacts the graph again, until a
 * round finds nothing.  Afterwards the job holds the cleaned graph: the
 * getters return its unitigs, chain nodes, branch points and edge outputs
 * (the .clean.* files).  Plain jobs only (BRUNO_BAD_INPUT for a batch);
 * BRUNO_STATE before bruno_compact or with table recycling on. */
typedef struct {
  uint32_t rounds;         /* rounds that erased edges */
  uint64_t deadends;       /* deadend chains found, summed over rounds */
  uint64_t bubble_chains;  /* chains in bubbles found, summed over rounds */
  uint64_t edges_erased;   /* queried edges below the threshold, erased */
} bruno_clean_stats;
bruno_status bruno_clean(bruno_job* job, int variant, uint64_t k1_lo, bruno_clean_stats* out);

bruno_status bruno_get_stats(bruno_job* job, bruno_stats* out);
bruno_status bruno_get_phase_ms(bruno_job* job, bruno_phase_ms* out);

/* Getters (two-call).  Counts: valid after parse_count and before filter
 * (unthresholded node counters of the parsed k-mers; the P1 table). */
bruno_status bruno_get_counts(bruno_job* job, bruno_node* buf, uint64_t* n);
/* Distinct canonical k2-mers with uint32 counts (the k2 counter); valid
 * before filter.  kmers: words_per_kmer words each. */
bruno_status bruno_get_k2(bruno_job* job, uint64_t* kmers, uint8_t* biedges, uint32_t* counts,
                          uint64_t* n);
/* Filtered graph (after filter). */
bruno_status bruno_get_nodes(bruno_job* job, bruno_node* buf, uint64_t* n);
/* Branch k-mers, sorted (after compact). */
bruno_status bruno_get_branches(bruno_job* job, uint64_t* kmers, uint64_t* n);
/* Listranked chain nodes sorted by (rep, |dist|) (after compact). */
bruno_status bruno_get_chain_nodes(bruno_job* job, bruno_chain_node* buf, uint64_t* n);
/* .chain.fasta / .branch.fasta bytes (after compact). */
bruno_status bruno_get_unitigs(bruno_job* job, char* buf, uint64_t* n);
bruno_status bruno_get_branch_fasta(bruno_job* job, char* buf, uint64_t* n);
/* Non-benchmark outputs (after compact; not after a compact that recycles
 * the table, which returns BRUNO_STATE).  Branch nodes sorted by k-mer with
 * their counts: the .branch.edges lines (print_branch_edge_frequencies,
 * compact_dbg_stats.hpp:219-262, print_graph_node order: in A C G T, out
 * A C G T, self).  Per-unitig end edges and self-count summary: the
 * .chain.edges lines (compute_freq_map + print_chain_frequencies,
 * compact_dbg_freq_construct_clean.cpp:759-829). */
bruno_status bruno_get_branch_nodes(bruno_job* job, bruno_node* buf, uint64_t* n);
bruno_status bruno_get_chain_edges(bruno_job* job, bruno_chain_edge* buf, uint64_t* n);

/* ---- Batches of independent jobs ---------------------------------------
 * A batch runs n_jobs independent assembly jobs (same k and thresholds,
 * different read sets of any size and read length) through every phase
 * with one set of launches: the jobs share one HBM table in which job i
 * owns capacities[i] slots, so k-mers of different jobs never meet, and
 * each job's results are bit-identical to running it alone (including its
 * own list-ranking round count).  The phase calls are the ones above
 * (bruno_filter, bruno_build_chains, bruno_list_rank, bruno_compact,
 * bruno_run_pipeline, bruno_job_reset, bruno_job_destroy); the whole-job
 * getters return BRUNO_BAD_INPUT on a batch of more than one job, and the
 * per-job getters below take the job index.  A batch of one job is a
 * plain job (its table may take the region layout and count by regions,
 * as bruno_job_create's).  The reference runs one assembly per MPI job
 * (compact_dbg_freq_construct_clean.cpp:431-840); it has no batched
 * counterpart. */
#define BRUNO_MAX_BATCH_JOBS 65536u

bruno_status bruno_batch_create(const bruno_job_desc* desc, uint32_t n_jobs, const uint64_t* capacities,
                                bruno_job** out);
/* d_bases / d_offsets as for bruno_parse_count; job_first_read is a HOST
 * array of n_jobs + 1 read indices: job i owns reads
 * [job_first_read[i], job_first_read[i+1]) of this call (first entry 0,
 * last n_reads, non-decreasing).  May be called repeatedly. */
bruno_status bruno_batch_parse_count(bruno_job* job, const char* d_bases, const uint64_t* d_offsets,
                                     uint64_t n_reads, const uint64_t* job_first_read);
/* Per-job statistics: n_reads, n_instances, n_reads_dropped, capacity, and
 * after the phases that produce them iterations, n_cycle_nodes, n_isolated,
 * n_branches, n_chain_nodes, n_nodes, n_unitig_nodes, n_unitigs.  The other
 * fields are batch-wide only (bruno_get_stats) and read 0 here. */
bruno_status bruno_batch_get_stats(bruno_job* job, uint32_t idx, bruno_stats* out);
/* Per-job forms of the getters above, same contents and order.  For
 * counts, k2 and nodes a NULL buffer makes the call count job idx's entries
 * with a pass over its table segment. */
bruno_status bruno_batch_get_counts(bruno_job* job, uint32_t idx, bruno_node* buf, uint64_t* n);
bruno_status bruno_batch_get_k2(bruno_job* job, uint32_t idx, uint64_t* kmers, uint8_t* biedges, uint32_t* counts,
                                uint64_t* n);
bruno_status bruno_batch_get_nodes(bruno_job* job, uint32_t idx, bruno_node* buf, uint64_t* n);
bruno_status bruno_batch_get_branches(bruno_job* job, uint32_t idx, uint64_t* kmers, uint64_t* n);
bruno_status bruno_batch_get_chain_nodes(bruno_job* job, uint32_t idx, bruno_chain_node* buf, uint64_t* n);
bruno_status bruno_batch_get_unitigs(bruno_job* job, uint32_t idx, char* buf, uint64_t* n);
bruno_status bruno_batch_get_branch_fasta(bruno_job* job, uint32_t idx, char* buf, uint64_t* n);
bruno_status bruno_batch_get_branch_nodes(bruno_job* job, uint32_t idx, bruno_node* buf, uint64_t* n);
bruno_status bruno_batch_get_chain_edges(bruno_job* job, uint32_t idx, bruno_chain_edge* buf, uint64_t* n);

/* Debug entry for the P0 parity object: parse only, write each canonical
 * (k-mer, biedge) record in read order (device pointers; out_kmers holds
 * words_per_kmer words per record).  *n_out receives the record count;
 * pass out_kmers == NULL to count only. */
bruno_status bruno_parse_records(int k, const char* d_bases, const uint64_t* d_offsets,
                                 uint64_t n_reads, uint64_t* d_out_kmers, uint8_t* d_out_biedges,
                                 uint64_t* n_out, void* stream);

/* ---- Device memory helpers ----------------------------------------------
 * For callers that hold no HIP runtime of their own (Python tests, the
 * benchmark): allocate, copy and synchronise through this library's runtime.
 * No counterpart in the reference (its buffers are host vectors). */
bruno_status bruno_dev_alloc(uint64_t bytes, void** out);
bruno_status bruno_dev_free(void* p);
/* kind: 0 host->device, 1 device->host, 2 device->device */
bruno_status bruno_dev_copy(void* dst, const void* src, uint64_t bytes, int kind);
bruno_status bruno_dev_sync(void);
/* free / total HBM bytes of the current device (memory held by the pool
 * below counts as used) */
bruno_status bruno_dev_mem_info(uint64_t* free_bytes, uint64_t* total_bytes);

/* ---- Tuning knobs ---------------------------------------------------------
 * Process-wide launch parameters, picked by the autotuner
 * (bruno_amd/autotune.py) on the box that runs the work.  They change speed
 * only, never results.  No counterpart in the reference. */
typedef struct {
  uint32_t parse_block;    /* workgroup size of the parse + count kernel: 128, 256 or 512    */
  uint32_t parse_grid;     /* most waves of the parse + count launch (one read per wave)     */
  uint32_t scan_grid;      /* most workgroups of the filter's table scan                     */
  uint32_t ruler_sample;   /* list ranking: one sampled ruler per this many chain nodes on
                              average (power of two, 8..4096): the walk length between
                              rulers and the size of the contracted list                    */
  uint32_t count_regions;  /* 1 or 2: plain jobs (k <= 63) with tables of >= 2^22 slots lay
                              the table out in LDS-sized regions and count by writing
                              instance records grouped by region, then building each region
                              on chip; 0: one probe window and atomic counting in HBM.  Read
                              when a table is allocated and at each count.                   */
  uint32_t region_slots;   /* most slots of a region when a region table is allocated:
                              2560 (80 KB of LDS per build workgroup) or 1280 (40 KB)       */
  uint32_t fuse_list_cap;  /* a count with the filter fused gives each node list room for at
                              most this many nodes (0: every node the table can hold); the
                              nodes beyond it are left to bruno_filter                      */
  uint32_t count_groups;   /* a region count's second split pass and region builds run in
                              this many groups of sub-buckets, a group's builds beside the
                              next group's split (1: one after the other; at most 64);
                              0: 16 groups from 8 sub-buckets per CU (the largest tables),
                              else 1                                                         */
} bruno_tuning;
bruno_status bruno_set_tuning(const bruno_tuning* t);
bruno_status bruno_get_tuning(bruno_tuning* t);

/* ---- Device memory pool --------------------------------------------------
 * Every device allocation of the library (job tables, phase buffers,
 * bruno_dev_alloc) is a block of a per-device pool, so a steady-state run
 * makes no hipMalloc / hipFree calls (each costs tens of ms per GB).  The
 * pool grows by chunks on demand; a server or benchmark reserves one large
 * chunk up front.  Memory returns to HIP only through bruno_pool_trim (fully
 * free chunks).  No counterpart in the reference (host containers). */
bruno_status bruno_pool_reserve(uint64_t bytes);
bruno_status bruno_pool_trim(void);
bruno_status bruno_pool_stats(uint64_t* reserved, uint64_t* in_use, uint64_t* largest_free);

/* ---- Synthetic workload generator (the benchmark configurations) ------ */
typedef struct {
  uint64_t seed;
  uint64_t genome_len;       /* bases                                          */
  uint32_t n_families;       /* repeat families (0 = none)                     */
  uint32_t family_min_len, family_max_len;
  double repeat_fraction;    /* fraction of the genome covered by copies       */
  double family_divergence;  /* substitution rate between copies               */
  uint64_t n_reads;
  uint32_t read_len;
  double sub_rate;           /* per-base substitution error                    */
  double indel_rate;         /* per-base 1-bp insertion/deletion rate          */
} bruno_synth_desc;

/* Generate reads directly in device memory (d_bases: n_reads*read_len bytes,
 * d_offsets: n_reads+1). */
bruno_status bruno_synth_reads_device(const bruno_synth_desc* d, char* d_bases, uint64_t* d_offsets,
                                      void* stream);
/* Same on the host (bit-identical to the device version). */
bruno_status bruno_synth_reads_host(const bruno_synth_desc* d, char* bases, uint64_t* offsets);
/* n_jobs workloads back to back (a batch's reads): job i's reads follow job
 * i-1's, offsets are global (n_reads_total + 1 entries). */
bruno_status bruno_synth_batch_device(const bruno_synth_desc* descs, uint32_t n_jobs, char* d_bases,
                                      uint64_t* d_offsets, void* stream);
bruno_status bruno_synth_batch_host(const bruno_synth_desc* descs, uint32_t n_jobs, char* bases, uint64_t* offsets);

/* ---- Request loop (serving) ---------------------------------------------
 * A server admits independent assembly jobs as they arrive and runs them
 * in batches on its worker streams: the oldest waiting request and every
 * later waiting one with the same k and thresholds whose estimated HBM
 * footprint fits the free budget form the next batch, so requests that
 * arrive while others are on the device join the next batch and batches
 * overlap.  A job with a table of 2^26 slots or more runs on its own
 * (counted by regions).  A job whose table proves too small runs again on its own with
 * four times the slots, and then with a table that cannot fill (two slots
 * per base).  Each job's result equals what
 * bruno_run_pipeline gives for it alone.  No counterpart in the reference
 * (one assembly per MPI job); this serves its per-job outputs
 * (.chain.fasta / .branch.fasta, compact_dbg_io.hpp:254-363). */
typedef struct bruno_server bruno_server;

typedef struct {
  uint64_t hbm_budget;      /* HBM the server's batches may hold at once; 0 = 85% of free HBM */
  uint32_t n_workers;       /* batches in flight, one stream each; 0 = 2                      */
  uint32_t max_batch_jobs;  /* most jobs per batch; 0 = 64                                    */
  int variant;              /* list ranking: 1 = list_rank_min_update, 0 = list_rank          */
  int reserve_pool;         /* nonzero: reserve the budget in the memory pool at create       */
} bruno_server_desc;

typedef struct {
  bruno_job_desc job;       /* k, thresholds; capacity = hash slots (0: one slot per base at
                               load 1/2); stream ignored                                      */
  const char* bases;        /* host reads, copied by bruno_server_submit ...                  */
  const uint64_t* offsets;  /* ... n_reads + 1 byte offsets                                    */
  uint64_t n_reads;
  const bruno_synth_desc* synth; /* or, instead of reads: generate this workload on the device */
} bruno_request;

typedef struct {
  int32_t status;           /* bruno_status of the job                                        */
  uint32_t batch_jobs;      /* jobs in the batch that produced the result                     */
  uint64_t batch_id;
  uint64_t n_reads, n_instances, n_unitigs, n_unitig_nodes, n_branches;
  uint64_t unitig_bytes, branch_bytes;
  uint32_t iterations;      /* list-ranking rounds of this job                                */
  uint32_t attempts;
  double submit_s, start_s, end_s; /* seconds since bruno_server_create                       */
} bruno_result_info;

typedef struct {
  uint64_t submitted, jobs_done, jobs_failed, batches, retries, max_batch_jobs;
  uint64_t queued, reserved_now, peak_reserved, hbm_budget;
} bruno_server_stats;

bruno_status bruno_server_create(const bruno_server_desc* desc, bruno_server** out);
bruno_status bruno_server_submit(bruno_server* s, const bruno_request* req, uint64_t* ticket);
bruno_status bruno_server_poll(bruno_server* s, uint64_t ticket, int* done);
/* blocks until the job has finished (successfully or not) */
bruno_status bruno_server_wait(bruno_server* s, uint64_t ticket, bruno_result_info* info);
/* two-call getters of a finished job's .chain.fasta / .branch.fasta bytes */
bruno_status bruno_server_get_unitigs(bruno_server* s, uint64_t ticket, char* buf, uint64_t* n);
bruno_status bruno_server_get_branch_fasta(bruno_server* s, uint64_t ticket, char* buf, uint64_t* n);
/* forget a finished job's result */
bruno_status bruno_server_release(bruno_server* s, uint64_t ticket);
bruno_status bruno_server_get_stats(bruno_server* s, bruno_server_stats* out);
/* finishes every submitted job, then stops the workers */
bruno_status bruno_server_destroy(bruno_server* s);

#ifdef __cplusplus
}
#endif
#endif
