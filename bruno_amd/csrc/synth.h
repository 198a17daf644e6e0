/*
 * synth.h -- seeded synthetic genomes and reads (benchmark workloads).
 *
 * Every base is a pure function of (seed, position) or (seed, read, index),
 * built from a counter-based generator (fmix64 of a stream/counter mix), so
 * the device kernel and the host loop produce bit-identical reads in any
 * order and with any parallelism.
 *
 * Genome: i.i.d. uniform ACGT, with an optional repeat layer: n_copies copies
 * of n_families random templates (lengths uniform in [min,max]) are placed in
 * disjoint genome slots until repeat_fraction of the genome is covered; each
 * copy differs from its template by independent substitutions at
 * family_divergence.
 *
 * Reads: start uniform in [0, G-L], strand 50/50 (reverse complement), each
 * base substituted with probability sub_rate (uniformly to one of the three
 * other bases); with indel_rate > 0 a 1-bp insertion or deletion occurs at
 * each read position with probability indel_rate.
 */
#ifndef BRUNO_SYNTH_H_
#define BRUNO_SYNTH_H_

#include <stdint.h>

#include "kmer.h"

namespace br {

struct synth_params {
  uint64_t seed;
  uint64_t G;
  uint32_t n_families, fmin, fmax;
  uint64_t n_copies;      /* repeat copies placed                          */
  uint64_t spacing;       /* genome slot per copy (>= fmax)                */
  uint64_t div_thresh;    /* family divergence as a 64-bit threshold       */
  uint64_t n_reads;
  uint32_t L;
  uint64_t sub_thresh;    /* substitution probability threshold            */
  uint64_t indel_thresh;  /* indel probability threshold                   */
};

BR_HD uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return fmix64(seed ^ fmix64(stream * 0x9e3779b97f4a7c15ULL + 0x7f4a7c159e3779b9ULL) ^
                fmix64(ctr + 0xd1b54a32d192ed03ULL));
}
/* uniform integer in [0, n) from a 64-bit random value (multiply-shift). */
BR_HD uint64_t below(uint64_t r, uint64_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(r, n);
#else
  return (uint64_t)(((unsigned __int128)r * n) >> 64);
#endif
}

enum : uint64_t {
  S_GENOME = 1, S_FAM_LEN = 2, S_FAM_BASE = 3, S_COPY_FAM = 4, S_COPY_OFF = 5, S_COPY_DIV = 6,
  S_COPY_SUB = 7, S_READ_START = 8, S_READ_STRAND = 9, S_READ_ERR = 10, S_READ_SUB = 11,
  S_READ_INDEL = 12, S_READ_INDEL_KIND = 13
};

BR_HD uint32_t family_len(const synth_params& p, uint32_t f) {
  return p.fmin + (uint32_t)below(rnd(p.seed, S_FAM_LEN, f), (uint64_t)(p.fmax - p.fmin + 1));
}

/* Base code (0..3) of the genome at position pos. */
BR_HD int genome_base(const synth_params& p, uint64_t pos) {
  if (p.n_copies > 0) {
    uint64_t slot = pos / p.spacing;
    if (slot < p.n_copies) {
      uint32_t f = (uint32_t)below(rnd(p.seed, S_COPY_FAM, slot), p.n_families);
      uint32_t flen = family_len(p, f);
      uint64_t off = below(rnd(p.seed, S_COPY_OFF, slot), p.spacing - flen + 1);
      uint64_t start = slot * p.spacing + off;
      if (pos >= start && pos < start + flen) {
        uint64_t i = pos - start;
        int b = (int)(rnd(p.seed, S_FAM_BASE, ((uint64_t)f << 32) | i) & 3);
        uint64_t ctr = (slot << 32) | i;
        if (rnd(p.seed, S_COPY_DIV, ctr) < p.div_thresh)
          b = (b + 1 + (int)below(rnd(p.seed, S_COPY_SUB, ctr), 3)) & 3;
        return b;
      }
    }
  }
  return (int)(rnd(p.seed, S_GENOME, pos) & 3);
}

/* Write read r (p.L bases, ASCII) to out. */
BR_HD void make_read(const synth_params& p, uint64_t r, char* out) {
  const char A[4] = {'A', 'C', 'G', 'T'};
  uint32_t L = p.L;
  /* with indels a read may consume up to ~L(1+rate) genome bases: keep a
   * margin of L/4 so the walk never leaves the genome. */
  uint64_t span = (p.indel_thresh ? (uint64_t)L + L / 4 : (uint64_t)L);
  if (span > p.G) span = p.G;
  uint64_t start = below(rnd(p.seed, S_READ_START, r), p.G - span + 1);
  bool rev = (rnd(p.seed, S_READ_STRAND, r) & 1) != 0;
  uint64_t g = 0;  /* genome offset consumed along the read's strand */
  for (uint32_t i = 0; i < L; ++i) {
    uint64_t ctr = (r << 20) ^ i;
    if (p.indel_thresh && rnd(p.seed, S_READ_INDEL, ctr) < p.indel_thresh) {
      if (rnd(p.seed, S_READ_INDEL_KIND, ctr) & 1) {
        /* insertion: a random base, genome offset unchanged */
        out[i] = A[rnd(p.seed, S_READ_SUB, ctr) & 3];
        continue;
      }
      g += 1; /* deletion: skip one genome base */
    }
    if (g >= span) g = span - 1;
    int b = rev ? 3 - genome_base(p, start + span - 1 - g) : genome_base(p, start + g);
    g += 1;
    if (rnd(p.seed, S_READ_ERR, ctr) < p.sub_thresh)
      b = (b + 1 + (int)below(rnd(p.seed, S_READ_SUB, ctr), 3)) & 3;
    out[i] = A[b];
  }
}

}  // namespace br

#endif
