/*
 * kmer.h -- 2-bit k-mer arithmetic shared by the HIP kernels and the host.
 *
 * A k-mer is an unsigned integer v = sum_i c_i * 4^(K-1-i) with DNA codes
 * A0 C1 G2 T3: the first base is the most significant digit, so numeric
 * order is lexicographic order (bruno's Kmer<K, DNA, uint64_t> value order,
 * used by lex_less, debruijn_common.hpp:44-67).  KT is uint64_t for k <= 31
 * and unsigned __int128 for k <= 63.
 *
 * Edge bytes use DNA16 nibbles A1 C2 G4 T8, 0 = no base; byte = in<<4 | out
 * (debruijn_biedge_loader.hpp:47-50).
 */
#ifndef BRUNO_KMER_H_
#define BRUNO_KMER_H_

#include <stdint.h>

#ifdef __HIPCC__
#define BR_HD __host__ __device__ __forceinline__
#else
#define BR_HD inline
#endif

typedef unsigned __int128 u128;

namespace br {

template <typename KT> struct kt_traits;
template <> struct kt_traits<uint64_t> { static constexpr int words = 1; };
template <> struct kt_traits<u128> { static constexpr int words = 2; };

BR_HD uint64_t fmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

/* The sketch's hash of a one-word k-mer (two multiplies of fmix64 and one
 * more: every output bit depends on every input bit, as a HyperLogLog's
 * leading-zero counts need). */
BR_HD uint64_t sketch_hash(uint64_t v) { return fmix64(v * 0x9e3779b97f4a7c15ULL + 0x632be59bd9b4e019ULL); }

/* The table's hash: a bijection of 64-bit words with one multiply (a
 * xor-shift by 32 undoes itself, the multiplier is odd), so the k-mer comes
 * back from its hash; the product's top bits, which pick the bucket and the
 * region, depend on every input bit, and the last xor-shift folds them
 * into the bits below that pick the home slot.  Each instance is hashed
 * once by the parser and unhashed once by the region build, both bound by
 * instruction issue, where a 64-bit multiply is several quarter-rate ops. */
constexpr uint64_t KMIX_MUL = 0xd6e8feb86659fd93ULL;
constexpr uint64_t KMIX_INV = 0xcfee444d8b59a89bULL; /* KMIX_MUL^-1 mod 2^64 */
BR_HD uint64_t kmix(uint64_t x) {
  x ^= x >> 32;
  x *= KMIX_MUL;
  return x ^ (x >> 32);
}
BR_HD uint64_t kmix_inv(uint64_t h) {
  h ^= h >> 32;
  h *= KMIX_INV;
  return h ^ (h >> 32);
}
BR_HD uint64_t kmer_hash(uint64_t v) { return kmix(v); }
BR_HD uint64_t kmer_unhash(uint64_t h) { return kmix_inv(h); }
BR_HD uint64_t kmer_hash(u128 v) { return kmix((uint64_t)v ^ kmix((uint64_t)(v >> 64) + 0x9e3779b97f4a7c15ULL)); }
/* a two-word k-mer from its hash and its high word */
BR_HD u128 kmer_unhash(uint64_t h, uint64_t hi) {
  return ((u128)hi << 64) | (kmix_inv(h) ^ kmix(hi + 0x9e3779b97f4a7c15ULL));
}
BR_HD uint64_t sketch_hash(u128 v) {
  return fmix64((uint64_t)v ^ fmix64((uint64_t)(v >> 64) + 0x9e3779b97f4a7c15ULL));
}

/* Reverse complement of a 64-bit word holding 32 bases: complement is the
 * bitwise NOT of each 2-bit code (3 - c), the base order is reversed by
 * swapping 2-bit fields inside nibbles, nibbles inside bytes, then bytes. */
BR_HD uint64_t rc_word(uint64_t x) {
  x = ~x;
  x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
  return __builtin_bswap64(x);
}

template <typename KT>
struct kops {
  int K;
  KT mask;
  BR_HD explicit kops(int k) : K(k) {
    mask = (2 * k >= (int)(8 * sizeof(KT))) ? ~(KT)0 : (((KT)1 << (2 * k)) - 1);
  }
  BR_HD KT rc(KT v) const;
  /* base at sequence position p (0 = first). */
  BR_HD int base(KT v, int p) const { return (int)((v >> (2 * (K - 1 - p))) & 3); }
  BR_HD int first(KT v) const { return (int)((v >> (2 * (K - 1))) & 3); }
  BR_HD int last(KT v) const { return (int)(v & 3); }
  /* out neighbour with base c (nextFromChar), in neighbour (nextReverseFromChar)
   * (debruijn_graph_node.hpp:315-381). */
  BR_HD KT out_nbr(KT v, int c) const { return ((v << 2) | (KT)c) & mask; }
  BR_HD KT in_nbr(KT v, int c) const { return (v >> 2) | ((KT)c << (2 * (K - 1))); }
  BR_HD KT canon(KT v) const {
    KT r = rc(v);
    return (v < r) ? v : r;
  }
  /* rc-palindrome test of a len-base value x. */
  BR_HD bool is_pal_len(KT x, int len) const;
  /* kmer_traits (kmer_traits.hpp:50-127). */
  BR_HD bool is_rc_palindrome(KT v) const { return (K & 1) ? false : (v == rc(v)); }
  BR_HD bool km1_high(KT v) const { return (K & 1) ? is_pal_len(v >> 2, K - 1) : false; }
  BR_HD bool km1_low(KT v) const {
    return (K & 1) ? is_pal_len(v & (((KT)1 << (2 * (K - 1))) - 1), K - 1) : false;
  }
};

template <>
BR_HD uint64_t kops<uint64_t>::rc(uint64_t v) const {
  return rc_word(v) >> (64 - 2 * K);
}
template <>
BR_HD u128 kops<u128>::rc(u128 v) const {
  uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
  u128 r = ((u128)rc_word(lo) << 64) | (u128)rc_word(hi);
  return r >> (128 - 2 * K);
}
template <>
BR_HD bool kops<uint64_t>::is_pal_len(uint64_t x, int len) const {
  return (rc_word(x) >> (64 - 2 * len)) == x;
}
template <>
BR_HD bool kops<u128>::is_pal_len(u128 x, int len) const {
  uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  u128 r = (((u128)rc_word(lo) << 64) | (u128)rc_word(hi)) >> (128 - 2 * len);
  return r == x;
}

/* DNA16 nibble helpers. */
BR_HD int nib_to_code(int n) { return n == 1 ? 0 : n == 2 ? 1 : n == 4 ? 2 : n == 8 ? 3 : -1; }
BR_HD uint8_t dna16_comp(uint8_t n) {
  return (uint8_t)(((n & 1) << 3) | ((n & 2) << 1) | ((n & 4) >> 1) | ((n & 8) >> 3));
}
/* Kmer<2,DNA16>::reverse_complement of the biedge byte. */
BR_HD uint8_t biedge_rc(uint8_t e) { return (uint8_t)((dna16_comp(e & 0xF) << 4) | dna16_comp(e >> 4)); }

}  // namespace br

#endif
