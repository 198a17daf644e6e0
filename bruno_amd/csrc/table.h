/*
 * table.h -- HBM open-addressing k-mer table used for counting and as the
 * graph store.
 *
 * One slot per canonical k-mer.  During counting a slot holds the k-mer and
 * NSUB k2 sub-entries (biedge code + uint count), so the joint (in, out)
 * counts that bruno's k2 counter keeps (LocalK2CountMapType,
 * debruijn_graph_map.hpp:160-168) sit in the same cache line as the key.
 * A k-mer with more than NSUB distinct biedges spills the extra ones into a
 * side table keyed by (slot << 9 | code).  After the filter the sub-entry
 * area of a surviving slot is overwritten in place with the node record
 * (9 saturating counters, class, chain index), so lookups by k-mer keep
 * working on the same table.
 *
 * k <= 31 (KT = uint64_t): 32-byte slot {key, sub[3]}.  key = k-mer with
 *   bit 63 = spilled, bit 62 = present after filter; EMPTY = all ones (a
 *   canonical 31-mer is never all T, so a real key never reads as EMPTY).
 * k <= 63 (KT = u128): 64-byte slot {meta, lo, hi, sub[5]}.  The key is two
 *   words, so a slot is claimed by CAS on meta (EMPTY -> LOCK), the k-mer is
 *   written, then meta is published as its flag word (bit 63 spilled, bit 62
 *   present).
 * Sub-entry: (code << 48) | count with code = biedge + 1 (1..256); EMPTY =
 *   all ones.
 */
#ifndef BRUNO_TABLE_H_
#define BRUNO_TABLE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kmer.h"

namespace br {

constexpr uint64_t EMPTY64 = ~0ULL;
constexpr uint64_t LOCK64 = ~0ULL - 1;
constexpr uint64_t F_SPILL = 1ULL << 63;
constexpr uint64_t F_PRESENT = 1ULL << 62;
constexpr uint64_t KEY_BITS = (1ULL << 62) - 1;
constexpr uint32_t MAX_PROBE = 1u << 20;
constexpr uint32_t MAX_SPIN = 1u << 22;

/* error bits */
constexpr uint32_t ERR_TABLE_FULL = 1;
constexpr uint32_t ERR_SPILL_FULL = 2;
constexpr uint32_t ERR_SPIN = 4;
constexpr uint32_t ERR_LIST_FULL = 8;

/* node record stored in a slot's sub area after the filter */
struct NodeRec {
  uint16_t c[9];  /* out A,C,G,T, in A,C,G,T, self */
  uint8_t cls;    /* CLS_* */
  uint8_t pad;
  uint32_t idx;   /* chain / branch index */
};
constexpr uint8_t CLS_BRANCH = 1;
constexpr uint8_t CLS_CHAIN = 2;

/* Side table of k2 entries beyond a slot's NSUB sub-entries: records of
 * SPILL_SUB more entries, keyed by (slot << 2) | part.  Part p + 1 of a
 * slot is opened only once part p is full (entries are never removed), so
 * a reader stops at the first part with an empty entry; one record holds
 * every entry of almost every such slot (a k-mer has at most 25 biedges).
 * Empty record / entry = all ones, so one fill clears the table. */
constexpr int SPILL_SUB = 7;
constexpr int SPILL_PARTS = 4; /* NSUB 3 + 4 x 7 >= 25 */
struct Spill {
  uint64_t key;
  uint64_t sub[SPILL_SUB];      /* (code << 48) | count, as a slot's sub-entries */
};

template <typename KT> struct Slot;
template <> struct Slot<uint64_t> {
  uint64_t key;
  uint64_t sub[3];
  static constexpr int NSUB = 3;
};
template <> struct Slot<u128> {
  uint64_t meta;
  uint64_t lo, hi;
  uint64_t sub[5];
  static constexpr int NSUB = 5;
};
static_assert(sizeof(Slot<uint64_t>) == 32, "slot size");
static_assert(sizeof(Slot<u128>) == 64, "slot size");
static_assert(sizeof(NodeRec) == 24, "node size");
static_assert(sizeof(Spill) == 64, "side-table record size");

__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t cas64(uint64_t* p, uint64_t cmp, uint64_t val) {
  return atomicCAS((unsigned long long*)p, (unsigned long long)cmp, (unsigned long long)val);
}
__device__ __forceinline__ void add64(uint64_t* p, uint64_t v) {
  atomicAdd((unsigned long long*)p, (unsigned long long)v);
}
__device__ __forceinline__ void or64(uint64_t* p, uint64_t v) {
  atomicOr((unsigned long long*)p, (unsigned long long)v);
}

__device__ __forceinline__ uint64_t home_slot(uint64_t h, uint64_t cap) { return __umul64hi(h, cap); }

/* Probe window of a k-mer hash.  rbits == 0: the whole table of span slots
 * (home = fastrange(h, span), wrapping at span).  rbits > 0: the table is
 * 2^rbits regions of span slots each; the top rbits hash bits pick the
 * region, the rest the home slot inside it, and probes wrap inside the
 * region, so every slot a k-mer can occupy lies in one region. */
struct Probe {
  uint64_t lo, hi, s;
  __device__ __forceinline__ Probe(uint64_t h, uint64_t span, uint32_t rbits) {
    lo = rbits ? (h >> (64 - rbits)) * span : 0;
    hi = lo + span;
    s = lo + __umul64hi(h << rbits, span);
  }
  __device__ __forceinline__ void next() { s = (s + 1 == hi) ? lo : s + 1; }
};

/* ------------------------------------------------------------------ */
/* key access per slot type                                             */
/* ------------------------------------------------------------------ */
template <typename KT> struct SlotOps;

/* Counting-time access rules.  Every word a counting thread reads is read
 * with an untorn, device-coherent relaxed load, and only when it is needed
 * (the key first, then sub-entries in order).  A key, a k-mer word or a
 * biedge code is written once and never changes, and every transition out
 * of EMPTY is made by a CAS that returns the current value.  No acquire /
 * release fence is used: on a multi-XCD device an agent-scope fence writes
 * back or invalidates the L2 caches. */
template <typename KT> struct SlotOps;

template <> struct SlotOps<uint64_t> {
  typedef Slot<uint64_t> S;
  /* Find or claim the slot of canonical k-mer v.  Returns the slot index or
   * ~0 on a full table; *claimed = 1 when this call inserted the key. */
  static __device__ uint64_t find_or_insert(S* T, uint64_t span, uint32_t rbits, uint64_t v, uint32_t* err,
                                            int* claimed) {
    Probe q(kmer_hash(v), span, rbits);
    const uint64_t lim = span < MAX_PROBE ? span : MAX_PROBE;
    for (uint64_t probe = 0; probe < lim; ++probe, q.next()) {
      S* p = T + q.s;
      uint64_t k0 = ld_relaxed(&p->key);
      if (k0 == EMPTY64) {
        uint64_t old = cas64(&p->key, EMPTY64, v);
        if (old == EMPTY64) {
          *claimed = 1;
          return q.s;
        }
        k0 = old;
      }
      if ((k0 & KEY_BITS) == v) return q.s;
    }
    atomicOr(err, ERR_TABLE_FULL);
    return ~0ULL;
  }
  /* One probe of a lookup: 1 = v is at slot s, 2 = v is absent (an empty
   * slot, or present is required and v did not survive the filter), 0 =
   * probe on. */
  static __device__ __forceinline__ int probe_state(const S* T, uint64_t s, uint64_t v, bool require_present) {
    const uint64_t k0 = T[s].key;
    if (k0 == EMPTY64) return 2;
    if ((k0 & KEY_BITS) != v) return 0;
    return (require_present && !(k0 & F_PRESENT)) ? 2 : 1;
  }
  /* Lookup only (separate kernels after counting); ~0 when absent.
   * require_present: only nodes that survived the filter. */
  static __device__ uint64_t find(const S* T, uint64_t span, uint32_t rbits, uint64_t v, bool require_present) {
    Probe q(kmer_hash(v), span, rbits);
    const uint64_t lim = span < MAX_PROBE ? span : MAX_PROBE;
    for (uint64_t probe = 0; probe < lim; ++probe, q.next()) {
      uint64_t k0 = T[q.s].key;
      if (k0 == EMPTY64) return ~0ULL;
      if ((k0 & KEY_BITS) == v) {
        if (require_present && !(k0 & F_PRESENT)) return ~0ULL;
        return q.s;
      }
    }
    return ~0ULL;
  }
  static __device__ __forceinline__ uint64_t* flag_word(S* p) { return &p->key; }
  static __device__ __forceinline__ bool occupied(const S* p) { return p->key != EMPTY64; }
  static __device__ __forceinline__ uint64_t kmer(const S* p) { return p->key & KEY_BITS; }
  static __device__ __forceinline__ uint64_t flags(const S* p) { return p->key & ~KEY_BITS; }
};

/* Two-word keys: meta is claimed by CAS (EMPTY -> LOCK), lo and hi are
 * written with atomic exchanges, then meta is published with a third
 * exchange whose value depends on both returned words, so it cannot be
 * performed before them.  Readers spin (bounded) while meta is LOCK, then
 * read lo and hi. */
template <> struct SlotOps<u128> {
  typedef Slot<u128> S;
  static __device__ uint64_t find_or_insert(S* T, uint64_t span, uint32_t rbits, u128 v, uint32_t* err,
                                            int* claimed) {
    uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
    Probe q(kmer_hash(v), span, rbits);
    const uint64_t lim = span < MAX_PROBE ? span : MAX_PROBE;
    for (uint64_t probe = 0; probe < lim; ++probe, q.next()) {
      const uint64_t s = q.s;
      S* p = T + s;
      uint64_t m = ld_relaxed(&p->meta);
      if (m == EMPTY64) {
        uint64_t old = cas64(&p->meta, EMPTY64, LOCK64);
        if (old == EMPTY64) {
          uint64_t a = atomicExch((unsigned long long*)&p->lo, (unsigned long long)lo);
          uint64_t b = atomicExch((unsigned long long*)&p->hi, (unsigned long long)hi);
          /* An EMPTY slot holds all-ones words, so pub is 0; any other old
           * value would only set the spilled flag, which makes the filter
           * consult the side table and changes no result. */
          uint64_t pub = ((a & b) ^ EMPTY64) & F_SPILL;
          atomicExch((unsigned long long*)&p->meta, (unsigned long long)pub);
          *claimed = 1;
          return s;
        }
        m = old;
      }
      uint32_t spin = 0;
      while (m == LOCK64 && spin < MAX_SPIN) {
        m = ld_relaxed(&p->meta);
        ++spin;
      }
      if (m == LOCK64) {
        atomicOr(err, ERR_SPIN);
        return ~0ULL;
      }
      if (ld_relaxed(&p->lo) == lo && ld_relaxed(&p->hi) == hi) return s;
    }
    atomicOr(err, ERR_TABLE_FULL);
    return ~0ULL;
  }
  static __device__ __forceinline__ int probe_state(const S* T, uint64_t s, u128 v, bool require_present) {
    const uint64_t m = T[s].meta;
    if (m == EMPTY64) return 2;
    if (T[s].lo != (uint64_t)v || T[s].hi != (uint64_t)(v >> 64)) return 0;
    return (require_present && !(m & F_PRESENT)) ? 2 : 1;
  }
  static __device__ uint64_t find(const S* T, uint64_t span, uint32_t rbits, u128 v, bool require_present) {
    uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
    Probe q(kmer_hash(v), span, rbits);
    const uint64_t lim = span < MAX_PROBE ? span : MAX_PROBE;
    for (uint64_t probe = 0; probe < lim; ++probe, q.next()) {
      const uint64_t s = q.s;
      uint64_t m = T[s].meta;
      if (m == EMPTY64) return ~0ULL;
      if (T[s].lo == lo && T[s].hi == hi) {
        if (require_present && !(m & F_PRESENT)) return ~0ULL;
        return s;
      }
    }
    return ~0ULL;
  }
  static __device__ __forceinline__ uint64_t* flag_word(S* p) { return &p->meta; }
  static __device__ __forceinline__ bool occupied(const S* p) { return p->meta != EMPTY64; }
  static __device__ __forceinline__ u128 kmer(const S* p) { return ((u128)p->hi << 64) | p->lo; }
  static __device__ __forceinline__ uint64_t flags(const S* p) { return p->meta; }
};

/* The record of part `part` of slot s: found, or claimed when claim is
 * set; nullptr when absent (or the table is full, flagged in err). */
__device__ __forceinline__ Spill* spill_record(Spill* sp, uint64_t spcap, uint64_t s, int part, bool claim,
                                               uint32_t* err) {
  const uint64_t key = (s << 2) | (uint64_t)part;
  uint64_t h = home_slot(fmix64(key), spcap);
  const uint64_t lim = spcap < MAX_PROBE ? spcap : MAX_PROBE;
  for (uint64_t probe = 0; probe < lim; ++probe) {
    Spill* q = sp + h;
    uint64_t k0 = ld_relaxed(&q->key);
    if (k0 == EMPTY64) {
      if (!claim) return nullptr;
      const uint64_t old = cas64(&q->key, EMPTY64, key);
      if (old == EMPTY64) return q;
      k0 = old;
    }
    if (k0 == key) return q;
    h = (h + 1 == spcap) ? 0 : h + 1;
  }
  if (claim) atomicOr(err, ERR_SPILL_FULL);
  return nullptr;
}

/* Count n instances (one unless a table is being moved) of entry code
 * (biedge + 1) of slot s in the side table.  Returns 1 if the entry is new. */
__device__ __forceinline__ int spill_add(Spill* sp, uint64_t spcap, uint64_t s, uint64_t code, uint32_t* err,
                                         uint64_t n = 1) {
  for (int part = 0; part < SPILL_PARTS; ++part) {
    Spill* q = spill_record(sp, spcap, s, part, true, err);
    if (!q) return 0;
#pragma unroll
    for (int j = 0; j < SPILL_SUB; ++j) {
      uint64_t v = ld_relaxed(&q->sub[j]);
      if (v == EMPTY64) {
        const uint64_t old = cas64(&q->sub[j], EMPTY64, (code << 48) | n);
        if (old == EMPTY64) return 1;
        v = old;
      }
      if ((v >> 48) == code) {
        add64(&q->sub[j], n);
        return 0;
      }
    }
  }
  atomicOr(err, ERR_SPILL_FULL);
  return 0;
}

/* f(code, count) for every side-table entry of slot s (after counting). */
template <typename F>
__device__ __forceinline__ void spill_each(const Spill* sp, uint64_t spcap, uint64_t s, F&& f) {
  for (int part = 0; part < SPILL_PARTS; ++part) {
    const Spill* q = spill_record(const_cast<Spill*>(sp), spcap, s, part, false, nullptr);
    if (!q) return;
    for (int j = 0; j < SPILL_SUB; ++j) {
      const uint64_t v = q->sub[j];
      if (v == EMPTY64) return;
      f(v >> 48, v & 0xFFFFFFFFFFFFULL);
    }
  }
}

/* Add n instances (one unless a table is being moved) of biedge e to slot
 * s.  Returns 1 if a new k2 entry was created. */
template <typename KT>
__device__ int add_k2(Slot<KT>* T, uint64_t s, uint8_t e, Spill* sp, uint64_t spcap, uint32_t* err,
                      uint64_t n = 1) {
  Slot<KT>* p = T + s;
  const uint64_t code = (uint64_t)e + 1;
#pragma unroll
  for (int j = 0; j < Slot<KT>::NSUB; ++j) {
    uint64_t v = ld_relaxed(&p->sub[j]);
    if (v == EMPTY64) {
      uint64_t old = cas64(&p->sub[j], EMPTY64, (code << 48) | n);
      if (old == EMPTY64) return 1;
      v = old;
    }
    if ((v >> 48) == code) {
      add64(&p->sub[j], n);
      return 0;
    }
  }
  /* spill: mark the slot, then count in the side table */
  uint64_t* fw = SlotOps<KT>::flag_word(p);
  if (!(ld_relaxed(fw) & F_SPILL)) or64(fw, F_SPILL);
  return spill_add(sp, spcap, s, code, err, n);
}

}  // namespace br

#endif
