"""Sizing of jobs and batches for synthetic workloads.

A job's hash table needs one slot per distinct canonical k-mer at the chosen
load factor.  For reads sampled from a genome of G bases with substitution
rate e (and 1-bp indel rate d), the distinct k-mers are about the genome's G
plus the novel k-mers each error creates: an error at a base is covered by
min(k, L) * (L - k + 1) / L of the read's k-mers on average, and almost every
such k-mer is new.  Peak device memory of a job is its table and side table
plus the reads plus the per-node lists of the later phases.
"""
SLOT_BYTES = {1: 32, 2: 64}


def words(k):
    return 1 if k <= 31 else 2


def est_distinct(s, k):
    L = s.read_len
    per_read = max(0, L - k + 1)
    err = s.sub_rate + 2 * s.indel_rate
    cover = min(k, L) * per_read / max(1, L)
    return s.genome_len + s.n_reads * L * err * cover


def capacity(s, k, load_factor=0.5):
    return int(est_distinct(s, k) / load_factor) + 4096


def footprint(s, k, load_factor=0.5, cap=None):
    """Estimated peak HBM bytes of one job: table + side table (one 64-byte
    record per 32 slots) + present list (8 B per distinct k-mer) + reads +
    ~300 B per genome k-mer for the chain, list-ranking and output arrays."""
    cap = capacity(s, k, load_factor) if cap is None else cap
    table = cap * SLOT_BYTES[words(k)] + max(1024, cap // 32) * 64
    present = 8 * est_distinct(s, k)
    reads = s.n_bases + 8 * (s.n_reads + 1)
    chains = 300 * s.genome_len * words(k)
    return int(table + present + reads + chains)


# A job with a table of at least this many slots runs alone.  Alone, a job
# of 2^22 slots or more counts by regions (the batched count is the atomic
# one); measured on config-4 jobs, from about 5e7 k-mer instances (tables of
# about 2^26 slots) its region count and later phases alone cost less than
# its share of a batch.
SOLO_SLOTS = 1 << 26


def plan_groups(cfgs, k, budget_bytes, load_factor=0.5, solo_slots=SOLO_SLOTS):
    """Jobs packed in arrival order into launch groups whose estimated
    footprint stays within the budget (a job larger than the budget runs
    alone).  A job whose table reaches solo_slots (0: none) runs alone too,
    in its place: counted by regions it goes faster than in a batch, and a
    batch gains nothing on a job that fills the device by itself.  The
    other jobs keep filling the open group past it, which then runs when
    it is full (or at the end)."""
    groups, cur, used = [], [], 0
    for i, s in enumerate(cfgs):
        if solo_slots and capacity(s, k, load_factor) >= solo_slots:
            groups.append([i])
            continue
        f = footprint(s, k, load_factor)
        if cur and used + f > budget_bytes:
            groups.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += f
    if cur:
        groups.append(cur)
    return groups


def instances(s, k):
    """k-mer instances of a synthetic job's reads."""
    return s.n_reads * max(0, s.read_len - k + 1)


def short_first(groups, cfgs, k):
    """Launch groups ordered by k-mer instances per job, fewest first
    (Smith's rule: for jobs submitted together this minimises the sum of
    their completion times, taking a group's time as proportional to its
    instances; the total time is the same in any order)."""
    return sorted(groups, key=lambda g: (sum(instances(cfgs[i], k) for i in g) / len(g), g[0]))
