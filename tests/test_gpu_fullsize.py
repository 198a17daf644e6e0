"""Full-size workloads (100 Mbp / 33 M reads at k=31; 20 Mbp / 3.2 M reads at
k=63), where the oracle would take many minutes: size-independent
properties of the outputs, in the spirit of the reference's own
check_compacted_files.sh (first k-mers of the unitigs = the chain
representatives), plus independence from the order in which read batches
arrive."""
import numpy as np
import pytest

import bruno_amd
from bruno_amd.configs import CONFIGS
from bruno_amd.device import DeviceBuffer, synchronize
from bruno_amd.pipeline import synth_reads_device
from bruno_amd.planner import capacity

pytestmark = pytest.mark.gpu

COMP = bytes.maketrans(b"ACGT", b"TGCA")


def rc_words(lo, hi, k):
    """Reverse complement of packed k-mers (numpy uint64 words, k <= 63)."""
    def rcw(x):
        x = ~x
        x = ((x >> np.uint64(2)) & np.uint64(0x3333333333333333)) | ((x & np.uint64(0x3333333333333333)) << np.uint64(2))
        x = ((x >> np.uint64(4)) & np.uint64(0x0F0F0F0F0F0F0F0F)) | ((x & np.uint64(0x0F0F0F0F0F0F0F0F)) << np.uint64(4))
        return x.byteswap()
    if k <= 31:
        return rcw(lo) >> np.uint64(64 - 2 * k), np.zeros_like(lo)
    # 128-bit value (hi, lo): reversed = (rcw(lo), rcw(hi)) shifted right by 128 - 2k
    a, b = rcw(lo), rcw(hi)           # new hi word, new lo word before the shift
    sh = 128 - 2 * k                  # 2 <= sh < 64 for 32 <= k <= 63
    new_lo = (b >> np.uint64(sh)) | (a << np.uint64(64 - sh))
    new_hi = a >> np.uint64(sh)
    return new_lo, new_hi


def canonical(lo, hi, k):
    rlo, rhi = rc_words(lo, hi, k)
    take_rc = (rhi < hi) | ((rhi == hi) & (rlo < lo))
    return np.where(take_rc, rlo, lo), np.where(take_rc, rhi, hi)


def decode(lo, hi, k):
    v = (int(hi) << 64) | int(lo)
    return "".join("ACGT"[(v >> (2 * (k - 1 - i))) & 3] for i in range(k)).encode()


def run(cfg, split_batches):
    s, k = cfg["synth"], cfg["k"]
    bases, offs = DeviceBuffer(s.n_bases), DeviceBuffer((s.n_reads + 1) * 8)
    synth_reads_device(s, bases.ptr, offs.ptr)
    synchronize()
    job = bruno_amd.GraphJob(k, cfg["thresholds"], capacity=capacity(s, k))
    if split_batches:
        half = s.n_reads // 2
        job.parse_count_device(bases.ptr, offs.ptr + 8 * half, s.n_reads - half)  # second half first
        job.parse_count_device(bases.ptr, offs.ptr, half)
    else:
        job.parse_count_device(bases.ptr, offs.ptr, s.n_reads)
    job.run(1)
    out = dict(stats=job.stats(), fasta=job.unitigs(), branch=job.branch_fasta(), nodes=job.chain_nodes(),
               branches=job.branches())
    job.close()
    bases.free()
    offs.free()
    return out


@pytest.mark.parametrize("name", ["config2", "config3"])
def test_full_size_invariants(name):
    cfg = CONFIGS[name]()
    s, k = cfg["synth"], cfg["k"]
    a = run(cfg, split_batches=True)
    st = a["stats"]
    # every read is ACGT of the configured length
    assert st["n_reads"] == s.n_reads and st["n_reads_dropped"] == 0
    assert st["n_instances"] == s.n_reads * (s.read_len - k + 1)
    assert st["n_nodes"] == st["n_branches"] + st["n_chain_nodes"]
    assert st["n_chain_nodes"] == st["n_unitig_nodes"] + st["n_cycle_nodes"] + st["n_isolated"]
    assert st["n_conflicts"] == 0

    # .chain.fasta: header = canonical of the first k-mer, records sorted by
    # the first k-mer (the representative), one node per sequence position
    lines = a["fasta"].split(b"\n")
    heads, seqs = lines[0::2], lines[1::2]
    assert len(heads) == len(seqs) == st["n_unitigs"]
    firsts = [q[:k] for q in seqs]
    assert all(len(q) >= k for q in seqs)
    assert all(h[:1] == b">" for h in heads)
    assert all(h[1:] == min(f, f.translate(COMP)[::-1]) for h, f in zip(heads, firsts))
    assert all(firsts[i] < firsts[i + 1] for i in range(len(firsts) - 1))
    assert sum(len(q) - k + 1 for q in seqs) == st["n_unitig_nodes"]

    # .chain.components: per representative, distances 0..m-1 once each, m =
    # the unitig's node count; the representatives are the unitigs' first k-mers
    cn = a["nodes"]
    assert len(cn) == st["n_unitig_nodes"]
    rep_lo, rep_hi = cn["rep"][:, 0], cn["rep"][:, 1]
    change = np.flatnonzero((rep_lo[1:] != rep_lo[:-1]) | (rep_hi[1:] != rep_hi[:-1])) + 1
    starts = np.concatenate([[0], change])
    ends = np.concatenate([change, [len(cn)]])
    assert len(starts) == len(seqs)
    sizes = ends - starts
    assert np.array_equal(sizes, np.array([len(q) - k + 1 for q in seqs]))
    expect_dist = np.arange(len(cn)) - np.repeat(starts, sizes)
    assert np.array_equal(cn["dist"], expect_dist)
    for i in range(0, len(starts), max(1, len(starts) // 2000)):  # sampled: decode is per k-mer in Python
        assert decode(rep_lo[starts[i]], rep_hi[starts[i]], k) == firsts[i]

    # every chain node's canonical k-mer is unique and no branch k-mer is one
    clo, chi = canonical(cn["k"][:, 0].copy(), cn["k"][:, 1].copy(), k)
    br = a["branches"]
    blo = br[:, 0]
    bhi = br[:, 1] if br.shape[1] > 1 else np.zeros_like(blo)
    assert np.all((bhi[1:] > bhi[:-1]) | ((bhi[1:] == bhi[:-1]) & (blo[1:] > blo[:-1])))  # sorted, distinct
    lo = np.concatenate([clo, blo])
    hi = np.concatenate([chi, bhi])
    order = np.lexsort((lo, hi)) if k > 31 else np.argsort(lo, kind="stable")
    lo, hi = lo[order], hi[order]
    dup = (lo[1:] == lo[:-1]) & (hi[1:] == hi[:-1])
    assert not dup.any()
    assert len(a["branch"]) == len(blo) * (2 * k + 3)

    # the order in which read batches arrive does not matter
    b = run(cfg, split_batches=False)
    assert b["fasta"] == a["fasta"] and b["branch"] == a["branch"]
    assert b["stats"]["iterations"] == st["iterations"]
