"""Shared test helpers: FASTQ fixtures, read packing, random inputs."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DATA = os.path.join(ROOT, "tests", "golden", "ref_data")

DEFAULT_T = 65536
COMP = str.maketrans("ACGT", "TGCA")


def thresholds_L(t):
    """-T -L t: {k_lo,k_hi,k1_lo,k1_hi,k2_lo,k2_hi}
    (compact_dbg_freq_construct_clean.cpp:874, 940-945)."""
    return [0, DEFAULT_T, min(t, DEFAULT_T), DEFAULT_T, 0, DEFAULT_T]


def read_fastq(path):
    seqs, quals = [], []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i + 3 < len(lines) + 1 and i < len(lines):
        if lines[i].startswith("@") and i + 3 < len(lines):
            seqs.append(lines[i + 1].strip())
            quals.append(lines[i + 3].strip())
            i += 4
        else:
            i += 1
    return seqs, quals


def pack(seqs):
    b = "".join(seqs).encode()
    off = np.zeros(len(seqs) + 1, np.uint64)
    if seqs:
        off[1:] = np.cumsum([len(s) for s in seqs])
    return b, off


def rc(s):
    return s.translate(COMP)[::-1]


def random_reads(seed, genome_len, n_reads, read_len, err=0.01, alphabet="ACGT"):
    rng = np.random.default_rng(seed)
    g = "".join(rng.choice(list(alphabet), genome_len))
    reads = []
    for _ in range(n_reads):
        L = read_len if np.isscalar(read_len) else int(rng.integers(read_len[0], read_len[1] + 1))
        L = min(L, genome_len)
        st = int(rng.integers(0, genome_len - L + 1))
        s = list(g[st:st + L])
        for i in range(L):
            if rng.random() < err:
                s[i] = rng.choice([c for c in "ACGT" if c != s[i]])
        s = "".join(s)
        if rng.random() < 0.5:
            s = rc(s)
        reads.append(s)
    return g, reads
