"""FASTQ ingest on the device: the split equals a host reader on the
reference's fixtures and on edge-case files, malformed files are rejected,
and parse + count straight from FASTQ bytes gives the oracle's graph."""
import os

import numpy as np
import pytest

import bruno_amd
from bruno_amd.device import DeviceBuffer
from bruno_amd.pipeline import fastq_split_device
from gpu_common import compare_results, run_oracle
from helpers import REF_DATA, pack, random_reads, read_fastq, thresholds_L

pytestmark = pytest.mark.gpu

FIXTURES = sorted(f for f in os.listdir(REF_DATA) if f.endswith(".fastq"))


def to_fastq(seqs, crlf=False, final_newline=True, trailing_blank=0):
    nl = "\r\n" if crlf else "\n"
    recs = ["@r%d%s%s%s+%s%s" % (i, nl, s, nl, nl, "I" * len(s)) for i, s in enumerate(seqs)]
    txt = nl.join(recs)
    if final_newline:
        txt += nl
    txt += nl * trailing_blank
    return txt.encode()


def split(data):
    buf = DeviceBuffer.from_numpy(np.frombuffer(data, np.uint8))
    try:
        b, o = fastq_split_device(buf.ptr, len(data))
    finally:
        buf.free()
    return [bytes(b[int(o[i]):int(o[i + 1])]).decode() for i in range(len(o) - 1)]


def strict_fastq(data):
    """Four-line records with '@' and '+' lines where they belong."""
    lines = data.split(b"\n")
    while lines and lines[-1].strip() == b"":
        lines.pop()
    return len(lines) % 4 == 0 and all(lines[4 * r].startswith(b"@") and lines[4 * r + 2].startswith(b"+")
                                       for r in range(len(lines) // 4))


FASTQ_FIXTURES = [f for f in FIXTURES if strict_fastq(open(os.path.join(REF_DATA, f), "rb").read())]


@pytest.mark.parametrize("name", FIXTURES)
def test_split_matches_host_reader_on_fixtures(name):
    with open(os.path.join(REF_DATA, name), "rb") as f:
        data = f.read()
    seqs, _ = read_fastq(os.path.join(REF_DATA, name))
    if name not in FASTQ_FIXTURES:
        # debruijn.palindrome3.fastq holds diff output, not FASTQ records
        with pytest.raises(bruno_amd.BrunoError):
            split(data)
        assert seqs == []
        return
    assert split(data) == seqs


@pytest.mark.parametrize("crlf", [False, True])
@pytest.mark.parametrize("final_newline", [False, True])
@pytest.mark.parametrize("blank", [0, 2])
def test_split_line_ends_and_file_ends(crlf, final_newline, blank):
    _, reads = random_reads(4, 5000, 3000, (20, 300), 0.01)
    reads[7] = ""                          # empty sequence line
    reads[11] = reads[11][:40] + "N" + reads[11][41:]
    if not final_newline:
        blank = 0
    assert split(to_fastq(reads, crlf, final_newline, blank)) == reads


def test_split_many_chunks_and_empty():
    _, reads = random_reads(5, 100000, 120000, (30, 250), 0.01)
    assert split(to_fastq(reads)) == reads  # ~20 MB: records cross 16 KB chunk borders
    assert split(b"") == []
    assert split(b"\n\n") == []


@pytest.mark.parametrize("bad", [b"@r\nACGT\n-\nIIII\n", b"r\nACGT\n+\nIIII\n", b"@r\nACGT\n+\nIIII\n@r2\nAC\n",
                                 b"@r\nACGT\n+\nIIII\nACGT\n+\nIIII\n"])
def test_malformed_fastq_is_rejected(bad):
    with pytest.raises(bruno_amd.BrunoError) as e:
        split(bad)
    assert e.value.code == 1


@pytest.mark.parametrize("name", FASTQ_FIXTURES)
def test_graph_from_fastq_bytes_equals_oracle(name):
    with open(os.path.join(REF_DATA, name), "rb") as f:
        data = f.read()
    seqs, _ = read_fastq(os.path.join(REF_DATA, name))
    b, off = pack(seqs)
    for th in (None, thresholds_L(2)):
        job = bruno_amd.GraphJob(31, th)
        job.parse_count_fastq(data)
        res = {"counts": job.counts(), "k2": job.k2()}
        job.filter()
        res["nodes"] = job.nodes()
        job.build_chains()
        res["rounds"] = job.list_rank(1)
        job.compact()
        res.update(fasta=job.unitigs(), branch_fasta=job.branch_fasta(), chain_nodes=job.chain_nodes(),
                   branches=job.branches(), stats=job.stats())
        job.close()
        compare_results(31, res, run_oracle(31, th, b, off, 1), 1)


@pytest.mark.parametrize("crlf", [False, True])
def test_in_place_parse_equals_split_reads(crlf):
    """bruno_parse_count_fastq parses the sequence lines where they are in
    the file; the graph equals the one from the split (contiguous) reads."""
    _, reads = random_reads(12, 30000, 8000, (40, 260), 0.01)
    data = to_fastq(reads, crlf=crlf, final_newline=False)
    th = thresholds_L(2)
    job = bruno_amd.GraphJob(31, th)
    job.parse_count_fastq(data)
    job.run(1)
    got = (job.unitigs(), job.branch_fasta(), job.stats()["n_instances"])
    job.close()
    b, off = pack(split(data))
    o = run_oracle(31, th, b, off, 1)
    assert got[0] == o.chain_fasta() and got[1] == o.branch_fasta()
    assert got[2] == o.stat("n_instances")


@pytest.mark.parametrize("crlf", [False, True])
def test_in_place_parse_by_regions(crlf):
    """The region count reads the sequence lines in place too (read bounds
    from the start / end arrays), reads with N included."""
    from bruno_amd.autotune import get_tuning, set_tuning
    _, reads = random_reads(14, 40000, 6000, (40, 300), 0.01)
    reads[7] = reads[7][:20] + "N" + reads[7][21:]
    reads[9] = reads[9][:-1] + "N"
    data = to_fastq(reads, crlf=crlf)
    th = thresholds_L(2)
    old = get_tuning()
    set_tuning(count_regions=2)
    try:
        job = bruno_amd.GraphJob(31, th, capacity=1 << 22)
        job.parse_count_fastq(data)
        assert job.stats()["count_by_regions"] == 1
        job.run(1)
        got = (job.unitigs(), job.branch_fasta(), job.stats()["n_instances"], job.stats()["n_reads_dropped"])
        job.close()
    finally:
        set_tuning(**old)
    b, off = pack(split(data))
    o = run_oracle(31, th, b, off, 1)
    assert got[0] == o.chain_fasta() and got[1] == o.branch_fasta()
    assert got[2] == o.stat("n_instances")
    assert got[3] == 2
