"""Independent string-level de Bruijn builder used to pin the C++ oracle.

Everything here works on Python strings, with no 2-bit packing and no list
ranking, so a packing, orientation or pointer-jumping mistake in the oracle
cannot be repeated here:

* k+1-mer (edge) counts come from counting k+1-mer windows in the
  reads, not from aggregating k2 records by k-mer block;
* unitigs come from walking unique edges between chain nodes, not from
  list ranking.

The only formula shared with the reference restatement is the palindrome
halving of k+1-mer counts (debruijn_graph_map.hpp:1468-1507), which is a
definition rather than a derivation; it is exercised by crafted inputs whose
expected values are also worked out by hand in the tests.
"""
from collections import defaultdict

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}
IDX = {"A": 0, "C": 1, "G": 2, "T": 3}
DNA16 = {"A": 1, "C": 2, "G": 4, "T": 8, None: 0}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def canon(s):
    r = rc(s)
    return s if s < r else r


def is_pal(s):
    return s == rc(s)


def records(reads, k):
    """Canonical (k-mer, left, right) instances; reads with non-ACGT bases or
    shorter than k+2 are dropped."""
    out = []
    for s in reads:
        if any(c not in IDX for c in s) or len(s) < k + 2:
            continue
        L = len(s)
        for i in range(L - k + 1):
            x = s[i:i + k]
            a = s[i - 1] if i > 0 else None
            b = s[i + k] if i + k < L else None
            y = rc(x)
            if not (x < y):
                x, a, b = y, (COMP[b] if b else None), (COMP[a] if a else None)
            out.append((x, a, b))
    return out


def kp1_counts(reads, k):
    """Occurrences of each k+1-mer window, as written in the reads."""
    c = defaultdict(int)
    for s in reads:
        if any(ch not in IDX for ch in s) or len(s) < k + 2:
            continue
        for i in range(len(s) - k):
            c[s[i:i + k + 1]] += 1
    return c


def unfiltered(reads, k):
    """P1: {kmer: [outA,outC,outG,outT,inA,inC,inG,inT,self]} saturating."""
    g = {}
    for x, a, b in records(reads, k):
        n = g.setdefault(x, [0] * 9)
        if b:
            n[IDX[b]] = min(n[IDX[b]] + 1, 65535)
        if a:
            n[4 + IDX[a]] = min(n[4 + IDX[a]] + 1, 65535)
        n[8] = min(n[8] + 1, 65535)
    return g


def filtered(reads, k, th):
    """P2 from the k2 multiset with edge counts taken from k+1-mer windows."""
    t0lo, t0hi, t1lo, t1hi, t2lo, t2hi = th
    k2 = defaultdict(int)
    for x, a, b in records(reads, k):
        k2[(x, a, b)] += 1
    kp1 = kp1_counts(reads, k)
    by_k = defaultdict(list)
    for (x, a, b), c in k2.items():
        by_k[x].append((a, b, c))
    g = {}
    for x, lst in by_k.items():
        kf = min(sum(c for _, _, c in lst), 65535)

        def edge_count(e, side):
            # e: k+1-mer on x's strand.  Palindromic k+1-mers (odd k only
            # possible) use the reference's per-k2 halving.
            if is_pal(e) and not is_pal(x):
                tot = 0
                for a, b, c in lst:
                    ch = a if side == "in" else b
                    if ch and ((ch + x) if side == "in" else (x + ch)) == e:
                        tot += (c + 1) >> 1
                return min(tot, 65535)
            if is_pal(x):
                # lex_less flips ties (debruijn_common.hpp:55), so every
                # instance of an rc-palindromic k-mer is recorded from the
                # reverse strand: its edge e is only seen as the window rc(e).
                return min(kp1[rc(e)], 65535)
            return min(kp1[e] + kp1[rc(e)], 65535)

        node = [0] * 9
        have = False
        for a, b, c in lst:
            cnt = min(c, 65535)
            flag = 0
            if cnt < t2lo or cnt >= t2hi:
                flag |= 4
            if kf < t0lo or kf >= t0hi:
                flag |= 8
            kin = edge_count(a + x, "in") if a else t1hi
            if kin < t1lo or kin >= t1hi:
                flag |= 2
            kout = edge_count(x + b, "out") if b else t1hi
            if kout < t1lo or kout >= t1hi:
                flag |= 1
            if flag >= 3:
                continue
            if flag == 1:
                b = None
            if flag == 2:
                a = None
            if b:
                node[IDX[b]] = min(node[IDX[b]] + cnt, 65535)
            if a:
                node[4 + IDX[a]] = min(node[4 + IDX[a]] + cnt, 65535)
            node[8] = min(node[8] + cnt, 65535)
            have = True
        if have:
            g[x] = node
    return g


def edge_filtered(reads, k, lo, hi):
    """The block-construct driver's graph: canonical k+1-mer windows counted
    over the reads (saturating at 65535), an edge kept when its k+1-mer count
    is in [lo, hi), a record kept when one of its edges is; nodes are the
    k-mers of kept records with the kept edges (counts of kept records)."""
    kp1 = kp1_counts(reads, k)

    def count(e):
        c = kp1[e] + (kp1[rc(e)] if rc(e) != e else 0)
        return min(c, 65535)

    g = {}
    k2 = defaultdict(int)
    for x, a, b in records(reads, k):
        k2[(x, a, b)] += 1
    for (x, a, b), c in k2.items():
        a_ok = a is not None and lo <= count(a + x) < hi
        b_ok = b is not None and lo <= count(x + b) < hi
        if not a_ok and not b_ok:
            continue
        node = g.setdefault(x, [0] * 9)
        cnt = min(c, 65535)
        if b_ok:
            node[IDX[b]] = min(node[IDX[b]] + cnt, 65535)
        if a_ok:
            node[4 + IDX[a]] = min(node[4 + IDX[a]] + cnt, 65535)
        node[8] = min(node[8] + cnt, 65535)
    return g


def kplus1_pal(x, node):
    """IsKplus1Palindrome with string tests: an existing edge whose k+1-mer is
    an rc-palindrome, or the 'paired' form where the far k-1 bases are an
    rc-palindrome and the edge base equals the opposite end base."""
    k = len(x)
    if k % 2 == 0:
        return False
    for c in "ACGT":
        if node[IDX[c]] > 0:
            if is_pal(x + c) or (x[0] == c and is_pal(x[1:])):
                return True
        if node[4 + IDX[c]] > 0:
            if is_pal(c + x) or (x[-1] == c and is_pal(x[:-1])):
                return True
    return False


def classify(g):
    br, ch = set(), set()
    for x, n in g.items():
        nin = sum(1 for i in range(4, 8) if n[i] > 0)
        nout = sum(1 for i in range(4) if n[i] > 0)
        pal = is_pal(x) or kplus1_pal(x, n)
        if nin >= 2 or nout >= 2 or pal:
            br.add(x)
        else:
            ch.add(x)
    return br, ch


def walk(g):
    """Walk chains of non-branch nodes.  Returns (paths, cycle_nodes): every
    chain as the list of its oriented k-mers in walking order, and the set of
    canonical k-mers on closed walks.

    Works on oriented k-mers: the successor of oriented k-mer u is the unique
    out neighbour of u (in u's orientation) if u has exactly one, it is a
    chain node, and u is its unique predecessor; otherwise u ends its chain.
    Nodes with no edges at all are isolated and dropped; closed walks are
    cycles and dropped."""
    br, ch = classify(g)

    def succ(u):
        o = out_neighbours(g, u)
        if len(o) != 1:
            return None
        v = o[0]
        if canon(v) not in ch:
            return None
        return v

    seen = set()
    paths = []
    cycles = set()
    for x in sorted(ch):
        if x in seen:
            continue
        n = g[x]
        if sum(n[:8]) == 0:
            seen.add(x)
            continue  # isolated
        # walk backwards to the start of the chain (or detect a cycle)
        u = x
        steps = 0
        cyc = False
        while True:
            p = succ(rc(u))
            if p is None:
                break
            u = rc(p)
            steps += 1
            if canon(u) == x:
                cyc = True
                break
            if steps > 10 * len(g) + 10:
                raise RuntimeError("walk did not terminate")
        if cyc:
            comp = set()
            v = x
            while canon(v) not in comp:
                comp.add(canon(v))
                v = succ(v)
            seen |= comp
            cycles |= comp
            continue
        path = [u]
        while True:
            v = succ(path[-1])
            if v is None:
                break
            path.append(v)
        for p in path:
            seen.add(canon(p))
        paths.append(path)
    return paths, cycles


def out_neighbours(g, u):
    """The out neighbours of oriented k-mer u (on u's strand)."""
    x = canon(u)
    n = g[x]
    if u == x:
        return [u[1:] + c for c in "ACGT" if n[IDX[c]] > 0]
    return [u[1:] + COMP[c] for c in "ACGT" if n[4 + IDX[c]] > 0]


def unitigs(g):
    """(sorted [(rep, seq)], n_cycle_nodes) of the walked chains: each chain
    in the orientation that starts with its representative (the smaller of
    its first k-mer and its last k-mer's reverse complement)."""
    paths, cycles = walk(g)
    result = []
    for path in paths:
        k = len(path[0])
        s = path[0] + "".join(p[-1] for p in path[1:])
        rep = min(s[:k], rc(s)[:k])
        seq = s if s[:k] == rep else rc(s)
        result.append((rep, seq))
    result.sort()
    return result, len(cycles)


def chain_fasta(g):
    us, _ = unitigs(g)
    return "\n".join(">%s\n%s" % (canon(rep), seq) for rep, seq in us)


def branch_fasta(g):
    br, _ = classify(g)
    return "".join(">%s\n%s\n" % (b, b) for b in sorted(br))


def branch_edges(g):
    """.branch.edges from the string graph: branch k-mers sorted, each with
    in A C G T, out A C G T and self counts, tab separated."""
    br, _ = classify(g)
    return "".join("\t".join([b] + [str(g[b][i]) for i in (4, 5, 6, 7, 0, 1, 2, 3, 8)]) + "\n" for b in sorted(br))


def chain_edges(g, k):
    """.chain.edges from the walked unitigs: per unitig (in the orientation
    that starts with its rep) the first k-mer L, the last k-mer R, the
    in-edge counts of L and the out-edge counts of R on that strand, and the
    float32 mean (truncated), min and max of the self counts of its k-mers;
    sorted by L, every field followed by a tab."""
    import numpy as np
    us, _ = unitigs(g)

    def in_counts(u):
        x = canon(u)
        n = g[x]
        return [n[4 + IDX[c]] if u == x else n[IDX[COMP[c]]] for c in "ACGT"]

    def out_counts(u):
        x = canon(u)
        n = g[x]
        return [n[IDX[c]] if u == x else n[4 + IDX[COMP[c]]] for c in "ACGT"]

    lines = []
    for _, seq in us:
        L, R = seq[:k], seq[-k:]
        selfs = [g[canon(seq[i:i + k])][8] for i in range(len(seq) - k + 1)]
        mean = int(np.float32(sum(selfs)) / np.float32(len(selfs)))
        vals = in_counts(L) + out_counts(R) + [mean, min(selfs), max(selfs)]
        lines.append((L, L + "\t" + R + "\t" + "".join(str(v) + "\t" for v in vals) + "\n"))
    return "".join(t for _, t in sorted(lines))


def _edge_index(u, out, c):
    """(canonical k-mer, counter index) of oriented u's edge on side out
    (else in) with base c."""
    x = canon(u)
    if u == x:
        return x, (IDX[c] if out else 4 + IDX[c])
    return x, (4 + IDX[COMP[c]] if out else IDX[COMP[c]])


def _erase_edge_both(g, u, out, c):
    x, i = _edge_index(u, out, c)
    if x in g:
        g[x][i] = 0
    if out:
        y, j = _edge_index(u[1:] + c, False, u[0])
    else:
        y, j = _edge_index(c + u[:-1], True, u[-1])
    if y in g:
        g[y][j] = 0


def _drop_isolated(g):
    for x in [x for x, n in g.items() if sum(n[:8]) == 0]:
        del g[x]


def _drop_nodes(g, xs):
    for x in xs:
        n = g.get(x)
        if n is None:
            continue
        for c in "ACGT":
            if n[IDX[c]] > 0:
                _erase_edge_both(g, x, True, c)
            if n[4 + IDX[c]] > 0:
                _erase_edge_both(g, x, False, c)
        del g[x]


def _total(g, u, out):
    x = canon(u)
    n = g[x]
    idx = range(0, 4) if (out == (u == x)) else range(4, 8)
    return min(65535, sum(n[i] for i in idx))


def chain_summaries(g):
    """One summary per walked chain: (branch before the 5' end or all-A,
    5' k-mer, 3' k-mer, branch after the 3' end or all-A, length in k-mers
    minus one, total count of edges into the 5' end, out of the 3' end),
    oriented so the 5' k-mer is the chain's representative."""
    paths, _ = walk(g)
    out = []
    for path in paths:
        k = len(path[0])
        s = path[0] + "".join(p[-1] for p in path[1:])
        if s[:k] != min(s[:k], rc(s)[:k]):
            path = [rc(u) for u in reversed(path)]
        t5, t3 = path[0], path[-1]
        ins = [rc(v) for v in out_neighbours(g, rc(t5))]
        outs = out_neighbours(g, t3)
        b5 = ins[0] if ins else "A" * k
        b3 = outs[0] if outs else "A" * k
        out.append((b5, t5, t3, b3, len(path) - 1, _total(g, t5, False), _total(g, t3, True)))
    return out


def clean(g, k, th2):
    """The graph after bruno's deadend / bubble cleaning loop, restated on
    strings: chains summarised by walking, a deadend is a chain with exactly
    one open end, shorter than 2k - 1, whose end counts are both below
    th2 + 1; bubbles are runs of chains between the same two branch points
    (on the strand whose 5' branch is smaller, ordered by length) whose
    neighbours differ in length by at most one and have an end count below
    th2 + 1, when a run holds two or more chains.  The edges tying them to
    their branch points are erased (both directions) when weaker than
    th2 + 1, edgeless nodes are dropped, cycles are dropped, and the loop
    repeats until nothing is found (or a round erases nothing).  Returns
    (graph, rounds)."""
    g = {x: list(n) for x, n in g.items()}
    min_freq = th2 + 1
    _, cyc = walk(g)
    _drop_nodes(g, cyc)
    _drop_isolated(g)
    rounds = 0

    def find(g):
        summ = chain_summaries(g)
        dead = [x for x in summ if x[4] < 2 * k - 1 and x[5] < min_freq and x[6] < min_freq
                and ((x[5] == 0) != (x[6] == 0))]
        norm = []
        for x in summ:
            if x[4] > 0 and x[5] > 0 and x[6] > 0:
                if x[0] > rc(x[3]):
                    x = (rc(x[3]), rc(x[2]), rc(x[1]), rc(x[0]), x[4], x[6], x[5])
                norm.append(x)
        norm.sort(key=lambda x: (x[0], x[3], x[4], x[1], x[2]))
        bub = []
        first = 0
        for i in range(1, len(norm) + 1):
            if i < len(norm):
                a, b = norm[i - 1], norm[i]
                if a[0] == b[0] and a[3] == b[3] and abs(a[4] - b[4]) <= 1 and \
                        min(a[5], a[6], b[5], b[6]) < min_freq:
                    continue
            if i - first > 1:
                bub += norm[first:i]
            first = i
        return dead, bub

    def erase_weak(g, q):
        hit = []
        for u, out, c in q:
            x, i = _edge_index(u, out, c)
            if x in g and 0 < g[x][i] < min_freq:
                hit.append((u, out, c))
        for u, out, c in hit:
            _erase_edge_both(g, u, out, c)
        return len(hit)

    dead, bub = find(g)
    while dead or bub:
        q = []
        for b5, t5, t3, b3, _, f5, f6 in dead:
            if f5 > 0:
                q.append((b5, True, t5[-1]))
            if f6 > 0:
                q.append((b3, False, t3[0]))
        erased = erase_weak(g, q)
        _drop_isolated(g)
        q = []
        for b5, t5, t3, b3, _, _, _ in bub:
            q.append((b5, True, t5[-1]))
            q.append((b3, False, t3[0]))
        erased += erase_weak(g, q)
        _drop_isolated(g)
        if erased == 0:
            break  # nothing changed: the same chains would be found again
        _, cyc = walk(g)
        _drop_nodes(g, cyc)
        _drop_isolated(g)
        dead, bub = find(g)
        rounds += 1
    return g, rounds
