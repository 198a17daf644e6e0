"""The reference applies a list-ranking round's updates in hash-table (and
MPI rank) order.  Its outputs must not depend on that order: the oracle is
run with shuffled application orders and every output must be identical."""
import numpy as np
import pytest

from helpers import pack, random_reads, thresholds_L
from oracle.oracle import Oracle

CASES = [(1, 300, 40, 30, 0.0, 7), (3, 2000, 300, 50, 0.01, 11), (4, 200, 200, 25, 0.02, 5),
         (6, 4000, 400, 100, 0.01, 31), (8, 300, 60, 40, 0.01, 8), (9, 3000, 300, 120, 0.01, 63)]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("variant", [0, 1])
def test_outputs_do_not_depend_on_update_order(case, variant):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    for tl in (None, 2):
        outs = set()
        for order_seed in range(0, 12):
            o = Oracle(k, None if tl is None else np.array(thresholds_L(tl), np.uint64), 1)
            o.add_reads(b, off)
            o.set_order_seed(order_seed)
            o.run(variant)
            outs.add((o.chain_fasta(), o.chain_components(), o.stat("iterations"),
                      o.stat("n_cycle_nodes"), o.stat("n_isolated")))
        assert len(outs) == 1
