"""The oracle's call-level mirror (orc_dag_*) pinned against the golden fixtures: minipoa's
main loop driven call by call through it prints exactly what the upstream binary printed, and
a res that reuses an existing node at two separate positions gives the graph the upstream
add_path rules give it (one more in-edge per visit, weights summed)."""
from __future__ import annotations

import numpy as np
import pytest

import golden_util as gu
import oracle_lib as ol

CASES = gu.load_cases()


@pytest.mark.parametrize("case", CASES, ids=gu.case_id)
def test_call_level_loop_matches_fixture(case):
    p = gu.parse_args(case["args"])
    seq, off, names, comments = gu.parse_fasta(case["fasta"])
    dag = ol.Dag(p, seq, off)
    text, _ = ol.run_loop(dag, len(off) - 1, seq, off, ol.oparams(p).result, names, comments,
                          order=dag.order())
    assert text == case["stdout"]


def test_repeated_node_in_res():
    """res visiting node X, then Y, then X again: add_adj's rules (node.h:23-61) on each edge
    in res order, checked by hand."""
    dag = ol.Dag({})
    # read 0: source, A, C, G (ids 2, 3, 4)
    dag.add_path(0, np.array([[0, 4, -1], [-1, 0, -1], [-1, 1, -1], [-1, 2, -1]], np.int32))
    dag.topsort(0)
    # read 1 walks 2 -> 3 -> 2 -> 4: node 2 twice, with node 3 between
    dag.add_path(1, np.array([[0, 4, -1], [2, 0, -1], [3, 1, -1], [2, 0, -1], [4, 2, -1]],
                             np.int32))
    g = dag.export()
    ins = {v: list(zip(g["in_src"][g["in_off"][v]:g["in_off"][v + 1]],
                       g["in_w"][g["in_off"][v]:g["in_off"][v + 1]])) for v in range(5)}
    outs = {v: list(zip(g["out_dst"][g["out_off"][v]:g["out_off"][v + 1]],
                        g["out_w"][g["out_off"][v]:g["out_off"][v + 1]])) for v in range(5)}
    assert ins[2] == [(0, 2), (3, 1)]      # 0->2 twice, then the new edge 3->2
    assert outs[2] == [(3, 2), (4, 1)]     # 2->3 twice, then the new edge 2->4
    assert ins[3] == [(2, 2)] and outs[3] == [(4, 1), (2, 1)]
    assert ins[4] == [(3, 1), (2, 1)] and outs[4] == [(1, 2)]
