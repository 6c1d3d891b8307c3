"""The call-level C ABI (mpoa_graph_*, mpoa_poa, mpoa_alignment) against the oracle's mirror of
minipoa's functions: minipoa's main loop driven call by call on the device gives every read's
res_t exactly as the oracle's, and the text the upstream binary printed; single poa() calls on
rank windows match in res and score; a res that reuses a node at two separate positions builds
the same graph."""
from __future__ import annotations

import numpy as np
import pytest

import golden_util as gu
import oracle_lib as ol
from poa_amd import api

CASES = gu.load_cases()


def res_rows(res: np.ndarray) -> np.ndarray:
    """api RES_DTYPE records as an (n, 3) int32 array like the oracle's."""
    return np.stack([res["from"], res["base"], res["aligned_id"]], axis=1).astype(np.int32)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=gu.case_id)
def test_call_level_loop(ctx, case):
    print(f"calls {gu.case_id(case)}", flush=True)
    p = gu.parse_args(case["args"])
    params = api.Params(**p)
    seq, off, names, comments = gu.parse_fasta(case["fasta"])
    n = len(off) - 1
    dag = ol.Dag(p, seq, off)
    g = api.Graph(ctx, params)
    seeder = api.Seeder(params, seq, off)
    order = seeder.order()
    np.testing.assert_array_equal(order, dag.order())
    rid_to_ord = np.zeros(n, np.int32)
    for i in range(n):
        rid = int(order[i])
        rid_to_ord[rid] = i
        q = bytes(seq[off[rid]:off[rid + 1]])
        want = dag.alignment(rid, q)
        got = g.alignment(rid, q, seeder)
        np.testing.assert_array_equal(res_rows(got), want, err_msg=f"read {rid}")
        dag.add_path(i, want)
        g.add_path(i, got)
        op = 1 if i + 1 == n else 0
        dag.topsort(op)
        g.topsort(op)
    ro = rid_to_ord if params.progressive else None
    assert g.output(params.result, names, comments, ro) == case["stdout"]
    g.close()


def _graph_pair(ctx, p: dict, reads):
    """The oracle's and the device's graph after all reads of `reads`, sorted (op 0)."""
    params = api.Params(**p)
    dag, g = ol.Dag(p), api.Graph(ctx, params)
    for i in range(len(reads.off) - 1):
        q = bytes(reads.seq[reads.off[i]:reads.off[i + 1]])
        res = dag.alignment(i, q)
        np.testing.assert_array_equal(res_rows(g.alignment(i, q)), res)
        dag.add_path(i, res)
        g.add_path(i, res)
        dag.topsort(0)
        g.topsort(0)
    return dag, g


def _same_graph(a: dict, b: dict, ranks: bool = True) -> None:
    keys = ["base", "in_off", "in_src", "in_w", "out_off", "out_dst", "out_w", "par_id",
            "aligned"] + (["rank_to_id", "hlen", "tlen"] if ranks else [])
    for k in keys:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize("ab_band", [0, 1])
def test_poa_rank_windows(ctx, ab_band):
    from poa_amd import synth
    reads = synth.window(0xca11, 600, 10, 0.05, 0.03, 0.03, 0.7)
    dag, g = _graph_pair(ctx, {}, reads)
    want_g = dag.export()
    _same_graph(want_g, g.export())
    order = want_g["rank_to_id"]
    n = len(order)
    q = bytes(reads.seq[reads.off[3]:reads.off[4]])
    rng = np.random.default_rng(7)
    for t in range(12):
        r0, r1 = sorted(rng.integers(0, n, 2))
        a, b = sorted(rng.integers(0, len(q) + 1, 2))
        print(f"poa ranks [{r0}, {r1}] query [{a}, {b}) ab_band {ab_band}", flush=True)
        want, ws = dag.poa(int(order[r0]), int(order[r1]), q[a:b], bool(ab_band))
        got, gs = g.poa(int(order[r0]), int(order[r1]), q[a:b], bool(ab_band))
        assert gs == ws
        np.testing.assert_array_equal(res_rows(got), want)
    # a window that ends before it begins: minipoa writes to an empty vector; here an error
    with pytest.raises(api.MpoaError) as e:
        g.poa(int(order[n - 1]), int(order[1]), q[:10], bool(ab_band))
    assert e.value.rc == -5
    g.close()


@pytest.mark.gpu
def test_add_path_repeated_node(ctx):
    """Node 2 visited twice with node 3 between: the lane-parallel add_path must hand this
    read to the one-lane walk; the graph equals the oracle's."""
    dag, g = ol.Dag({"result": 1}), api.Graph(ctx, api.Params(result=1))
    r0 = np.array([[0, 4, -1], [-1, 0, -1], [-1, 1, -1], [-1, 2, -1]], np.int32)
    r1 = np.array([[0, 4, -1], [2, 0, -1], [3, 1, -1], [2, 0, -1], [4, 2, -1]], np.int32)
    for i, r in enumerate((r0, r1)):
        dag.add_path(i, r)
        g.add_path(i, r)
        dag.topsort(0)
        g.topsort(0)
    _same_graph(dag.export(), g.export(), ranks=False)
    g.close()


@pytest.mark.gpu
def test_add_path_rejects_unknown_nodes(ctx):
    g = api.Graph(ctx, api.Params())
    with pytest.raises(api.MpoaError) as e:
        g.add_path(0, np.array([[0, 4, -1], [7, 0, -1]], np.int32))
    assert e.value.rc == -1
    with pytest.raises(api.MpoaError):
        g.add_path(1, np.array([[0, 4, -1]], np.int32))  # reads are added in order
    g.close()
