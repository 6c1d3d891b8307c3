"""Read-set ingestion (mpoa_reads_load / mpoa_reads_parse, no GPU) against upstream's own parse
of the same files (tests/golden/reader_cases.json): upstream's -r 1 rows show every read's name,
comment and bases as codes, in input order, so each read the library returns must print the
same header and, without gaps, the same row. Also: many files as windows, on several threads,
equal the files parsed one by one; a missing file is an error."""
from __future__ import annotations

import gzip
import json
import os

import numpy as np
import pytest

from poa_amd import api

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "reader_cases.json")))["cases"]
CODE_CHAR = {ord(c): c for c in "ACGT"}
CODE_CHAR.update({ord(c): u for c, u in zip("acgtUu", "ACGTTT")})


def as_codes(seq: bytes) -> str:
    return "".join(CODE_CHAR.get(b, "N") for b in seq)


def upstream_rows(stdout: str):
    lines = stdout.split("\n")
    assert lines[-1] == ""
    return [(lines[i], lines[i + 1]) for i in range(0, len(lines) - 1, 2)]


def check_against(rs: api.ReadSet, stdout: str, lo: int = 0, hi: int | None = None) -> None:
    rows = upstream_rows(stdout)
    hi = rs.n_reads if hi is None else hi
    assert hi - lo == len(rows)
    for k, (head, row) in enumerate(rows):
        i = lo + k
        name, comment = rs.names[i], rs.comments[i]
        # output_rc_msa's header (graph.cpp:247-248)
        assert head == (f">{k + 1}" if not name else f">{name} {comment}"), k
        got = bytes(rs.seq[rs.seq_off[i]:rs.seq_off[i + 1]])
        assert as_codes(got) == row.replace("-", ""), k


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["case"])
def test_reader_matches_upstream(tmp_path, case):
    assert case["returncode"] == 0
    path = tmp_path / ("in.fq.gz" if case["gzip"] else "in.txt")
    data = case["text"].encode("latin-1")
    path.write_bytes(gzip.compress(data) if case["gzip"] else data)
    rs = api.ReadSet([str(path)])
    assert rs.n_windows == 1
    check_against(rs, case["stdout"])
    # the same text from memory
    mem = api.ReadSet(text=data)
    np.testing.assert_array_equal(mem.seq_off, rs.seq_off)
    np.testing.assert_array_equal(mem.seq, rs.seq)
    assert mem.names == rs.names and mem.comments == rs.comments


def test_reader_quirks():
    """Cases read from the fixtures' outputs, spelled out: a header ending in a blank and a
    carriage return keeps the lone \\r as its comment; a CRLF empty line before the first
    sequence line adds a \\r; blanks inside sequence lines are bases; an empty name; an empty
    read; a FASTQ record cut short is dropped with the file's end."""
    by = {c["case"]: c for c in CASES}
    rs = api.ReadSet(text=by["crlf"]["text"].encode())
    assert rs.names == ["r1", "r2", "r3", "r4"]
    assert rs.comments == ["comment\twith tabs", "\r", "", "plain comment"]
    r3 = bytes(rs.seq[rs.seq_off[2]:rs.seq_off[3]])
    assert r3.startswith(b"\r") and b"\r" not in r3[1:]
    rs = api.ReadSet(text=by["blanks"]["text"].encode())
    assert rs.names[:2] == ["b1", ""] and rs.comments[1] == "only a comment"
    assert b" " in bytes(rs.seq[rs.seq_off[0]:rs.seq_off[1]])
    assert rs.seq_off[5] - rs.seq_off[4] == 0  # b5 has no bases
    assert rs.names[5] == "b6" and rs.comments[5] == " two blanks"
    rs = api.ReadSet(text=by["fastq_cut"]["text"].encode())
    assert rs.names == ["q1", "q2", "q3", "q4"]


def test_many_files_as_windows(tmp_path):
    paths = []
    for k, c in enumerate(CASES * 3):
        p = tmp_path / f"w{k}.txt"
        p.write_bytes(c["text"].encode("latin-1"))
        paths.append(str(p))
    serial = api.ReadSet(paths, threads=1)
    pooled = api.ReadSet(paths, threads=8)
    assert serial.n_windows == pooled.n_windows == len(paths)
    for a in ("win_off", "seq_off", "seq", "name_off", "comment_off"):
        np.testing.assert_array_equal(getattr(serial, a), getattr(pooled, a))
    assert serial.names == pooled.names
    for w, c in enumerate(CASES * 3):
        check_against(pooled, c["stdout"], int(pooled.win_off[w]), int(pooled.win_off[w + 1]))


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(api.MpoaError) as e:
        api.ReadSet([str(tmp_path / "no_such.fa")])
    assert e.value.rc == -1 and "no_such.fa" in str(e.value)
