"""mpoa_align_reads_noalloc (the allocation-free entry point) through the C ABI against
mpoa_align_reads on the same windows: two calls in a row, the second with a smaller window,
in consensus and MSA output; a window too large for the fixed scratch is refused on the host.
Device buffers come from the HIP runtime the library itself uses (ctypes), not from the entry
point."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from poa_amd import api, synth

E_SCRATCH = -9


class Hip:
    def __init__(self):
        api.lib()  # load libmpoa.so first: the runtime below is the one it links
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, n: int) -> ctypes.c_void_p:
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(n, 16)) == 0
        return p

    def put(self, arr: np.ndarray) -> ctypes.c_void_p:
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0
        return p

    def get(self, p, n: int, dtype) -> np.ndarray:
        out = np.zeros(n, dtype)
        if out.nbytes:
            assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0
        return out


def run_noalloc(hip: Hip, ctx: api.Context, params: api.Params, reads, out_cap: int):
    L = api.lib()
    L.mpoa_align_reads_noalloc.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(api.Params), ctypes.c_int32, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
        ctypes.c_void_p]
    seq = np.ascontiguousarray(reads.seq, np.uint8)
    off = np.ascontiguousarray(reads.off, np.int64)
    d_seq, d_off = hip.put(seq), hip.put(off)
    d_out, d_res = hip.alloc(out_cap), hip.put(np.full(2, 12345, np.int64))
    try:
        rc = L.mpoa_align_reads_noalloc(ctx.h, ctypes.byref(params), len(off) - 1, d_seq, d_off,
                                        off.ctypes.data_as(ctypes.c_void_p), d_out, out_cap,
                                        d_res, None)
        if rc != 0:
            return rc, None, None
        assert hip.h.hipDeviceSynchronize() == 0
        status, n = (int(v) for v in hip.get(d_res, 2, np.int64))
        body = hip.get(d_out, n, np.uint8).tobytes().decode() if status == 0 else None
        return rc, status, body
    finally:
        for p in (d_seq, d_off, d_out, d_res):
            hip.h.hipFree(p)


def expected_body(text: str, result: int) -> str:
    lines = text.splitlines()
    if result == 0:
        assert lines[0] == ">consensus sequence"
        return lines[1]
    return "".join(l for l in lines if not l.startswith(">"))


@pytest.mark.gpu
@pytest.mark.parametrize("result", [0, 1])
def test_noalloc_matches_align_reads(ctx, result):
    hip = Hip()
    params = api.Params(result=result)
    windows = [synth.window(0x51a7, 900, 30, 0.05, 0.03, 0.03, 0.7),
               synth.window(0x51a8, 300, 12, 0.08, 0.04, 0.04, 0.7)]
    for i, reads in enumerate(windows):
        print(f"noalloc result={result} call {i}: {len(reads.off) - 1} reads", flush=True)
        want = expected_body(ctx.align_reads(params, reads.seq, reads.off), result)
        rc, status, body = run_noalloc(hip, ctx, params, reads, 1 << 20)
        assert rc == 0 and status == 0
        assert body == want


@pytest.mark.gpu
def test_noalloc_refuses_oversized_window(ctx):
    hip = Hip()
    reads = synth.config1()  # ~67k estimated nodes x 1,092 columns: beyond the fixed scratch
    rc, _, _ = run_noalloc(hip, ctx, api.Params(), reads, 1 << 16)
    assert rc == E_SCRATCH


@pytest.mark.gpu
def test_noalloc_small_output_buffer(ctx):
    hip = Hip()
    reads = synth.window(0x51a9, 400, 8, 0.05, 0.03, 0.03, 0.7)
    rc, status, _ = run_noalloc(hip, ctx, api.Params(), reads, 16)
    assert rc == 0 and status == -8  # the consensus does not fit 16 bytes
