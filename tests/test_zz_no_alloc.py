"""sxc_align_pair_static (no device allocation) against sxc_align_pair, through the C ABI:
two calls in a row, the second with a smaller pair, both equal to the allocating entry
point's records and to the oracle's."""
import ctypes as C

import numpy as np
import pytest

import oracle
from satsuma2_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _call(ctx, fn, pair, prm):
    from satsuma2_amd.api import _pairs_array

    arr, keep = _pairs_array([pair])
    cap = 8192
    out = np.zeros(cap, _lib.RESULT_DTYPE)
    n = C.c_int32(0)
    _lib.check(getattr(ctx.L, fn)(ctx.handle, arr, C.byref(prm), out.ctypes.data_as(C.POINTER(_lib.Result)), cap,
                                  C.byref(n)))
    return out[: n.value].copy()


def test_static_pair_equals_allocating_pair(golden):
    from satsuma2_amd.api import ChunkPair, Context, SlaveParams

    ctx = Context(0)
    g = golden("single_pair.npz")
    rng = np.random.default_rng(31)
    small_t = synth.random_seq(rng, 1000, 0.41)
    small_q = synth.random_seq(rng, 800, 0.41)
    small_q[100:600] = synth.mutate(rng, small_t[300:800], 0.9)
    b = golden("bdhv_pairs.npz")
    cases = [
        (ChunkPair(g["t"].tobytes(), g["q"].tobytes()), SlaveParams(target_total=4096.0)),   # N = 8192
        (ChunkPair(small_t.tobytes(), small_q.tobytes()), SlaveParams(t_chunk=1024, target_total=1e5)),  # N = 2048
        (ChunkPair(b["bdhv3_t"].tobytes(), b["bdhv3_q"].tobytes()), SlaveParams(t_chunk=2048, target_total=4096.0 * 64)),
    ]
    for rep in range(2):
        for pair, sp in cases:
            prm = sp.to_c()
            got = _call(ctx, "sxc_align_pair_static", pair, prm)
            want = _call(ctx, "sxc_align_pair", pair, prm)
            assert len(want) > 0
            assert got.tobytes() == want.tobytes(), (rep, sp)
            ref = oracle.align_batch([dict(tseq=pair.tseq, qseq=pair.qseq)],
                                     oracle.make_params(N=prm.N, target_total=sp.target_total))[0]
            for f in ("qstart", "tstart", "len", "reverse", "ident"):
                assert np.array_equal(got[f], ref[f]), f
    ctx.close()
