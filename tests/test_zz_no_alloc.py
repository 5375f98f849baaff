"""rtmhip_batch_steps_noalloc (no device allocation, the shot table staged in
the library's fixed device array) against a batch created with
rtmhip_batch_create over the same shots and chunks, bit for bit: wavefields,
UPB history and image of every shot after every chunk.

Two runs in a row share the staging array: the first with three shots, the
second with a smaller input (two shots, the smaller grids), so stale table
entries of the first run must not matter.  Too many shots are rejected.
"""
import numpy as np
import pytest

from test_gpu_batch import BWD, FWD, Shots, _compare, _run_batch

pytestmark = pytest.mark.gpu


def _run_noalloc(gpu, sh, members, fwd_chunks, bwd_chunks):
    states = []
    for it0, n in fwd_chunks:
        gpu.batch_steps_noalloc(members, sh.coefs, sh.dg, +1, it0, n)
        states.append([Shots.state(m) for m in members])
    for hi, n in bwd_chunks:
        gpu.batch_steps_noalloc(members, sh.coefs, sh.dg, -1, hi, n)
        states.append([Shots.state(m) for m in members])
    return states


@pytest.mark.parametrize("fmt", ["f32", "f16"])
def test_noalloc_matches_batch_two_calls(gpu, fmt):
    for keep in ([0, 1, 2], [1, 0]):  # the second run: fewer shots, smaller grids
        print("noalloc", fmt, "shots", keep, flush=True)
        sh = Shots(gpu, fmt)
        sh.host = [sh.host[k] for k in keep]
        ref = sh.device()
        want = _run_batch(gpu, sh, ref, FWD, BWD, (0, 0, 0), 8)
        sh.free(ref)
        mem = sh.device()
        got = _run_noalloc(gpu, sh, mem, FWD, BWD)
        sh.free(mem)
        assert np.abs(want[-1][0][5]).max() > 0
        _compare(want, got)


def test_noalloc_rejects_too_many_shots(gpu):
    sh = Shots(gpu, "f32")
    mem = sh.device()
    try:
        many = [mem[0]] * 65
        with pytest.raises(gpu.RtmError, match="at most 64"):
            gpu.batch_steps_noalloc(many, sh.coefs, sh.dg, +1, 0, 1)
        with pytest.raises(gpu.RtmError, match="direction"):
            gpu.batch_steps_noalloc(mem, sh.coefs, sh.dg, 0, 0, 1)
    finally:
        sh.free(mem)
