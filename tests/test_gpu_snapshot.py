"""Snapshot RTM through the C ABI (rtmhip_snap_forward_steps /
rtmhip_snap_backward_steps) against the oracle's snapshot RTM
(pyoracle.snapshot_migrate_shot), bit for bit, for every wavefield format,
both snapshot formats and several snapshot intervals, with both passes split
into chunks.  The border is odd (11), so the upper taper's edge and the
imaging range start inside a z pair of the two-z kernels.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

B, NT = 11, 26
GRID = (36, 42, 70)
SRC = (17, 19, 2)                      # interior coordinates
RCV = (1, 2, 3, 3, 4, 4, 5)           # start x, y, depth, dist x, y, count x, y


def _inputs(seed=3, b=B, grid=GRID):
    rng = np.random.default_rng(seed)
    V = rng.uniform(0.0, 0.15, grid).astype(np.float32)
    coefs = np.zeros(27, np.float32)
    c8 = np.array([-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    for d, h in enumerate((10.0, 11.0, 9.0)):
        coefs[d * 9:(d + 1) * 9] = (c8 / (h * h)).astype(np.float32)
    g = np.exp(-(np.sqrt(-np.log(0.9)) / b * np.arange(b)) ** 2).astype(np.float32)
    ricker = rng.uniform(-1, 1, NT).astype(np.float32)
    rcv = rng.uniform(-1, 1, (RCV[5], RCV[6], NT)).astype(np.float32)
    return V, coefs, g, ricker, rcv


@pytest.mark.parametrize("fusion", [-1, 1, 0], ids=["tdefault", "tfused", "tsep"])
@pytest.mark.parametrize("sfmt, k", [("f32", 1), ("f16", 1), ("f32", 3), ("f16", 4)])
@pytest.mark.parametrize("wfmt", ["f32", "f16", "bf16"])
def test_snapshot_rtm_matches_oracle(gpu, oracle, wfmt, sfmt, k, fusion):
    print("snapshot rtm", wfmt, sfmt, k, fusion, flush=True)
    gpu.set_taper_fusion(wfmt, fusion)
    V, coefs, g, ricker, rcv = _inputs()
    wd = gpu.WAVE_FORMATS[wfmt][1]
    shot = gpu.make_shot(*GRID, B, 8, NT, NT, tuple(c + B for c in SRC), gpu.RcvGeom(*RCV), wfmt)
    dV, dg, dsrc, drcv = (gpu.DeviceArray.from_host(a) for a in (V, g, ricker, rcv))
    S, R = gpu.FieldPair(GRID, wd), gpu.FieldPair(GRID, wd)
    img = gpu.DeviceArray(GRID).zero()
    snaps = gpu.snap_buffer(shot, sfmt, k)
    assert snaps.shape[0] == (NT - 1) // k + 1
    try:
        for it0, n in ((0, 11), (11, NT - 11)):
            gpu.snap_forward_steps(shot, dV, coefs, dg, dsrc, S, snaps, sfmt, k, it0, n)
        for hi, n in ((NT - 1, 12), (NT - 13, NT - 12)):
            gpu.snap_backward_steps(shot, dV, coefs, dg, R, snaps, sfmt, k, drcv, img, hi, n)
        got = img.download()
    finally:
        gpu.set_taper_fusion(wfmt, -1)
        for d in (dV, dg, dsrc, drcv, S, R, img, snaps):
            d.free()
    want = oracle.snapshot_migrate_shot(V, coefs, g, ricker, B, 8, NT, SRC,
                                        oracle.RcvGeom(*RCV), rcv, k, sfmt,
                                        None if wfmt == "f32" else wfmt)
    assert np.abs(want).max() > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("sfmt", ["f32", "f16"])
@pytest.mark.parametrize("wfmt", ["f32", "f16", "bf16"])
def test_snapshot_rtm_aligned_rows_matches_oracle(gpu, oracle, wfmt, sfmt):
    """Border 12 and nz 72: every snapshot row starts on a 4-cell boundary,
    so the store takes its four-cells-per-lane form."""
    print("snapshot rtm aligned rows", wfmt, sfmt, flush=True)
    b, grid, k = 12, (36, 42, 72), 2
    assert b % 4 == 0 and (grid[2] - 2 * b) % 4 == 0 and grid[2] % 4 == 0
    V, coefs, g, ricker, rcv = _inputs(5, b, grid)
    wd = gpu.WAVE_FORMATS[wfmt][1]
    shot = gpu.make_shot(*grid, b, 8, NT, NT, tuple(c + b for c in SRC), gpu.RcvGeom(*RCV), wfmt)
    dV, dg, dsrc, drcv = (gpu.DeviceArray.from_host(a) for a in (V, g, ricker, rcv))
    S, R = gpu.FieldPair(grid, wd), gpu.FieldPair(grid, wd)
    img = gpu.DeviceArray(grid).zero()
    snaps = gpu.snap_buffer(shot, sfmt, k)
    try:
        gpu.snap_forward_steps(shot, dV, coefs, dg, dsrc, S, snaps, sfmt, k, 0, NT)
        gpu.snap_backward_steps(shot, dV, coefs, dg, R, snaps, sfmt, k, drcv, img, NT - 1, NT)
        got = img.download()
    finally:
        for d in (dV, dg, dsrc, drcv, S, R, img, snaps):
            d.free()
    want = oracle.snapshot_migrate_shot(V, coefs, g, ricker, b, 8, NT, SRC,
                                        oracle.RcvGeom(*RCV), rcv, k, sfmt,
                                        None if wfmt == "f32" else wfmt)
    assert np.abs(want).max() > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_snapshot_rejects_bad_arguments(gpu):
    V, coefs, g, ricker, rcv = _inputs()
    shot = gpu.make_shot(*GRID, B, 8, NT, NT, tuple(c + B for c in SRC), gpu.RcvGeom(*RCV))
    dV, dg, dsrc = (gpu.DeviceArray.from_host(a) for a in (V, g, ricker))
    S = gpu.FieldPair(GRID)
    snaps = gpu.snap_buffer(shot, "f32", 2)
    try:
        with pytest.raises(gpu.RtmError, match="interval"):
            gpu.snap_forward_steps(shot, dV, coefs, dg, dsrc, S, snaps, "f32", 0, 0, 1)
        with pytest.raises(gpu.RtmError, match="time range"):
            gpu.snap_forward_steps(shot, dV, coefs, dg, dsrc, S, snaps, "f32", 2, NT - 1, 2)
        c = np.ascontiguousarray(coefs)
        rc = gpu.load().rtmhip_snap_forward_steps(
            gpu.C.byref(shot), dV.ptr, c.ctypes.data_as(gpu.C.c_void_p), dg.ptr, dsrc.ptr,
            gpu.C.byref(S.p), gpu.C.byref(S.pp), snaps.ptr, 7, 2, 0, 1)
        assert rc == gpu.ERR_UNSUPPORTED
    finally:
        for d in (dV, dg, dsrc, S, snaps):
            d.free()
