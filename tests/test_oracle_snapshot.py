"""The snapshot-RTM oracle (pyoracle.snapshot_migrate_shot) pinned without
restating it.

With an impulse source (non-zero at step 0 only) and no taper, the HBC
backward reconstruction of the source field is exact in exact arithmetic: the
reverse step p(t) = 2 p(t+1) - p(t+2) + v lap p(t+1) omits only source terms,
and there are none after step 0.  The reference's HBC image
(orc_migrate_shot, pinned to the reference program) is then the imaging
condition on the forward source field at every step -- exactly what snapshot
RTM with k = 1 computes.  In float64 the two must agree to rounding; a
snapshot taken one step early or late, or a receiver half out of order,
leaves an O(1) difference.  With k > 1 the image is the every-k-th-step sum
weighted by k: close to, but not equal to, the k = 1 image.
"""
import numpy as np
import pytest

NX, NY, NZ, B, NT = 30, 34, 40, 8, 60


def _case(oracle, dtype):
    rng = np.random.default_rng(1)
    vel = np.full((NX, NY, NZ), 2000.0, np.float32)
    vel[:, :, 20:] = 3000.0
    ve = oracle.extend_velocity(vel, B, "hbc")
    dt = np.float32(0.4 * 10 / 3000)
    coefs = oracle.stencil_coefs(8, 10.0, 10.0, 10.0, dtype)
    v2 = oracle.v2dt2(ve, float(dt) if dtype == np.float64 else dt, dtype)
    geom = oracle.RcvGeom(0, 0, 0, 2, 2, NX // 2, NY // 2)
    rcv = rng.uniform(-1, 1, (NX // 2, NY // 2, NT)).astype(dtype)
    return v2, coefs, geom, rcv, (NX // 2, NY // 2, 2)


def test_snapshot_k1_equals_exact_hbc_image_f64(oracle):
    v2, coefs, geom, rcv, shot = _case(oracle, np.float64)
    src = np.zeros(NT)
    src[0] = 1.0
    g = np.ones(B)
    hbc = oracle.migrate_shot(v2, coefs, g, src, B, 8, NT, NT, shot, geom, rcv, True)
    img = oracle.snapshot_migrate_shot(v2, coefs, g, src, B, 8, NT, shot, geom, rcv, 1)
    assert np.abs(hbc).max() > 0.1
    assert np.linalg.norm(img - hbc) / np.linalg.norm(hbc) < 1e-12


@pytest.mark.parametrize("k", [2, 3])
def test_snapshot_interval_weights_every_kth_step(oracle, k):
    """k > 1: the k-weighted sum over every k-th step approximates the k = 1
    image of a smooth source (a 25 Hz Ricker) but does not equal it."""
    v2, coefs, geom, rcv, shot = _case(oracle, np.float64)
    src = oracle.ricker(NT, float(np.float32(0.4 * 10 / 3000)), 25.0, np.float64)
    g = oracle.taper_table(B, 0.9, np.float64)
    rcv = np.zeros_like(rcv)
    rcv[:, :, NT // 2:] = 1.0  # a step in time: the receiver field is smooth too
    full = oracle.snapshot_migrate_shot(v2, coefs, g, src, B, 8, NT, shot, geom, rcv, 1)
    dec = oracle.snapshot_migrate_shot(v2, coefs, g, src, B, 8, NT, shot, geom, rcv, k)
    rel = np.linalg.norm(dec - full) / np.linalg.norm(full)
    assert 1e-6 < rel < 0.5, rel


def test_snapshot_f16_rounds_the_stored_field(oracle):
    """FP16 snapshots: the image changes by about FP16 rounding, no more."""
    v2, coefs, geom, rcv, shot = _case(oracle, np.float32)
    src = oracle.ricker(NT, np.float32(0.4 * 10 / 3000), np.float32(25.0))
    g = oracle.taper_table(B, np.float32(0.9))
    a = oracle.snapshot_migrate_shot(v2, coefs, g, src, B, 8, NT, shot, geom, rcv, 1)
    h = oracle.snapshot_migrate_shot(v2, coefs, g, src, B, 8, NT, shot, geom, rcv, 1, "f16")
    rel = np.linalg.norm(h.astype(np.float64) - a) / np.linalg.norm(a)
    assert 1e-6 < rel < 2e-3, rel
