"""The oracle against the reference program's own outputs (golden fixtures
made by tests/golden/make_golden.py from the reference compiled from its
sources): bit-exact records, images and randomized border velocities."""
import ctypes

import numpy as np
import pytest


@pytest.mark.parametrize("name", ["small_mod", "o4_mod"])
def test_modeling_records_bit_exact(oracle, cases, golden, name):
    p, fx = cases[name], golden(name)
    rec = oracle.pipeline_modeling(p, fx["vel"])
    assert rec.shape == fx["records"].shape
    assert np.array_equal(rec.view(np.uint32), fx["records"].view(np.uint32))


@pytest.mark.parametrize("bc", ["hbc", "rbc"])
def test_border_velocity_bit_exact(oracle, cases, golden, bc):
    vel = golden("small_mod")["vel"]
    ve = oracle.extend_velocity(vel, cases["small_" + bc]["blen"], bc)
    assert np.array_equal(ve, golden("small_" + bc)["vpe"])


@pytest.mark.parametrize("bc", ["hbc", "rbc"])
def test_migration_image_bit_exact(oracle, cases, golden, bc):
    p = cases["small_" + bc]
    img = oracle.pipeline_migration(p, golden("small_" + bc)["vpe"], golden("small_mod")["records"])
    ref = golden("small_" + bc)["image"]
    assert np.abs(ref).max() > 0.1
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_glibc_rand_matches_libc(oracle):
    libc = ctypes.CDLL("libc.so.6")
    for seed in (1, 7):
        r = oracle.GlibcRand(seed)
        libc.srand(seed)
        assert all(r.next() == libc.rand() for _ in range(20000))


def test_oracle_fixture_differs_from_f64_only_by_rounding(oracle, cases, golden):
    """The FP64 oracle is the same algorithm: its records agree with the
    reference FP32 records to single-precision accuracy."""
    p, fx = cases["small_mod"], golden("small_mod")
    r64 = oracle.pipeline_modeling(p, fx["vel"], np.float64)
    r32 = fx["records"].astype(np.float64)
    rel = np.linalg.norm(r64 - r32) / np.linalg.norm(r64)
    assert rel < 1e-4, rel
