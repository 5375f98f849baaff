"""Transform-coded snapshots (rtmhip_snap_dct_*, rtmhip_dct_roundtrip).

* The coder against a float64 numpy restatement of it (8 x 8 x 32 bricks of
  the grid, orthonormal 3D DCT-II, the lowest power-of-two threshold keeping
  at most K coefficients of the whole snapshot, FP16 values over the brick's
  largest magnitude): same kept count and decoded fields within FP16
  quantisation.
* Exact properties: scaling a field by 2^k scales the decoded field by 2^k
  bit for bit; keeping every coefficient reproduces the field within FP16
  quantisation; cells outside the imaging range are never written.
* Snapshot RTM with coded snapshots: one call equals the same run split
  into chunks bit for bit (decoding does not depend on the order in which
  bricks were packed), for every wave format, and the image stays close to
  the exact-FP32-snapshot image.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dct_matrix(n):
    i = np.arange(n)
    d = np.cos(np.pi * (2 * i[None, :] + 1) * i[:, None] / (2 * n)) * np.sqrt(2.0 / n)
    d[0] /= np.sqrt(2.0)
    return d


def _bricks(f, b):
    """[nbr, 8, 8, 32] bricks of f (x, y, z order of bricks; cells outside
    z in [b, nz-b) or past the grid are 0)."""
    nx, ny, nz = f.shape
    px, py, pz = -(-nx // 8) * 8, -(-ny // 8) * 8, -(-nz // 32) * 32
    g = np.zeros((px, py, pz))
    g[:nx, :ny, b:nz - b] = f[:, :, b:nz - b]
    g = g.reshape(px // 8, 8, py // 8, 8, pz // 32, 32).transpose(0, 2, 4, 1, 3, 5)
    return g.reshape(-1, 8, 8, 32), (px, py, pz)


def _unbricks(br, shape, padded, b):
    nx, ny, nz = shape
    px, py, pz = padded
    g = br.reshape(px // 8, py // 8, pz // 32, 8, 8, 32).transpose(0, 3, 1, 4, 2, 5)
    g = g.reshape(px, py, pz)[:nx, :ny, :nz]
    out = np.zeros(shape)
    out[:, :, b:nz - b] = g[:, :, b:nz - b]
    return out


def coder(f, b, keep_div, basis=None):
    """float64 restatement of the device coder: (decoded field, kept).
    basis: (z 32x32, x 8x8, y 8x8) rows, default the DCT-II."""
    if basis is None:
        dz, dx, dy = _dct_matrix(32), _dct_matrix(8), _dct_matrix(8)
    else:
        dz, dx, dy = basis
    br, padded = _bricks(np.asarray(f, np.float64), b)
    c = np.einsum("ui,vj,wk,nijk->nuvw", dx, dy, dz, br)
    c32 = c.astype(np.float32)
    bits = c32.view(np.uint32) & 0x7FFFFFFF
    K = max(1, f.shape[0] * f.shape[1] * (f.shape[2] - 2 * b) // keep_div)
    hist = np.bincount((bits[bits != 0] >> 23).astype(np.int64), minlength=256)
    cum, thr = 0, 1
    for e in range(255, -1, -1):
        if cum + hist[e] > K:
            thr = (e + 1) << 23
            break
        cum += hist[e]
    keep = (bits >= thr) & (bits != 0)
    scale = np.abs(c32).reshape(len(c32), -1).max(axis=1)
    safe = np.where(scale > 0, scale, np.float32(1.0))[:, None, None, None]
    q = (c32 / safe).astype(np.float16).astype(np.float64)
    q = np.where(keep, q * scale[:, None, None, None].astype(np.float64), 0.0)
    rec = np.einsum("ui,vj,wk,nuvw->nijk", dx, dy, dz, q)
    return _unbricks(rec, f.shape, padded, b), int(keep.sum())


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _wavefield(shape, seed):
    """A smooth field with a few energetic packets (wavefield-like sparsity)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal(shape), 1.2) * 0.05
    for _ in range(4):
        c = [rng.integers(8, n - 8) for n in shape]
        x, y, z = np.ogrid[:shape[0], :shape[1], :shape[2]]
        r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        f += np.cos(0.9 * np.sqrt(r2)) * np.exp(-r2 / 40.0) * rng.uniform(1, 20)
    return f.astype(np.float32)


# the last case: the imaging range is exactly one layer of bricks and the
# budget covers every coefficient, so only FP16 quantisation remains
@pytest.mark.parametrize("shape, b, keep_div", [((40, 36, 90), 11, 16), ((70, 66, 100), 16, 32),
                                                ((48, 48, 96), 32, 1)])
def test_coder_matches_restatement(gpu, shape, b, keep_div):
    f = _wavefield(shape, 3)
    d = gpu.DeviceArray.from_host(f)
    try:
        got, kept = gpu.dct_roundtrip(d, "f32", b, keep_div)
    finally:
        d.free()
    want, kept_np = coder(f, b, keep_div)
    inside = np.s_[:, :, b:shape[2] - b]
    assert np.isfinite(got).all()
    assert abs(kept - kept_np) <= max(2, kept_np // 10000), (kept, kept_np)
    assert _rel(got[inside], want[inside]) < 2e-4
    assert (got[:, :, :b] == 0).all() and (got[:, :, shape[2] - b:] == 0).all()
    err = _rel(got[inside], f[inside])
    print("keep 1/%d: kept %d, field rel-L2 %.2e" % (keep_div, kept, err))
    if keep_div == 1:
        assert err < 2e-3


def _random_basis(seed):
    """A random orthonormal separable basis, packed (z, x, y rows)."""
    rng = np.random.default_rng(seed)
    qs = [np.linalg.qr(rng.standard_normal((n, n)))[0] for n in (32, 8, 8)]
    return np.concatenate([q.ravel() for q in qs]).astype(np.float32)


@pytest.mark.parametrize("keep_div", [1, 16])
def test_coder_with_caller_basis_matches_restatement(gpu, keep_div):
    """A caller's basis (random orthonormal rows, nothing like a cosine)
    replaces the DCT in both transforms: the coder agrees with the
    restatement using the same matrices, and with every coefficient kept
    only FP16 quantisation remains."""
    from rtm3d_demo_amd import basis as bb
    shape, b = (48, 48, 96), 32 if keep_div == 1 else 11
    f = _wavefield(shape, 4)
    rb = _random_basis(11)
    d = gpu.DeviceArray.from_host(f)
    try:
        got, kept = gpu.dct_roundtrip(d, "f32", b, keep_div, basis=rb)
        dflt, _ = gpu.dct_roundtrip(d, "f32", b, keep_div)
    finally:
        d.free()
    want, kept_np = coder(f, b, keep_div, bb.unpack(rb))
    inside = np.s_[:, :, b:shape[2] - b]
    assert abs(kept - kept_np) <= max(2, kept_np // 10000), (kept, kept_np)
    assert _rel(got[inside], want[inside]) < 2e-4
    # the default basis is not what ran
    assert _rel(got[inside], dflt[inside]) > 1e-3 or keep_div == 1
    if keep_div == 1:
        assert _rel(got[inside], f[inside]) < 2e-3


def test_coder_rejects_a_basis_that_is_not_orthonormal(gpu):
    rb = _random_basis(12)
    rb[32 * 32 + 3] += 0.01          # one x row entry
    d = gpu.DeviceArray.from_host(_wavefield((40, 36, 90), 1))
    try:
        with pytest.raises(Exception, match="orthonormal"):
            gpu.dct_roundtrip(d, "f32", 11, 16, basis=rb)
    finally:
        d.free()


def test_coder_subnormal_bricks(gpu):
    """Bricks of subnormal magnitude beside normal ones (FP32 wavefields reach
    it far from the source) with a budget that keeps them: their scales are
    subnormal, whose reciprocals overflow; the codes stay finite.  (All of a
    snapshot's subnormal coefficients share the lowest exponent class, so a
    budget that cannot hold them all drops them together.)"""
    shape, b = (48, 48, 96), 32
    f = _wavefield(shape, 9)
    f[:16, :16, :] *= np.float32(1e-41)
    d = gpu.DeviceArray.from_host(f)
    try:
        got, kept = gpu.dct_roundtrip(d, "f32", b, 1)
    finally:
        d.free()
    assert np.isfinite(got).all()
    inside = np.s_[:, :, b:shape[2] - b]
    assert _rel(got[inside], f[inside]) < 2e-3
    tiny = np.abs(got[:16, :16, b:shape[2] - b])
    assert tiny.max() < 1e-38 and tiny.max() > 0


@pytest.mark.parametrize("fmt", ["f32", "f16", "bf16"])
def test_coder_power_of_two_scaling_is_exact(gpu, oracle, fmt):
    shape, b = (40, 44, 96), 12
    f = _wavefield(shape, 5)
    # values exact in the format, so that every scaled copy below is too
    if fmt == "f16":
        f = f.astype(np.float16).astype(np.float32)
    elif fmt == "bf16":
        f = (oracle.to_format(f, fmt).astype(np.uint32) << 16).view(np.float32)
    outs = []
    for k in (0, 4, 8):
        g = f * np.float32(2.0 ** k)
        if fmt != "f32":
            g = oracle.to_format(g, fmt)
        d = gpu.DeviceArray.from_host(g, gpu.WAVE_FORMATS[fmt][1])
        try:
            outs.append(gpu.dct_roundtrip(d, fmt, b, 16))
        finally:
            d.free()
    (o0, k0), (o4, k4), (o8, k8) = outs
    assert k0 == k4 == k8 and k0 > 0
    assert np.array_equal(o4, o0 * np.float32(16.0))
    assert np.array_equal(o8, o0 * np.float32(256.0))


# v^2 dt^2 of 2.0 .. 4.0 m^2 on ~10 m cells: Courant 0.14 .. 0.2, so the
# source and receiver fields travel far enough in NT steps to overlap (an
# image of fields that never meet is all coder noise)
B, NT = 12, 100
GRID = (40, 44, 96)
SRC = (20, 22, 2)
RCV = (1, 2, 3, 3, 4, 4, 5)


def _inputs(seed=7):
    rng = np.random.default_rng(seed)
    V = rng.uniform(2.0, 4.0, GRID).astype(np.float32)
    coefs = np.zeros(27, np.float32)
    c8 = np.array([-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    for d, h in enumerate((10.0, 11.0, 9.0)):
        coefs[d * 9:(d + 1) * 9] = (c8 / (h * h)).astype(np.float32)
    g = np.exp(-(np.sqrt(-np.log(0.9)) / B * np.arange(B)) ** 2).astype(np.float32)
    t = np.arange(NT) * 0.3
    ricker = ((1 - 2 * (t - 3) ** 2) * np.exp(-(t - 3) ** 2)).astype(np.float32) * 50
    rcv = rng.uniform(-1, 1, (RCV[5], RCV[6], NT)).astype(np.float32)
    return V, coefs, g, ricker, rcv


def _run(gpu, wfmt, k, keep_div, splits_f, splits_b, coded=True, basis=None, src=SRC,
         snaps_out=None):
    V, coefs, g, ricker, rcv = _inputs()
    wd = gpu.WAVE_FORMATS[wfmt][1]
    shot = gpu.make_shot(*GRID, B, 8, NT, NT, tuple(c + B for c in src), gpu.RcvGeom(*RCV), wfmt)
    dV, dg, dsrc, drcv = (gpu.DeviceArray.from_host(a) for a in (V, g, ricker, rcv))
    S, R = gpu.FieldPair(GRID, wd), gpu.FieldPair(GRID, wd)
    for f in (S.a, S.b, R.a, R.b):
        f.zero()
    img = gpu.DeviceArray(GRID).zero()
    store = (gpu.DctSnapshots(shot, k, keep_div, basis) if coded
             else gpu.snap_buffer(shot, "f32", k))
    try:
        it0 = 0
        for n in splits_f:
            if coded:
                store.forward_steps(dV, coefs, dg, dsrc, S, it0, n)
            else:
                gpu.snap_forward_steps(shot, dV, coefs, dg, dsrc, S, store, "f32", k, it0, n)
            it0 += n
        if snaps_out is not None:
            snaps_out.append(store.download())
        hi = NT - 1
        for n in splits_b:
            if coded:
                store.backward_steps(dV, coefs, dg, R, drcv, img, hi, n)
            else:
                gpu.snap_backward_steps(shot, dV, coefs, dg, R, store, "f32", k, drcv, img, hi, n)
            hi -= n
        return img.download()
    finally:
        for d in (dV, dg, dsrc, drcv, S, R, img, store):
            d.free()


@pytest.mark.parametrize("wfmt", ["f32", "f16", "bf16"])
def test_coded_snapshot_rtm_split_equals_whole_and_tracks_exact(gpu, wfmt):
    print("dct snapshot rtm", wfmt, flush=True)
    whole = _run(gpu, wfmt, 2, 8, (NT,), (NT,))
    split = _run(gpu, wfmt, 2, 8, (33, NT - 33), (45, NT - 45))
    assert np.array_equal(whole.view(np.uint32), split.view(np.uint32))
    exact = _run(gpu, "f32", 2, 8, (NT,), (NT,), coded=False)
    err = _rel(whole[:, :, B:-B], exact[:, :, B:-B])
    print("%s wavefields, DCT snapshots keep 1/8: image rel-L2 %.2e vs exact FP32 snapshots"
          % (wfmt, err))
    assert np.abs(exact).max() > 0
    # the 16-bit runs carry their wavefield storage error as well (the
    # yardstick propagates in FP32); measured 6.9e-3 / 8.4e-3 / 2.2e-2
    assert err < {"f32": 1e-2, "f16": 2e-2, "bf16": 6e-2}[wfmt]


def test_learned_basis_snapshot_rtm(gpu):
    """A basis learned from another shot's FP32 snapshots (the principal
    directions of their brick lines) codes this shot's snapshots: the image
    stays within the DCT's bound of the exact-snapshot image, and the learned
    basis compacts this shot's snapshot energy into the leading z
    coefficients at least as well as the DCT (the property it is trained
    for, measured on unseen data)."""
    from rtm3d_demo_amd import basis as bb
    other = []
    _run(gpu, "f32", 2, 8, (NT,), (1,), coded=False, src=(8, 30, 6), snaps_out=other)
    learner = bb.BasisLearner()
    for snap in other[0][NT // 4:]:
        learner.add(snap, B)
    klt = learner.basis()
    mine = []
    exact = _run(gpu, "f32", 2, 8, (NT,), (NT,), coded=False, snaps_out=mine)
    test = bb.BasisLearner()
    for snap in mine[0][NT // 4:]:
        test.add(snap, B)
    cz = test.cz
    dz, kz = bb.unpack(bb.dct_basis())[0], bb.unpack(klt)[0]
    e_dct = np.cumsum(np.einsum("ui,ij,uj->u", dz, cz, dz))
    e_klt = np.cumsum(np.einsum("ui,ij,uj->u", kz, cz, kz))
    print("leading-8 z energy on the unseen shot: DCT %.4f, learned %.4f"
          % (e_dct[7] / e_dct[-1], e_klt[7] / e_klt[-1]))
    assert e_klt[7] >= 0.98 * e_dct[7]
    errs = {}
    for name, basis in (("dct", None), ("learned", klt)):
        img = _run(gpu, "f32", 2, 8, (NT,), (NT,), basis=basis)
        errs[name] = _rel(img[:, :, B:-B], exact[:, :, B:-B])
    print("keep 1/8 image rel-L2 vs exact: DCT %.2e, learned %.2e" % (errs["dct"],
                                                                       errs["learned"]))
    assert errs["learned"] < 1e-2
