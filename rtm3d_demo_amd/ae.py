"""Learned snapshot compression (rtmhip_ae_*): binding, training on snapshots
the propagator generates, and snapshot-based RTM with exact or compressed
source snapshots.

No reference counterpart: the reference keeps the source wavefield by HBC
reconstruction (rtm-3d/src/kernel/RTMFiniteDifferencesKernel_HBC.cpp:212-365)
or as exact snapshots.  Here a snapshot of the imaging region (all x, y;
z in [blen, nze - blen)) is cut into 32^3 bricks and each brick is stored as
2048 FP16 codes plus one FP32 scale (32x smaller than the FP32 snapshot).
Exact FP32 snapshots remain the fallback and the yardstick of the image error.
"""
import ctypes as C

import numpy as np

from . import hip

BRICK = 32768
CODES = 2048


def region_of(shape, blen):
    """The imaging region of an extended grid: all x, y; z in [blen, nze-blen)."""
    nx, ny, nz = shape
    return (0, 0, blen, nx, ny, nz - blen)


def region_bricks(region):
    x0, y0, z0, x1, y1, z1 = region
    return (-(-(x1 - x0) // 32)) * (-(-(y1 - y0) // 32)) * (-(-(z1 - z0) // 32))


class Codec:
    """The autoencoder on the device (weights, Adam state, workspace for
    `max_bricks` bricks per pass)."""

    def __init__(self, max_bricks=256, seed=1):
        L = hip.load()
        h = C.c_void_p()
        hip.check(L.rtmhip_ae_create(C.byref(h), max_bricks, seed))
        self.h = h
        self.nparams = L.rtmhip_ae_param_count()

    def params(self, which=0):
        out = np.empty(self.nparams, np.float32)
        hip.check(hip.load().rtmhip_ae_get_params(self.h, out.ctypes.data_as(C.c_void_p), which))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        hip.check(hip.load().rtmhip_ae_set_params(self.h, p.ctypes.data_as(C.c_void_p)))

    def gradient(self, bricks, n):
        loss = C.c_float(0.0)
        hip.check(hip.load().rtmhip_ae_gradient(self.h, bricks.ptr, n, C.byref(loss)))
        return loss.value, self.params(1)

    def train(self, bricks, nbricks, iters, batch=64, lr=1e-3, seed=7):
        losses = np.zeros(iters, np.float32)
        hip.check(hip.load().rtmhip_ae_train(self.h, bricks.ptr, nbricks, iters, batch,
                                             float(lr), seed,
                                             losses.ctypes.data_as(C.c_void_p)))
        return losses

    def encode(self, bricks, n, codes):
        hip.check(hip.load().rtmhip_ae_encode(self.h, bricks.ptr, n, codes.ptr))

    def decode(self, codes, scales, n, field, region):
        hip.check(hip.load().rtmhip_ae_decode(self.h, codes.ptr, scales.ptr, n, field.ptr,
                                              *field.shape, *region))

    def destroy(self):
        if self.h:
            hip.check(hip.load().rtmhip_ae_destroy(self.h))
            self.h = C.c_void_p()


def cut_bricks(field, region, bricks, scales):
    """Bricks (FP16, device, [n][32768]) and scales of a region of field."""
    hip.check(hip.load().rtmhip_ae_bricks(field.ptr, *field.shape, *region, bricks.ptr,
                                          scales.ptr))


def collect_bricks(snaps, slots, pool, row, scales):
    """Cuts the given slots of an F32 snapshot buffer
    ([nsnap][nxe][nye][nze-2blen], device) into FP16 bricks at rows row, ...
    of pool ([n][32768], device); scales ([bricks], device) is scratch.
    Returns the next free row."""
    region = (0, 0, 0) + tuple(snaps.shape[1:])
    nb = region_bricks(region)
    for j in slots:
        cut_bricks(snaps.slot(j), region, pool.view(row, nb), scales)
        row += nb
    return row


class SnapshotStore:
    """Source snapshots every `k` steps: 'exact' (FP32 copies of the imaging
    region's grid), 'ae' (codes + scales), or 'decimate' (the same storage
    without learning: 2x2x4 box averages in FP16, linear reconstruction;
    host-side, for comparison)."""

    def __init__(self, kind, shape, blen, codec=None):
        self.kind, self.shape, self.blen, self.codec = kind, shape, blen, codec
        self.region = region_of(shape, blen)
        self.nb = region_bricks(self.region)
        self.items = {}
        if kind == "ae":
            self.bricks = hip.DeviceArray((self.nb, BRICK), np.uint16)

    def put(self, it, field):
        if self.kind == "exact":
            d = hip.DeviceArray(self.shape)
            hip.check(hip.load().rtmhip_copy_device(d.ptr, field.ptr, d.nbytes))
            self.items[it] = d
        elif self.kind == "ae":
            scales = hip.DeviceArray((self.nb,))
            codes = hip.DeviceArray((self.nb, CODES), np.uint16)
            cut_bricks(field, self.region, self.bricks, scales)
            self.codec.encode(self.bricks, self.nb, codes)
            self.items[it] = (codes, scales)
        else:
            from_ = _region_view(field.download(), self.region)
            self.items[it] = _decimate(from_)

    def get(self, it, out):
        """Writes the snapshot of step it into out (FP32 device field)."""
        if self.kind == "exact":
            d = self.items[it]
            hip.check(hip.load().rtmhip_copy_device(out.ptr, d.ptr, d.nbytes))
        elif self.kind == "ae":
            codes, scales = self.items[it]
            self.codec.decode(codes, scales, self.nb, out, self.region)
        else:
            full = np.zeros(self.shape, np.float32)
            x0, y0, z0, x1, y1, z1 = self.region
            full[x0:x1, y0:y1, z0:z1] = self.items[it]
            out.upload(full)

    def nbytes(self):
        if self.kind == "exact":
            return sum(d.nbytes for d in self.items.values())
        if self.kind == "ae":
            return sum(c.nbytes + s.nbytes for c, s in self.items.values())
        return sum(a.size // 16 * 2 for a in self.items.values())

    def free(self):
        if self.kind == "exact":
            for d in self.items.values():
                d.free()
        elif self.kind == "ae":
            for c, s in self.items.values():
                c.free()
                s.free()
            self.bricks.free()
        self.items = {}


def _region_view(a, region):
    x0, y0, z0, x1, y1, z1 = region
    return a[x0:x1, y0:y1, z0:z1]


def _decimate(region_field, factors=(2, 2, 4)):
    """2x2x4 box averages stored in FP16, separable linear reconstruction at
    cell centres (16x fewer values than the region: the AE's storage)."""
    f = np.asarray(region_field, np.float32)
    shp = f.shape
    pad = [(0, (-s) % k) for s, k in zip(shp, factors)]
    g = np.pad(f, pad, mode="edge")
    c = g.reshape(g.shape[0] // factors[0], factors[0], g.shape[1] // factors[1], factors[1],
                  g.shape[2] // factors[2], factors[2]).mean(axis=(1, 3, 5))
    out = c.astype(np.float16).astype(np.float32)
    for ax, k in enumerate(factors):
        n = shp[ax]
        pos = (np.arange(n) + 0.5) / k - 0.5
        i0 = np.clip(np.floor(pos).astype(np.int64), 0, out.shape[ax] - 1)
        i1 = np.clip(i0 + 1, 0, out.shape[ax] - 1)
        w = np.clip(pos - i0, 0.0, 1.0).astype(np.float32)
        sh = [1, 1, 1]
        sh[ax] = n
        w = w.reshape(sh)
        out = np.take(out, i0, axis=ax) * (1 - w) + np.take(out, i1, axis=ax) * w
    return out.astype(np.float32)


class CodedSnapshots:
    """Snapshot RTM of one shot with coded snapshots through the library's
    loops (rtmhip_snap_ae_forward_steps / rtmhip_snap_ae_backward_steps): the
    snapshot steps are nt-1, nt-1-k, ... and the backward pass images
    img += k * (decoded snapshot * PR).  Owns the codes ([nsnap][bricks][2048]
    FP16), scales ([nsnap][bricks] FP32) and the work buffer; FP32 wavefields."""

    def __init__(self, shot, k, codec):
        nb = C.c_int(0)
        cb, sb, wb = C.c_size_t(0), C.c_size_t(0), C.c_size_t(0)
        hip.check(hip.load().rtmhip_snap_ae_sizes(C.byref(shot), k, C.byref(nb), C.byref(cb),
                                                  C.byref(sb), C.byref(wb)))
        self.shot, self.k, self.codec, self.nb = shot, k, codec, nb.value
        self.nsnap = (shot.nt - 1) // k + 1
        self.codes = hip.DeviceArray((self.nsnap, self.nb, CODES), np.uint16)
        self.scales = hip.DeviceArray((self.nsnap, self.nb))
        self.work = hip.DeviceArray((wb.value // 4,))
        assert self.codes.nbytes == cb.value and self.scales.nbytes == sb.value

    def forward_steps(self, V, coefs, g, ricker, fields, it0, nsteps):
        c = hip._f32(coefs)
        hip.check(hip.load().rtmhip_snap_ae_forward_steps(
            C.byref(self.shot), V.ptr, hip._host_ptr(c), g.ptr, ricker.ptr, C.byref(fields.p),
            C.byref(fields.pp), self.codec.h, self.codes.ptr, self.scales.ptr, self.work.ptr,
            self.k, it0, nsteps))

    def backward_steps(self, V, coefs, g, rcv_fields, rcv_chunk, img, it_hi, nsteps):
        c = hip._f32(coefs)
        hip.check(hip.load().rtmhip_snap_ae_backward_steps(
            C.byref(self.shot), V.ptr, hip._host_ptr(c), g.ptr, C.byref(rcv_fields.p),
            C.byref(rcv_fields.pp), self.codec.h, self.codes.ptr, self.scales.ptr,
            self.work.ptr, self.k, rcv_chunk.ptr, img.ptr, it_hi, nsteps))

    def nbytes(self):
        """Bytes kept per shot (codes and scales)."""
        return self.codes.nbytes + self.scales.nbytes

    def free(self):
        for d in (self.codes, self.scales, self.work):
            d.free()


def forward_snapshots(shot, V, coefs, g, src, nt, k, stores, upb):
    """HBC forward propagation of one shot; after every k-th step the source
    field at that time (PP after the source injection) goes to each store."""
    F = hip.FieldPair((shot.nxe, shot.nye, shot.nze))
    for it in range(nt):
        hip.forward_steps(shot, True, V, coefs, g, src, F, upb, it, 1)
        if it % k == 0:
            for s in stores:
                s.put(it, F.PP)
    F.free()


def backward_image(shot, V, coefs, g, rcv, nt, k, store, order=8):
    """Receiver back-propagation (the reference's receiver half of the HBC
    backward step: upper tapers, step, receiver injection) with the imaging
    condition at every k-th step from the stored source snapshot.  Returns
    the extended-grid image (host)."""
    shape = (shot.nxe, shot.nye, shot.nze)
    b = shot.blen
    R = hip.FieldPair(shape)
    S = hip.DeviceArray(shape).zero()
    img = hip.DeviceArray(shape).zero()
    Pr, PPr = R.a, R.b
    for it in range(nt - 1, -1, -1):
        hip.taper(Pr, g, b, True)
        hip.taper(PPr, g, b, True)
        hip.step(Pr, PPr, V, coefs, order)
        hip.receivers(PPr, rcv, b, shot.rcv, nt, it, True)
        if it % k == 0:
            store.get(it, S)
            hip.image_condition(img, S, PPr, b)
        Pr, PPr = PPr, Pr
    out = img.download()
    for d in (R, S, img):
        d.free()
    return out
