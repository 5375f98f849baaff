"""ctypes binding of librtm3d_hip.so (the C ABI in include/rtm_hip.h).

Host-side plumbing for tests, the benchmark and scripting.  There is no CPU
fallback: if the library is missing or no GPU is present, every entry raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "lib", "librtm3d_hip.so")

OK, ERR_INVALID, ERR_HIP, ERR_DISPERSION, ERR_NOT_INIT, ERR_UNSUPPORTED = range(6)

EXPORTED = [
    "rtmhip_init", "rtmhip_finalize", "rtmhip_last_error", "rtmhip_device_count",
    "rtmhip_synchronize", "rtmhip_version", "rtmhip_alloc", "rtmhip_free",
    "rtmhip_copy_to_device", "rtmhip_copy_from_device", "rtmhip_copy_device", "rtmhip_fill",
    "rtmhip_fill16",
    "rtmhip_step", "rtmhip_taper", "rtmhip_apply_source", "rtmhip_receivers",
    "rtmhip_upper_border", "rtmhip_image_condition", "rtmhip_image_finish", "rtmhip_stack",
    "rtmhip_model_steps", "rtmhip_forward_steps", "rtmhip_backward_steps",
    "rtmhip_snap_buffer_bytes", "rtmhip_snap_forward_steps", "rtmhip_snap_backward_steps",
    "rtmhip_set_tiling", "rtmhip_get_tiling", "rtmhip_timing_enable", "rtmhip_timing_read",
    "rtmhip_set_taper_fusion", "rtmhip_get_taper_fusion",
    "rtmhip_batch_create", "rtmhip_batch_destroy", "rtmhip_batch_forward_steps",
    "rtmhip_batch_backward_steps", "rtmhip_batch_fields", "rtmhip_batch_set_stencil",
    "rtmhip_batch_set_graph", "rtmhip_batch_steps_noalloc",
    "rtmhip_ae_param_count", "rtmhip_ae_create", "rtmhip_ae_destroy", "rtmhip_ae_get_params",
    "rtmhip_ae_set_params", "rtmhip_ae_gradient", "rtmhip_ae_train", "rtmhip_ae_bricks",
    "rtmhip_ae_encode", "rtmhip_ae_decode", "rtmhip_ae_gemm_test", "rtmhip_freq_image",
    "rtmhip_freq_steps",
    "rtmhip_snap_ae_sizes", "rtmhip_snap_ae_forward_steps", "rtmhip_snap_ae_backward_steps",
    "rtmhip_snap_dct_sizes", "rtmhip_snap_dct_forward_steps", "rtmhip_snap_dct_backward_steps",
    "rtmhip_dct_roundtrip",
]


class RtmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rtmhip error %d: %s" % (code, msg))
        self.code = code


class RcvGeom(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("start_x", "start_y", "depth_z", "dist_x", "dist_y", "count_x", "count_y")]


class Shot(C.Structure):
    _fields_ = [("nxe", C.c_int), ("nye", C.c_int), ("nze", C.c_int), ("blen", C.c_int),
                ("order", C.c_int), ("nt", C.c_int), ("ntstep", C.c_int), ("sxe", C.c_int),
                ("sye", C.c_int), ("sze", C.c_int), ("rcv", RcvGeom), ("wave_format", C.c_int)]


class BatchShotDesc(C.Structure):
    """rtmhip_batch_shot."""
    _fields_ = [("shot", Shot), ("V2DT2", C.c_void_p), ("ricker", C.c_void_p),
                ("fs", C.c_void_p * 2), ("fr", C.c_void_p * 2), ("upb", C.c_void_p),
                ("rcv", C.c_void_p), ("img", C.c_void_p)]


# Wavefield storage formats (RTMHIP_WAVE_*): name -> (code, numpy storage dtype).
# 16-bit wavefields are held as raw uint16 bits on the host.
WAVE_FORMATS = {"f32": (0, np.float32), "f16": (1, np.uint16), "bf16": (2, np.uint16)}


_lib = None


def load():
    """Load the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError("librtm3d_hip.so not built (%s); run __graft_entry__.build()" % LIBPATH)
    L = C.CDLL(LIBPATH)
    i, vp, fp = C.c_int, C.c_void_p, C.c_void_p
    sz = C.c_size_t
    pp = C.POINTER(C.c_void_p)
    sig = {
        "rtmhip_init": [i], "rtmhip_finalize": [], "rtmhip_device_count": [C.POINTER(C.c_int)],
        "rtmhip_synchronize": [], "rtmhip_version": [],
        "rtmhip_alloc": [pp, sz], "rtmhip_free": [vp],
        "rtmhip_copy_to_device": [vp, vp, sz], "rtmhip_copy_from_device": [vp, vp, sz],
        "rtmhip_copy_device": [vp, vp, sz], "rtmhip_fill": [fp, sz, C.c_float],
        "rtmhip_fill16": [vp, sz, C.c_ushort],
        "rtmhip_step": [fp, fp, fp, fp, i, i, i, i],
        "rtmhip_taper": [fp, fp, i, i, i, i, i],
        "rtmhip_apply_source": [fp, i, i, i, i, i, i, C.c_float],
        "rtmhip_receivers": [fp, fp, i, i, i, i, C.POINTER(RcvGeom), i, i, i],
        "rtmhip_upper_border": [fp, fp, i, i, i, i, i, i, i],
        "rtmhip_image_condition": [fp, fp, fp, i, i, i, i],
        "rtmhip_image_finish": [fp, fp, i, i, i, i, fp],
        "rtmhip_stack": [fp, fp, sz],
        "rtmhip_model_steps": [C.POINTER(Shot), fp, fp, fp, fp, pp, pp, fp, i, i],
        "rtmhip_forward_steps": [C.POINTER(Shot), i, fp, fp, fp, fp, pp, pp, fp, i, i],
        "rtmhip_backward_steps": [C.POINTER(Shot), i, fp, fp, fp, pp, pp, pp, pp, fp, fp, fp,
                                  i, i, i],
        "rtmhip_snap_buffer_bytes": [C.POINTER(Shot), i, i, C.POINTER(C.c_size_t)],
        "rtmhip_snap_forward_steps": [C.POINTER(Shot), fp, fp, fp, fp, pp, pp, vp, i, i, i, i],
        "rtmhip_snap_backward_steps": [C.POINTER(Shot), fp, fp, fp, pp, pp, vp, i, i, fp, fp, i,
                                       i],
        "rtmhip_set_tiling": [i, i, i, i, i],
        "rtmhip_set_taper_fusion": [i, i],
        "rtmhip_get_taper_fusion": [i, C.POINTER(C.c_int)],
        "rtmhip_get_tiling": [i] + [C.POINTER(C.c_int)] * 4,
        "rtmhip_timing_enable": [i],
        "rtmhip_timing_read": [i, C.POINTER(C.c_longlong), C.POINTER(C.c_double)],
        "rtmhip_batch_create": [pp, C.POINTER(BatchShotDesc), i, fp, fp],
        "rtmhip_batch_destroy": [vp],
        "rtmhip_batch_forward_steps": [vp, i, i],
        "rtmhip_batch_backward_steps": [vp, i, i],
        "rtmhip_batch_fields": [vp, i, pp, pp, pp, pp],
        "rtmhip_batch_set_stencil": [vp, i, i],
        "rtmhip_batch_set_graph": [vp, i],
        "rtmhip_batch_steps_noalloc": [C.POINTER(BatchShotDesc), i, fp, fp, i, i, i],
        "rtmhip_freq_image": [fp] * 10 + [i] * 9,
        "rtmhip_freq_steps": [C.POINTER(Shot), fp, fp, fp, fp, pp, pp, pp, pp, fp, fp, fp, fp,
                              fp, fp, fp, fp, fp, i, i, i, i, i],
        "rtmhip_ae_param_count": [],
        "rtmhip_ae_create": [pp, i, C.c_ulonglong],
        "rtmhip_ae_destroy": [vp],
        "rtmhip_ae_get_params": [vp, fp, i],
        "rtmhip_ae_set_params": [vp, fp],
        "rtmhip_ae_gradient": [vp, vp, i, C.POINTER(C.c_float)],
        "rtmhip_ae_train": [vp, vp, i, i, i, C.c_float, C.c_ulonglong, fp],
        "rtmhip_ae_bricks": [fp, i, i, i, i, i, i, i, i, i, vp, fp],
        "rtmhip_ae_encode": [vp, vp, i, vp],
        "rtmhip_ae_decode": [vp, vp, fp, i, fp, i, i, i, i, i, i, i, i, i],
        "rtmhip_ae_gemm_test": [vp, C.c_longlong, C.c_longlong, i, vp, C.c_longlong,
                                C.c_longlong, i, fp, C.c_longlong, fp, C.c_longlong, i,
                                C.c_longlong, i, i],
        "rtmhip_snap_ae_sizes": [C.POINTER(Shot), i, C.POINTER(C.c_int),
                                 C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_size_t)],
        "rtmhip_snap_ae_forward_steps": [C.POINTER(Shot), fp, fp, fp, fp, pp, pp, vp, vp, fp,
                                         vp, i, i, i],
        "rtmhip_snap_ae_backward_steps": [C.POINTER(Shot), fp, fp, fp, pp, pp, vp, vp, fp, vp,
                                          i, fp, fp, i, i],
        "rtmhip_snap_dct_sizes": [C.POINTER(Shot), i, i, C.POINTER(C.c_size_t),
                                  C.POINTER(C.c_size_t)],
        "rtmhip_snap_dct_forward_steps": [C.POINTER(Shot), fp, fp, fp, fp, pp, pp, vp, vp, fp, i,
                                          i, i, i],
        "rtmhip_snap_dct_backward_steps": [C.POINTER(Shot), fp, fp, fp, pp, pp, vp, i, i, fp, fp,
                                           i, i],
        "rtmhip_dct_roundtrip": [vp, i, i, i, i, i, i, vp, vp, fp, fp, C.POINTER(C.c_uint)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.rtmhip_last_error.argtypes = []
    L.rtmhip_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise RtmError(rc, load().rtmhip_last_error().decode(errors="replace"))


def _host_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A device buffer owned by the caller (the C ABI only borrows it)."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(load().rtmhip_alloc(C.byref(p), self.nbytes))
        self.ptr = p

    @property
    def size(self):
        return int(np.prod(self.shape))

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        check(load().rtmhip_copy_to_device(self.ptr, _host_ptr(a), self.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(load().rtmhip_copy_from_device(_host_ptr(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        if self.dtype.itemsize == 2:
            check(load().rtmhip_fill16(self.ptr, self.size, 0))
        else:
            check(load().rtmhip_fill(self.ptr, self.size, 0.0))
        return self

    def free(self):
        if self.ptr and not getattr(self, "base", None):
            check(load().rtmhip_free(self.ptr))
        self.ptr = C.c_void_p()

    def view(self, start, count):
        """Non-owning view of elements [start, start+count) of the leading
        axis (valid while this array lives; free() on it releases nothing)."""
        assert 0 <= start and count > 0 and start + count <= self.shape[0]
        v = DeviceArray.__new__(DeviceArray)
        v.shape, v.dtype, v.base = (count,) + self.shape[1:], self.dtype, self
        row = self.nbytes // self.shape[0]
        v.nbytes = row * count
        v.ptr = C.c_void_p(self.ptr.value + start * row)
        return v

    def slot(self, i):
        """Non-owning view of element i of the leading axis, without that axis."""
        v = self.view(i, 1)
        v.shape = v.shape[1:]
        return v

    @classmethod
    def from_host(cls, a, dtype=np.float32):
        d = cls(np.shape(a), dtype)
        return d.upload(a)


def init(device=0):
    check(load().rtmhip_init(device))


def finalize():
    check(load().rtmhip_finalize())


def synchronize():
    check(load().rtmhip_synchronize())


def device_count():
    n = C.c_int(0)
    check(load().rtmhip_device_count(C.byref(n)))
    return n.value


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


# ---------------- per-step operators ----------------
def step(P, PP, V, coefs, order):
    c = _f32(coefs)
    check(load().rtmhip_step(P.ptr, PP.ptr, V.ptr, _host_ptr(c), *P.shape, order))


def taper(P, g, blen, upper_only):
    check(load().rtmhip_taper(P.ptr, g.ptr, *P.shape, blen, int(upper_only)))


def apply_source(PP, sxe, sye, sze, value):
    check(load().rtmhip_apply_source(PP.ptr, *PP.shape, sxe, sye, sze, float(value)))


def receivers(PP, rcv, blen, geom, ntstep, lt, restore):
    check(load().rtmhip_receivers(PP.ptr, rcv.ptr, *PP.shape, blen, C.byref(geom), ntstep, lt,
                                  int(restore)))


def upper_border(PP, upb, blen, order, lt, restore):
    check(load().rtmhip_upper_border(PP.ptr, upb.ptr, *PP.shape, blen, order, lt, int(restore)))


def image_condition(img, PS, PR, blen):
    check(load().rtmhip_image_condition(img.ptr, PS.ptr, PR.ptr, *img.shape, blen))


def image_finish(img, out, blen, fcoefs):
    c = _f32(fcoefs)
    check(load().rtmhip_image_finish(img.ptr, out.ptr, *img.shape, blen, _host_ptr(c)))


def freq_image(img, PS, PR, acc, kre, kim, w2, blen, nwstep, iw0, nw, nt, st):
    """Frequency contributions of iw0 .. iw0+nwstep-1 at time st and the
    imaging condition (acc: four DeviceArrays sRe, sIm, rRe, rIm)."""
    check(load().rtmhip_freq_image(img.ptr, PS.ptr, PR.ptr, *[a.ptr for a in acc], kre.ptr,
                                   kim.ptr, w2.ptr, *img.shape, blen, nwstep, iw0, nw, nt, st))


def freq_steps(shot, V, coefs, g, ricker, S, R, rcv_chunk, img, acc, kre, kim, w2, nwstep, iw0,
               nw, it_hi, nsteps):
    """rtmhip_freq_steps: the frequency-domain migration loop of one group
    (S, R: FieldPairs of the source and receiver fields, swapped in place;
    acc: four DeviceArrays sRe, sIm, rRe, rIm)."""
    c = _f32(coefs)
    check(load().rtmhip_freq_steps(C.byref(shot), V.ptr, _host_ptr(c), g.ptr, ricker.ptr,
                                   C.byref(S.p), C.byref(S.pp), C.byref(R.p), C.byref(R.pp),
                                   rcv_chunk.ptr, img.ptr, *[a.ptr for a in acc], kre.ptr,
                                   kim.ptr, w2.ptr, nwstep, iw0, nw, it_hi, nsteps))


def stack(dst, src):
    check(load().rtmhip_stack(dst.ptr, src.ptr, dst.size))


# ---------------- time loops ----------------
def make_shot(nxe, nye, nze, blen, order, nt, ntstep, src_ext, geom, fmt="f32"):
    return Shot(nxe, nye, nze, blen, order, nt, ntstep, src_ext[0], src_ext[1], src_ext[2], geom,
                WAVE_FORMATS[fmt][0])


class FieldPair:
    """Two wavefield buffers swapped in place by the time loops (dtype:
    float32, or uint16 for the 16-bit formats)."""

    def __init__(self, shape, dtype=np.float32):
        self.a = DeviceArray(shape, dtype).zero()
        self.b = DeviceArray(shape, dtype).zero()
        self.p = C.c_void_p(self.a.ptr.value)
        self.pp = C.c_void_p(self.b.ptr.value)

    def by_ptr(self, v):
        return self.a if v.value == self.a.ptr.value else self.b

    @property
    def P(self):
        return self.by_ptr(self.p)

    @property
    def PP(self):
        return self.by_ptr(self.pp)

    def free(self):
        self.a.free()
        self.b.free()


def model_steps(shot, V, coefs, g, ricker, fields, rcv_chunk, it0, nsteps):
    c = _f32(coefs)
    check(load().rtmhip_model_steps(C.byref(shot), V.ptr, _host_ptr(c), g.ptr, ricker.ptr,
                                    C.byref(fields.p), C.byref(fields.pp), rcv_chunk.ptr, it0,
                                    nsteps))


def forward_steps(shot, hbc, V, coefs, g, ricker, fields, upb, it0, nsteps):
    """rtmhip_forward_steps.  hbc: 0 RBC; 1 HBC with upb holding every
    step's upper-border layer [nt][nxe][nye][order/2]; 2 HBC with upb
    holding one receiver chunk's layers [ntstep][nxe][nye][order/2] (the
    steps of the call must lie in one chunk)."""
    c = _f32(coefs)
    check(load().rtmhip_forward_steps(C.byref(shot), int(hbc), V.ptr, _host_ptr(c),
                                      g.ptr if g is not None else None, ricker.ptr,
                                      C.byref(fields.p), C.byref(fields.pp),
                                      upb.ptr if upb is not None else None, it0, nsteps))


def backward_steps(shot, hbc, V, coefs, g, src_fields, rcv_fields, upb, rcv_chunk, img, it_hi,
                   nsteps, fused=True):
    """rtmhip_backward_steps; hbc and upb as for forward_steps (with hbc 2,
    upb holds the layers of the receiver chunk being processed)."""
    c = _f32(coefs)
    check(load().rtmhip_backward_steps(C.byref(shot), int(hbc), V.ptr, _host_ptr(c),
                                       g.ptr if g is not None else None,
                                       C.byref(src_fields.p), C.byref(src_fields.pp),
                                       C.byref(rcv_fields.p), C.byref(rcv_fields.pp),
                                       upb.ptr if upb is not None else None, rcv_chunk.ptr,
                                       img.ptr, it_hi, nsteps, int(fused)))


# ---------------- snapshot RTM ----------------
# Snapshot formats (RTMHIP_SNAP_*): name -> (code, numpy storage dtype).
SNAP_FORMATS = {"f32": (0, np.float32), "f16": (1, np.float16)}


def snap_buffer(shot, fmt, k):
    """A device buffer for one shot's snapshots ([nsnap][nxe][nye][nze-2blen])."""
    n = C.c_size_t(0)
    check(load().rtmhip_snap_buffer_bytes(C.byref(shot), SNAP_FORMATS[fmt][0], k, C.byref(n)))
    cells = shot.nxe * shot.nye * (shot.nze - 2 * shot.blen)
    dt = np.dtype(SNAP_FORMATS[fmt][1])
    return DeviceArray((n.value // (cells * dt.itemsize), shot.nxe, shot.nye,
                        shot.nze - 2 * shot.blen), dt)


def snap_forward_steps(shot, V, coefs, g, ricker, fields, snaps, fmt, k, it0, nsteps):
    c = _f32(coefs)
    check(load().rtmhip_snap_forward_steps(C.byref(shot), V.ptr, _host_ptr(c), g.ptr, ricker.ptr,
                                           C.byref(fields.p), C.byref(fields.pp), snaps.ptr,
                                           SNAP_FORMATS[fmt][0], k, it0, nsteps))


def snap_backward_steps(shot, V, coefs, g, rcv_fields, snaps, fmt, k, rcv_chunk, img, it_hi,
                        nsteps):
    c = _f32(coefs)
    check(load().rtmhip_snap_backward_steps(C.byref(shot), V.ptr, _host_ptr(c), g.ptr,
                                            C.byref(rcv_fields.p), C.byref(rcv_fields.pp),
                                            snaps.ptr, SNAP_FORMATS[fmt][0], k, rcv_chunk.ptr,
                                            img.ptr, it_hi, nsteps))


def _basis_ptr(basis):
    if basis is None:
        return None, None
    b = np.ascontiguousarray(basis, np.float32).ravel()
    assert b.size == 32 * 32 + 2 * 64, b.size
    return b, _host_ptr(b)


class DctSnapshots:
    """Transform-coded snapshots of one shot (rtmhip_snap_dct_*): the store of
    every snapshot and the work buffer; keep_div sets the coefficient budget
    (nxe*nye*(nze-2*blen) / keep_div per snapshot); basis: None for the
    DCT-II, or a learned separable basis (basis.learn_brick_basis)."""

    def __init__(self, shot, k, keep_div=32, basis=None):
        sb, wb = C.c_size_t(0), C.c_size_t(0)
        check(load().rtmhip_snap_dct_sizes(C.byref(shot), k, keep_div, C.byref(sb), C.byref(wb)))
        self.shot, self.k, self.keep_div, self.basis = shot, k, keep_div, basis
        self.store = DeviceArray((sb.value,), np.uint8)
        self.work = DeviceArray((wb.value,), np.uint8)

    def forward_steps(self, V, coefs, g, ricker, fields, it0, nsteps):
        c = _f32(coefs)
        keep, bp = _basis_ptr(self.basis)
        check(load().rtmhip_snap_dct_forward_steps(
            C.byref(self.shot), V.ptr, _host_ptr(c), g.ptr, ricker.ptr, C.byref(fields.p),
            C.byref(fields.pp), self.store.ptr, self.work.ptr, bp, self.k, self.keep_div, it0,
            nsteps))

    def backward_steps(self, V, coefs, g, rcv_fields, rcv_chunk, img, it_hi, nsteps):
        c = _f32(coefs)
        check(load().rtmhip_snap_dct_backward_steps(
            C.byref(self.shot), V.ptr, _host_ptr(c), g.ptr, C.byref(rcv_fields.p),
            C.byref(rcv_fields.pp), self.store.ptr, self.k, self.keep_div, rcv_chunk.ptr, img.ptr,
            it_hi, nsteps))

    def nbytes(self):
        return self.store.nbytes

    def free(self):
        self.store.free()
        self.work.free()


def dct_roundtrip(field, fmt, blen, keep_div=32, basis=None):
    """Encodes the imaging range of a device field and decodes it: returns
    (decoded FP32 host array with the other cells 0, kept coefficients)."""
    nxe, nye, nze = field.shape
    shot = make_shot(nxe, nye, nze, blen, 8, 1, 1, (0, 0, 0), RcvGeom(0, 0, 0, 1, 1, 1, 1))
    sb, wb = C.c_size_t(0), C.c_size_t(0)
    check(load().rtmhip_snap_dct_sizes(C.byref(shot), 1, keep_div, C.byref(sb), C.byref(wb)))
    store, work = DeviceArray((sb.value,), np.uint8), DeviceArray((wb.value,), np.uint8)
    out = DeviceArray(field.shape).zero()
    kept = C.c_uint(0)
    try:
        keep, bp = _basis_ptr(basis)
        check(load().rtmhip_dct_roundtrip(field.ptr, WAVE_FORMATS[fmt][0], nxe, nye, nze, blen,
                                          keep_div, store.ptr, work.ptr, bp, out.ptr,
                                          C.byref(kept)))
        return out.download(), kept.value
    finally:
        for d in (store, work, out):
            d.free()


def set_tiling(kernel_class, tz=0, by=0, ry=0, xchunk=0):
    check(load().rtmhip_set_tiling(kernel_class, tz, by, ry, xchunk))


def set_taper_fusion(fmt, mode):
    """mode 1: upper tapers fused into the stencil; 2: also the modeling
    loop's every-border tapers, applied as values are loaded; 3: those
    applied to each step's output instead (needs the source and receivers
    off the borders, else the separate pass runs); 4: also the forward
    loops' upper tapers applied to each step's output (backward loops fused
    on load); 0: separate in-place passes; -1: the format's default (F32:
    mode 4; 16-bit formats: upper tapers in separate passes, every-border
    tapers on the output)."""
    check(load().rtmhip_set_taper_fusion(WAVE_FORMATS[fmt][0], mode))


def get_taper_fusion(fmt):
    v = C.c_int()
    check(load().rtmhip_get_taper_fusion(WAVE_FORMATS[fmt][0], C.byref(v)))
    return v.value


def get_tiling(kernel_class):
    v = [C.c_int() for _ in range(4)]
    check(load().rtmhip_get_tiling(kernel_class, *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


TILINGS = [(1, 4, 4), (1, 2, 6), (1, 2, 7), (2, 4, 3), (64, 1, 6), (64, 1, 8), (128, 1, 4),
           (128, 1, 6), (128, 1, 8), (64, 4, 4), (64, 4, 2), (64, 8, 2), (32, 8, 4), (32, 8, 2)]


def timing_enable(on=True):
    check(load().rtmhip_timing_enable(int(on)))


def timing_read(kernel_class):
    n = C.c_longlong(0)
    ms = C.c_double(0.0)
    check(load().rtmhip_timing_read(kernel_class, C.byref(n), C.byref(ms)))
    return n.value, ms.value


# ---------------- batches of shots ----------------
class Batch:
    """Many HBC shots advanced in lockstep (rtmhip_batch_*).  `members` is a
    list of dicts with keys shot (Shot), V, ricker, S (FieldPair, source),
    R (FieldPair, receiver), upb, rcv, img (DeviceArrays); the FieldPairs'
    current roles are taken as (P, PP) and are kept in step with the batch."""

    def __init__(self, members, coefs, g):
        self.members = members
        n = len(members)
        arr = (BatchShotDesc * n)()
        for k, m in enumerate(members):
            d = arr[k]
            d.shot = m["shot"]
            d.V2DT2 = m["V"].ptr.value
            d.ricker = m["ricker"].ptr.value
            d.fs[0], d.fs[1] = m["S"].p.value, m["S"].pp.value
            d.fr[0], d.fr[1] = m["R"].p.value, m["R"].pp.value
            d.upb = m["upb"].ptr.value
            d.rcv = m["rcv"].ptr.value
            d.img = m["img"].ptr.value
        c = _f32(coefs)
        h = C.c_void_p()
        check(load().rtmhip_batch_create(C.byref(h), arr, n, _host_ptr(c), g.ptr))
        self.h = h

    def _sync_roles(self):
        L = load()
        for k, m in enumerate(self.members):
            check(L.rtmhip_batch_fields(self.h, k, C.byref(m["S"].p), C.byref(m["S"].pp),
                                        C.byref(m["R"].p), C.byref(m["R"].pp)))

    def forward_steps(self, it0, nsteps):
        check(load().rtmhip_batch_forward_steps(self.h, it0, nsteps))
        self._sync_roles()

    def backward_steps(self, it_hi, nsteps):
        check(load().rtmhip_batch_backward_steps(self.h, it_hi, nsteps))
        self._sync_roles()

    def set_stencil(self, kernel_class, mode):
        check(load().rtmhip_batch_set_stencil(self.h, kernel_class, mode))

    def set_graph(self, steps):
        check(load().rtmhip_batch_set_graph(self.h, steps))

    def destroy(self):
        if self.h:
            check(load().rtmhip_batch_destroy(self.h))
            self.h = C.c_void_p()


def batch_steps_noalloc(members, coefs, g, direction, it, nsteps):
    """rtmhip_batch_steps_noalloc over `members` (the dicts of Batch): the
    FieldPairs' current roles go in as (P, PP) and are updated from the
    roles the call returns."""
    n = len(members)
    arr = (BatchShotDesc * n)()
    for k, m in enumerate(members):
        d = arr[k]
        d.shot = m["shot"]
        d.V2DT2 = m["V"].ptr.value
        d.ricker = m["ricker"].ptr.value
        d.fs[0], d.fs[1] = m["S"].p.value, m["S"].pp.value
        d.fr[0], d.fr[1] = m["R"].p.value, m["R"].pp.value
        d.upb = m["upb"].ptr.value
        d.rcv = m["rcv"].ptr.value
        d.img = m["img"].ptr.value
    c = _f32(coefs)
    check(load().rtmhip_batch_steps_noalloc(arr, n, _host_ptr(c), g.ptr, direction, it, nsteps))
    for k, m in enumerate(members):
        m["S"].p, m["S"].pp = C.c_void_p(arr[k].fs[0]), C.c_void_p(arr[k].fs[1])
        m["R"].p, m["R"].pp = C.c_void_p(arr[k].fr[0]), C.c_void_p(arr[k].fr[1])
