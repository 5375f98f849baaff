"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported by the tests, by the smoke check and by the CPU-baseline leg of the
benchmark -- never by the product.  ``pipeline_*`` reproduce the reference
controller around the per-shot loops (velocity extension, border fill, v2dt2,
shot loop, image stacking, Laplacian filter, border removal):
rtm-3d/src/controller/RTMController.cpp:27-166 and
rtm-3d/src/controller/RTMAcousticController.cpp:129-269.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "librtmoracle.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")

# Wavefield storage formats: suffix -> (arithmetic pointer, wavefield pointer,
# scalar type, wavefield dtype).  f16/bf16 wavefields are raw uint16 bits.
FORMATS = ("f32", "f64", "f16", "bf16")


class RcvGeom(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("start_x", "start_y", "depth_z", "dist_x", "dist_y", "count_x", "count_y")]


class RandState(C.Structure):
    _fields_ = [("r", C.c_int32 * 34), ("tbl", C.c_int32 * 31), ("f", C.c_int), ("b", C.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        i, f, d, vp = C.c_int, C.c_float, C.c_double, C.c_void_p
        L.orc_laplacian_taps_f32.argtypes = [i, _f32p]
        L.orc_stencil_coefs_f32.argtypes = [i, f, f, f, _f32p]
        L.orc_ricker_f32.argtypes = [i, f, f, _f32p]
        L.orc_taper_table_f32.argtypes = [i, f, _f32p]
        L.orc_glibc_rand_seed.argtypes = [C.POINTER(RandState), C.c_uint]
        L.orc_glibc_rand_next.argtypes = [C.POINTER(RandState)]
        L.orc_extend_borders_f32.argtypes = [_f32p, i, i, i, i, _f32p]
        L.orc_borders_abc_f32.argtypes = [_f32p, i, i, i, i]
        L.orc_borders_rbc_f32.argtypes = [_f32p, i, i, i, i, C.POINTER(RandState)]
        L.orc_borders_hbc_f32.argtypes = [_f32p, i, i, i, i, C.POINTER(RandState)]
        L.orc_v2dt2_f32.argtypes = [_f32p, C.c_size_t, f]
        L.orc_stencil_coefs_f64.argtypes = [i, d, d, d, _f64p]
        L.orc_ricker_f64.argtypes = [i, d, d, _f64p]
        L.orc_taper_table_f64.argtypes = [i, d, _f64p]
        L.orc_v2dt2_f64.argtypes = [_f32p, C.c_size_t, d, _f64p]
        for fn in ("orc_to_f16", "orc_to_bf16"):
            getattr(L, fn).argtypes = [_f32p, _u16p, C.c_size_t]
        for fn in ("orc_from_f16", "orc_from_bf16"):
            getattr(L, fn).argtypes = [_u16p, _f32p, C.c_size_t]
        for s, p, w, r in (("f32", _f32p, _f32p, f), ("f64", _f64p, _f64p, d),
                           ("f16", _f32p, _u16p, f), ("bf16", _f32p, _u16p, f)):
            getattr(L, "orc_step_" + s).argtypes = [w, w, p, p, i, i, i, i]
            getattr(L, "orc_taper_all_" + s).argtypes = [w, p, i, i, i, i]
            getattr(L, "orc_taper_upper_" + s).argtypes = [w, p, i, i, i, i]
            getattr(L, "orc_apply_source_" + s).argtypes = [w, i, i, i, i, i, i, r]
            getattr(L, "orc_receivers_" + s).argtypes = [w, p, i, i, i, i, C.POINTER(RcvGeom),
                                                         i, i, i]
            getattr(L, "orc_upb_" + s).argtypes = [w, w, i, i, i, i, i, i, i]
            getattr(L, "orc_image_" + s).argtypes = [p, w, w, i, i, i, i]
            getattr(L, "orc_image_finish_" + s).argtypes = [p, p, i, i, i, i, p]
            getattr(L, "orc_model_shot_" + s).argtypes = [p, p, p, p, i, i, i, i, i, i, i, i, i,
                                                          i, C.POINTER(RcvGeom), p]
            getattr(L, "orc_migrate_shot_" + s).argtypes = [p, p, p, p, i, i, i, i, i, i, i, i,
                                                            i, i, C.POINTER(RcvGeom), p, i, p,
                                                            vp]
            getattr(L, "orc_freq_image_" + s).argtypes = [p, w, w, p, p, p, p, p, p, p, i, i, i,
                                                          i, i, i, i, i, i]
            getattr(L, "orc_fmig_shot_" + s).argtypes = [p, p, p, p, i, i, i, i, i, i, i, i, i,
                                                         i, C.POINTER(RcvGeom), p, p, p, p, i, i,
                                                         p]
        L.orc_freq_tables.argtypes = [i, f, f, f, vp, vp, vp, i]
        _lib = L
    return _lib


def _sfx(dtype, fmt=None):
    if fmt is not None:
        assert fmt in FORMATS, fmt
        return fmt
    dtype = np.dtype(dtype)
    if dtype == np.uint16:
        raise ValueError("16-bit wavefields need fmt='f16' or 'bf16'")
    return "f32" if dtype == np.float32 else "f64"


def to_format(a, fmt):
    """float32 array -> wavefield storage of format fmt (uint16 bits for
    f16/bf16; nearest-even)."""
    a = np.ascontiguousarray(a, np.float32)
    if fmt == "f32":
        return a.copy()
    out = np.empty(a.shape, np.uint16)
    getattr(lib(), "orc_to_" + fmt)(a.reshape(-1), out.reshape(-1), a.size)
    return out


def from_format(a, fmt):
    """Wavefield storage of format fmt -> float32 (exact)."""
    if fmt == "f32":
        return np.array(a, np.float32)
    a = np.ascontiguousarray(a, np.uint16)
    out = np.empty(a.shape, np.float32)
    getattr(lib(), "orc_from_" + fmt)(a.reshape(-1), out.reshape(-1), a.size)
    return out


# ---------------- set-up helpers ----------------
def stencil_coefs(order, dx, dy, dz, dtype=np.float32):
    out = np.zeros(3 * (order + 1), dtype)
    if dtype == np.float32:
        lib().orc_stencil_coefs_f32(order, dx, dy, dz, out)
    else:
        lib().orc_stencil_coefs_f64(order, dx, dy, dz, out)
    return out


def laplacian_taps(order):
    out = np.zeros(order + 1, np.float32)
    lib().orc_laplacian_taps_f32(order, out)
    return out


def ricker(nt, dt, fpeak, dtype=np.float32):
    out = np.zeros(nt, dtype)
    (lib().orc_ricker_f32 if dtype == np.float32 else lib().orc_ricker_f64)(nt, dt, fpeak, out)
    return out


def taper_table(blen, factor, dtype=np.float32):
    out = np.zeros(blen, dtype)
    (lib().orc_taper_table_f32 if dtype == np.float32 else lib().orc_taper_table_f64)(
        blen, factor, out)
    return out


class GlibcRand:
    def __init__(self, seed=1):
        self.st = RandState()
        lib().orc_glibc_rand_seed(C.byref(self.st), seed)

    def next(self):
        return lib().orc_glibc_rand_next(C.byref(self.st))


def extend_velocity(vel, blen, bc, rng=None):
    """Extend a [nx][ny][nz] model and fill its borders ('abc'|'hbc'|'rbc')."""
    nx, ny, nz = vel.shape
    ve = np.zeros((nx + 2 * blen, ny + 2 * blen, nz + 2 * blen), np.float32)
    lib().orc_extend_borders_f32(np.ascontiguousarray(vel, np.float32), nx, ny, nz, blen, ve)
    dims = ve.shape + (blen,)
    if bc == "abc":
        lib().orc_borders_abc_f32(ve, *dims)
    else:
        rng = rng or GlibcRand(1)
        fn = lib().orc_borders_rbc_f32 if bc == "rbc" else lib().orc_borders_hbc_f32
        fn(ve, *dims, C.byref(rng.st))
    return ve


def v2dt2(ve, dt, dtype=np.float32):
    if dtype == np.float32:
        out = np.array(ve, np.float32, copy=True)
        lib().orc_v2dt2_f32(out, out.size, dt)
    else:
        out = np.zeros(ve.shape, np.float64)
        lib().orc_v2dt2_f64(np.ascontiguousarray(ve, np.float32), ve.size, dt, out)
    return out


# ---------------- per-step operators ----------------
def step(P, PP, V2, coefs, order, fmt=None):
    """In place on PP; returns 0 or -1 on NaN."""
    nxe, nye, nze = P.shape
    return getattr(lib(), "orc_step_" + _sfx(P.dtype, fmt))(P, PP, V2, coefs, nxe, nye, nze,
                                                           order)


def taper(P, g, blen, upper_only, fmt=None):
    fn = "orc_taper_upper_" if upper_only else "orc_taper_all_"
    getattr(lib(), fn + _sfx(P.dtype, fmt))(P, g, *P.shape, blen)


def apply_source(PP, sxe, sye, sze, val, fmt=None):
    getattr(lib(), "orc_apply_source_" + _sfx(PP.dtype, fmt))(PP, *PP.shape, sxe, sye, sze, val)


def receivers(PP, rcv, blen, geom, ntstep, lt, restore, fmt=None):
    getattr(lib(), "orc_receivers_" + _sfx(PP.dtype, fmt))(PP, rcv, *PP.shape, blen,
                                                           C.byref(geom), ntstep, lt,
                                                           int(restore))


def upb(PP, buf, blen, order, lt, restore, fmt=None):
    getattr(lib(), "orc_upb_" + _sfx(PP.dtype, fmt))(PP, buf, *PP.shape, blen, order, lt,
                                                     int(restore))


def image(img, PS, PR, blen, fmt=None):
    getattr(lib(), "orc_image_" + _sfx(img.dtype, fmt))(img, PS, PR, *img.shape, blen)


def image_finish(img, blen, dx, dy, dz):
    dt = img.dtype
    fc = stencil_coefs(2, dx, dy, dz, dt)
    nxe, nye, nze = img.shape
    out = np.zeros((nxe - 2 * blen, nye - 2 * blen, nze - 2 * blen), dt)
    getattr(lib(), "orc_image_finish_" + _sfx(dt))(img, out, nxe, nye, nze, blen, fc)
    return out


def geom_from_params(p):
    return RcvGeom(p["receiver_start_x"], p["receiver_start_y"], p["receiver_depth_z"],
                   p["receiver_distance_x"], p["receiver_distance_y"],
                   p["receiver_count_x"], p["receiver_count_y"])


def shots_from_params(p):
    return [(ix * p["source_distance_x"] + p["source_start_x"],
             iy * p["source_distance_y"] + p["source_start_y"], p["source_depth_z"])
            for ix in range(p["source_count_x"]) for iy in range(p["source_count_y"])]


def model_shot(v2, coefs, g, src, blen, order, nt, ntstep, shot, geom, fmt=None):
    dt = v2.dtype
    nch = nt // ntstep
    rcv = np.zeros((nch, geom.count_x, geom.count_y, ntstep), dt)
    rc = getattr(lib(), "orc_model_shot_" + _sfx(dt, fmt))(v2, coefs, g, src, *v2.shape, blen,
                                                           order, nt, ntstep, *shot,
                                                           C.byref(geom), rcv)
    if rc != 0:
        raise RuntimeError("oracle modeling failed (%d)" % rc)
    return rcv


def migrate_shot(v2, coefs, g, src, blen, order, nt, ntstep, shot, geom, rcv, hbc=True,
                 fmt=None, fwd_last=None):
    """Image of one shot on the extended grid; fwd_last (optional, an array of
    the wavefield storage type and grid shape) receives the forward pass's
    last source wavefield."""
    dt = v2.dtype
    img = np.zeros(v2.shape, dt)
    rcv = np.ascontiguousarray(rcv, dt)
    last = None
    if fwd_last is not None:
        assert fwd_last.shape == v2.shape and fwd_last.flags["C_CONTIGUOUS"]
        last = fwd_last.ctypes.data_as(C.c_void_p)
    rc = getattr(lib(), "orc_migrate_shot_" + _sfx(dt, fmt))(v2, coefs, g, src, *v2.shape, blen,
                                                             order, nt, ntstep, *shot,
                                                             C.byref(geom), rcv, int(hbc), img,
                                                             last)
    if rc != 0:
        raise RuntimeError("oracle migration failed (%d)" % rc)
    return img


def snapshot_migrate_shot(v2, coefs, g, src, blen, order, nt, shot, geom, rcv, k,
                          snap_fmt="f32", fmt=None):
    """Snapshot RTM of one shot (no reference counterpart as a loop): the HBC
    forward step sequence of orc_migrate_shot (upper tapers, step, source)
    keeping the source wavefield of the steps nt-1, nt-1-k, ... over the
    imaging range z in [blen, nze-blen) -- rounded to FP16 (nearest-even)
    for snap_fmt='f16' -- instead of the upper-border history; then the
    receiver half of its backward loop (upper tapers, step, receivers set)
    with img += k * (snapshot * PR) at those steps, in the arithmetic type of
    v2 (numpy, that operation order).  rcv: [count_x][count_y][nt].
    fmt: wavefield storage format ('f16' / 'bf16' with float32 v2)."""
    dt = v2.dtype
    nxe, nye, nze = v2.shape
    wdt = np.uint16 if fmt in ("f16", "bf16") else dt
    sx, sy, sz = shot
    z0, z1 = blen, nze - blen

    def widen(a):
        return from_format(a, fmt) if wdt == np.uint16 else np.array(a, dt)

    P, PP = np.zeros(v2.shape, wdt), np.zeros(v2.shape, wdt)
    snaps = {}
    for it in range(nt):
        P, PP = PP, P
        taper(P, g, blen, True, fmt)
        taper(PP, g, blen, True, fmt)
        if step(P, PP, v2, coefs, order, fmt) != 0:
            raise RuntimeError("oracle step produced a NaN")
        apply_source(PP, sx + blen, sy + blen, sz + blen, src[it], fmt)
        if (nt - 1 - it) % k == 0:
            a = widen(PP[:, :, z0:z1])
            if snap_fmt == "f16":
                a = a.astype(np.float16).astype(dt)
            snaps[it] = a
    img = np.zeros(v2.shape, dt)
    Pr, PPr = np.zeros(v2.shape, wdt), np.zeros(v2.shape, wdt)
    rcv = np.ascontiguousarray(rcv, dt)
    w = dt.type(k)
    for it in range(nt - 1, -1, -1):
        taper(Pr, g, blen, True, fmt)
        taper(PPr, g, blen, True, fmt)
        if step(Pr, PPr, v2, coefs, order, fmt) != 0:
            raise RuntimeError("oracle step produced a NaN")
        receivers(PPr, rcv, blen, geom, nt, it, True, fmt)
        if it in snaps:
            v = snaps[it] * widen(PPr[:, :, z0:z1])
            img[:, :, z0:z1] = img[:, :, z0:z1] + w * v
        Pr, PPr = PPr, Pr
    return img


def freq_tables(nt, dt, fmin, fmax):
    """(w2[nw], kre[nw][nt], kim[nw][nt]) of the frequency-domain migration."""
    nw = lib().orc_freq_tables(nt, dt, fmin, fmax, None, None, None, 0)
    w2 = np.zeros(max(nw, 1), np.float32)
    kre = np.zeros((max(nw, 1), nt), np.float32)
    kim = np.zeros((max(nw, 1), nt), np.float32)
    lib().orc_freq_tables(nt, dt, fmin, fmax, w2.ctypes.data_as(C.c_void_p),
                          kre.ctypes.data_as(C.c_void_p), kim.ctypes.data_as(C.c_void_p), nw)
    return w2[:nw], kre[:nw], kim[:nw]


def freq_image(img, PS, PR, acc, kre, kim, w2, blen, nwstep, iw0, nt, st, fmt=None):
    """One time step of frequency contributions + imaging, in place
    (acc: four [nwstep][nx][ny][nz] arrays sRe, sIm, rRe, rIm)."""
    getattr(lib(), "orc_freq_image_" + _sfx(img.dtype, fmt))(
        img, PS, PR, *acc, np.ascontiguousarray(kre), np.ascontiguousarray(kim), w2,
        *img.shape, blen, nwstep, iw0, len(w2), nt, st)


def fmig_shot(v2, coefs, g, src, blen, order, nt, ntstep, shot, geom, rcv, tables, nwstep,
              img=None):
    """Frequency-domain migration of one shot, accumulated into img (a new
    zero image when None) and returned."""
    w2, kre, kim = tables
    if img is None:
        img = np.zeros(v2.shape, v2.dtype)
    rcv = np.ascontiguousarray(rcv, v2.dtype)
    rc = getattr(lib(), "orc_fmig_shot_" + _sfx(v2.dtype))(
        v2, coefs, g, src, *v2.shape, blen, order, nt, ntstep, *shot, C.byref(geom), rcv,
        np.ascontiguousarray(kre, v2.dtype), np.ascontiguousarray(kim, v2.dtype),
        np.ascontiguousarray(w2, v2.dtype), len(w2), nwstep, img)
    if rc != 0:
        raise RuntimeError("oracle frequency-domain migration failed (%d)" % rc)
    return img


# ---------------- controller-level pipelines ----------------
def _setup(p, ve, dtype):
    order, blen = p["stencil_order"], p["blen"]
    if dtype == np.float32:
        coefs = stencil_coefs(order, p["dx"], p["dy"], p["dz"])
        v2 = v2dt2(ve, np.float32(p["dt"]))
        src = ricker(p["nt"], np.float32(p["dt"]), np.float32(p["fpeak"]))
        g = taper_table(blen, np.float32(p["taper_factor"]))
    else:
        coefs = stencil_coefs(order, float(np.float32(p["dx"])), float(np.float32(p["dy"])),
                              float(np.float32(p["dz"])), np.float64)
        v2 = v2dt2(ve, float(np.float32(p["dt"])), np.float64)
        src = ricker(p["nt"], float(np.float32(p["dt"])), float(np.float32(p["fpeak"])),
                     np.float64)
        g = taper_table(blen, float(np.float32(p["taper_factor"])), np.float64)
    return coefs, v2, src, g


def pipeline_modeling(p, vel, dtype=np.float32, fmt=None):
    """Records per shot, [nshot][nchunk][nrx][nry][ntstep] (modeling uses ABC).
    fmt = 'f16' / 'bf16' runs the float pipeline with 16-bit wavefields."""
    ve = extend_velocity(vel, p["blen"], "abc")
    coefs, v2, src, g = _setup(p, ve, dtype)
    geom = geom_from_params(p)
    return np.stack([model_shot(v2, coefs, g, src, p["blen"], p["stencil_order"], p["nt"],
                                p["ntstep"], s, geom, fmt) for s in shots_from_params(p)])


def pipeline_migration(p, vpe, records, dtype=np.float32, fmt=None):
    """Final [nx][ny][nz] image from an already-extended velocity model."""
    coefs, v2, src, g = _setup(p, vpe, dtype)
    geom = geom_from_params(p)
    bc = p["boundary_condition"]
    hbc = bc != "rbc"
    total = np.zeros(vpe.shape, dtype)
    if p.get("fmig"):
        # the CPU reference processes every frequency in one group
        tables = freq_tables(p["nt"], np.float32(p["dt"]), np.float32(p["fmig_min_freq"]),
                             np.float32(p["fmig_max_freq"]))
        # RTMFreqDomainKernel never clears its image grid between shots
        # (no imgGrid->reset(), unlike RTMFiniteDifferencesKernel_HBC.cpp:257):
        # every shot image it hands to the stack is the running sum so far
        running = np.zeros(vpe.shape, dtype)
    for k, s in enumerate(shots_from_params(p)):
        if p.get("fmig"):
            fmig_shot(v2, coefs, g, src, p["blen"], p["stencil_order"], p["nt"], p["ntstep"], s,
                      geom, records[k], tables, len(tables[0]), running)
            shot_img = running.copy()
        else:
            shot_img = migrate_shot(v2, coefs, g, src, p["blen"], p["stencil_order"], p["nt"],
                                    p["ntstep"], s, geom, records[k], hbc, fmt)
        total = total + shot_img
    return image_finish(total, p["blen"], float(np.float32(p["dx"])), float(np.float32(p["dy"])),
                        float(np.float32(p["dz"])))
