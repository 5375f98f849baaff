"""ctypes binding of librtm3d_host.so: the host set-up arithmetic of the
driver (stencil taps, Ricker samples, taper table, velocity extension and
border fills, v^2 dt^2) and an in-process run of a JSON input."""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "lib", "librtm3d_host.so")
_lib = None
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
BC = {"abc": 0, "hbc": 1, "rbc": 2}


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError("librtm3d_host.so not built; run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        i, f = C.c_int, C.c_float
        L.rtm3d_stencil_coefs.argtypes = [i, f, f, f, _f32p]
        L.rtm3d_stencil_coefs.restype = i
        L.rtm3d_ricker.argtypes = [i, f, f, _f32p]
        L.rtm3d_taper_table.argtypes = [i, f, _f32p]
        L.rtm3d_extend_velocity.argtypes = [_f32p, i, i, i, i, i, _f32p]
        L.rtm3d_v2dt2.argtypes = [_f32p, C.c_size_t, f]
        L.rtm3d_freq_tables.argtypes = [i, f, f, f, C.c_void_p, C.c_void_p, C.c_void_p, i]
        L.rtm3d_freq_tables.restype = i
        L.rtm3d_run_json.argtypes = [C.c_char_p, i]
        L.rtm3d_run_json.restype = i
        _lib = L
    return _lib


def stencil_coefs(order, dx, dy, dz):
    out = np.zeros(3 * (order + 1), np.float32)
    if load().rtm3d_stencil_coefs(order, dx, dy, dz, out) != 0:
        raise ValueError("stencil order %d not supported (2, 4, 6 or 8)" % order)
    return out


def ricker(nt, dt, fpeak):
    out = np.zeros(nt, np.float32)
    load().rtm3d_ricker(nt, dt, fpeak, out)
    return out


def taper_table(blen, factor):
    out = np.zeros(max(blen, 1), np.float32)
    load().rtm3d_taper_table(blen, factor, out)
    return out[:blen] if blen > 0 else out


def extend_velocity(vel, blen, bc):
    nx, ny, nz = vel.shape
    out = np.zeros((nx + 2 * blen, ny + 2 * blen, nz + 2 * blen), np.float32)
    load().rtm3d_extend_velocity(np.ascontiguousarray(vel, np.float32), nx, ny, nz, blen,
                                 BC.get(bc, 1), out)
    return out


def v2dt2(ve, dt):
    out = np.array(ve, np.float32, copy=True)
    load().rtm3d_v2dt2(out, out.size, dt)
    return out


def freq_tables(nt, dt, fmin, fmax):
    """Frequency-domain migration tables: (w2[nw], kre[nw][nt], kim[nw][nt])."""
    L = load()
    nw = L.rtm3d_freq_tables(nt, dt, fmin, fmax, None, None, None, 0)
    w2 = np.zeros(max(nw, 1), np.float32)
    kre = np.zeros((max(nw, 1), nt), np.float32)
    kim = np.zeros((max(nw, 1), nt), np.float32)
    L.rtm3d_freq_tables(nt, dt, fmin, fmax, w2.ctypes.data_as(C.c_void_p),
                        kre.ctypes.data_as(C.c_void_p), kim.ctypes.data_as(C.c_void_p), nw)
    return w2[:nw], kre[:nw], kim[:nw]


def run_json(path, device=0):
    return load().rtm3d_run_json(path.encode(), device)
