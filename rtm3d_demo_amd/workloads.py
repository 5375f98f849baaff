"""Synthetic velocity models and shot geometries of the benchmark workloads.

There are no field data here: every model is generated from a seed.

  three_layer_128   128^3 at 10 m, 1500/2500/3500 m/s, interfaces at z = 40, 80,
                    15 Hz Ricker 2 cells below the surface at the centre,
                    receivers every 2 cells on the surface, border 32, 1000 steps.
  salt_512          512x512x256 at 12.5 m: v(z) = 1500 + 0.8 z[m] capped at 4500,
                    Gaussian-smoothed +-5 % perturbation (seed 2024), then a
                    seeded random ellipsoidal salt body at 4500 m/s; 25 Hz Ricker
                    at (256, 256, 2), receivers every 4 cells on the full surface
                    (128 x 128), HBC border 40, 3000 steps, dt from CFL 0.4.
  salt_1024         1024x1024x320 at 12.5 m, built like salt_512 (v(z) gradient,
                    smoothed +-5 % perturbation, seeded ellipsoidal salt bodies);
                    64 shots with uneven apertures: nx = ny and nz drawn
                    seeded-uniform from {256, 288, ..., 448} and {256, 288, 320},
                    each window cut around its source; HBC border 40, 2500 steps,
                    receivers every 4 cells on the window's surface.
"""
import numpy as np


def three_layer(nx=128, ny=128, nz=128, z1=40, z2=80, v=(1500.0, 2500.0, 3500.0)):
    m = np.empty((nx, ny, nz), np.float32)
    m[:, :, :z1] = v[0]
    m[:, :, z1:z2] = v[1]
    m[:, :, z2:] = v[2]
    return m


def salt_model(nx=512, ny=512, nz=256, d=12.5, seed=2024, vmax=4500.0, smooth=4.0, pert=0.05):
    """v(z) gradient, smoothed random perturbation, ellipsoidal salt body."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float64) * d
    vz = np.minimum(1500.0 + 0.8 * z, vmax).astype(np.float32)
    v = np.empty((nx, ny, nz), np.float32)
    v[:] = vz[None, None, :]
    noise = rng.standard_normal((nx, ny, nz)).astype(np.float32)
    noise = gaussian_filter(noise, smooth, mode="reflect")
    noise *= np.float32(pert / np.abs(noise).max())
    v *= 1.0 + noise
    # salt: centre in the middle third laterally, upper-middle in depth
    cx = rng.uniform(0.35, 0.65) * nx
    cy = rng.uniform(0.35, 0.65) * ny
    cz = rng.uniform(0.45, 0.6) * nz
    ax = rng.uniform(0.12, 0.22) * nx
    ay = rng.uniform(0.12, 0.22) * ny
    az = rng.uniform(0.08, 0.15) * nz
    for x in range(nx):
        dx2 = ((x - cx) / ax) ** 2
        if dx2 >= 1.0:
            continue
        yy = (np.arange(ny)[:, None] - cy) / ay
        zz = (np.arange(nz)[None, :] - cz) / az
        inside = dx2 + yy ** 2 + zz ** 2 < 1.0
        v[x][inside] = vmax
    np.minimum(v, vmax, out=v)
    return v


def cfl_dt(vmax, d, courant=0.4):
    return float(np.float32(courant * d / vmax))


def params(name, nx, ny, nz, d, blen, nt, dt, fpeak, src, rcv, order=8, taper=0.9,
           bc="hbc"):
    """Reference-schema parameters (the keys of the reference JSON)."""
    sx, sy, sz = src
    rsx, rsy, rdz, rdx, rdy, rnx, rny = rcv
    return {
        "mname": name, "vpfile": "vel/model.bin", "vpefile": "vel/model_ext.bin",
        "save_vpe_file": False, "load_vpe_from_file": False,
        "outdir": "out", "tmpdir": "tmp", "datdir": "dat",
        "nx": nx, "ny": ny, "nz": nz, "blen": blen,
        "width_m": int(nx * d), "height_m": int(ny * d), "depth_m": int(nz * d),
        "nt": nt, "dz": d, "dy": d, "dx": d, "dt": dt, "fpeak": fpeak, "fmax": 2 * fpeak,
        "taper_factor": taper, "stencil_order": order, "modeling": True, "migration": False,
        "rtm_type": "acoustic", "boundary_condition": bc,
        "source_total": 1, "source_start_x": sx, "source_start_y": sy, "source_depth_z": sz,
        "source_distance_x": 1, "source_distance_y": 1, "source_count_x": 1,
        "source_count_y": 1, "receiver_total": rnx * rny, "receiver_start_x": rsx,
        "receiver_start_y": rsy, "receiver_depth_z": rdz, "receiver_distance_x": rdx,
        "receiver_distance_y": rdy, "receiver_count_x": rnx, "receiver_count_y": rny,
        "ntstep": nt, "nthreads": 16, "save_snapshots": False, "distributed_grid": False,
        "snapshot_step": 10, "fmig": False, "fmig_max_freq": 2 * fpeak, "fmig_min_freq": 1.0,
        "fmig_nwstep": 1, "fmig_distributed_imaging": False, "fpga_xclbin": "",
        "fpga_fwd_kernel": "", "fpga_bwd_kernel": "", "verbose": True,
    }


def salt_512_params(nt=3000):
    d = 12.5
    return params("salt512", 512, 512, 256, d, 40, nt, cfl_dt(4500.0, d), 25.0, (256, 256, 2),
                  (0, 0, 0, 4, 4, 128, 128))


def three_layer_128_params(nt=1000):
    d = 10.0
    return params("l3_128", 128, 128, 128, d, 32, nt, cfl_dt(3500.0, d), 15.0, (64, 64, 2),
                  (0, 0, 0, 2, 2, 64, 64))


# ---------------- the 1024 x 1024 x 320 model and its shot windows ----------------
SALT1024 = (1024, 1024, 320)
_salt1024_cache = {}


def _salt1024_fields(seed):
    """Perturbation on a coarse lattice (every 4th cell, smoothed; linearly
    interpolated to full resolution) and the salt ellipsoids."""
    if seed in _salt1024_cache:
        return _salt1024_cache[seed]
    from scipy.ndimage import gaussian_filter

    nx, ny, nz = SALT1024
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((nx // 4 + 1, ny // 4 + 1, nz // 4 + 1)).astype(np.float32)
    noise = gaussian_filter(noise, 1.0, mode="reflect")
    noise *= np.float32(0.05 / np.abs(noise).max())
    bodies = []
    for _ in range(4):
        bodies.append((rng.uniform(0.15, 0.85) * nx, rng.uniform(0.15, 0.85) * ny,
                       rng.uniform(0.45, 0.65) * nz, rng.uniform(0.06, 0.12) * nx,
                       rng.uniform(0.06, 0.12) * ny, rng.uniform(0.08, 0.15) * nz))
    _salt1024_cache[seed] = (noise, bodies)
    return noise, bodies


def _lerp_matrix(start, n, step, ncoarse):
    """Rows: full-resolution cells start .. start+n-1; columns: coarse nodes."""
    pos = (start + np.arange(n)) / step
    i0 = np.minimum(np.floor(pos).astype(np.int64), ncoarse - 2)
    w = (pos - i0).astype(np.float32)
    m = np.zeros((n, ncoarse), np.float32)
    m[np.arange(n), i0] = 1.0 - w
    m[np.arange(n), i0 + 1] = w
    return m


def salt_1024_model(seed=2025, vmax=4500.0, d=12.5):
    """The whole 1024 x 1024 x 320 model (1.3 GB, float32), cell for cell the
    windows of salt_1024_window put side by side."""
    nx, ny, nz = SALT1024
    v = np.empty(SALT1024, np.float32)
    step = 128
    for x0 in range(0, nx, step):
        v[x0:x0 + step] = salt_1024_window(x0, 0, step, ny, nz, seed, vmax, d)
    return v


def salt_1024_window(x0, y0, nx, ny, nz, seed=2025, vmax=4500.0, d=12.5):
    """Cells [x0, x0+nx) x [y0, y0+ny) x [0, nz) of the 1024^2 x 320 model."""
    noise, bodies = _salt1024_fields(seed)
    wx = _lerp_matrix(x0, nx, 4.0, noise.shape[0])
    wy = _lerp_matrix(y0, ny, 4.0, noise.shape[1])
    wz = _lerp_matrix(0, nz, 4.0, noise.shape[2])
    pert = np.tensordot(wx, noise, axes=(1, 0))           # nx, Cy, Cz
    pert = np.einsum("yc,xcz->xyz", wy, pert, optimize=True)
    pert = np.tensordot(pert, wz, axes=(2, 1))            # nx, ny, nz
    vz = np.minimum(1500.0 + 0.8 * np.arange(nz) * d, vmax).astype(np.float32)
    v = (vz[None, None, :] * (1.0 + pert)).astype(np.float32)
    yy = np.arange(y0, y0 + ny)[:, None]
    zz = np.arange(nz)[None, :]
    for cx, cy, cz, ax, ay, az in bodies:
        for x in range(max(x0, int(cx - ax)), min(x0 + nx, int(cx + ax) + 1)):
            dx2 = ((x - cx) / ax) ** 2
            if dx2 >= 1.0:
                continue
            inside = dx2 + ((yy - cy) / ay) ** 2 + ((zz - cz) / az) ** 2 < 1.0
            v[x - x0][inside] = vmax
    np.minimum(v, vmax, out=v)
    return v


def salt_1024_shots(n=64, seed=64):
    """n shots: (x0, y0, nxy, nz, source (interior coordinates of the window))."""
    rng = np.random.default_rng(seed)
    sizes = np.arange(256, 449, 32)
    depths = np.array([256, 288, 320])
    out = []
    for _ in range(n):
        nxy = int(rng.choice(sizes))
        nz = int(rng.choice(depths))
        sx = int(rng.integers(0, SALT1024[0]))
        sy = int(rng.integers(0, SALT1024[1]))
        x0 = min(max(sx - nxy // 2, 0), SALT1024[0] - nxy)
        y0 = min(max(sy - nxy // 2, 0), SALT1024[1] - nxy)
        out.append((x0, y0, nxy, nz, (sx - x0, sy - y0, 2)))
    return out


def salt_1024_params(nxy, nz, src, nt=2500):
    d = 12.5
    return params("salt1024_window", nxy, nxy, nz, d, 40, nt, cfl_dt(4500.0, d), 25.0, src,
                  (0, 0, 0, 4, 4, nxy // 4, nxy // 4))
