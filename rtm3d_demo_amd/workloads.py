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
