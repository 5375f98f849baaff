"""Fixed launch sequence for hardware-counter profiling of the stencil.

Run under rocprofv3 (one counter pass per run), e.g.
    rocprofv3 --pmc FETCH_SIZE -d OUT -o run --output-format csv -- \
        python3 rtm3d_demo_amd/profile_stencil.py

On the 592x592x336 grid of the 512x512x256 + 40 salt workload it launches:
  * k_stack over the grid 3 times: a calibration kernel that reads 8 B and
    writes 4 B per cell, so its counters give the scale between the counter
    and true bytes;
  * the plain stencil step 5 times (rtmhip_step);
  * 5 HBC forward steps (tapered stencil class) and 5 fused backward steps;
  * 7 backward steps with the source and receiver steps in one pass
    (mode 2: two seeding steps, then 5 dual steps).
Fields hold random values; counters do not depend on them.

    python3 rtm3d_demo_amd/profile_stencil.py --summarize FETCH_DIR WRITE_DIR [--tiling ...]

turns the counter_collection.csv files of a FETCH_SIZE pass and a WRITE_SIZE
pass into per-kernel bytes per dispatch.  Each counter is scaled by what it
reads for k_stack, whose true traffic is known (8 B read + 4 B written per
cell), and the first kernel of each stencil class is printed as
"traffic_bytes_per_launch[class C, tile T]: N", the lines bench.py reads (0
plain, 1 fused receiver, 2 tapered forward, 3 dual backward; T the class's
tile shape TZxBYxRY, or "default").

--tiling C:TZ,BY,RY (repeatable) sets the tile shape of stencil class C for
the sequence; give the same options to --summarize so that the lines carry
the shapes.

--wave-format f16 (or bf16) runs the 16-bit sequence instead: k_stack, then
5 HBC forward steps, 7 fused backward steps (mode 1) and 5 + 5 snapshot-RTM
steps (16-bit snapshots) with 16-bit wavefields and the library's defaults
(separate upper-taper pass, two-z waves).  Its summary (--summarize with the
same option) gives per-kernel bytes and the 16-bit algorithmic bytes per
updated cell: 10 for a step (V stays FP32), 20 for the imaging step.
"""
import csv
import glob
import os
import re
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


SHAPE = (592, 592, 336)


def _per_dispatch(d, counter):
    """Mean counter value (KiB) per dispatch, keyed by full kernel name."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % d)
    per = {}
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row["Counter_Name"] != counter:
                    continue
                key = (row["Kernel_Name"], row["Dispatch_Id"])
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    out = {}
    for (name, _), v in per.items():
        out.setdefault(name, []).append(v)
    return {k: (sum(v) / len(v), len(v)) for k, v in out.items()}


_MANGLED_TYPES = {"f": "float", "DF16_": "_Float16", "DF16b": "__bf16", "u6__bf16": "__bf16"}


def _demangle_args(args):
    """Template arguments of an Itanium-mangled kernel name (the subset the
    library's kernels use: f, DF16_, DF16b / u6__bf16, Li<n>E, Lb<0|1>E)."""
    out = []
    i = 0
    while i < len(args):
        for code, name in _MANGLED_TYPES.items():
            if args.startswith(code, i):
                out.append(name)
                i += len(code)
                break
        else:
            m = re.match(r"L([ib])(\d+)E", args[i:])
            if not m:
                break
            out.append(("true" if m.group(2) == "1" else "false") if m.group(1) == "b"
                       else m.group(2))
            i += m.end()
    return out


def _short(name):
    if name.startswith("_Z"):
        # names the profiler left mangled (16-bit template arguments)
        m = re.search(r"\d+(k_\w+?)I(.*?)EEv", name)
        if m:
            return "%s<%s>" % (m.group(1), ", ".join(_demangle_args(m.group(2))))
        return name
    name = name.replace("rtmhip::(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def parse_tilings(argv):
    """{class: (tz, by, ry)} from repeated --tiling C:TZ,BY,RY options."""
    out = {}
    for i, a in enumerate(argv):
        if a == "--tiling":
            c, shape = argv[i + 1].split(":")
            out[int(c)] = tuple(int(v) for v in shape.split(","))
    return out


def summarize(fetch_dir, write_dir, tilings=None, wave_format="f32"):
    tilings = tilings or {}
    fetch = _per_dispatch(fetch_dir, "FETCH_SIZE")
    write = _per_dispatch(write_dir, "WRITE_SIZE")
    cells = SHAPE[0] * SHAPE[1] * SHAPE[2]
    stack = [k for k in fetch if _short(k) == "k_stack"]
    if not stack:
        raise SystemExit("k_stack not found: calibration impossible")
    kf, kw = fetch[stack[0]][0] * 1024.0, write[stack[0]][0] * 1024.0
    f_scale, w_scale = 8.0 * cells / kf, 4.0 * cells / kw
    interior = (SHAPE[0] - 8) * (SHAPE[1] - 8) * (SHAPE[2] - 8)
    print("calibration: k_stack reads %.1f MB and writes %.1f MB per dispatch;"
          % (8e-6 * cells, 4e-6 * cells))
    print("FETCH_SIZE reads %.1f MB (scale %.3f), WRITE_SIZE %.1f MB (scale %.3f)"
          % (kf * 1e-6, f_scale, kw * 1e-6, w_scale))
    print()
    print("| kernel | dispatches | read MB | write MB | total MB | B per updated cell |")
    print("|---|---|---|---|---|---|")
    first = {}
    for name in sorted(fetch, key=lambda k: -fetch[k][0]):
        if name not in write:
            continue
        r = fetch[name][0] * 1024.0 * f_scale
        w = write[name][0] * 1024.0 * w_scale
        short = _short(name)
        print("| %s | %d | %.1f | %.1f | %.1f | %.2f |" % (
            short, fetch[name][1], r * 1e-6, w * 1e-6, (r + w) * 1e-6, (r + w) / interior))
        cls = None
        if short.startswith("k_stencil_dual"):
            cls = 3
        elif short.startswith("k_stencil"):
            tail = short.rsplit(",", 2)[-2:]  # the IMG and TAPER arguments
            img, tap = (t.strip(" >") not in ("false", "0") for t in tail)
            cls = 1 if img else (2 if tap else 0)
        if cls is not None and cls not in first:
            first[cls] = (short, r + w)
    print()
    if wave_format != "f32":
        # 16-bit wavefields: P, PP read, PP written (2 B each), V (4 B); the
        # imaging step adds PS (2 B) and the FP32 image read and written
        for cls in sorted(first):
            short, t = first[cls]
            bpc = 20.0 if cls == 1 else 10.0
            print("%s class %d (%s): algorithmic %.0f B per launch (%.0f B x %d cells), "
                  "traffic %.0f B (%.2f x)" % (wave_format, cls, short, bpc * interior, bpc,
                                               interior, t, t / (bpc * interior)))
        return
    algo = {0: 16.0, 1: 28.0, 2: 16.0, 3: 36.0}
    for cls in sorted(first):
        short, t = first[cls]
        print("class %d (%s): algorithmic %.0f B per launch (%.0f B x %d updated cells)"
              % (cls, short, algo[cls] * interior, algo[cls], interior))
        tile = "%dx%dx%d" % tilings[cls] if cls in tilings else "default"
        print("traffic_bytes_per_launch[class %d, tile %s]: %.0f" % (cls, tile, t))


def parse_wave_format(argv):
    return argv[argv.index("--wave-format") + 1] if "--wave-format" in argv else "f32"


def run16(hip, host, fmt):
    """The 16-bit launch sequence (see the module docstring)."""
    shape = SHAPE
    b, order, nt = 40, 8, 16
    rng = np.random.default_rng(0)

    def rand(lo, hi):
        out = np.empty(shape, np.float32)
        for x in range(shape[0]):
            out[x] = rng.uniform(lo, hi, shape[1:])
        return out

    wd = hip.WAVE_FORMATS[fmt][1]
    V = hip.DeviceArray.from_host(rand(0.0, 0.2))
    img = hip.DeviceArray(shape).zero()
    coefs = host.stencil_coefs(order, 12.5, 12.5, 12.5)
    g = hip.DeviceArray.from_host(host.taper_table(b, 0.9))
    src = hip.DeviceArray.from_host(host.ricker(nt, 1.1e-3, 25.0))
    F, R = hip.FieldPair(shape, wd), hip.FieldPair(shape, wd)
    for f in (F.a, F.b, R.a, R.b):
        f.zero()
    upb = hip.DeviceArray((nt, shape[0], shape[1], 4), wd).zero()
    rcv = hip.DeviceArray((128, 128, nt)).zero()
    geom = hip.RcvGeom(0, 0, 0, 4, 4, 128, 128)
    shot = hip.make_shot(*shape, b, order, nt, nt, (296, 296, 42), geom, fmt)
    for _ in range(3):
        hip.stack(img, V)
    hip.forward_steps(shot, True, V, coefs, g, src, F, upb, 0, 5)
    hip.backward_steps(shot, True, V, coefs, g, F, R, upb, rcv, img, nt - 1, 7, True)
    upb.free()
    snaps = hip.snap_buffer(shot, "f16", 1)
    hip.snap_forward_steps(shot, V, coefs, g, src, F, snaps, "f16", 1, 0, 5)
    hip.snap_backward_steps(shot, V, coefs, g, R, snaps, "f16", 1, rcv, img, nt - 1, 5)
    hip.synchronize()
    print("profile sequence done", flush=True)


def main():
    tilings = parse_tilings(sys.argv)
    wfmt = parse_wave_format(sys.argv)
    if len(sys.argv) >= 4 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2], sys.argv[3], tilings, wfmt)
        return 0
    from rtm3d_demo_amd import hip, host
    if wfmt != "f32":
        hip.init(0)
        run16(hip, host, wfmt)
        hip.finalize()
        return 0

    shape = SHAPE
    b, order, nt = 40, 8, 16
    hip.init(0)
    for cls, t in tilings.items():
        hip.set_tiling(cls, *t, 0)
    rng = np.random.default_rng(0)

    def rand(lo, hi):
        out = np.empty(shape, np.float32)
        for x in range(shape[0]):
            out[x] = rng.uniform(lo, hi, shape[1:])
        return out

    V = hip.DeviceArray.from_host(rand(0.0, 0.2))
    coefs = host.stencil_coefs(order, 12.5, 12.5, 12.5)
    g = hip.DeviceArray.from_host(host.taper_table(b, 0.9))
    src = hip.DeviceArray.from_host(host.ricker(nt, 1.1e-3, 25.0))
    F, R = hip.FieldPair(shape), hip.FieldPair(shape)
    F.a.upload(rand(-1, 1))
    F.b.upload(rand(-1, 1))
    img = hip.DeviceArray(shape).zero()
    upb = hip.DeviceArray((nt, shape[0], shape[1], 4)).zero()
    rcv = hip.DeviceArray((128, 128, nt)).zero()
    geom = hip.RcvGeom(0, 0, 0, 4, 4, 128, 128)
    shot = hip.make_shot(*shape, b, order, nt, nt, (296, 296, 42), geom)
    for _ in range(3):
        hip.stack(img, V)
    for _ in range(5):
        hip.step(F.a, F.b, V, coefs, order)
    hip.forward_steps(shot, True, V, coefs, g, src, F, upb, 0, 5)
    hip.backward_steps(shot, True, V, coefs, g, F, R, upb, rcv, img, nt - 1, 7, True)
    hip.backward_steps(shot, True, V, coefs, g, F, R, upb, rcv, img, nt - 1, 7, 2)
    hip.synchronize()
    print("profile sequence done", flush=True)
    hip.finalize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
