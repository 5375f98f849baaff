"""Batches of shots (rtmhip_batch_*) against the same shots run one at a time
through rtmhip_forward_steps / rtmhip_backward_steps (fused receiver step),
bit for bit: wavefields, UPB history and image of every shot, after every
chunk of a run split into chunks (so every captured-graph replay is checked,
with host synchronisation and downloads between replays).

The shots have uneven grids (nz not a multiple of 64, ny - 8 not a multiple
of the wave rows) and put the source in each special place of the per-step
small operators: below the upper border, inside the UPB layer, and in the
frame the stencil does not update.  The single-shot loops are themselves
pinned to the oracle (test_gpu_loops, test_gpu_16bit).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

B, ORDER, NT = 12, 8, 26
# (grid, source cell (extended), receiver geometry (interior coordinates))
SHOTS = [
    ((40, 44, 90), (20, 22, B + 2), (1, 2, 3, 3, 4, 7, 6)),      # source below the border
    ((34, 38, 70), (15, 17, B - 2), (0, 0, 0, 2, 2, 5, 7)),       # source in the UPB layer
    ((46, 40, 130), (2, 19, B - 6), (4, 3, 2, 0, 5, 1, 3)),       # source in the frame
]


def _wave(rng, shape, fmt):
    a = rng.uniform(-1, 1, shape).astype(np.float32)
    return a if fmt == "f32" else a.astype(np.float16).view(np.uint16)


class Shots:
    """Host inputs of the batch and device copies for one run."""

    def __init__(self, gpu, fmt, seed=5):
        self.gpu, self.fmt = gpu, fmt
        self.wd = gpu.WAVE_FORMATS[fmt][1]
        rng = np.random.default_rng(seed)
        self.coefs = np.zeros(27, np.float32)
        c8 = np.array([-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315,
                       -1 / 560])
        for d, h in enumerate((10.0, 11.0, 9.0)):
            self.coefs[d * 9:(d + 1) * 9] = (c8 / (h * h)).astype(np.float32)
        self.g = np.exp(-(np.sqrt(-np.log(0.9)) / B * np.arange(B)) ** 2).astype(np.float32)
        self.host = []
        for grid, src, rg in SHOTS:
            nx, ny, nz = grid
            h = dict(grid=grid, src=src, rg=rg,
                     V=rng.uniform(0.0, 0.15, grid).astype(np.float32),
                     ricker=rng.uniform(-1, 1, NT).astype(np.float32),
                     rcv=rng.uniform(-1, 1, (rg[5], rg[6], NT)).astype(np.float32),
                     img=rng.uniform(-1, 1, grid).astype(np.float32),
                     S=(_wave(rng, grid, fmt), _wave(rng, grid, fmt)),
                     R=(_wave(rng, grid, fmt), _wave(rng, grid, fmt)))
            self.host.append(h)

    def device(self):
        gpu = self.gpu
        out = []
        for h in self.host:
            nx, ny, nz = h["grid"]
            m = dict(shot=gpu.make_shot(nx, ny, nz, B, ORDER, NT, NT, h["src"],
                                        gpu.RcvGeom(*h["rg"]), self.fmt),
                     V=gpu.DeviceArray.from_host(h["V"]),
                     ricker=gpu.DeviceArray.from_host(h["ricker"]),
                     rcv=gpu.DeviceArray.from_host(h["rcv"]),
                     img=gpu.DeviceArray.from_host(h["img"]),
                     upb=gpu.DeviceArray((NT, nx, ny, ORDER // 2), self.wd).zero(),
                     S=gpu.FieldPair(h["grid"], self.wd), R=gpu.FieldPair(h["grid"], self.wd))
            m["S"].a.upload(h["S"][0])
            m["S"].b.upload(h["S"][1])
            m["R"].a.upload(h["R"][0])
            m["R"].b.upload(h["R"][1])
            out.append(m)
        self.dg = gpu.DeviceArray.from_host(self.g)
        return out

    @staticmethod
    def state(m):
        return [m["S"].P.download(), m["S"].PP.download(), m["R"].P.download(),
                m["R"].PP.download(), m["upb"].download(), m["img"].download()]

    def free(self, members):
        for m in members:
            for k in ("V", "ricker", "rcv", "img", "upb", "S", "R"):
                m[k].free()
        self.dg.free()


def _run_single(gpu, sh, members, fwd_chunks, bwd_chunks):
    """Per-shot reference; returns the state of every shot after each chunk."""
    states = []
    for it0, n in fwd_chunks:
        for m in members:
            gpu.forward_steps(m["shot"], True, m["V"], sh.coefs, sh.dg, m["ricker"], m["S"],
                              m["upb"], it0, n)
        states.append([Shots.state(m) for m in members])
    for hi, n in bwd_chunks:
        for m in members:
            gpu.backward_steps(m["shot"], True, m["V"], sh.coefs, sh.dg, m["S"], m["R"],
                               m["upb"], m["rcv"], m["img"], hi, n, True)
        states.append([Shots.state(m) for m in members])
    return states


def _run_batch(gpu, sh, members, fwd_chunks, bwd_chunks, modes, graph_steps):
    batch = gpu.Batch(members, sh.coefs, sh.dg)
    try:
        for cls, mode in enumerate(modes):
            batch.set_stencil(cls, mode)
        batch.set_graph(graph_steps)
        states = []
        for it0, n in fwd_chunks:
            batch.forward_steps(it0, n)
            states.append([Shots.state(m) for m in members])
        for hi, n in bwd_chunks:
            batch.backward_steps(hi, n)
            states.append([Shots.state(m) for m in members])
    finally:
        batch.destroy()
    return states


def _compare(want, got):
    names = ["P", "PP", "Pr", "PPr", "upb", "img"]
    for c, (ws, gs) in enumerate(zip(want, got)):
        for k, (w, g) in enumerate(zip(ws, gs)):
            for n, a, b in zip(names, w, g):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), \
                    "chunk %d shot %d field %s differs" % (c, k, n)


FWD = [(0, 11), (11, 15)]            # 1 eager + 2 graph replays (G = 4) + 2 eager, ...
BWD = [(NT - 1, 12), (NT - 13, 13)]  # the two seeding steps, then batched steps


@pytest.mark.parametrize("modes, graph_steps", [
    ((0, 0, 0), 4),   # per-shot stencil launches inside the graph
    ((8, 8, 8), 4),   # one launch for all shots, 8 rows per wave
    ((6, 6, 6), 2),
    ((8, 0, 6), 0),   # mixed, no graph
    ((0, 0, 0), 8),   # the defaults
], ids=["pershot-g4", "ry8-g4", "ry6-g2", "mixed-eager", "default-g8"])
@pytest.mark.parametrize("fmt", ["f32", "f16"])
def test_batch_matches_single_shots(gpu, fmt, modes, graph_steps):
    print("batch", fmt, modes, graph_steps, flush=True)
    sh = Shots(gpu, fmt)
    ref = sh.device()
    want = _run_single(gpu, sh, ref, FWD, BWD)
    sh.free(ref)
    mem = sh.device()
    got = _run_batch(gpu, sh, mem, FWD, BWD, modes, graph_steps)
    sh.free(mem)
    assert np.abs(want[-1][0][5]).max() > 0
    _compare(want, got)


def test_batch_rejects_mismatched_shots(gpu):
    sh = Shots(gpu, "f32")
    mem = sh.device()
    try:
        s = mem[1]["shot"]
        s.ntstep = NT // 2
        with pytest.raises(gpu.RtmError, match="ntstep"):
            gpu.Batch(mem, sh.coefs, sh.dg)
        s.ntstep = NT
        s.blen = B + 1
        with pytest.raises(gpu.RtmError, match="blen"):
            gpu.Batch(mem, sh.coefs, sh.dg)
        s.blen = B
        b = gpu.Batch(mem, sh.coefs, sh.dg)
        with pytest.raises(gpu.RtmError):
            b.set_graph(3)
        with pytest.raises(gpu.RtmError):
            b.set_stencil(1, 7)
        with pytest.raises(gpu.RtmError):
            b.forward_steps(NT - 2, 5)
        b.destroy()
    finally:
        sh.free(mem)
