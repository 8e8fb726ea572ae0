"""Ragged batch (iirx_batch_*): many channels of uneven length, order and
coefficients in one persistent launch.  No counterpart in the reference; every
channel is checked against the FP64 cascade with the reference's accuracy
gate (ref_generate.py golden + test_accuracy.cu tolerance), as a single
channel is."""
import numpy as np
import pytest

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O
from workloads import ragged_channels, ragged_input


def _explicit_channels():
    """Edge lengths around the chunk sizes (64 x 16/32/64) and mixed orders."""
    lens = [1, 2, 3, 63, 64, 65, 1023, 1024, 1025, 2047, 2048, 2049, 4095, 4096, 4097, 5000, 70001, 0, 12345]
    orders = [2, 4, 8, 16]
    wns = [0.02, 0.1, 0.3, 0.45]
    offsets = np.zeros(len(lens) + 1, np.int64)
    offsets[1:] = np.cumsum(lens)
    sections, gains = [], []
    for i in range(len(lens)):
        s, g = iirx.butterworth_sos(orders[i % 4], wns[(i // 4) % 4])
        sections.append(s)
        gains.append(g)
    return offsets, sections, np.asarray(gains)


def _check_channels(y, x, offsets, sections, chans=None):
    chans = range(len(sections)) if chans is None else chans
    for c in chans:
        a, b = int(offsets[c]), int(offsets[c + 1])
        if b == a:
            continue
        ref = O.golden(sections[c], x[a:b])
        ok, err, mref, rel = O.gate(y[a:b], ref)
        assert ok, (c, b - a, sections[c].shape[0], rel)


# --- host-side validation (no device needed: it is checked before any HIP call)

def _create_status(offsets, ns, sos, n_blocks=16, algo=None, bs_min_sections=0, bs_tiles=0):
    import ctypes
    from cuda_recursive_filtering_amd import _lib
    cfg = _lib.BatchConfig()
    cfg.n_blocks = n_blocks
    cfg.bs_min_sections = bs_min_sections
    cfg.bs_tiles = bs_tiles
    off = np.ascontiguousarray(offsets, np.int64)
    nsa = np.ascontiguousarray(ns, np.int32)
    s = np.ascontiguousarray(sos, np.float32)
    h = ctypes.c_void_p()
    al = None if algo is None else np.ascontiguousarray(algo, np.int32)
    st = _lib.lib().iirx_batch_create_ex(ctypes.byref(cfg), nsa.size, off.ctypes.data, nsa.ctypes.data,
                                         s.ctypes.data, None if al is None else al.ctypes.data, ctypes.byref(h))
    if st == 0:
        _lib.lib().iirx_batch_destroy(h)
    return st


def test_batch_rejects_bad_channels():
    sos, _ = iirx.butterworth_sos(4, 0.1)
    good = np.concatenate([sos, sos])
    assert _create_status([0, 10, 5], [2, 2], good) == 1          # decreasing offsets
    assert _create_status([-1, 10, 20], [2, 2], good) == 1        # negative start
    assert _create_status([0, 10, 20], [2, 9], np.zeros((11, 6))) == 1  # 9 sections
    bad = good.copy()
    bad[1, 4] = np.nan
    assert _create_status([0, 10, 20], [2, 2], bad) == 1          # non-finite coefficient
    assert _create_status([0, 10, 20], [2, 2], good, n_blocks=8) == 2  # no instance for N=8
    assert _create_status([0, 10, 20], [2, 2], good, bs_tiles=48) == 2  # nor for 48 block-state tiles
    assert _create_status([0, 0, 0], [2, 2], good) == 1           # no samples at all
    assert _create_status([0, 10, 20], [2, 2], good, algo=[0, 2]) == 1   # unknown algorithm
    assert _create_status([0, 10, 20], [2, 2], good, algo=[-2, 0]) == 1
    assert _create_status([0, 10, 20], [2, 2], good, bs_min_sections=-1) == 1


def _algos(kind, n):
    """Per-channel algorithm list of a test case: all PH, all block-state, or
    alternating."""
    if kind == "mixed":
        return ["ph" if c % 2 else "blockstate" for c in range(n)]
    return kind


# --- GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("n_blocks", [16, 32, 64])
@pytest.mark.parametrize("schedule", ["persistent", "ticket"])
@pytest.mark.parametrize("algo", ["ph", "blockstate", "mixed"])
def test_batch_edge_lengths(device, n_blocks, schedule, algo):
    import torch
    print(f"case batch edge lengths N={n_blocks} {schedule} {algo}", flush=True)
    offsets, sections, gains = _explicit_channels()
    x = ragged_input(offsets, gains, seed=40)
    xin = torch.from_numpy(x).to(device)
    out = torch.full_like(xin, float("nan"))
    with iirx.Batch(offsets, sections, n_blocks=n_blocks, schedule=schedule, algo=_algos(algo, len(sections)),
                    bs_tiles=n_blocks) as b:
        for _ in range(3):  # launch-versioned lookback: repeated launches on one batch
            b.execute(xin, out)
            b.check()
        y = out.cpu().numpy()
        info = b.info()
    assert info["n_launches"] == (2 if algo == "mixed" else 1)
    assert info["n_blockstate"] == {"ph": 0, "blockstate": len(sections), "mixed": (len(sections) + 1) // 2}[algo]
    _check_channels(y, x, offsets, sections)
    # no write past any channel: the packed output is covered exactly once
    assert np.isfinite(y).all()


@pytest.mark.gpu
def test_batch_many_tiny_channels(device):
    """20000 channels of 0-300 samples (mostly one partial chunk each, some
    empty), orders 2-16 with order 16 on the block-state form: every channel
    starts from zero state and no sample of a neighbour leaks in."""
    import torch
    print("case batch 20000 tiny channels", flush=True)
    rng = np.random.default_rng(12)
    n_ch = 20000
    lens = rng.integers(0, 301, n_ch)
    offsets = np.zeros(n_ch + 1, np.int64)
    offsets[1:] = np.cumsum(lens)
    designs = {o: iirx.butterworth_sos(o, 0.1) for o in (2, 4, 8, 16)}
    orders = rng.choice([2, 4, 8, 16], n_ch)
    sections = [designs[o][0] for o in orders]
    gains = np.array([designs[o][1] for o in orders])
    x = ragged_input(offsets, gains, seed=1200)
    xin = torch.from_numpy(x).to(device)
    out = torch.full_like(xin, float("nan"))
    with iirx.Batch(offsets, sections) as b:
        b.execute(xin, out)
        b.check()
        info = b.info()
    assert info["n_blockstate"] == int((orders == 16).sum()) and info["n_launches"] == 2
    y = out.cpu().numpy()
    assert np.isfinite(y).all()
    _check_channels(y, x, offsets, sections, chans=range(0, n_ch, 37))


@pytest.mark.gpu
@pytest.mark.parametrize("bs_tiles", [16, 64])
def test_batch_device_tables_equal_host(device, bs_tiles):
    """Block-state channel tables are built in FP64 on the GPU, one workgroup
    per channel; each must equal the host precompute for tiles of 64 samples
    and bs_tiles tiles per chunk, bit for bit."""
    print(f"case batch device tables G={bs_tiles}", flush=True)
    offsets, sections, gains = ragged_channels(12, seed=9, hi_log2=14)
    with iirx.Batch(offsets, sections, algo="blockstate", bs_tiles=bs_tiles) as b:
        dev = b.device_tables()
    off = 0
    for c, sec in enumerate(sections):
        host, _ = iirx.plan_tables("BS", sec, 64 * bs_tiles, block_size=bs_tiles, n_blocks=64)
        got = dev[off:off + host.size]
        assert np.array_equal(got.view(np.uint32), host.view(np.uint32)), c
        off += host.size
    assert off == dev.size


@pytest.mark.gpu
def test_batch_algorithms_agree(device):
    """The same channels run as PH and as the block-state form: both pass the
    FP64 gate, and the block-state RMS error stays at PH's level (an operand
    or slab mistake that stays within the max gate shows up there)."""
    import torch
    print("case batch PH vs block-state", flush=True)
    offsets, sections, gains = ragged_channels(48, seed=6, hi_log2=15)
    x = ragged_input(offsets, gains, seed=600)
    xin = torch.from_numpy(x).to(device)
    ys = {}
    for algo in ("ph", "blockstate"):
        out = torch.empty_like(xin)
        with iirx.Batch(offsets, sections, algo=algo) as b:
            b.execute(xin, out)
            b.check()
        ys[algo] = out.cpu().numpy()
        _check_channels(ys[algo], x, offsets, sections)
    rms = {k: [] for k in ys}
    for c in range(len(sections)):
        a, e = int(offsets[c]), int(offsets[c + 1])
        ref = O.golden(sections[c], x[a:e])
        scale = np.sqrt(np.mean(ref * ref))
        for k in ys:
            rms[k].append(np.sqrt(np.mean((ys[k][a:e] - ref) ** 2)) / scale)
    print({k: float(np.median(v)) for k, v in rms.items()})
    assert np.median(rms["blockstate"]) <= 2.0 * np.median(rms["ph"])


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["ph", "blockstate"])
def test_batch_16_byte_staging_any_offset(device, algo):
    """Full chunks of 16-byte aligned buffers move with 16-byte accesses
    shifted by the chunk's offset mod 4; buffers that are not 16-byte aligned
    take 4-byte accesses.  Channels start at every offset mod 4 and span
    several full chunks; both paths pass the gate per channel and agree to
    8 float32 ulps of each channel's peak, and nothing outside the channels
    is written."""
    import torch
    print(f"case batch 16-byte staging {algo}", flush=True)
    lengths = [3 * 2048 + 1, 2 * 2048 + 2, 4 * 2048 + 3, 2048 + 7, 5 * 2048, 2 * 2048 + 5, 3 * 2048 + 6, 4096]
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    assert {int(o) % 4 for o in offsets[:-1]} == {0, 1, 2, 3}
    sections = [iirx.butterworth_sos([2, 4, 8, 16][c % 4], 0.05 + 0.04 * c)[0] for c in range(len(lengths))]
    n = int(offsets[-1])
    x = (np.random.default_rng(7).standard_normal(n) * 0.01).astype(np.float32)
    outs = {}
    for shift in (0, 1):  # float offset 1: not 16-byte aligned
        xin = torch.zeros(n + 4, dtype=torch.float32, device=device)
        xin[shift:shift + n] = torch.from_numpy(x).to(device)
        out = torch.full((n + 8,), 7.0, dtype=torch.float32, device=device)
        with iirx.Batch(offsets, sections, algo=algo, n_blocks=32, bs_tiles=32) as b:
            b.execute(xin[shift:shift + n], out[4 + shift:4 + shift + n])
            b.check()
        o = out.cpu().numpy()
        assert np.all(o[:4 + shift] == 7.0) and np.all(o[4 + shift + n:] == 7.0), "write outside the channels"
        outs[shift] = o[4 + shift:4 + shift + n]
    # the same arithmetic per chunk; a carry taken from a full record or from
    # folded partials (timing) may differ in the last bits at lookback depth >= 2
    for c in range(len(lengths)):
        a, e = int(offsets[c]), int(offsets[c + 1])
        peak = float(np.max(np.abs(outs[0][a:e])))
        assert np.max(np.abs(outs[0][a:e] - outs[1][a:e])) <= 8 * np.finfo(np.float32).eps * peak, c
    _check_channels(outs[0], x, offsets, sections)
    _check_channels(outs[1], x, offsets, sections)


@pytest.mark.gpu
def test_batch_ragged_workload(device):
    """256 channels of the ragged workload's distribution (lengths 2^10..2^16)."""
    import torch
    print("case batch ragged 256 channels", flush=True)
    offsets, sections, gains = ragged_channels(256, seed=4, hi_log2=16)
    x = ragged_input(offsets, gains, seed=400)
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    with iirx.Batch(offsets, sections, bs_min_sections=4) as b:
        b.execute(xin, out)
        b.check()
        info = b.info()
    assert info["n_channels"] == 256 and info["num_chunks"] > 256
    assert info["n_blockstate"] == int(sum(s.shape[0] >= 4 for s in sections)) > 0
    _check_channels(out.cpu().numpy(), x, offsets, sections)


@pytest.mark.gpu
def test_batch_graph_replay(device):
    """One captured launch replayed with new input contents each time; every
    replay is checked, with eager work and synchronisation between replays."""
    import torch
    print("case batch graph replay", flush=True)
    offsets, sections, gains = ragged_channels(64, seed=5, hi_log2=15)
    xin = torch.empty(int(offsets[-1]), dtype=torch.float32, device=device)
    out = torch.empty_like(xin)
    s = torch.cuda.Stream()
    with iirx.Batch(offsets, sections, bs_min_sections=4) as b:
        b.capture(xin, out, s)
        assert b.info()["graph_ready"] == 1
        for r in range(4):
            x = ragged_input(offsets, gains, seed=500 + 97 * r)
            xin.copy_(torch.from_numpy(x))
            torch.cuda.synchronize()
            b.replay(s)
            s.synchronize()
            b.check()
            _check_channels(out.cpu().numpy(), x, offsets, sections)
            # eager launch between replays shares the lookback state
            b.execute(xin, out)
            b.check()
            _check_channels(out.cpu().numpy(), x, offsets, sections, chans=range(0, 64, 7))


@pytest.mark.gpu
@pytest.mark.slow
def test_batch_full_ragged_workload(device):
    """The full ragged workload: 4096 channels, lengths 2^10..2^20 (~6.2e8
    samples); 64 channels spread over the batch are checked against FP64."""
    import torch
    print("case batch full ragged workload", flush=True)
    offsets, sections, gains = ragged_channels(4096, seed=4)
    n = int(offsets[-1])
    chans = list(range(0, 4096, 64))
    xin = torch.randn(n, dtype=torch.float32, device=device)
    g = torch.from_numpy(np.repeat(gains, np.diff(offsets)).astype(np.float32)).to(device)
    xin.mul_(g)
    del g
    out = torch.empty_like(xin)
    with iirx.Batch(offsets, sections) as b:
        b.execute(xin, out)
        b.check()
        info = b.info()
    # by default the order-16 channels run the block-state form, in a second launch
    assert info["n_blockstate"] == sum(s.shape[0] == 8 for s in sections) and info["n_launches"] == 2
    x_cpu = {c: xin[int(offsets[c]):int(offsets[c + 1])].cpu().numpy() for c in chans}
    y_cpu = {c: out[int(offsets[c]):int(offsets[c + 1])].cpu().numpy() for c in chans}
    for c in chans:
        ref = O.golden(sections[c], x_cpu[c])
        ok, err, mref, rel = O.gate(y_cpu[c], ref)
        assert ok, (c, rel)


@pytest.mark.gpu
def test_batch_autotuned(device):
    """The batch autotuner picks one of its searched geometries and the batch
    it builds filters every channel within the accuracy gate."""
    import torch
    print("case batch autotune 64 channels", flush=True)
    offsets, sections, gains = ragged_channels(64, seed=6, hi_log2=15)
    x = ragged_input(offsets, gains, seed=600)
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    with iirx.Batch.autotuned(offsets, sections, reps=3) as b:
        assert b.tuned_ms > 0
        assert b._cfg.n_blocks in (16, 32, 64) and b._cfg.bs_tiles in (16, 32, 64)
        assert b._cfg.bs_min_sections in (4, 8, 9)
        b.execute(xin, out)
        b.check()
    _check_channels(out.cpu().numpy(), x, offsets, sections)


def test_batch_autotune_rejects_bad_arguments():
    import ctypes
    from cuda_recursive_filtering_amd import _lib
    cfg = _lib.BatchConfig()
    best = _lib.BatchConfig()
    ms = ctypes.c_float()
    off = np.array([0, 10], np.int64)
    ns = np.array([1], np.int32)
    sos = np.array([[1, 0, 0, 1, -0.5, 0]], np.float32)
    st = _lib.lib().iirx_batch_autotune(ctypes.byref(cfg), 1, off.ctypes.data, ns.ctypes.data, sos.ctypes.data, 0,
                                        ctypes.byref(best), ctypes.byref(ms))
    assert st == _lib.IIRX_ERR_INVALID
