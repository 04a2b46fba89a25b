"""Randomized parity against the oracle: batches of random items, random
frames through the per-frame stages, random viewer windows, and random
streams served to several threads at once.

Each batch mixes 24 items of random kind (16-colour art, alpha sprites,
curves, dense crossings), size (1..160 x 1..120), scale (1..15),
subdivision, byte order, row order, output flip, input row padding, output
row padding and output alignment; every item must equal the oracle and no
byte outside the images may change.  Seeds are fixed, so a failure names
its batch and item.
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

SENTINEL = 0xA5


def random_batch(D, seed, n_items=24):
    rng = np.random.default_rng(seed)
    items, blobs, meta = [], [], []
    in_off, out_off = 0, int(rng.choice([0, 4, 8, 12]))
    for k in range(n_items):
        kind = int(rng.choice([0, 1, 2, 4]))
        ch = 4 if kind == 1 else 3
        W, H = int(rng.integers(1, 161)), int(rng.integers(1, 121))
        S = int(rng.integers(1, 16))
        while W * H * S * S > 3_000_000:
            S -= 1
        sub = int(rng.random() < 0.25)
        flags = 0
        if rng.random() < 0.4:
            flags |= D.TOPDOWN
        if rng.random() < 0.3:
            flags |= D.RGB
        if kind == 1:
            flags |= D.ALPHA
        flip = rng.random() < 0.25
        if flip:
            flags |= D.FLIP_OUT
        pad = int(rng.integers(0, 8))
        img = D.synth_frame(kind, int(seed * 1000 + k), W, H, ch, W * ch + pad)
        out_stride = W * S * 4 + int(rng.choice([0, 0, 0, 4, 8, 20]))
        items.append(D.Item(in_off, out_off, W, H, img.shape[1], ch, S, sub, flags, out_stride))
        blobs.append(img.reshape(-1))
        in_off += img.size
        out_off += out_stride * H * S + int(rng.choice([0, 4, 16, 64]))
        meta.append((kind, W, H, ch, S, sub, flags, flip, img))
    return items, np.concatenate(blobs), out_off + 16, meta


@pytest.mark.parametrize("seed", list(range(30)))
def test_random_batch_matches_oracle(gpu, D, O, seed):
    torch = gpu
    items, blob, out_bytes, meta = random_batch(D, 5000 + seed)
    print("batch seed", 5000 + seed, flush=True)
    d_in = torch.from_numpy(blob).to("cuda:0")
    d_out = torch.full((out_bytes,), SENTINEL, dtype=torch.uint8, device="cuda:0")
    ctx = D.Context(0)
    ctx.run(d_in, d_out, items)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    ctx.close()
    covered = np.zeros(out.size, dtype=bool)
    for k, (it, (kind, W, H, ch, S, sub, flags, flip, img)) in enumerate(zip(items, meta)):
        ref = O.depixelize(img, W, H, scale=S, subdivide=bool(sub), channels=ch, flags=flags & ~D.FLIP_OUT)["rgba"]
        if flip:
            ref = ref[::-1]
        rows = out[it.out_offset:it.out_offset + it.out_stride * H * S].reshape(H * S, it.out_stride)
        got = rows[:, :W * S * 4].reshape(H * S, W * S, 4)
        for y in range(H * S):
            covered[it.out_offset + y * it.out_stride:it.out_offset + y * it.out_stride + W * S * 4] = True
        what = f"item {k}: kind {kind} {W}x{H} ch {ch} x{S} sub {sub} flags {flags} out_off%16 {it.out_offset % 16}"
        bad = np.argwhere((got != ref).any(-1))
        assert len(bad) == 0, f"{what}: {len(bad)} pixels differ, first at {bad[:3].tolist()}"
    assert (out[~covered] == SENTINEL).all(), "bytes outside the images were written"


CS = 45


@pytest.mark.parametrize("seed", list(range(80)))
def test_random_frame_stages_match_oracle(gpu, D, O, seed):
    """The plain per-frame path on a random frame (kind, size 1..120 x 1..90,
    scale 1..8, subdivision, byte and row order): pixels, graph, edge counts
    and the defined slots of the triangle diagram equal the oracle's, and so
    do the component labels and the border chains of its graph."""
    torch = gpu
    rng = np.random.default_rng(9000 + seed)
    kind = int(rng.choice([0, 1, 2, 4]))
    ch = 4 if kind == 1 else 3
    W, H = int(rng.integers(1, 121)), int(rng.integers(1, 91))
    S, sub = int(rng.integers(1, 9)), bool(rng.random() < 0.4)
    flags = (D.TOPDOWN if rng.random() < 0.4 else 0) | (D.RGB if rng.random() < 0.3 else 0) | (D.ALPHA if kind == 1 else 0)
    img = D.synth_frame(kind, 9000 + seed, W, H, ch, W * ch + int(rng.integers(0, 8)))
    what = f"case {seed}: kind {kind} {W}x{H} ch {ch} x{S} sub {sub} flags {flags}"
    print(what, flush=True)
    ref = O.depixelize(img, W, H, scale=S, subdivide=sub, channels=ch, flags=flags, want_diagram=True)
    rgba, g, ec, diag = D.frame(img, W, H, ch, scale=S, subdivide=sub, flags=flags, want_diagram=True)
    assert np.array_equal(g, ref["graph"]), what + ": graph"
    assert np.array_equal(ec, ref["edge_count"]), what + ": edge counts"
    dd, rd = diag.reshape(-1, CS, 2), ref["diagram"].reshape(-1, CS, 2)
    for k, m in enumerate(np.asarray(ec).reshape(-1)):
        w = max(0, (int(m) - 2) * 3)
        assert np.array_equal(dd[k, :w], rd[k, :w]), f"{what}: triangles of cell {k}"
    assert np.array_equal(rgba, ref["rgba"]), what + ": pixels"
    s = torch.cuda.current_stream()
    d_g = torch.from_numpy(np.ascontiguousarray(ref["graph"]).reshape(-1)).to("cuda:0")
    d_lab = torch.full((W * H,), -1, dtype=torch.int32, device="cuda:0")
    D.check(D.lib.dpx_stage_components(d_g.data_ptr(), W, H, d_lab.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    assert np.array_equal(d_lab.cpu().numpy().reshape(H, W), O.components(ref["graph"])), what + ": labels"
    from test_gpu_parity import gpu_border_chains
    chains, nn, status = gpu_border_chains(torch, D, ref["graph"])
    assert status == 0 and chains == O.border_chains(ref["graph"]), what + ": border chains"


@pytest.mark.parametrize("seed", list(range(60)))
def test_random_view_matches_oracle(gpu, D, O, seed):
    """dpx_stage_present on a random frame and view (cells or quads, scale
    0.50..16.00, translation -40..40, window 1..256 x 1..192, antialiasing,
    row order, with and without subdivision) equals orc_present."""
    import ctypes
    torch = gpu
    rng = np.random.default_rng(7000 + seed)
    kind = int(rng.choice([0, 1, 2, 4]))
    ch = 4 if kind == 1 else 3
    W, H = int(rng.integers(1, 49)), int(rng.integers(1, 41))
    flags = (D.TOPDOWN if rng.random() < 0.4 else 0) | (D.ALPHA if kind == 1 else 0)
    sub = bool(rng.random() < 0.5)
    mode, sc = int(rng.integers(0, 2)), int(rng.integers(50, 1601))
    tx, ty = int(rng.integers(-40, 41)), int(rng.integers(-40, 41))
    ww, wh = int(rng.integers(1, 257)), int(rng.integers(1, 193))
    vf = (D.VIEW_AA if rng.random() < 0.5 else 0) | (D.VIEW_TOPDOWN if rng.random() < 0.3 else 0)
    what = f"case {seed}: kind {kind} {W}x{H} sub {sub} mode {mode} x{sc / 100} t ({tx},{ty}) {ww}x{wh} view {vf}"
    print(what, flush=True)
    img = D.synth_frame(kind, 7000 + seed, W, H, ch)
    ref = O.depixelize(img, W, H, scale=1, subdivide=sub, channels=ch, flags=flags, want_diagram=True)
    _, _, ec, diag = D.frame(img, W, H, ch, scale=1, subdivide=sub, flags=flags, want_diagram=True)
    d_img = torch.from_numpy(img).to("cuda:0")
    d_diag = torch.from_numpy(np.ascontiguousarray(diag)).to("cuda:0")
    d_ec = torch.from_numpy(np.ascontiguousarray(ec)).to("cuda:0")
    d_out = torch.zeros(wh * ww * 4, dtype=torch.uint8, device="cuda:0")
    view = D.View(mode, sc, tx, ty, ww, wh, vf, 0)
    D.check(D.lib.dpx_stage_present(d_img.data_ptr(), W, H, img.shape[1], ch, flags, d_diag.data_ptr(),
                                    d_ec.data_ptr(), ctypes.byref(view), d_out.data_ptr(), ww * 4,
                                    torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    want = O.present(img, W, H, ref["diagram"], ref["edge_count"], mode=mode, scale_x100=sc, tx=tx, ty=ty,
                     width=ww, height=wh, view_flags=vf, channels=ch, flags=flags)
    got = d_out.cpu().numpy().reshape(wh, ww, 4)
    bad = np.argwhere((got != want).any(-1))
    assert len(bad) == 0, f"{what}: {len(bad)} pixels differ, first at {bad[:3].tolist()}"


def test_random_streams_from_threads_match_oracle(gpu, D, O):
    """Four threads, each with four streams of random size, scale, subdivision
    and layout on one server, submit random frames in bursts of 1..8 and wait
    for them; every result equals the oracle's (continuous batching mixes the
    threads' requests on the device)."""
    import threading
    srv = D.Server(budget_bytes=8 << 30, max_batch=32)
    errors, done = [], []
    lock = threading.Lock()

    def worker(wid):
        rng = np.random.default_rng(300 + wid)
        streams = []
        for _ in range(4):
            kind = int(rng.choice([0, 1, 2, 4]))
            ch = 4 if kind == 1 else 3
            W, H = int(rng.integers(1, 97)), int(rng.integers(1, 81))
            S, sub = int(rng.integers(1, 9)), bool(rng.random() < 0.3)
            flags = (D.TOPDOWN if rng.random() < 0.4 else 0) | (D.ALPHA if kind == 1 else 0)
            streams.append((srv.open_stream(W, H, ch, S, sub, flags), kind, W, H, ch, S, sub, flags))
        n = 0
        for burst in range(25):
            pend = []
            for _ in range(int(rng.integers(1, 9))):
                sid, kind, W, H, ch, S, sub, flags = streams[int(rng.integers(0, len(streams)))]
                img = D.synth_frame(kind, int(rng.integers(0, 1 << 30)), W, H, ch)
                out = np.zeros((H * S, W * S * 4), dtype=np.uint8)
                pend.append((srv.submit(sid, img, out), img, out, (kind, W, H, ch, S, sub, flags)))
            for t, img, out, (kind, W, H, ch, S, sub, flags) in pend:
                rc = srv.wait(t)
                ref = O.depixelize(img, W, H, scale=S, subdivide=sub, channels=ch, flags=flags)["rgba"]
                if rc != 0 or not np.array_equal(out.reshape(H * S, W * S, 4), ref):
                    with lock:
                        errors.append((wid, burst, rc, kind, W, H, S, sub, flags))
                n += 1
        for st in streams:
            srv.close_stream(st[0])
        with lock:
            done.append(n)

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    srv.close()
    print("frames", sum(done), flush=True)
    assert len(done) == 4 and not errors, f"{len(errors)} failed requests, first {errors[:3]}"
