"""GPU parity of the format adapters and the presentation stage.

* palette-indexed frames (DPX_INDEXED) on the batched and serving paths: an
  indexed frame must give exactly the output and graph of the 4-channel
  frame of its palette entries (the oracle runs on that expanded frame);
* DPX_FLIP_OUT: the output rows of the same frame in the opposite order;
* dpx_stage_present: the viewer window against the oracle's orc_present.
Bar: bit-exact."""
import ctypes

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


def assert_equal(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = np.argwhere(a != b)
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first at {bad[:3].tolist()}"


def to_indexed(img, W, H, ch, pad, seed):
    """(index rows (H, W + pad), palette (256, 4), expanded 4-channel frame (H, 4W))."""
    px = np.ascontiguousarray(img[:, :W * ch].reshape(H, W, ch))
    if ch == 3:
        px = np.concatenate([px, np.full((H, W, 1), 255, np.uint8)], axis=2)
    words = px.reshape(-1, 4).copy().view(np.uint32).reshape(-1)
    uniq, inv = np.unique(words, return_inverse=True)
    assert len(uniq) <= 256, f"{len(uniq)} colours"
    rng = np.random.default_rng(seed)
    order = rng.permutation(256)[:len(uniq)]  # scattered entries; the rest random and never read
    palette = rng.integers(0, 256, size=(256, 4), dtype=np.uint8)
    palette[order] = uniq.view(np.uint8).reshape(-1, 4)
    idx = np.zeros((H, W + pad), dtype=np.uint8)
    idx[:, :W] = order[inv].reshape(H, W)
    return idx, palette, np.ascontiguousarray(px.reshape(H, W * 4))


def run_items(torch, D, blobs_items):
    """blobs_items: list of (frame rows, palette or None, W, H, ch, scale, sub, flags)."""
    chunks, items, in_off, out_off = [], [], 0, 0

    def put(a):
        nonlocal in_off
        off = in_off
        chunks.append(a.reshape(-1))
        in_off += a.size
        pad = (in_off + 255) // 256 * 256 - in_off
        chunks.append(np.zeros(pad, dtype=np.uint8))
        in_off += pad
        return off

    for rows, pal, W, H, ch, s, sub, flags in blobs_items:
        fo = put(rows)
        po = put(pal) if pal is not None else 0
        it = D.Item(fo, out_off, W, H, rows.shape[1], ch, s, int(sub), flags, W * s * 4)
        it.palette_offset = po
        items.append(it)
        out_off += (W * s * 4 * H * s + 255) // 256 * 256
    d_in = torch.from_numpy(np.concatenate(chunks)).to("cuda:0")
    d_out = torch.zeros(out_off, dtype=torch.uint8, device="cuda:0")
    ctx = D.Context(0)
    ctx.keep_graph(True)
    ctx.run(d_in, d_out, items)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    gptr, goffs = ctx.graph(len(items))
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    res, graphs = [], []
    for it, go in zip(items, goffs):
        n = it.out_stride * it.height * it.scale
        res.append(out[it.out_offset:it.out_offset + n].reshape(it.height * it.scale, it.width * it.scale, 4))
        g = np.zeros((it.height, it.width), dtype=np.uint8)
        assert hip.hipMemcpy(g.ctypes.data, gptr + go, g.size, 2) == 0  # hipMemcpyDeviceToHost
        graphs.append(g)
    ctx.close()
    return res, graphs


# (kind, seed, W, H, channels of the source frame, flags, scale, subdivide, index-row padding)
INDEXED_CASES = [
    (0, 201, 256, 224, 3, 0, 4, False, 0),
    (0, 202, 128, 96, 3, 0, 4, True, 0),
    (2, 203, 61, 47, 3, 1, 3, True, 5),       # RGB order, ragged, padded rows
    (2, 204, 80, 72, 3, 2, 2, False, 0),      # top-down
    (1, 205, 24, 24, 4, 4, 6, False, 0),      # sprite, transparent class
    (1, 206, 48, 64, 4, 7, 6, True, 3),       # RGB + top-down + alpha, subdivided
    (0, 207, 1, 1, 3, 0, 5, False, 0),
    (0, 208, 7, 3, 3, 0, 8, True, 1),
]


def test_indexed_frames_match_their_expanded_frames(gpu, D, O):
    torch = gpu
    blobs, refs = [], []
    for kind, seed, W, H, ch, flags, s, sub, pad in INDEXED_CASES:
        img = D.synth_frame(kind, seed, W, H, ch)
        idx, pal, exp = to_indexed(img, W, H, ch, pad, seed)
        blobs.append((idx, pal, W, H, 1, s, sub, flags | D.INDEXED))
        blobs.append((exp, None, W, H, 4, s, sub, flags))  # the same frame as 4 channels, same batch
        refs.append(O.depixelize(exp, W, H, scale=s, subdivide=sub, channels=4, flags=flags))
    outs, graphs = run_items(torch, D, blobs)
    for k, c in enumerate(INDEXED_CASES):
        print(f"indexed case {c}", flush=True)
        assert_equal(graphs[2 * k], refs[k]["graph"], f"{c} indexed graph")
        assert_equal(outs[2 * k], refs[k]["rgba"], f"{c} indexed rgba")
        assert_equal(outs[2 * k + 1], refs[k]["rgba"], f"{c} expanded rgba")


def test_flip_out_reverses_output_rows(gpu, D, O):
    torch = gpu
    cases = [(0, 301, 96, 64, 3, 0, 4, False), (2, 302, 61, 40, 3, 2, 3, True), (1, 303, 32, 32, 4, 4, 6, True)]
    blobs, refs = [], []
    for kind, seed, W, H, ch, flags, s, sub in cases:
        img = D.synth_frame(kind, seed, W, H, ch)
        blobs.append((img, None, W, H, ch, s, sub, flags | D.FLIP_OUT))
        refs.append(O.depixelize(img, W, H, scale=s, subdivide=sub, channels=ch, flags=flags))
        rgba, _, _ = D.frame(img, W, H, ch, scale=s, subdivide=sub, flags=flags | D.FLIP_OUT)
        assert_equal(rgba, refs[-1]["rgba"][::-1], f"plain path {kind},{seed}")
    outs, _ = run_items(torch, D, blobs)
    for o, r, c in zip(outs, refs, cases):
        assert_equal(o, r["rgba"][::-1], f"batched {c}")


def test_indexed_server_stream_uses_the_palette_of_each_submit(gpu, D, O):
    W, H, s = 40, 30, 4
    srv = D.Server(64 << 20, max_batch=8)
    sid = srv.open_stream(W, H, 1, s, False, D.INDEXED)
    img_a = D.synth_frame(0, 401, W, H, 3)
    idx, pal_a, exp_a = to_indexed(img_a, W, H, 3, 0, 401)
    out = np.zeros((H * s, W * s * 4), dtype=np.uint8)
    with pytest.raises(D.DepixelError):
        srv.submit(sid, idx, out)  # no palette yet
    srv.set_palette(sid, pal_a)
    outs, tickets = [], []
    pal_b = pal_a.copy()
    pal_b[:, [0, 2]] = pal_b[:, [2, 0]]  # the same indices under swapped channels
    exp_b = np.ascontiguousarray(pal_b[idx].reshape(H, W * 4))
    for k in range(6):
        if k == 3:
            srv.set_palette(sid, pal_b)
        o = np.zeros((H * s, W * s * 4), dtype=np.uint8)
        tickets.append(srv.submit(sid, idx, o))
        outs.append(o)
    for t in tickets:
        assert srv.wait(t) == 0
    ref_a = O.depixelize(exp_a, W, H, scale=s, channels=4)["rgba"]
    ref_b = O.depixelize(exp_b, W, H, scale=s, channels=4)["rgba"]
    for k, o in enumerate(outs):
        assert_equal(o.reshape(H * s, W * s, 4), ref_a if k < 3 else ref_b, f"frame {k}")
    # frames and results in pinned blocks move by DMA; the palette stays resident
    # on the device (only the current one: the replaced palette was released)
    assert srv.stats().hbm_in_use == 1024
    pin_in = D.pinned_empty(idx.shape)
    pin_in[:] = idx
    pins = [D.pinned_empty((H * s, W * s * 4)) for _ in range(4)]
    tickets = [srv.submit(sid, pin_in, p) for p in pins]
    for t in tickets:
        assert srv.wait(t) == 0
    for k, p in enumerate(pins):
        assert_equal(p.reshape(H * s, W * s, 4), ref_b, f"pinned frame {k}")
    srv.close()


# (kind, seed, W, H, ch, flags, subdivide, mode, scale_x100, tx, ty, width, height, view flags)
VIEW_CASES = [
    (0, 501, 256, 224, 3, 0, False, 0, 800, 1, 1, 1200, 800, 1),   # the viewer's defaults
    (0, 502, 256, 224, 3, 0, True, 0, 800, 1, 1, 1200, 800, 1),
    (0, 503, 64, 48, 3, 0, True, 0, 400, 0, 0, 256, 192, 0),       # = the 4x raster
    (2, 504, 96, 64, 3, 1, False, 1, 850, 1, 1, 900, 600, 3),      # quads, AA, top-down window
    (2, 505, 61, 47, 3, 2, True, 0, 333, -5, 7, 500, 300, 0),
    (1, 506, 32, 32, 4, 4, True, 0, 650, 1, 1, 240, 230, 1),       # alpha sprite
    (0, 507, 9, 7, 3, 0, False, 0, 25, 0, 0, 4, 3, 1),             # scale 1/4
]


@pytest.mark.parametrize("case", VIEW_CASES, ids=lambda c: f"k{c[0]}_{c[2]}x{c[3]}m{c[7]}s{c[8]}f{c[13]}")
def test_present_matches_oracle(gpu, D, O, case):
    torch = gpu
    kind, seed, W, H, ch, flags, sub, mode, sc, tx, ty, ww, wh, vf = case
    print(f"view case {case}", flush=True)
    img = D.synth_frame(kind, seed, W, H, ch)
    ref = O.depixelize(img, W, H, scale=1, subdivide=sub, channels=ch, flags=flags, want_diagram=True)
    _, _, ec, diag = D.frame(img, W, H, ch, scale=1, subdivide=sub, flags=flags, want_diagram=True)
    dev = "cuda:0"
    d_img = torch.from_numpy(img).to(dev)
    d_diag = torch.from_numpy(np.ascontiguousarray(diag)).to(dev)
    d_ec = torch.from_numpy(np.ascontiguousarray(ec)).to(dev)
    d_out = torch.zeros(wh * ww * 4, dtype=torch.uint8, device=dev)
    view = D.View(mode, sc, tx, ty, ww, wh, vf, 0)
    D.check(D.lib.dpx_stage_present(d_img.data_ptr(), W, H, img.shape[1], ch, flags, d_diag.data_ptr(),
                                    d_ec.data_ptr(), ctypes.byref(view), d_out.data_ptr(), ww * 4,
                                    torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    want = O.present(img, W, H, ref["diagram"], ref["edge_count"], mode=mode, scale_x100=sc, tx=tx, ty=ty,
                     width=ww, height=wh, view_flags=vf, channels=ch, flags=flags)
    assert_equal(d_out.cpu().numpy().reshape(wh, ww, 4), want, "window")
