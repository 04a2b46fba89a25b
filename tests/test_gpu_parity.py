"""GPU parity: every HIP path against the oracle on the same seeded inputs.

Bar: bit-exact graph, edge counts, cell/triangle vertices and output pixels.
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

CS = 45
DX8, DY8 = [-1, 0, 1, -1, 1, -1, 0, 1], [1, 1, 1, 0, 0, -1, -1, -1]  # link e: (dx, dy)

# (kind, seed, W, H, channels, flags, scale, subdivide, row padding)
FRAME_CASES = [
    (0, 1, 256, 224, 3, 0, 4, False, 0),     # the 256x224 -> 4x workload
    (0, 2, 256, 224, 3, 0, 4, True, 0),
    (0, 3, 61, 47, 3, 0, 3, True, 8),        # ragged size, padded rows
    (2, 4, 96, 64, 3, 1, 4, False, 4),       # RGB order
    (2, 5, 80, 72, 3, 2, 2, True, 0),        # top-down rows
    (4, 6, 120, 90, 3, 0, 8, True, 4),
    (1, 7, 24, 24, 4, 4, 6, False, 0),       # sprite with alpha
    (1, 8, 48, 64, 4, 6, 6, True, 0),        # sprite, alpha + top-down
    (0, 9, 1, 1, 3, 0, 4, False, 0),         # single pixel
    (0, 10, 7, 1, 3, 0, 5, True, 3),         # single row
    (2, 11, 1, 9, 3, 2, 3, True, 0),         # single column
]


def make_frame(D, kind, seed, W, H, ch, pad):
    stride = W * ch + pad
    return D.synth_frame(kind, seed, W, H, ch, stride)


def assert_equal(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = np.argwhere(a != b)
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first at {bad[:3].tolist()}"


def defined_slots(diag, ec, width_fn):
    """{cell: slots} of the slots a stage defines."""
    d = diag.reshape(-1, CS, 2)
    return [d[k, :max(0, width_fn(int(n)))] for k, n in enumerate(np.asarray(ec).reshape(-1))]


@pytest.mark.parametrize("case", FRAME_CASES, ids=lambda c: f"k{c[0]}s{c[1]}_{c[2]}x{c[3]}c{c[4]}f{c[5]}x{c[6]}sub{int(c[7])}")
def test_plain_frame_matches_oracle(gpu, D, O, case):
    kind, seed, W, H, ch, flags, scale, sub, pad = case
    print(f"case {case}", flush=True)
    img = make_frame(D, kind, seed, W, H, ch, pad)
    ref = O.depixelize(img, W, H, scale=scale, subdivide=sub, channels=ch, flags=flags, want_diagram=True)
    rgba, graph, ec, diag = D.frame(img, W, H, ch, scale=scale, subdivide=sub, flags=flags, want_diagram=True)
    assert_equal(graph, ref["graph"], "graph")
    assert_equal(ec, ref["edge_count"], "edge_count")
    got = defined_slots(diag, ec, lambda n: (n - 2) * 3)
    exp = defined_slots(ref["diagram"], ref["edge_count"], lambda n: (n - 2) * 3)
    for k, (a, b) in enumerate(zip(got, exp)):
        assert np.array_equal(a, b), f"triangles of cell {k} differ"
    assert_equal(rgba, ref["rgba"], "rgba")


def test_stage_entry_points_match_oracle(gpu, D, O):
    torch = gpu
    W, H, S = 83, 61, 4
    img = D.synth_frame(2, 21, W, H, 3, W * 3 + 4)
    stride = img.shape[1]
    dev = "cuda:0"
    s = torch.cuda.current_stream()
    d_img = torch.from_numpy(img.reshape(-1)).to(dev)
    d_g = torch.zeros(W * H, dtype=torch.uint8, device=dev)
    d_ga = torch.zeros(W * H, dtype=torch.uint8, device=dev)
    D.check(D.lib.dpx_stage_similarity_graph(d_img.data_ptr(), W, H, stride, 3, 0, d_g.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    g0 = O.similarity_graph(img, W, H)
    assert_equal(d_g.cpu().numpy().reshape(H, W), g0, "similarity graph")
    D.check(D.lib.dpx_stage_trivial_cross(d_g.data_ptr(), W, H, s.cuda_stream))
    torch.cuda.synchronize()
    g1 = O.trivial_cross(g0)
    assert_equal(d_g.cpu().numpy().reshape(H, W), g1, "trivial cross")
    D.check(D.lib.dpx_stage_heuristic_cross(d_g.data_ptr(), d_ga.data_ptr(), W, H, s.cuda_stream))
    torch.cuda.synchronize()
    g2 = O.heuristic_cross(g1)
    assert_equal(d_g.cpu().numpy().reshape(H, W), g2, "heuristic cross")
    assert_equal(d_ga.cpu().numpy().reshape(H, W), g1, "heuristic snapshot")

    d_diag = torch.zeros(W * H * CS * 2, dtype=torch.float32, device=dev)
    d_daux = torch.zeros_like(d_diag)
    d_ec = torch.zeros(W * H, dtype=torch.int32, device=dev)
    d_eaux = torch.zeros_like(d_ec)
    D.check(D.lib.dpx_stage_cells(d_g.data_ptr(), W, H, d_diag.data_ptr(), d_ec.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    od, oec = O.cells(g2)
    assert_equal(d_ec.cpu().numpy(), oec, "cell counts")
    for k, (a, b) in enumerate(zip(defined_slots(d_diag.cpu().numpy(), oec, lambda n: n + 1),
                                   defined_slots(od.reshape(-1, 2), oec, lambda n: n + 1))):
        assert np.array_equal(a, b), f"hull of cell {k}"

    D.check(D.lib.dpx_stage_subdivide(d_img.data_ptr(), W, H, stride, 3, 0, d_g.data_ptr(), d_diag.data_ptr(),
                                      d_daux.data_ptr(), d_ec.data_ptr(), d_eaux.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    sd, sec, _, _ = O.subdivide(img, W, H, g2, od, oec)
    assert_equal(d_ec.cpu().numpy(), sec, "subdivided counts")
    for k, (a, b) in enumerate(zip(defined_slots(d_diag.cpu().numpy(), sec, lambda n: n),
                                   defined_slots(sd.reshape(-1, 2), sec, lambda n: n))):
        assert np.array_equal(a, b), f"subdivided polygon of cell {k}"

    D.check(D.lib.dpx_stage_triangulate(d_diag.data_ptr(), d_ec.data_ptr(), W, H, s.cuda_stream))
    torch.cuda.synchronize()
    td, _ = O.triangulate(sd, sec, W, H)
    assert_equal(d_diag.cpu().numpy().reshape(-1), td.reshape(-1), "triangulated diagram (all 45 slots)")

    d_pos = torch.zeros(W * H * CS * 2, dtype=torch.float32, device=dev)
    d_col = torch.zeros(W * H * CS * 4, dtype=torch.uint8, device=dev)
    D.check(D.lib.dpx_stage_export_vbo(d_img.data_ptr(), W, H, stride, 3, 0, d_diag.data_ptr(), d_ec.data_ptr(),
                                       d_pos.data_ptr(), d_col.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    pos, col = O.export_vbo(img, W, H, td, sec)
    assert_equal(d_pos.cpu().numpy().reshape(-1, 2), pos, "VBO positions")
    assert_equal(d_col.cpu().numpy().reshape(-1, 4), col, "VBO colours")

    d_out = torch.zeros(H * S * W * S * 4, dtype=torch.uint8, device=dev)
    D.check(D.lib.dpx_stage_rasterize(d_img.data_ptr(), W, H, stride, 3, 0, d_diag.data_ptr(), d_ec.data_ptr(), S,
                                      d_out.data_ptr(), W * S * 4, s.cuda_stream))
    torch.cuda.synchronize()
    ras = O.rasterize(img, W, H, td, sec, S)
    assert_equal(d_out.cpu().numpy().reshape(H * S, W * S, 4), ras, "raster")

    d_lab = torch.zeros(W * H, dtype=torch.int32, device=dev)
    D.check(D.lib.dpx_stage_components(d_g.data_ptr(), W, H, d_lab.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    assert_equal(d_lab.cpu().numpy().reshape(H, W), O.components(g2), "component labels")


def test_heuristic_stage_on_graphs_with_links_off_the_frame(gpu, D, O):
    """A caller's graph may hold links that leave the frame (the similarity
    stage never makes one): the curve walks of the heuristic stage then find
    a node without links past the edge, as in the same graph inside a border
    of nodes without links (the oracle's run on the padded graph, cropped).
    Border pixels with one link get a second one leading off the frame, so
    that walks along line-heavy frames' curves step past the edge."""
    torch = gpu
    dev = "cuda:0"
    s = torch.cuda.current_stream()
    stepped = changed = 0
    for seed in range(12):
        W, H = 24 + 3 * seed, 20 + 2 * seed
        img = D.synth_frame(2, 4300 + seed, W, H, 3)
        g = O.trivial_cross(O.similarity_graph(img, W, H)).copy()
        g0 = g.copy()
        rng = np.random.default_rng(seed)
        for j in range(H):
            for i in range(W):
                if bin(int(g[j, i])).count("1") != 1:
                    continue
                bits = [e for e in range(8) if not (0 <= i + DX8[e] < W and 0 <= j + DY8[e] < H)]
                if bits:
                    g[j, i] |= np.uint8(1 << int(rng.choice(bits)))
                    stepped += 1
        pad = np.zeros((H + 2, W + 2), dtype=np.uint8)
        pad[1:-1, 1:-1] = g
        want = O.heuristic_cross(pad)[1:-1, 1:-1]
        pad[1:-1, 1:-1] = g0
        changed += int(((want ^ g) != (O.heuristic_cross(pad)[1:-1, 1:-1] ^ g0)).any())
        d_g = torch.from_numpy(g.reshape(-1).copy()).to(dev)
        d_ga = torch.zeros(W * H, dtype=torch.uint8, device=dev)
        print("frame", seed, W, H, flush=True)
        D.check(D.lib.dpx_stage_heuristic_cross(d_g.data_ptr(), d_ga.data_ptr(), W, H, s.cuda_stream))
        torch.cuda.synchronize()
        assert_equal(d_g.cpu().numpy().reshape(H, W), want, f"heuristic cross, frame {seed}")
    # links were added, and walks past the edge changed some decisions
    assert stepped > 50 and changed > 0


def test_subdivide_stage_leaves_oversized_cells(gpu, D, O):
    """dpx_stage_subdivide takes the edge counts from the caller: a cell with
    17..31 points (more than subdivision's output slots hold at two points
    per vertex; real hulls have at most 8) is left as it is."""
    torch = gpu
    W, H = 24, 20
    img = D.synth_frame(0, 5, W, H, 3)
    g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H)))
    od, oec = O.cells(g)
    dev = "cuda:0"
    diag = od.reshape(W * H, CS, 2).copy()
    ec = oec.reshape(-1).copy()
    big = [k for k in range(W * H) if g.reshape(-1)[k] != 90][:15]
    rng = np.random.default_rng(3)
    for n, k in zip(range(17, 32), big):
        ec[k] = n
        diag[k, :n] = rng.integers(-1, 6, size=(n, 2)) * 0.25
    d_img = torch.from_numpy(img.reshape(-1)).to(dev)
    d_g = torch.from_numpy(g.reshape(-1).copy()).to(dev)
    d_diag = torch.from_numpy(diag.reshape(-1).astype(np.float32)).to(dev)
    d_daux = torch.zeros_like(d_diag)
    d_ec = torch.from_numpy(ec.astype(np.int32)).to(dev)
    d_eaux = torch.zeros_like(d_ec)
    s = torch.cuda.current_stream()
    D.check(D.lib.dpx_stage_subdivide(d_img.data_ptr(), W, H, img.shape[1], 3, 0, d_g.data_ptr(), d_diag.data_ptr(),
                                      d_daux.data_ptr(), d_ec.data_ptr(), d_eaux.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    got_d = d_diag.cpu().numpy().reshape(W * H, CS, 2)
    got_e = d_ec.cpu().numpy()
    for n, k in zip(range(17, 32), big):
        assert got_e[k] == n, f"cell {k}: count {got_e[k]} != {n}"
        np.testing.assert_array_equal(got_d[k], diag[k], err_msg=f"cell {k} with {n} points changed")


def test_launch_kernel_dropin(gpu, D, O):
    torch = gpu
    W, H = 64, 40
    img = D.synth_frame(0, 31, W, H, 3, W * 3)
    for sub in (False, True):
        pos = torch.zeros(W * H * CS * 2, dtype=torch.float32, device="cuda:0")
        col = torch.zeros(W * H * CS * 4, dtype=torch.uint8, device="cuda:0")
        diag, ec, graph = D.launch_kernel(img, W, H, sub, pos, col)
        ref = O.depixelize(img, W, H, scale=1, subdivide=sub, want_diagram=True)
        assert_equal(graph.reshape(H, W), ref["graph"], "graph_h")
        assert_equal(ec.reshape(H, W), ref["edge_count"], "edge_count_h")
        assert_equal(diag.reshape(-1), ref["diagram"].reshape(-1), "returned diagram")
        opos, ocol = O.export_vbo(img, W, H, ref["diagram"], ref["edge_count"].reshape(-1))
        assert_equal(pos.cpu().numpy().reshape(-1, 2), opos, "pos VBO")
        assert_equal(col.cpu().numpy().reshape(-1, 4), ocol, "colorPos VBO")


_hip = None


def hip_d2h(host, dev_ptr, n):
    """hipMemcpy device -> host of a raw device pointer (the context's graph arena)."""
    global _hip
    import ctypes
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip.hipMemcpy(host.ctypes.data, dev_ptr, n, 2)  # hipMemcpyDeviceToHost


def run_batch(torch, D, frames, items_spec):
    """frames: list of (img, W, H, ch); items_spec: list of (scale, sub, flags)."""
    in_off, out_off, items = 0, 0, []
    blobs = []
    for (img, W, H, ch), (scale, sub, flags) in zip(frames, items_spec):
        out_stride = W * scale * 4
        items.append(D.Item(in_off, out_off, W, H, img.shape[1], ch, scale, int(sub), flags, out_stride))
        blobs.append(img.reshape(-1))
        in_off += img.size
        in_off = (in_off + 255) // 256 * 256
        blobs.append(np.zeros(in_off - sum(b.size for b in blobs), dtype=np.uint8))
        out_off += out_stride * H * scale
        out_off = (out_off + 255) // 256 * 256
    d_in = torch.from_numpy(np.concatenate(blobs)).to("cuda:0")
    d_out = torch.zeros(out_off, dtype=torch.uint8, device="cuda:0")
    ctx = D.Context(0)
    ctx.keep_graph(True)
    ctx.run(d_in, d_out, items)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    gptr, goffs = ctx.graph(len(items))
    graphs = []
    for it, go in zip(items, goffs):
        n = it.width * it.height
        host = np.zeros(n, dtype=np.uint8)
        assert hip_d2h(host, gptr + go, n) == 0
        graphs.append(host.reshape(it.height, it.width))
    res = []
    for it in items:
        o = out[it.out_offset:it.out_offset + it.out_stride * it.height * it.scale]
        res.append(o.reshape(it.height * it.scale, it.out_stride)[:, :it.width * it.scale * 4]
                   .reshape(it.height * it.scale, it.width * it.scale, 4))
    ctx.close()
    return res, graphs


def test_tall_graph_tiles_match_oracle(gpu, D, O):
    """A batch large enough for full-height graph tiles, with heights for
    which the plan takes tiles of 56 rows rather than 32 (224 = 4 x 56; 100,
    110 and 50 rows end in a partial tile), in BGR and RGB, bottom-up and
    top-down rows, and 4-channel sprites with transparent pixels: sampled
    items' graphs and pixels equal the oracle's."""
    torch = gpu
    sizes = [(256, 224), (100, 100), (70, 110), (33, 50)]
    layouts = [0, D.RGB, D.TOPDOWN, D.RGB | D.TOPDOWN]
    frames, spec = [], []
    for k in range(160):
        W, H = sizes[k % 4]
        flags = layouts[(k // 4) % 4]
        if k % 5 == 4:  # a sprite with an alpha background
            frames.append((D.synth_frame(1, 900 + k, W, H, 4), W, H, 4))
            flags |= D.ALPHA
        else:
            frames.append((D.synth_frame((0, 2, 4)[k % 3], 900 + k, W, H, 3), W, H, 3))
        spec.append((2, False, flags))
    # the plan's rule, restated: enough 32-row tiles for full-height tiles,
    # and fewer rows converted (tiles x (rows + 4)) with 56-row tiles
    t32 = sum(-(-W // 64) * -(-H // 32) for _, W, H, _ in frames)
    t56 = sum(-(-W // 64) * -(-H // 56) for _, W, H, _ in frames)
    assert t32 >= 512 and t56 * 60 < t32 * 36
    outs, graphs = run_batch(torch, D, frames, spec)
    for k in range(0, 160, 7):
        img, W, H, ch = frames[k]
        print("item", k, W, H, ch, spec[k][2], flush=True)
        ref = O.depixelize(img, W, H, scale=2, channels=ch, flags=spec[k][2])
        assert_equal(graphs[k], ref["graph"], f"graph {k}")
        assert_equal(outs[k], ref["rgba"], f"rgba {k}")


def test_batched_mixed_items_match_oracle(gpu, D, O):
    torch = gpu
    frames, spec, refs = [], [], []
    cases = [
        (0, 41, 256, 224, 3, 0, 4, False, 0),
        (0, 42, 256, 224, 3, 0, 4, True, 0),
        (2, 43, 320, 224, 3, 0, 4, False, 0),
        (2, 44, 240, 160, 3, 2, 4, True, 4),
        (4, 45, 160, 120, 3, 0, 8, True, 0),
        (0, 46, 63, 37, 3, 1, 5, False, 8),
        (0, 47, 2, 2, 3, 0, 15, True, 0),
        (0, 48, 50, 30, 3, 0, 1, False, 0),
        (0, 49, 50, 30, 3, 0, 7, True, 0),
    ]
    for s in range(10):
        w, h = D.sprite_size(5, s)
        cases.append((1, 100 + s, w, h, 4, 4 | (2 if s % 3 == 0 else 0), 6, s % 2 == 1, 0))
    for kind, seed, W, H, ch, flags, scale, sub, pad in cases:
        img = make_frame(D, kind, seed, W, H, ch, pad)
        frames.append((img, W, H, ch))
        spec.append((scale, sub, flags))
        refs.append(O.depixelize(img, W, H, scale=scale, subdivide=sub, channels=ch, flags=flags))
    outs, graphs = run_batch(torch, D, frames, spec)
    for k, (o, g, r, c) in enumerate(zip(outs, graphs, refs, cases)):
        print(f"item {k}: {c}", flush=True)
        assert_equal(g, r["graph"], f"item {k} graph")
        assert_equal(o, r["rgba"], f"item {k} rgba")


def test_batch_result_independent_of_batch_mix(gpu, D):
    """The same frame alone and inside a mixed batch gives identical bytes."""
    torch = gpu
    img = D.synth_frame(0, 77, 256, 224, 3)
    alone, _ = run_batch(torch, D, [(img, 256, 224, 3)], [(4, False, 0)])
    others = [(D.synth_frame(2, 78 + k, 96 + 8 * k, 64 + 4 * k, 3), 96 + 8 * k, 64 + 4 * k, 3) for k in range(5)]
    frames = others[:2] + [(img, 256, 224, 3)] + others[2:]
    spec = [(4, k % 2 == 0, 0) for k in range(len(frames))]
    spec[2] = (4, False, 0)
    mixed, _ = run_batch(torch, D, frames, spec)
    assert_equal(mixed[2], alone[0], "frame alone vs in a batch")


def test_graph_output_is_opt_in(gpu, D):
    """Without dpx_context_keep_graph the batch writes the same images and
    no graphs, and dpx_batch_graph says so."""
    torch = gpu
    W, H, S = 96, 64, 4
    img = D.synth_frame(0, 91, W, H, 3)
    d_in = torch.from_numpy(img.reshape(-1)).to("cuda:0")
    outs = []
    for keep in (False, True):
        d_out = torch.zeros(H * S * W * S * 4, dtype=torch.uint8, device="cuda:0")
        ctx = D.Context(0)
        ctx.keep_graph(keep)
        ctx.run(d_in, d_out, [D.Item(0, 0, W, H, img.shape[1], 3, S, 0, 0, W * S * 4)])
        torch.cuda.synchronize()
        outs.append(d_out.cpu().numpy())
        if keep:
            ctx.graph(1)
        else:
            with pytest.raises(D.DepixelError):
                ctx.graph(1)
        ctx.close()
    assert_equal(outs[0], outs[1], "images with and without the graph output")


def test_unaligned_and_padded_output_layouts(gpu, D, O):
    """Output offsets that are 4 mod 16 and padded output rows take the
    narrower store paths (per-cell rows, 8- or 4-byte stores); the pixels
    and the padding around them must come out exactly."""
    torch = gpu
    cases = [(0, 131, 40, 24, 3, 0), (4, 132, 37, 21, 6, 2), (2, 133, 33, 17, 8, 8), (4, 134, 30, 18, 12, 0),
             (1, 135, 24, 24, 6, 4 | 8)]
    frames, items, refs = [], [], []
    in_off, out_off = 0, 4  # every output starts 4 bytes past a 16-byte boundary
    for kind, seed, W, H, S, flags in cases:
        ch = 4 if kind == 1 else 3
        img = D.synth_frame(kind, seed, W, H, ch)
        out_stride = W * S * 4 + 20  # padded rows
        items.append(D.Item(in_off, out_off, W, H, img.shape[1], ch, S, 0, flags, out_stride))
        frames.append(img.reshape(-1))
        in_off += img.size
        out_off += out_stride * H * S + 16
        refs.append(O.depixelize(img, W, H, scale=S, channels=ch, flags=flags & ~8)["rgba"])
    d_in = torch.from_numpy(np.concatenate(frames)).to("cuda:0")
    sentinel = 0xA5
    d_out = torch.full((out_off,), sentinel, dtype=torch.uint8, device="cuda:0")
    ctx = D.Context(0)
    ctx.run(d_in, d_out, items)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    ctx.close()
    covered = np.zeros(out_off, dtype=bool)
    for it, ref, c in zip(items, refs, cases):
        W, H, S = it.width, it.height, it.scale
        rows = out[it.out_offset:it.out_offset + it.out_stride * H * S].reshape(H * S, it.out_stride)
        got = rows[:, :W * S * 4].reshape(H * S, W * S, 4)
        want = ref[::-1] if c[5] & 8 else ref
        assert_equal(got, want, f"case {c}")
        for y in range(H * S):
            covered[it.out_offset + y * it.out_stride:it.out_offset + y * it.out_stride + W * S * 4] = True
    assert (out[~covered] == sentinel).all(), "bytes outside the images were written"


def test_repeated_and_changed_batches_on_one_context(gpu, D, O):
    """A context reuses the last batch's plan only for identical items: a
    repeat, an item pointed at another frame, another scale, and a reserve
    that replaces the descriptor buffer in between all give the oracle's
    output."""
    torch = gpu
    W, H = 64, 40
    frames = [D.synth_frame(0, 120 + k, W, H, 3) for k in range(3)]
    fb = frames[0].size
    d_in = torch.from_numpy(np.concatenate([f.reshape(-1) for f in frames])).to("cuda:0")
    ctx = D.Context(0)

    def run(items, S):
        d_out = torch.zeros(len(items) * H * S * W * S * 4, dtype=torch.uint8, device="cuda:0")
        ctx.run(d_in, d_out, items)
        torch.cuda.synchronize()
        return d_out.cpu().numpy().reshape(len(items), H * S, W * S, 4)

    def items_for(order, S):
        return [D.Item(k * fb, i * H * S * W * S * 4, W, H, W * 3, 3, S, 0, 0, W * S * 4) for i, k in enumerate(order)]

    refs = {(k, S): O.depixelize(frames[k], W, H, scale=S)["rgba"] for k in range(3) for S in (3, 4)}
    for order, S, reserve in (((0, 1), 4, False), ((0, 1), 4, False), ((2, 1), 4, False), ((2, 1), 3, False),
                              ((2, 1), 3, True), ((0, 2), 4, False)):
        if reserve:
            D.check(D.lib.dpx_context_reserve(ctx.h, 4096, W * H, 3))
        got = run(items_for(order, S), S)
        for i, k in enumerate(order):
            assert_equal(got[i], refs[(k, S)], f"order {order} scale {S} item {i}")
    ctx.close()


def gpu_border_chains(torch, D, g, max_nodes=None, max_chains=None):
    H, W = g.shape
    dev = "cuda:0"
    max_nodes = 8 * W * H + 8 if max_nodes is None else max_nodes
    max_chains = W * H + 1 if max_chains is None else max_chains
    d_g = torch.from_numpy(np.ascontiguousarray(g).reshape(-1)).to(dev)
    d_scr = torch.zeros(W * H, dtype=torch.uint8, device=dev)
    d_nodes = torch.full((max(1, max_nodes),), -1, dtype=torch.int32, device=dev)
    d_sizes = torch.zeros(max(1, max_chains), dtype=torch.int32, device=dev)
    d_counts = torch.zeros(3, dtype=torch.int32, device=dev)
    s = torch.cuda.current_stream()
    D.check(D.lib.dpx_stage_border_chains(d_g.data_ptr(), W, H, d_scr.data_ptr(), d_nodes.data_ptr(), max_nodes,
                                          d_sizes.data_ptr(), max_chains, d_counts.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    nc, nn, status = d_counts.cpu().tolist()
    nodes, sizes = d_nodes.cpu().numpy(), d_sizes.cpu().numpy()
    out, k = [], 0
    for sz in sizes[:nc]:
        out.append(nodes[k:k + sz].tolist())
        k += sz
    return out, nn, status


def test_component_labels_repeat_on_larger_graphs(gpu, D, O):
    """Union-find labels (canonical: the smallest index of the component) on
    resolved graphs of whole frames, each run three times: the final pass
    must not disturb labels other threads have already written."""
    torch = gpu
    s = torch.cuda.current_stream()
    for kind, seed, W, H in ((0, 41, 256, 224), (2, 42, 320, 224), (4, 43, 200, 150), (2, 44, 1, 57)):
        img = D.synth_frame(kind, seed, W, H, 3)
        g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H)))
        ref = O.components(g)
        d_g = torch.from_numpy(np.ascontiguousarray(g).reshape(-1)).to("cuda:0")
        for rep in range(3):
            print("components", kind, seed, W, H, "run", rep, flush=True)
            d_lab = torch.full((W * H,), -1, dtype=torch.int32, device="cuda:0")
            D.check(D.lib.dpx_stage_components(d_g.data_ptr(), W, H, d_lab.data_ptr(), s.cuda_stream))
            torch.cuda.synchronize()
            assert_equal(d_lab.cpu().numpy().reshape(H, W), ref, f"labels {kind}/{seed} run {rep}")


def test_border_chains_match_oracle(gpu, D, O):
    torch = gpu
    import json
    import os
    from conftest import GOLDEN
    d = json.load(open(os.path.join(GOLDEN, "reference_graphs.json")))["alex"]
    alex = np.array(d["graph"], dtype=np.uint8).reshape(d["height"], d["width"])
    expected = json.load(open(os.path.join(GOLDEN, "alex_chains.json")))["chains"]
    print("alex", flush=True)
    got, nn, status = gpu_border_chains(torch, D, alex)
    assert status == 0 and got == expected and nn == sum(map(len, expected))
    # (1024x700: the scan's marks do not fit a block's LDS and live in d_scratch)
    for kind, seed, W, H in ((0, 31, 256, 224), (2, 32, 97, 61), (4, 33, 120, 90), (0, 34, 1, 1), (2, 35, 7, 1),
                             (2, 36, 320, 288), (4, 37, 640, 480), (0, 38, 1024, 700), (2, 39, 33, 1),
                             (2, 40, 1, 40)):
        print("frame", kind, seed, W, H, flush=True)
        img = D.synth_frame(kind, seed, W, H, 3)
        g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H)))
        got, nn, status = gpu_border_chains(torch, D, g)
        assert status == 0
        assert got == O.border_chains(g), f"frame {kind}/{seed}"
    # every orthogonal link of a 2000x200 grid: the frame's edge is one chain
    # of 4,396 nodes, longer than the scan's output ring of 4,096
    g = np.full((200, 2000), 90, dtype=np.uint8)
    g[0, :] &= ~np.uint8(1 << 6)
    g[-1, :] &= ~np.uint8(1 << 1)
    g[:, 0] &= ~np.uint8(1 << 3)
    g[:, -1] &= ~np.uint8(1 << 4)
    print("grid 2000x200", flush=True)
    got, nn, status = gpu_border_chains(torch, D, g)
    ref = O.border_chains(g)
    assert status == 0 and max(map(len, ref)) > 4096 and got == ref


def random_symmetric_graph(seed):
    """A random graph of links in both directions (not one a frame would
    give): many pinched faces, where a node's own walk and the chain that
    lists it differ."""
    rng = np.random.default_rng(100000 + seed)
    W, H = int(rng.integers(9, 20)), int(rng.integers(8, 16))
    p = rng.random() * 0.6 + 0.1
    g = np.zeros((H, W), dtype=np.uint8)
    dx, dy = [-1, 0, 1, -1, 1, -1, 0, 1], [1, 1, 1, 0, 0, -1, -1, -1]
    for j in range(H):
        for i in range(W):
            for e in (1, 2, 4, 7):
                ni, nj = i + dx[e], j + dy[e]
                if 0 <= ni < W and 0 <= nj < H and rng.random() < p:
                    g[j, i] |= 1 << e
                    g[nj, ni] |= 1 << (7 - e)
    return g


def test_border_chains_of_random_graphs(gpu, D, O):
    """Random symmetric graphs, frames of 72..285 nodes (a last group of
    64 nodes that the frame does not fill; seeds 102, 1166 and 1698 once
    gave an extra chain from nodes past the frame's end)."""
    torch = gpu
    for seed in [102, 1166, 1698] + list(range(60)):
        g = random_symmetric_graph(seed)
        print("graph", seed, g.shape, flush=True)
        got, nn, status = gpu_border_chains(torch, D, g)
        assert status == 0 and got == O.border_chains(g), f"graph {seed}"


def test_border_chains_refuse_invalid_graphs(gpu, D, O):
    """A link that leaves the frame or has no reverse link: status
    DPX_EINVAL, no chains."""
    torch = gpu
    g = np.zeros((5, 6), dtype=np.uint8)
    g[0, 0] = 1 << 0  # up-left of (0, 0): outside the frame
    _, nn, status = gpu_border_chains(torch, D, g)
    assert status == D.EINVAL and nn == 0
    g = np.zeros((5, 6), dtype=np.uint8)
    g[2, 2] = 1 << 4  # right, without the left link back
    _, nn, status = gpu_border_chains(torch, D, g)
    assert status == D.EINVAL and nn == 0
    g[2, 3] = 1 << 3  # now symmetric: one two-node chain
    chains, nn, status = gpu_border_chains(torch, D, g)
    assert status == 0 and chains == O.border_chains(g)


def test_border_chains_report_small_outputs(gpu, D, O):
    torch = gpu
    import json
    import os
    from conftest import GOLDEN
    d = json.load(open(os.path.join(GOLDEN, "reference_graphs.json")))["alex"]
    alex = np.array(d["graph"], dtype=np.uint8).reshape(d["height"], d["width"])
    _, _, status = gpu_border_chains(torch, D, alex, max_nodes=10)
    assert status == D.ENOMEM
    _, _, status = gpu_border_chains(torch, D, alex, max_chains=3)
    assert status == D.ENOMEM


def test_large_same_size_batch_matches_oracle(gpu, D, O):
    """150 items of one size (a 3-D launch grid over items) with alternating
    row orders: every item equals the oracle (graph and pixels)."""
    torch = gpu
    n, W, H, S = 150, 40, 24, 3
    frames = [(D.synth_frame(2 if k % 2 else 0, 500 + k, W, H, 3), W, H, 3) for k in range(n)]
    spec = [(S, False, D.TOPDOWN if k % 3 == 0 else 0) for k in range(n)]
    outs, graphs = run_batch(torch, D, frames, spec)
    for k in range(n):
        img = frames[k][0]
        ref = O.depixelize(img, W, H, scale=S, flags=spec[k][2])
        assert_equal(graphs[k], ref["graph"], f"item {k} graph")
        assert_equal(outs[k], ref["rgba"], f"item {k} rgba")


def test_full_headline_batch(gpu, D, O):
    """The measured workload at its full size: 1024 frames of 256x224 -> 4x in
    one batch.  Sampled frames equal the oracle; every frame equals the same
    frame run in a batch of its own (each frame's result depends on that frame
    only); two runs of the batch are identical."""
    torch = gpu
    W, H, S, B = 256, 224, 4, 1024
    inb, outb = W * H * 3, W * S * H * S * 4
    host = np.stack([D.synth_frame(0, k, W, H, 3, W * 3) for k in range(B)])
    d_in = torch.from_numpy(host.reshape(-1)).to("cuda:0")
    items = D.items_array([D.Item(k * inb, k * outb, W, H, W * 3, 3, S, 0, 0, W * S * 4) for k in range(B)])
    ctx = D.Context(0)
    d_out = torch.empty(B * outb, dtype=torch.uint8, device="cuda:0")
    ctx.run(d_in, d_out, items)
    first = d_out.clone()
    ctx.run(d_in, d_out, items)
    torch.cuda.synchronize()
    assert torch.equal(first, d_out), "two runs of the same batch differ"
    out = d_out.view(B, H * S, W * S, 4)
    for k in (0, 1, 511, 777, 1023):
        print(f"frame {k}", flush=True)
        assert_equal(out[k].cpu().numpy(), O.depixelize(host[k], W, H, scale=S)["rgba"], f"frame {k} vs oracle")
    one = torch.empty(outb, dtype=torch.uint8, device="cuda:0")
    ctx2 = D.Context(0)
    for k in range(0, B, 97):
        ctx2.run(d_in, one, [D.Item(k * inb, 0, W, H, W * 3, 3, S, 0, 0, W * S * 4)])
        torch.cuda.synchronize()
        assert torch.equal(one.view(H * S, W * S, 4), out[k]), f"frame {k} alone vs in the batch"
    ctx.close()
    ctx2.close()


def test_large_frame_at_8x(gpu, D, O):
    """One 640x480 frame of the dense-crossing kind at 8x (5120x3840 out)."""
    torch = gpu
    W, H, S = 640, 480, 8
    img = D.synth_frame(4, 7, W, H, 3, W * 3)
    outs, graphs = run_batch(torch, D, [(img, W, H, 3)], [(S, False, 0)])
    ref = O.depixelize(img, W, H, scale=S)
    assert_equal(graphs[0], ref["graph"], "graph")
    assert_equal(outs[0], ref["rgba"], "rgba")


def test_full_sprite_batch_at_6x(gpu, D, O):
    """The 4096-sprite batch at its real size: sprites of six sizes with alpha
    backgrounds, one launch at 6x; every sprite's graph and pixels equal the
    oracle's."""
    torch = gpu
    frames, spec = [], []
    for k in range(4096):
        w, h = D.sprite_size(0, k)
        frames.append((D.synth_frame(1, k, w, h, 4), w, h, 4))
        spec.append((6, False, D.ALPHA))
    print("batch of", len(frames), "sprites", flush=True)
    outs, graphs = run_batch(torch, D, frames, spec)
    for k, (img, w, h, ch) in enumerate(frames):
        ref = O.depixelize(img, w, h, scale=6, channels=4, flags=D.ALPHA)
        assert_equal(graphs[k], ref["graph"], f"sprite {k} ({w}x{h}) graph")
        assert_equal(outs[k], ref["rgba"], f"sprite {k} ({w}x{h}) rgba")


def test_full_console_batch_at_4x(gpu, D, O):
    """The 64-frame batch of five console sizes (one-pixel curves, 16-64
    colours) at its real size, one launch at 4x; every frame equals the
    oracle."""
    torch = gpu
    sizes = [(256, 240)] * 13 + [(256, 224)] * 13 + [(320, 224)] * 13 + [(384, 224)] * 13 + [(240, 160)] * 12
    frames = [(D.synth_frame(2, k, w, h, 3), w, h, 3) for k, (w, h) in enumerate(sizes)]
    print("batch of", len(frames), "frames", flush=True)
    outs, graphs = run_batch(torch, D, frames, [(4, False, 0)] * len(frames))
    for k, (img, w, h, ch) in enumerate(frames):
        ref = O.depixelize(img, w, h, scale=4)
        assert_equal(graphs[k], ref["graph"], f"frame {k} ({w}x{h}) graph")
        assert_equal(outs[k], ref["rgba"], f"frame {k} ({w}x{h}) rgba")


def test_large_frames_at_8x_subdivided(gpu, D, O):
    """Two 640x480 dense-crossing frames at 8x with spline subdivision in one
    batch (5120x3840 each): graphs and pixels equal the oracle's, whose
    subdivision is exact lattice arithmetic (no per-vertex error to bound)."""
    torch = gpu
    W, H, S = 640, 480, 8
    frames = [(D.synth_frame(4, 70 + k, W, H, 3, W * 3), W, H, 3) for k in range(2)]
    outs, graphs = run_batch(torch, D, frames, [(S, True, 0), (S, True, D.TOPDOWN)])
    for k, flags in enumerate((0, D.TOPDOWN)):
        print("frame", k, flush=True)
        ref = O.depixelize(frames[k][0], W, H, scale=S, subdivide=True, flags=flags)
        assert_equal(graphs[k], ref["graph"], f"frame {k} graph")
        assert_equal(outs[k], ref["rgba"], f"frame {k} rgba")


@pytest.mark.parametrize("S", list(range(1, 16)))
def test_dense_tiles_every_scale(gpu, D, O, S):
    """Tiles where nearly every cell needs the candidate search: a frame of
    random pixels from three colours (every 32x8-cell tile dense, so the
    searched pixels overflow a tile's colour buffer and, at 12x and up, its
    candidate-code buffer too), a dense-crossing frame and a ragged sprite,
    in both row orders, at every scale 1 to 15."""
    torch = gpu
    rng = np.random.default_rng(1000 + S)
    pal = np.array([[20, 30, 200], [240, 240, 240], [90, 160, 40]], dtype=np.uint8)
    noise = np.ascontiguousarray(pal[rng.integers(0, 3, size=(48, 96))].reshape(48, 96 * 3))
    frames = [(noise, 96, 48, 3), (D.synth_frame(4, 60 + S, 128, 64, 3), 128, 64, 3),
              (D.synth_frame(1, 80 + S, 37, 29, 4), 37, 29, 4)]
    flags = [0, D.TOPDOWN, D.ALPHA | D.TOPDOWN]
    print("scale", S, flush=True)
    outs, graphs = run_batch(torch, D, frames, [(S, False, f) for f in flags])
    for k, ((img, W, H, ch), f) in enumerate(zip(frames, flags)):
        ref = O.depixelize(img, W, H, scale=S, channels=ch, flags=f)
        assert_equal(graphs[k], ref["graph"], f"scale {S} item {k} graph")
        assert_equal(outs[k], ref["rgba"], f"scale {S} item {k} rgba")
