"""Plans captured into HIP graphs, and plans launched as the library's own
HIP graph: every replay equals the oracle.

`dpx_plan_run` only launches kernels on the given stream, with descriptors
and graph scratch that belong to the plan, so the caller may capture it and
replay it while the context runs other batches (other items, grown buffers)
in between.  With DPX_PLAN_GRAPH the library builds the graph itself and
updates its input and output pointers when they change.
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

# (kind, W, H, channels, flags, scale): three scales, so three raster launches
SHAPES = [(0, 64, 48, 3, 0, 4), (1, 37, 29, 4, 4 | 2, 3), (4, 96, 40, 3, 0, 6)]


def frames_for(D, seed0):
    return [D.synth_frame(k, seed0 + i, W, H, ch) for i, (k, W, H, ch, f, S) in enumerate(SHAPES)]


def pack(D, frames):
    """Input blob, output size and items for the frames (offsets aligned to 256)."""
    blobs, items, in_off, out_off = [], [], 0, 0
    for img, (k, W, H, ch, f, S) in zip(frames, SHAPES):
        out_stride = W * S * 4
        items.append(D.Item(in_off, out_off, W, H, img.shape[1], ch, S, 0, f, out_stride))
        blobs.append(img.reshape(-1))
        pad = (-(in_off + img.size)) % 256
        blobs.append(np.zeros(pad, dtype=np.uint8))
        in_off += img.size + pad
        out_off += (out_stride * H * S + 255) // 256 * 256
    return np.concatenate(blobs), out_off, items


def unpack(items, out):
    res = []
    for it in items:
        o = out[it.out_offset:it.out_offset + it.out_stride * it.height * it.scale]
        res.append(o.reshape(it.height * it.scale, it.width * it.scale, 4))
    return res


def other_batch(D, torch):
    """A batch of different items (and a larger graph scratch) for the same context."""
    W, H, S = 200, 150, 5
    img = D.synth_frame(2, 99, W, H, 3)
    items = [D.Item(0, 0, W, H, img.shape[1], 3, S, 1, 0, W * S * 4)]
    return torch.from_numpy(img.reshape(-1)).to("cuda:0"), torch.zeros(W * S * H * S * 4, dtype=torch.uint8,
                                                                         device="cuda:0"), items


def test_captured_plan_replays_match_oracle(gpu, D, O):
    torch = gpu
    sets = [frames_for(D, 700), frames_for(D, 800)]
    blob0, out_bytes, items = pack(D, sets[0])
    blob1, _, _ = pack(D, sets[1])
    refs = [[O.depixelize(img, W, H, scale=S, channels=ch, flags=f)["rgba"]
             for img, (k, W, H, ch, f, S) in zip(fr, SHAPES)] for fr in sets]
    d_in = torch.from_numpy(blob0).to("cuda:0")
    d_out = torch.zeros(out_bytes, dtype=torch.uint8, device="cuda:0")
    ctx = D.Context(0)
    plan = D.Plan(ctx, items, out_align=d_out.data_ptr() % 16)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        plan.run(d_in, d_out, stream=side)
    torch.cuda.synchronize()
    for k, got in enumerate(unpack(items, d_out.cpu().numpy())):
        np.testing.assert_array_equal(got, refs[0][k], err_msg=f"eager item {k}")

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run(d_in, d_out, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    o_in, o_out, o_items = other_batch(D, torch)
    for rep in range(4):
        which = rep % 2
        d_in.copy_(torch.from_numpy(blob1 if which else blob0).to("cuda:0"))
        d_out.zero_()
        torch.cuda.synchronize()
        print("replay", rep, "input set", which, flush=True)
        g.replay()
        torch.cuda.synchronize()
        for k, got in enumerate(unpack(items, d_out.cpu().numpy())):
            np.testing.assert_array_equal(got, refs[which][k], err_msg=f"replay {rep} item {k}")
        # other work on the same context between replays: a batch of other
        # items through dpx_batch_run, which rewrites the context's
        # descriptors, and (first time) grows its graph scratch
        ctx.run(o_in, o_out, o_items, stream=side)
        if rep == 1:
            ctx.reserve(64, 1 << 20, 7)
        plan.run(d_in, d_out, stream=side)
        torch.cuda.synchronize()
        for k, got in enumerate(unpack(items, d_out.cpu().numpy())):
            np.testing.assert_array_equal(got, refs[which][k], err_msg=f"eager after replay {rep} item {k}")
    del g
    plan.close()
    ctx.close()


def test_library_graph_plan_matches_oracle(gpu, D, O):
    """DPX_PLAN_GRAPH: the plan's kernels as one HIP graph; runs on two input
    and output buffer pairs (the graph's pointers updated in place), with
    other batches on the context in between."""
    torch = gpu
    sets = [frames_for(D, 900), frames_for(D, 1000)]
    blobs, refs = [], []
    for fr in sets:
        blob, out_bytes, items = pack(D, fr)
        blobs.append(blob)
        refs.append([O.depixelize(img, W, H, scale=S, channels=ch, flags=f)["rgba"]
                     for img, (k, W, H, ch, f, S) in zip(fr, SHAPES)])
    ins = [torch.from_numpy(b).to("cuda:0") for b in blobs]
    outs = [torch.zeros(out_bytes, dtype=torch.uint8, device="cuda:0") for _ in blobs]
    assert outs[0].data_ptr() % 16 == outs[1].data_ptr() % 16
    ctx = D.Context(0)
    plan = D.Plan(ctx, items, out_align=outs[0].data_ptr() % 16, flags=D.PLAN_GRAPH)
    o_in, o_out, o_items = other_batch(D, torch)
    stream = torch.cuda.current_stream()
    for rep in range(6):
        which = (rep // 2) % 2 if rep < 4 else rep % 2
        outs[which].zero_()
        torch.cuda.synchronize()
        print("graph run", rep, "buffers", which, flush=True)
        plan.run(ins[which], outs[which], stream=stream)
        torch.cuda.synchronize()
        for k, got in enumerate(unpack(items, outs[which].cpu().numpy())):
            np.testing.assert_array_equal(got, refs[which][k], err_msg=f"run {rep} item {k}")
        ctx.run(o_in, o_out, o_items, stream=stream)
    plan.close()
    ctx.close()


def test_plan_rejects_other_alignment(gpu, D):
    torch = gpu
    img = D.synth_frame(0, 1, 16, 8, 3)
    items = [D.Item(0, 0, 16, 8, img.shape[1], 3, 2, 0, 0, 16 * 2 * 4)]
    ctx = D.Context(0)
    plan = D.Plan(ctx, items, out_align=0)
    d_in = torch.from_numpy(img.reshape(-1)).to("cuda:0")
    d_out = torch.zeros(16 * 2 * 8 * 2 * 4 + 64, dtype=torch.uint8, device="cuda:0")
    base = d_out.data_ptr() % 16
    with pytest.raises(D.DepixelError):
        plan.run(d_in, d_out[(16 - base) % 16 + 4:])
    plan.close()
    ctx.close()
