"""A batch captured into a HIP graph: every replay equals the oracle.

`dpx_batch_run` with items equal to the context's cached plan (same items,
same output alignment) only launches its kernels on the given stream: no
allocation, no copy, no synchronisation.  After one eager call with those
items it can therefore be captured into a graph by the caller and replayed;
the graph reads the input buffer at replay time.
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


def test_captured_batch_replays_match_oracle(gpu, D, O):
    torch = gpu
    sets = [frames_for(D, 700), frames_for(D, 800)]
    blob0, out_bytes, items = pack(D, sets[0])
    blob1, _, _ = pack(D, sets[1])
    refs = [[O.depixelize(img, W, H, scale=S, channels=ch, flags=f)["rgba"]
             for img, (k, W, H, ch, f, S) in zip(fr, SHAPES)] for fr in sets]
    arr = D.items_array(items)
    d_in = torch.from_numpy(blob0).to("cuda:0")
    d_out = torch.zeros(out_bytes, dtype=torch.uint8, device="cuda:0")
    ctx = D.Context(0)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ctx.run(d_in, d_out, arr, stream=side)  # builds the plan and the tables
    torch.cuda.synchronize()
    for k, got in enumerate(unpack(items, d_out.cpu().numpy())):
        np.testing.assert_array_equal(got, refs[0][k], err_msg=f"eager item {k}")

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.run(d_in, d_out, arr, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    # replays with eager work in between: clear the output, swap the input
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
        # an eager batch on the same context between replays
        ctx.run(d_in, d_out, arr, stream=side)
        torch.cuda.synchronize()
        for k, got in enumerate(unpack(items, d_out.cpu().numpy())):
            np.testing.assert_array_equal(got, refs[which][k], err_msg=f"eager after replay {rep} item {k}")
    del g
    ctx.close()
