"""Every raster tile shape (16x16, 32x8 and 64x4 cells per block), with and
without the walk window, against the oracle: random batches (random kinds,
sizes including items narrower than a tile and ragged edges, scales 1..15,
subdivision, layouts) with the context's picks forced to one shape and one
walk mode for every size class and scale, then the workload shapes
(sprites, console frames, a large frame) whose picks the autotuner chooses
between."""
import numpy as np
import pytest

from test_gpu_random import SENTINEL, random_batch

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


def force_tile(D, ctx, tx, walk=None):
    for cls in (D.CLASS_SPRITE, D.CLASS_FRAME, D.CLASS_LARGE):
        for s in range(1, 16):
            for sub in (False, True):
                ctx.set_tile(cls, s, sub, tx)
                if walk is not None:
                    ctx.set_walk(cls, s, sub, walk)


def check_batch(torch, D, O, ctx, items, blob, out_bytes, meta, label):
    d_in = torch.from_numpy(blob).to("cuda:0")
    d_out = torch.full((out_bytes,), SENTINEL, dtype=torch.uint8, device="cuda:0")
    ctx.run(d_in, d_out, items)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    touched = np.zeros(out_bytes, dtype=bool)
    for k, (it, (kind, W, H, ch, S, sub, flags, flip, img)) in enumerate(zip(items, meta)):
        rows = out[it.out_offset:it.out_offset + it.out_stride * H * S].reshape(H * S, it.out_stride)
        got = rows[:, :W * S * 4].reshape(H * S, W * S, 4)
        ref = O.depixelize(img, W, H, scale=S, subdivide=bool(sub), channels=ch, flags=flags & ~D.FLIP_OUT)["rgba"]
        if flip:
            ref = ref[::-1]
        assert np.array_equal(got, ref), f"{label} item {k}: kind {kind} {W}x{H} x{S} sub {sub} flags {flags}"
        for r in range(H * S):
            touched[it.out_offset + r * it.out_stride:it.out_offset + r * it.out_stride + W * S * 4] = True
    assert np.all(out[~touched] == SENTINEL), f"{label}: bytes outside the images were written"


@pytest.mark.parametrize("walk", [1, 2])  # DPX_WALK_ALWAYS, DPX_WALK_NEVER
@pytest.mark.parametrize("tx", [16, 32, 64])
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_tile_shape_random_batches(gpu, D, O, tx, seed, walk):
    torch = gpu
    print(f"tile {tx} walk mode {walk} batch seed {7100 + seed}", flush=True)
    ctx = D.Context(0)
    force_tile(D, ctx, tx, walk)
    assert ctx.get_walk(D.CLASS_FRAME, 3, True) == walk
    items, blob, out_bytes, meta = random_batch(D, 7100 + seed)
    check_batch(torch, D, O, ctx, items, blob, out_bytes, meta, f"tile {tx} walk {walk} seed {7100 + seed}")
    ctx.close()


def workload_batch(D):
    """Sprites of the sprite mix at 6x, console frames at 4x, one 320x240 frame at 8x with subdivision."""
    items, blobs, meta = [], [], []
    in_off = out_off = 0
    spec = [(1, k, *D.sprite_size(0, k), 4, D.ALPHA, 6, 0) for k in range(24)]
    spec += [(2, 40 + k, w, h, 3, 0, 4, k % 2) for k, (w, h) in enumerate([(256, 240), (240, 160), (384, 224)])]
    spec += [(4, 50, 520, 120, 3, D.TOPDOWN, 8, 1)]
    for kind, seed, W, H, ch, flags, S, sub in spec:
        img = D.synth_frame(kind, seed, W, H, ch)
        items.append(D.Item(in_off, out_off, W, H, img.shape[1], ch, S, sub, flags, W * S * 4))
        blobs.append(img.reshape(-1))
        in_off += img.size
        out_off += W * S * 4 * H * S
        meta.append((kind, W, H, ch, S, sub, flags, False, img))
    return items, np.concatenate(blobs), out_off, meta


@pytest.mark.parametrize("walk", [1, 2])
@pytest.mark.parametrize("tx", [16, 32, 64])
def test_tile_shape_workload_items(gpu, D, O, tx, walk):
    torch = gpu
    print(f"tile {tx} walk mode {walk} workload items", flush=True)
    ctx = D.Context(0)
    force_tile(D, ctx, tx, walk)
    for cls in (D.CLASS_SPRITE, D.CLASS_FRAME, D.CLASS_LARGE):
        assert ctx.get_tile(cls, 4, False) == tx
    items, blob, out_bytes, meta = workload_batch(D)
    check_batch(torch, D, O, ctx, items, blob, out_bytes, meta, f"tile {tx} workload")
    ctx.close()
