"""Randomized parity: batches of random items against the oracle.

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
