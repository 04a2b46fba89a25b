"""Parity fuzzing of the batched path (run by hand on a GPU; not collected
by pytest): random batches of the randomized parity test's generator (24
items of random kind, size, scale 1..15, subdivision, byte and row order,
flip, paddings and alignment) for the seeds start .. end-1, each under random
raster tile shapes and walk-window modes for every size class and scale,
against the oracle; bytes outside the images must stay untouched.  Stops at
the first mismatch and names its seed and item.

    python tests/fuzz_batches.py 6000 7200
"""
import sys, os, time
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "pixel-art-remaster-gpu_amd"))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import numpy as np, torch
import depixel as D
import oracle as O
from test_gpu_random import random_batch, SENTINEL
start, end = int(sys.argv[1]), int(sys.argv[2])
t0 = time.time()
for seed in range(start, end):
    items, blob, out_bytes, meta = random_batch(D, seed)
    rng = np.random.default_rng(seed + 77)
    d_in = torch.from_numpy(blob).to("cuda:0")
    d_out = torch.full((out_bytes,), SENTINEL, dtype=torch.uint8, device="cuda:0")
    ctx = D.Context(0)
    for cls in (D.CLASS_SPRITE, D.CLASS_FRAME, D.CLASS_LARGE):
        for s in range(1, 16):
            for sub in (False, True):
                ctx.set_tile(cls, s, sub, int(rng.choice([16, 32, 64])))
                ctx.set_walk(cls, s, sub, int(rng.choice([D.WALK_AUTO, D.WALK_ALWAYS, D.WALK_NEVER])))
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
        bad = np.argwhere((got != ref).any(-1))
        if len(bad):
            print(f"MISMATCH seed {seed} item {k}: kind {kind} {W}x{H} ch {ch} x{S} sub {sub} flags {flags}: {len(bad)} px, first {bad[:3].tolist()}", flush=True)
            sys.exit(1)
    if not (out[~covered] == SENTINEL).all():
        print(f"MISMATCH seed {seed}: bytes outside the images written", flush=True)
        sys.exit(1)
    print(f"seed {seed} ok ({time.time() - t0:.0f} s)", flush=True)
print("all ok")
