"""Parity fuzzing of dpx_stage_border_chains (run by hand on a GPU; not
collected by pytest): for seeds start .. end-1, every tenth graph is the
resolved graph of a random frame (kind, size up to 199x159), the others are
random symmetric graphs (up to 39x39, random link density; a third with
internal nodes, value 90, made symmetric with their neighbours).  The chains
must equal the oracle's, and a graph the oracle refuses must be refused.

    python tests/fuzz_border_chains.py 0 2200
"""
import sys, os, time
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "pixel-art-remaster-gpu_amd"))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import numpy as np, torch
import depixel as D
import oracle as O
from test_gpu_parity import gpu_border_chains
t0 = time.time()
DX, DY = [-1, 0, 1, -1, 1, -1, 0, 1], [1, 1, 1, 0, 0, -1, -1, -1]
n = 0
for seed in range(int(sys.argv[1]), int(sys.argv[2])):
    rng = np.random.default_rng(300000 + seed)
    if seed % 10 == 0:  # a frame's resolved graph
        kind = int(rng.choice([0, 1, 2, 4]))
        ch = 4 if kind == 1 else 3
        W, H = int(rng.integers(1, 200)), int(rng.integers(1, 160))
        img = D.synth_frame(kind, seed, W, H, ch)
        g = O.depixelize(img, W, H, scale=1, channels=ch, flags=D.ALPHA if ch == 4 else 0)["graph"]
    else:  # a random symmetric graph
        W, H = int(rng.integers(2, 40)), int(rng.integers(2, 40))
        p = rng.random() * 0.8 + 0.05
        g = np.zeros((H, W), dtype=np.uint8)
        for j in range(H):
            for i in range(W):
                for e in (1, 2, 4, 7):
                    ni, nj = i + DX[e], j + DY[e]
                    if 0 <= ni < W and 0 <= nj < H and rng.random() < p:
                        g[j, i] |= 1 << e
                        g[nj, ni] |= 1 << (7 - e)
        if rng.random() < 0.3 and W > 2 and H > 2:  # some internal nodes (90: the four orthogonal links), symmetric
            for _ in range(int(rng.integers(1, 2 + W * H // 10))):  # (1 .. W*H/10 + 1: a 3x3 frame gets one)
                i, j = int(rng.integers(1, W - 1)), int(rng.integers(1, H - 1))
                old = int(g[j, i])
                for e in range(8):  # drop the old links' reverses
                    if (old >> e) & 1:
                        g[j + DY[e], i + DX[e]] &= np.uint8(~(1 << (7 - e)) & 0xFF)
                g[j, i] = 90
                for e in (1, 3, 4, 6):
                    g[j + DY[e], i + DX[e]] |= np.uint8(1 << (7 - e))
    try:
        want = O.border_chains(g)
    except RuntimeError:
        want = None
    got, nn, status = gpu_border_chains(torch, D, g)
    if want is None:
        if status == 0:
            print(f"seed {seed} {g.shape}: the oracle refuses the graph, the GPU does not", flush=True)
            sys.exit(1)
    elif status != 0 or got != want:
        print(f"seed {seed} {g.shape}: status {status}, chains {len(got) if got else None} vs {len(want)}", flush=True)
        sys.exit(1)
    n += 1
    if seed % 200 == 0:
        print(f"seed {seed} ok ({time.time() - t0:.0f} s)", flush=True)
print("all ok", n)
