"""Autotuner of the batched path's raster tile shape.

For each workload class (sprites, console-size frames, large frames) at the
scales it is measured on, with and without subdivision, every tile shape
(16x16, 32x8 and 64x4 cells per block) is timed on the device with the
context's own kernel timing (HIP events on the launch stream), the outputs
of the three shapes are checked to be identical, and the fastest shape per
(size class, scale, subdivision) is written to tuning.json next to this
file, which every dpx_context reads when it is created.  A shape replaces
the default (32x8) only when it is faster by more than `--margin`.

    python pixel-art-remaster-gpu_amd/autotune.py [--out tuning.json] [--runs 10]

It replaces the reference's fixed 2x2-thread blocks (kernel.cu:392-394).
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

CLASS_NAMES = {0: "sprite", 1: "frame", 2: "large"}


def workloads(D):
    """(label, size class, scale, subdivide, [(kind, seed, W, H, channels, flags)])"""
    sprites = []
    for k in range(4096):
        w, h = D.sprite_size(0, k)
        sprites.append((1, k, w, h, 4, D.ALPHA))
    frames = [(0, k, 256, 224, 3, 0) for k in range(1024)]
    console = []
    sizes = [(256, 240)] * 13 + [(256, 224)] * 13 + [(320, 224)] * 13 + [(384, 224)] * 13 + [(240, 160)] * 12
    for k, (w, h) in enumerate(sizes):
        console.append((2, k, w, h, 3, 0))
    large = [(4, k, 640, 480, 3, 0) for k in range(16)]
    out = []
    for s in (3, 6, 8):
        out.append((f"4096 sprites of the sprite mix at {s}x", D.CLASS_SPRITE, s, 0, sprites))
    out.append(("1024 sprites at 6x, subdivided", D.CLASS_SPRITE, 6, 1, sprites[:1024]))
    for s in (2, 3, 6, 8):
        out.append((f"256 frames of 256x224 at {s}x", D.CLASS_FRAME, s, 0, frames[:256]))
    out.append(("1024 frames of 256x224 at 4x (the headline batch)", D.CLASS_FRAME, 4, 0, frames))
    out.append(("64 frames of 256x224 at 4x, subdivided", D.CLASS_FRAME, 4, 1, frames[:64]))
    out.append(("64 console frames of five sizes at 4x, subdivided", D.CLASS_FRAME, 4, 1, console))
    for s in (4, 8):
        out.append((f"16 frames of 640x480 at {s}x", D.CLASS_LARGE, s, 0, large))
    out.append(("16 frames of 640x480 at 8x, subdivided", D.CLASS_LARGE, 8, 1, large))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    ap.add_argument("--runs", type=int, default=10)
    ap.add_argument("--margin", type=float, default=0.02, help="minimum gain over the default shape")
    args = ap.parse_args()
    import torch
    import depixel as D

    assert torch.cuda.is_available(), "the autotuner times kernels on the GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    arch = torch.cuda.get_device_properties(0).gcnArchName
    picks, table = {}, []
    for label, cls, scale, sub, spec in workloads(D):
        hosts, items = [], []
        in_off = out_off = 0
        for kind, seed, w, h, ch, flags in spec:
            img = D.synth_frame(kind, seed, w, h, ch)
            hosts.append(img.reshape(-1))
            items.append(D.Item(in_off, out_off, w, h, img.shape[1], ch, scale, sub, flags, w * scale * 4))
            in_off += img.size
            out_off += w * scale * 4 * h * scale
        d_in = torch.from_numpy(np.concatenate(hosts)).to(dev)
        d_out = torch.empty(out_off, dtype=torch.uint8, device=dev)
        arr = D.items_array(items)
        times, digests = {}, {}
        for tx in D.TILE_WIDTHS:
            ctx = D.Context(0)
            ctx.set_tile(cls, scale, bool(sub), tx)
            for _ in range(3):
                ctx.run(d_in, d_out, arr, stream)
            torch.cuda.synchronize()
            ctx.timing(True)
            ctx.timing_read()
            for _ in range(args.runs):
                ctx.run(d_in, d_out, arr, stream)
            t = ctx.timing_read()
            times[tx] = t.raster_ms / max(1, t.batches)
            digests[tx] = hashlib.sha1(d_out.cpu().numpy().tobytes()).hexdigest()
            ctx.close()
        assert len(set(digests.values())) == 1, f"{label}: the tile shapes disagree: {digests}"
        best = min(times, key=times.get)
        if times[best] > times[32] * (1 - args.margin):
            best = 32
        key = (cls, scale, sub)
        # several workloads of one key: keep the pick of the first (the listed order puts the headline first)
        if key not in picks:
            picks[key] = (best, label, times)
        table.append((label, CLASS_NAMES[cls], scale, sub, times, best))
        print(f"{label}: " + ", ".join(f"{tx}: {times[tx]:.4f} ms" for tx in D.TILE_WIDTHS) + f" -> {best}",
              flush=True)
    doc = {
        "device": arch,
        "generated_by": "pixel-art-remaster-gpu_amd/autotune.py",
        "generated_at": time.strftime("%Y-%m-%d %H:%M:%S"),
        "metric": "smoothing + raster kernel time per batch (ms), HIP events on the launch stream",
        "picks": [
            {"class": CLASS_NAMES[c], "scale": s, "subdivide": sb, "tile": best, "workload": label,
             "ms": {str(k): round(v, 5) for k, v in times.items()},
             "gain_vs_default": round(1 - times[best] / times[32], 4)}
            for (c, s, sb), (best, label, times) in sorted(picks.items())
        ],
    }
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print("| workload | class | scale | subdivide | 16x16 ms | 32x8 ms | 64x4 ms | pick | gain vs 32x8 |")
    print("|---|---|---|---|---|---|---|---|---|")
    for label, c, s, sb, times, best in table:
        print(f"| {label} | {c} | {s} | {sb} | {times[16]:.4f} | {times[32]:.4f} | {times[64]:.4f} | "
              f"{best} | {1 - times[best] / times[32]:.1%} |")
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
