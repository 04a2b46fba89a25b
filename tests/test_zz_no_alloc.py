"""dpx_batch_run_noalloc (every buffer beyond the input and output in the
library's static device array; items without subdivision) against
dpx_batch_run on the same batches:
two calls in a row on one stream, the second with a smaller batch, each
output bit-exact equal to the context path's (which the other GPU tests
check against the oracle), plus the oracle on a few items directly."""
import ctypes

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

# (kind, seed, W, H, channels, flags, scale, subdivide)
BATCH_A = [(0, 11, 256, 224, 3, 0, 4, 0), (0, 12, 256, 224, 3, 0, 4, 0), (2, 13, 97, 61, 3, 2, 3, 0),
           (1, 14, 24, 24, 4, 4, 6, 0), (4, 15, 128, 96, 3, 1, 8, 0), (0, 16, 1, 9, 3, 0, 2, 0),
           (2, 17, 320, 224, 3, 0, 4, 0), (1, 18, 48, 64, 4, 4 | 8, 5, 0)]
BATCH_B = [(2, 21, 64, 48, 3, 0, 4, 0), (0, 22, 33, 17, 3, 2, 7, 0), (1, 23, 16, 32, 4, 4, 6, 0)]


def pack(D, spec, torch):
    blobs, items, in_off, out_off = [], [], 0, 0
    for kind, seed, W, H, ch, flags, s, sub in spec:
        img = D.synth_frame(kind, seed, W, H, ch)
        out_stride = W * s * 4
        items.append(D.Item(in_off, out_off, W, H, img.shape[1], ch, s, sub, flags, out_stride))
        pad = (-(in_off + img.size)) % 256
        blobs += [img.reshape(-1), np.zeros(pad, np.uint8)]
        in_off += img.size + pad
        out_off += (out_stride * H * s + 255) // 256 * 256
    d_in = torch.from_numpy(np.concatenate(blobs)).to("cuda:0")
    return d_in, out_off, items


def run_noalloc(D, d_in, d_out, items, stream):
    arr = D.items_array(items)
    D.check(D.lib.dpx_batch_run_noalloc(d_in.data_ptr(), d_out.data_ptr(), arr, len(arr), stream.cuda_stream))


def test_noalloc_matches_context_path(gpu, D, O):
    torch = gpu
    stream = torch.cuda.current_stream()
    ctx = D.Context(0)
    for name, spec in (("A", BATCH_A), ("B", BATCH_B)):
        print(f"no-alloc batch {name}: {len(spec)} items", flush=True)
        d_in, out_bytes, items = pack(D, spec, torch)
        ref = torch.zeros(out_bytes, dtype=torch.uint8, device="cuda:0")
        got = torch.full((out_bytes,), 7, dtype=torch.uint8, device="cuda:0")
        ctx.run(d_in, ref, items, stream)
        run_noalloc(D, d_in, got, items, stream)
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        for k, it in enumerate(items):
            n = it.out_stride * it.height * it.scale
            a = g[it.out_offset:it.out_offset + n]
            np.testing.assert_array_equal(a, r[it.out_offset:it.out_offset + n], err_msg=f"batch {name} item {k}")
        # and the oracle on the first and last items
        for k in (0, len(items) - 1):
            kind, seed, W, H, ch, flags, s, sub = spec[k]
            img = D.synth_frame(kind, seed, W, H, ch)
            want = O.depixelize(img, W, H, scale=s, subdivide=bool(sub), channels=ch, flags=flags & ~8)["rgba"]
            if flags & 8:
                want = want[::-1]
            it = items[k]
            a = g[it.out_offset:it.out_offset + it.out_stride * H * s].reshape(H * s, W * s, 4)
            np.testing.assert_array_equal(a, want, err_msg=f"batch {name} item {k} against the oracle")
    ctx.close()


def test_noalloc_refuses_batches_beyond_its_limits(gpu, D):
    torch = gpu
    d = torch.zeros(1024, dtype=torch.uint8, device="cuda:0")
    too_many = [D.Item(0, 0, 1, 1, 3, 3, 1, 0, 0, 4)] * 4097
    with pytest.raises(D.DepixelError):
        run_noalloc(D, d, d, too_many, torch.cuda.current_stream())
    too_big = [D.Item(0, 0, 8192, 8192, 8192 * 3, 3, 1, 0, 0, 8192 * 4)]
    with pytest.raises(D.DepixelError) as e:
        run_noalloc(D, d, d, too_big, torch.cuda.current_stream())
    assert "graph bytes" in str(e.value)
    smoothed = [D.Item(0, 0, 4, 4, 12, 3, 2, 1, 0, 32)]
    with pytest.raises(D.DepixelError) as e:
        run_noalloc(D, d, d, smoothed, torch.cuda.current_stream())
    assert "subdivision" in str(e.value)
