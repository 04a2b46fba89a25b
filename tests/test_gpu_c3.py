"""The served-stream workload on the GPU: 128 concurrent game streams of
320x224 palette-indexed frames at 4x through one dpx_server, palettes
resident on the device, an HBM budget of 64 GiB.  Each stream is a sequence
of the stream generator (sprites moving 1-4 px per frame, about a tenth of
the 16x16 tiles changing between frames); four client threads submit the
streams' frames interleaved, and every one of the 1,024 results is checked
bit-exact against the oracle run on the 4-channel frame of the palette
entries.  The per-frame call it replaces is simpleVBO.cpp:131-160 (runCuda,
once per display, callbacks.cpp:62-69)."""
import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = [pytest.mark.timeout(600)]

W, H, S = 320, 224, 4
STREAMS, FRAMES = 128, 8


@pytest.mark.gpu
def test_c3_streams_match_oracle(gpu, D, O):
    srv = D.Server(budget_bytes=64 << 30, max_batch=64)
    sids, pals = [], []
    for k in range(STREAMS):
        sid = srv.open_stream(W, H, 1, S, False, D.INDEXED)
        _, pal = D.stream_frame(k, 0, W, H)
        srv.set_palette(sid, pal)
        sids.append(sid)
        pals.append(pal)
    # the device holds one resident palette per stream
    assert srv.stats().hbm_in_use == STREAMS * 1024
    results = {}
    errors = []
    lock = threading.Lock()

    def client(c):
        rng = np.random.default_rng(c)
        mine = list(range(c, STREAMS, 4))
        nxt = {k: 0 for k in mine}
        pending = []
        while mine:
            k = mine[rng.integers(len(mine))]
            t = nxt[k]
            idx, _ = D.stream_frame(k, t, W, H)
            # half of the results land in pinned blocks, half in ordinary memory
            out = D.pinned_empty((H * S, W * S * 4)) if (k + t) % 2 else np.zeros((H * S, W * S * 4), np.uint8)
            pending.append((k, t, idx, out, srv.submit(sids[k], idx, out)))
            nxt[k] += 1
            if nxt[k] == FRAMES:
                mine.remove(k)
            if len(pending) >= 24 or not mine:
                for kk, tt, ii, oo, tk in pending:
                    rc, ms = srv.wait_timed(tk)
                    with lock:
                        if rc != 0:
                            errors.append((kk, tt, rc))
                        results[(kk, tt)] = (ii, oo, ms)
                pending = []

    threads = [threading.Thread(target=client, args=(c,)) for c in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[:5]
    assert len(results) == STREAMS * FRAMES
    st = srv.stats()
    assert st.completed == STREAMS * FRAMES and st.failed == 0
    assert st.hbm_peak <= st.hbm_budget
    print(f"c3 test: {st.completed} frames, {st.batches} batches, mean batch {st.mean_batch:.1f}, "
          f"p50 {st.p50_ms:.3f} ms, p99 {st.p99_ms:.3f} ms", flush=True)
    srv.close()

    def check(key):
        k, t = key
        idx, out, _ = results[key]
        ref = O.depixelize(D.expand_indexed(idx, pals[k]), W, H, scale=S, channels=4)["rgba"]
        return key, bool(np.array_equal(out.reshape(H * S, W * S, 4), ref))

    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(workers) as ex:
        bad = [key for key, ok in ex.map(check, sorted(results)) if not ok]
    assert not bad, f"{len(bad)} frames differ from the oracle, e.g. {bad[:5]}"


def test_stream_generator_changes_a_tenth_of_the_tiles(D):
    """About 10% of the 16x16 tiles change between consecutive frames."""
    fracs = []
    for seed in range(8):
        frames = [D.stream_frame(seed, t, W, H)[0] for t in range(40)]
        for a, b in zip(frames, frames[1:]):
            fracs.append((a != b).reshape(H // 16, 16, W // 16, 16).any(axis=(1, 3)).mean())
    assert 0.07 <= float(np.mean(fracs)) <= 0.14
