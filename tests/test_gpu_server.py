"""Serving loop on the GPU: continuous batching across streams, HBM-budget
admission, per-request failures, cancellation and time-outs.  Every completed
frame is checked against the oracle (bit-exact)."""
import threading

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


def out_buf(W, H, s):
    return np.zeros((H * s, W * s * 4), dtype=np.uint8)


def expect(O, img, W, H, s, ch=3, flags=0, sub=False):
    return O.depixelize(img, W, H, scale=s, channels=ch, flags=flags, subdivide=sub)["rgba"].reshape(H * s, W * s * 4)


def test_many_streams_from_many_threads(gpu, D, O):
    srv = D.Server(budget_bytes=256 << 20, max_batch=16)
    specs = [(64, 48, 3, 4, False, 0), (80, 40, 3, 2, True, D.TOPDOWN), (24, 24, 4, 6, False, D.ALPHA),
             (50, 33, 3, 3, True, D.RGB)]
    sids = [srv.open_stream(W, H, ch, s, sub, fl) for (W, H, ch, s, sub, fl) in specs]
    results, errors = [], []

    def client(k):
        W, H, ch, s, sub, fl = specs[k % len(specs)]
        for f in range(12):
            img = D.synth_frame(1 if ch == 4 else 2, 1000 * k + f, W, H, ch)
            out = out_buf(W, H, s)
            t = srv.submit(sids[k % len(specs)], img, out)
            rc = srv.wait(t)
            if rc != 0:
                errors.append((k, f, rc))
            results.append((img, out, (W, H, ch, s, sub, fl)))

    threads = [threading.Thread(target=client, args=(k,)) for k in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors
    assert len(results) == 96
    for img, out, (W, H, ch, s, sub, fl) in results:
        assert np.array_equal(out, expect(O, img, W, H, s, ch, fl, sub))
    st = srv.stats()
    assert st.completed == 96 and st.failed == 0 and st.batches >= 1
    assert st.p99_ms >= st.p50_ms > 0 and st.hbm_peak <= st.hbm_budget
    srv.close()


def test_tiny_budget_admits_in_order_and_drains(gpu, D, O):
    W, H, s = 64, 48, 4
    per_frame = W * H * 3 + W * s * H * s * 4
    srv = D.Server(budget_bytes=2 * per_frame + 1024, max_batch=8)
    sid = srv.open_stream(W, H, 3, s)
    imgs = [D.synth_frame(0, 50 + k, W, H, 3) for k in range(20)]
    outs = [out_buf(W, H, s) for _ in imgs]
    tickets = [srv.submit(sid, im, o) for im, o in zip(imgs, outs)]
    assert all(srv.wait(t) == 0 for t in tickets)
    for im, o in zip(imgs, outs):
        assert np.array_equal(o, expect(O, im, W, H, s))
    st = srv.stats()
    assert st.hbm_peak <= st.hbm_budget and st.completed == 20
    srv.close()


def test_oversize_request_is_rejected_and_others_finish(gpu, D, O):
    srv = D.Server(budget_bytes=2 << 20, max_batch=8)
    small = srv.open_stream(32, 32, 3, 4)
    big = srv.open_stream(512, 512, 3, 4)
    img_big = D.synth_frame(0, 1, 512, 512, 3)
    with pytest.raises(D.DepixelError) as e:
        srv.submit(big, img_big, out_buf(512, 512, 4))
    assert e.value.code == D.EBUDGET
    img = D.synth_frame(0, 2, 32, 32, 3)
    out = out_buf(32, 32, 4)
    assert srv.wait(srv.submit(small, img, out)) == 0
    assert np.array_equal(out, expect(O, img, 32, 32, 4))
    with pytest.raises(D.DepixelError) as e:  # malformed: stride too small
        srv.submit(small, img[:, :40].copy(), out)
    assert e.value.code == D.EINVAL
    assert srv.stats().rejected == 1
    srv.close()


def test_injected_failure_hits_only_its_request(gpu, D, O):
    srv = D.Server(budget_bytes=64 << 20, max_batch=8)
    sid = srv.open_stream(40, 30, 3, 4)
    srv.inject_failure(5)
    imgs = [D.synth_frame(2, 70 + k, 40, 30, 3) for k in range(12)]
    outs = [out_buf(40, 30, 4) for _ in imgs]
    tickets = [srv.submit(sid, im, o) for im, o in zip(imgs, outs)]
    codes = [srv.wait(t) for t in tickets]
    assert codes[5] == D.EINJECTED and all(c == 0 for k, c in enumerate(codes) if k != 5)
    for k, (im, o) in enumerate(zip(imgs, outs)):
        if k != 5:
            assert np.array_equal(o, expect(O, im, 40, 30, 4))
    st = srv.stats()
    assert st.failed == 1 and st.completed == 11
    srv.close()


def test_cancel_and_timeout_of_a_queued_request(gpu, D, O):
    W, H, s = 256, 224, 4
    per_frame = W * H * 3 + W * s * H * s * 4
    srv = D.Server(budget_bytes=per_frame + 4096, max_batch=1)  # one frame in flight at a time
    sid = srv.open_stream(W, H, 3, s)
    imgs = [D.synth_frame(0, 90 + k, W, H, 3) for k in range(6)]
    outs = [out_buf(W, H, s) for _ in imgs]
    tickets = [srv.submit(sid, im, o) for im, o in zip(imgs, outs)]
    last = tickets[-1]
    rc = srv.cancel(last)
    if rc == 0:  # still queued behind the others: cancelled
        assert srv.wait(last) == D.ECANCELLED
    else:  # the loop was faster than the cancel
        assert srv.wait(last) == 0
    assert all(srv.wait(t) == 0 for t in tickets[:-1])
    for im, o in zip(imgs[:-1], outs[:-1]):
        assert np.array_equal(o, expect(O, im, W, H, s))
    srv.close()


def test_wait_timeout_returns_while_pending(gpu, D):
    W, H, s = 256, 224, 4
    per_frame = W * H * 3 + W * s * H * s * 4
    srv = D.Server(budget_bytes=per_frame + 4096, max_batch=1)
    sid = srv.open_stream(W, H, 3, s)
    imgs = [D.synth_frame(0, k, W, H, 3) for k in range(8)]
    outs = [out_buf(W, H, s) for _ in imgs]
    tickets = [srv.submit(sid, im, o) for im, o in zip(imgs, outs)]
    rc = srv.wait(tickets[-1], timeout_ms=0)
    assert rc in (D.ETIMEOUT, 0)
    assert all(srv.wait(t) == 0 for t in tickets if not (t == tickets[-1] and rc == 0))
    srv.close()


def test_pinned_frames_and_results_move_by_dma(gpu, D, O):
    """Frames and results in dpx_host_alloc blocks skip the staging copies;
    the output is the same, also with a padded result stride."""
    srv = D.Server(budget_bytes=256 << 20, max_batch=8)
    W, H, s = 72, 40, 4
    sid = srv.open_stream(W, H, 3, s)
    cases = []
    for k in range(6):
        img = D.synth_frame(2, 300 + k, W, H, 3)
        pin = D.pinned_empty(img.shape)
        pin[:] = img
        pad = 64 if k % 2 else 0
        out = D.pinned_empty((H * s, W * s * 4 + pad))
        out[:] = 7
        cases.append((img, pin, out, pad, srv.submit(sid, pin, out)))
    for img, pin, out, pad, t in cases:
        assert srv.wait(t) == 0
        assert np.array_equal(out[:, :W * s * 4], expect(O, img, W, H, s))
        if pad:
            assert (out[:, W * s * 4:] == 7).all(), "the padding of a result row is not written"
    srv.close()
