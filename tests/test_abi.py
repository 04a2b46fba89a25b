"""The C-ABI library loads on a machine without a GPU, exports every function
include/depixel.h declares, and refuses work cleanly when no GPU is visible."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "depixel.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", src)
    keywords = {"if", "while", "for", "return", "sizeof", "defined"}
    return sorted(set(n for n in names if n not in keywords))


def test_header_declares_the_drop_in_and_the_stages():
    names = declared_functions()
    for n in ("launch_kernel", "dpx_frame", "dpx_stage_similarity_graph", "dpx_stage_trivial_cross",
              "dpx_stage_heuristic_cross", "dpx_stage_cells", "dpx_stage_subdivide", "dpx_stage_triangulate",
              "dpx_stage_export_vbo", "dpx_stage_rasterize", "dpx_stage_components",
              "dpx_stage_border_chains", "dpx_batch_run",
              "dpx_server_submit"):
        assert n in names


def test_library_exports_every_declared_function(D):
    lib = ctypes.CDLL(D.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert D.lib.dpx_version().startswith(b"depixel")


def test_synthetic_generator_is_deterministic(D):
    import numpy as np
    a = D.synth_frame(0, 5, 256, 224, 3)
    b = D.synth_frame(0, 5, 256, 224, 3)
    c = D.synth_frame(0, 6, 256, 224, 3)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    sizes = {D.sprite_size(9, k) for k in range(200)}
    assert sizes == {(8, 8), (16, 16), (16, 32), (24, 24), (32, 32), (48, 64)}
    s = D.synth_frame(1, 3, 24, 24, 4)
    alpha = s.reshape(24, 24, 4)[..., 3]
    assert 0.15 < (alpha == 0).mean() < 0.45


def test_synthetic_generator_takes_one_pixel_sides(D):
    """Every kind at sizes with a side of 1 to 3 pixels (the sprite's detail
    rectangles are clipped to the sprite; a 1-pixel side once divided by
    zero), written inside the frame only."""
    import numpy as np
    for kind in (0, 1, 2, 4):
        ch = 4 if kind == 1 else 3
        for W, H in ((1, 1), (1, 5), (5, 1), (3, 1), (1, 3), (148, 1), (1, 120), (2, 2), (3, 3)):
            stride = W * ch + 3
            f = D.synth_frame(kind, 11, W, H, ch, stride)
            assert f.shape == (H, stride)
            assert not f[:, W * ch:].any(), f"kind {kind} {W}x{H} wrote past the row"


def test_argument_checks_need_no_gpu(D):
    import numpy as np
    img = np.zeros((4, 12), dtype=np.uint8)
    p = D.Params(4, 0, 0, 0)
    # bad channel count and bad stride are rejected before any device call
    assert D.lib.dpx_frame(img.ctypes.data, 4, 4, 12, 2, ctypes.byref(p), None, 0, None, None, None) == D.EINVAL
    assert D.lib.dpx_frame(img.ctypes.data, 4, 4, 8, 3, ctypes.byref(p), None, 0, None, None, None) == D.EINVAL
    assert "stride" in D.last_error()
    # unknown flags, and indexed frames outside the batched and serving paths
    for flags, msg in ((64, "unknown flags"), (D.INDEXED, "indexed")):
        q = D.Params(4, 0, flags, 0)
        assert D.lib.dpx_frame(img.ctypes.data, 4, 4, 12, 3, ctypes.byref(q), None, 0, None, None, None) == D.EINVAL
        assert msg in D.last_error()
        assert D.lib.dpx_stage_similarity_graph(img.ctypes.data, 4, 4, 12, 3, flags, img.ctypes.data, None) == D.EINVAL
    # views outside the documented ranges
    one = ctypes.c_void_p(img.ctypes.data)
    for view, msg in ((D.View(2, 800, 1, 1, 64, 64, 0, 0), "view mode"),
                      (D.View(0, 800, 1, 1, 64, 64, 0, 0), "triangulated diagram"),
                      (D.View(1, 24, 1, 1, 64, 64, 0, 0), "scale_x100"),
                      (D.View(1, 800, 1, 1, 0, 64, 0, 0), "window"),
                      (D.View(1, 800, 1, 1, 64, 64, 4, 0), "view flags"),
                      (D.View(1, 800, 70000, 1, 64, 64, 0, 0), "translation")):
        rc = D.lib.dpx_stage_present(one, 4, 4, 12, 3, 0, None, None, ctypes.byref(view), one, 256, None)
        assert rc == D.EINVAL and msg in D.last_error(), (msg, D.last_error())
    ok = D.View(1, 800, 1, 1, 64, 64, 0, 0)
    assert D.lib.dpx_stage_present(one, 4, 4, 12, 3, 0, None, None, ctypes.byref(ok), one, 252, None) == D.EINVAL
    assert "out_stride" in D.last_error()


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"),
                    reason="a GPU may be visible")
def test_no_device_is_reported(D):
    import numpy as np
    img = np.zeros((4, 12), dtype=np.uint8)
    p = D.Params(4, 0, 0, 0)
    rc = D.lib.dpx_frame(img.ctypes.data, 4, 4, 12, 3, ctypes.byref(p), None, 0, None, None, None)
    assert rc == D.ENODEV
    h = ctypes.c_void_p()
    assert D.lib.dpx_context_create(0, ctypes.byref(h)) == D.ENODEV
    assert D.lib.dpx_server_create(0, 1 << 20, 8, ctypes.byref(h)) == D.ENODEV
