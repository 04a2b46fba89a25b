"""Spline-vertex error of the GPU against the FP64 oracle on the large-output
workload (one 640x480 kind-4 frame, subdivision on): the subdivided cell
polygons of dpx_stage_subdivide against the oracle's (which evaluates the
reference's corner cuts and midpoints with double temporaries, and whose
vertices equal the exact rationals: test_oracle_exact_subdivision.py).
Tolerance 1e-5 relative; the error found is printed (it is 0: every vertex
lies on the 1/64 lattice, which both hold exactly)."""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

W, H, CS = 640, 480, 45


def test_subdivided_vertices_against_fp64_oracle(gpu, D, O):
    torch = gpu
    img = D.synth_frame(4, 0, W, H, 3)
    g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H)))
    od, oec = O.cells(g)
    sd, sec, _, _ = O.subdivide(img, W, H, g, od, oec)
    dev = "cuda:0"
    s = torch.cuda.current_stream()
    d_img = torch.from_numpy(img.reshape(-1)).to(dev)
    d_g = torch.from_numpy(g.reshape(-1).copy()).to(dev)
    d_diag = torch.from_numpy(od.reshape(-1)).to(dev)
    d_daux = torch.zeros_like(d_diag)
    d_ec = torch.from_numpy(oec.astype(np.int32)).to(dev)
    d_eaux = torch.zeros_like(d_ec)
    D.check(D.lib.dpx_stage_subdivide(d_img.data_ptr(), W, H, img.shape[1], 3, 0, d_g.data_ptr(), d_diag.data_ptr(),
                                      d_daux.data_ptr(), d_ec.data_ptr(), d_eaux.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    got = d_diag.cpu().numpy().reshape(W * H, CS, 2)
    ec = d_ec.cpu().numpy()
    np.testing.assert_array_equal(ec, sec, err_msg="vertex counts")
    ref = sd.reshape(W * H, CS, 2)
    mask = np.arange(CS)[None, :] < ec[:, None]
    a = got[mask].astype(np.float64)
    b = ref[mask].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1.0)
    print(f"640x480 with subdivision: {a.shape[0]} polygon vertices, max relative error vs the FP64 oracle "
          f"{rel.max():.3g}", flush=True)
    assert rel.max() <= 1e-5
    assert np.all(np.floor(a * 64) == a * 64), "a vertex off the 1/64 lattice"
