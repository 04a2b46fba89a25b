"""Whole-polygon coverage rows (poly_rows_xor in geom.hpp, used by the
smoothing kernel) against the union of the ear-clipped triangles' rows they
replace, on the host: the smoothed cells of synthetic frames (the oracle's
subdivision) and random, partly self-intersecting polygons, at every
scale.  The winding-sum check must never pass on rows that differ from the
union, and must pass on every smoothed cell."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "pixel-art-remaster-gpu_amd", "csrc")


def smoothed_cells(O, D):
    out = []
    for kind, seed, W, H, ch in ((0, 1, 96, 64, 3), (2, 2, 96, 64, 3), (4, 3, 96, 64, 3), (1, 4, 32, 32, 4),
                                 (1, 5, 24, 16, 4)):
        img = D.synth_frame(kind, 4100 + seed, W, H, ch)
        flags = D.ALPHA if ch == 4 else 0
        g = O.depixelize(img, W, H, scale=1, channels=ch, flags=flags)["graph"]
        d, ec = O.cells(g)
        d2, e2, _, _ = O.subdivide(img, W, H, g, d, ec, channels=ch, flags=flags)
        for k in np.nonzero(g.reshape(-1) != 90)[0]:
            m = int(e2[k])
            pts = np.round(d2[k, :m].astype(np.float64) * 64).astype(np.int64)
            assert np.array_equal(pts / 64.0, d2[k, :m])  # the 1/64 lattice
            out.append(np.concatenate([[m], pts.reshape(-1)]))
    return np.concatenate(out).astype(np.int16)


@pytest.mark.timeout(600)
def test_poly_rows_equal_union_of_ears(tmp_path, O, D):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    cells = tmp_path / "cells.bin"
    smoothed_cells(O, D).tofile(cells)
    exe = tmp_path / "poly_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "native", "poly_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe), str(cells)], capture_output=True, text=True, timeout=500)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 30
    refused_random = 0
    for line in lines:
        v = dict(re.findall(r"(polygons|incomplete|refused|unsound) (\d+)", line))
        assert int(v["unsound"]) == 0, line
        assert int(v["polygons"]) > 1000, line
        if line.startswith("cells"):
            assert int(v["refused"]) == 0 and int(v["incomplete"]) == 0, line
        else:
            refused_random += int(v["refused"])
    assert refused_random > 0  # the random set does reach the check's refusals
