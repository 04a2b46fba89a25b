"""Component labels (A18): the oracle's canonical labelling (smallest node
index of each component) against scipy's connected components."""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

EDX = [-1, 0, 1, -1, 1, -1, 0, 1]
EDY = [1, 1, 1, 0, 0, -1, -1, -1]


def scipy_labels(g):
    H, W = g.shape
    rows, cols = [], []
    for j in range(H):
        for i in range(W):
            for e in range(8):
                if (int(g[j, i]) >> e) & 1:
                    rows.append(j * W + i)
                    cols.append((j + EDY[e]) * W + i + EDX[e])
    n = W * H
    m = coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    _, lab = connected_components(m, directed=False)
    canon = np.full(lab.max() + 1, n, dtype=np.int64)
    np.minimum.at(canon, lab, np.arange(n))
    return canon[lab].reshape(H, W)


def test_components_match_scipy(D, O):
    for kind, seed, W, H in ((0, 1, 40, 30), (2, 2, 33, 21), (4, 3, 25, 25)):
        img = D.synth_frame(kind, seed, W, H, 3)
        g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H)))
        assert np.array_equal(O.components(g), scipy_labels(g))


def test_components_of_reference_graphs(O):
    import json
    import os
    from conftest import GOLDEN
    d = json.load(open(os.path.join(GOLDEN, "reference_graphs.json")))
    for name in ("c_pattern", "alex"):
        g = np.array(d[name]["graph"], dtype=np.uint8).reshape(d[name]["height"], d[name]["width"])
        lab = O.components(g)
        assert np.array_equal(lab, scipy_labels(g))
        # every border chain of alex_png.txt stays inside one component
        if name == "alex":
            ch = json.load(open(os.path.join(GOLDEN, "alex_chains.json")))["chains"]
            flat = lab.reshape(-1)
            for c in ch:
                assert len(set(flat[c])) == 1
