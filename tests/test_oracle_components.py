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


# ---- border chains (extractBorderPoints cc_functions.cu:348-503) ----

def _alex():
    import json
    import os
    from conftest import GOLDEN
    d = json.load(open(os.path.join(GOLDEN, "reference_graphs.json")))["alex"]
    g = np.array(d["graph"], dtype=np.uint8).reshape(d["height"], d["width"])
    ch = json.load(open(os.path.join(GOLDEN, "alex_chains.json")))["chains"]
    return g, ch


def test_border_chains_match_alex_png(O):
    """All 21 chains of alex_png.txt:1-21, node for node and in order."""
    g, expected = _alex()
    assert O.border_chains(g) == expected


def linked(g, a, b):
    H, W = g.shape
    ai, aj, bi, bj = a % W, a // W, b % W, b // W
    for e in range(8):
        if ai + EDX[e] == bi and aj + EDY[e] == bj:
            return bool((int(g.reshape(-1)[a]) >> e) & 1)
    return False


def check_chain_invariants(g, chains):
    flat = g.reshape(-1)
    seen = set()
    for c in chains:
        assert all(int(flat[v]) != 90 for v in c), "a chain never holds an internal node"
        for a, b in zip(c, c[1:] + c[:1]):  # consecutive nodes, and the closing step
            assert linked(g, a, b)
        assert not (set(c) & seen), "chains are disjoint"
        seen |= set(c)
    starts = [c[0] for c in chains]
    assert starts == sorted(starts), "chains come in scan order of their first node"


def test_border_chain_invariants_on_synthetic_frames(D, O):
    for kind, seed, W, H in ((0, 5, 48, 40), (2, 6, 40, 32), (4, 7, 36, 36), (1, 8, 24, 24)):
        ch = 4 if kind == 1 else 3
        img = D.synth_frame(kind, seed, W, H, ch)
        g = O.heuristic_cross(O.trivial_cross(O.similarity_graph(img, W, H, channels=ch, flags=4 if ch == 4 else 0)))
        chains = O.border_chains(g)
        check_chain_invariants(g, chains)
    g, expected = _alex()
    check_chain_invariants(g, expected)


def test_border_chain_of_a_2x2_block(O):
    """A 2x2 block of one colour inside another: its four pixels keep only
    their orthogonal links (trivial crossing removal), and the walk from its
    lower-left pixel leaves by the first link clockwise from up-left (up) and
    turns clockwise: lower-left, upper-left, upper-right, lower-right."""
    W, H = 12, 10
    img = np.zeros((H, W * 3), dtype=np.uint8)
    x, y = 5, 4
    for j in (y, y + 1):
        img[j, 3 * x:3 * x + 6] = 200
    g = O.trivial_cross(O.similarity_graph(img, W, H))
    chains = O.border_chains(g)
    ll = y * W + x
    assert [ll, ll + W, ll + W + 1, ll + 1] in chains
    check_chain_invariants(g, chains)
