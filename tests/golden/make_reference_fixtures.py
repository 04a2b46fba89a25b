"""Writes reference_graphs.json and alex_chains.json from the data the
reference ships (run in the build container, where the reference sources are
mounted; the JSON files are committed and used without it).

  c_pattern graph  kernel.cu:291-293 (8x8, commented-out graph_dump)
  alex graph       kernel.cu:296     (24x24, commented-out graph_dump)
  alex chains      alex_png.txt:1-21 (border chains "Spline: k - nodes")
                   alex_png.txt:24-44 (chain-code walk from node 5)
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def graph_dumps():
    lines = open(os.path.join(REF, "kernel.cu"), encoding="latin-1").read().splitlines()
    c_pattern = " ".join(lines[290:293])   # kernel.cu:291-293
    alex = lines[295]                        # kernel.cu:296
    def nums(s):
        body = s[s.index("{") + 1:s.rindex("}")]
        return [int(v) & 0xFF for v in re.findall(r"-?\d+", re.sub(r"\(char\)", "", body))]
    return nums(c_pattern), nums(alex)


def chains():
    text = open(os.path.join(REF, "alex_png.txt"), encoding="latin-1").read().splitlines()
    out = []
    for ln in text:
        m = re.match(r"Spline:\s*(\d+)\s*-\s*(.*)$", ln)
        if m:
            out.append([int(v) for v in re.findall(r"\d+", m.group(2))])
    walk = [[int(a), int(b)] for a, b in re.findall(r"index:\s*(\d+),\s*edge:\s*(\d+)", "\n".join(text))]
    return out, walk


if __name__ == "__main__":
    cp, al = graph_dumps()
    assert len(cp) == 64 and len(al) == 576, (len(cp), len(al))
    json.dump({"source": "kernel.cu:291-296 (graph_dump arrays, row-major, row 0 = bottom)",
               "c_pattern": {"width": 8, "height": 8, "graph": cp},
               "alex": {"width": 24, "height": 24, "graph": al}},
              open(os.path.join(OUT, "reference_graphs.json"), "w"), indent=0)
    ch, walk = chains()
    json.dump({"source": "alex_png.txt:1-44 (border chains of the alex graph and a chain-code walk)",
               "chains": ch, "walk_from_5": walk},
              open(os.path.join(OUT, "alex_chains.json"), "w"), indent=0)
    print(len(cp), len(al), len(ch), len(walk))
