"""Writes tests/golden/reference_tables.json: the numeric tables the
reference holds for the tracking path, read as text from its sources.

  swap32 .. swap2  cuslines/cuda_c/cuwsort.cuh:37-80 (warp sort networks)
  DISC_2_VERT/FACE cuslines/cuda_c/disc.h:20-84 (PTT disc, quality 2)

Run in a container that has the reference mounted at /root/reference.
"""

import json
import os
import re

REF = "/root/reference/cuslines/cuda_c"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")


def ints(text):
    return [int(x) for x in re.findall(r"-?\d+", text)]


def floats(text):
    return [float(x) for x in re.findall(r"-?\d+\.\d+", text)]


def main():
    src = open(os.path.join(REF, "cuwsort.cuh")).read()
    out = {}
    for n in (32, 16, 8, 4, 2):
        m = re.search(r"swap%d\[(\d+)\]\[(\d+)\]\s*=\s*\{(.*?)\};" % n, src, re.S)
        r, c = int(m.group(1)), int(m.group(2))
        v = ints(m.group(3))
        out[f"swap{n}"] = [v[i * c:(i + 1) * c] for i in range(r)]
    disc = open(os.path.join(REF, "disc.h")).read()
    mv = re.search(r"#define DISC_2_VERT \{(.*?)\}", disc, re.S)
    mf = re.search(r"#define DISC_2_FACE \{(.*?)\}", disc, re.S)
    out["disc2_vert"] = floats(mv.group(1))
    out["disc2_face"] = ints(mf.group(1))
    with open(OUT, "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
