"""Writes bvh_order_table.json: the BVH4 child visit-order table.

Regenerates TRAVERSAL_TABLE by following the steps of the reference's build
script (sub_crates/bvh_order/build.rs:5-59): for each ray octant, 27 codes
for four children (top/left/right split axes), 9 + 9 codes for three
children (left or right pair split) and 3 codes for two children; each
entry packs the visit order two bits per slot, first child in the low bits.

    python tests/golden/gen_bvh_order_table.py
"""

import json
import os


def table():
    out = []
    for raydir in range(8):
        ray = [raydir & 1, (raydir >> 1) & 1, (raydir >> 2) & 1]
        row = []
        for s2 in range(3):
            for s1 in range(3):
                for s0 in range(3):
                    p = [0, 1, 2, 3]
                    if ray[s1]:
                        p[0], p[1] = p[1], p[0]
                    if ray[s2]:
                        p[2], p[3] = p[3], p[2]
                    if ray[s0]:
                        p[0], p[2] = p[2], p[0]
                        p[1], p[3] = p[3], p[1]
                    row.append(p[0] + (p[1] << 2) + (p[2] << 4) + (p[3] << 6))
        for s1 in range(3):
            for s0 in range(3):
                p = [0, 1, 2]
                if ray[s1]:
                    p[0], p[1] = p[1], p[0]
                if ray[s0]:
                    p[0], p[1] = p[1], p[0]
                    p[0], p[2] = p[2], p[0]
                row.append(p[0] + (p[1] << 2) + (p[2] << 4))
        for s1 in range(3):
            for s0 in range(3):
                p = [0, 1, 2]
                if ray[s1]:
                    p[1], p[2] = p[2], p[1]
                if ray[s0]:
                    p[0], p[2] = p[2], p[0]
                    p[0], p[1] = p[1], p[0]
                row.append(p[0] + (p[1] << 2) + (p[2] << 4))
        for s0 in range(3):
            p = [0, 1]
            if ray[s0]:
                p[0], p[1] = p[1], p[0]
            row.append(p[0] + (p[1] << 2))
        out.append(row)
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bvh_order_table.json")
    with open(path, "w") as f:
        json.dump({"source": "sub_crates/bvh_order/build.rs:5-59", "table": table()}, f)
    print("wrote", path)
