"""Writes tiny10.json: the hand-derived 10x10 fixture.

Matrix (n = 10): row 0 = {0}; row 1 = {0..9}; row 2 empty; row 3 = {2,3,4};
rows 4..9 = {r}.  a(i,j) = 10 i + j + 1, x_j = j + 1.  Every array below was
derived by hand from the reference's builder and kernel definitions
(getArraysLenAXT / getArraysAxSecAXT_* / int_axt_com_h1, intSpmv_main.c:
1118-1260, 1517-1566), not by running code, and is written out literally.
"""
import json
import os

rows = [[0], list(range(10)), [], [2, 3, 4]] + [[r] for r in range(4, 10)]
fx = {
    "n": 10,
    "rows": rows,
    "value_rule": "a(i,j) = 10*i + j + 1",
    "x_rule": "x_j = j + 1",
    "nnz": 20,
    "row_ptr": [0, 1, 11, 11, 14, 15, 16, 17, 18, 19, 20],
    "y_exact": [1, 935, 0, 410, 225, 336, 469, 624, 801, 1000],
    "noc_th1": {
        "tileN": 10, "lenAX": 160, "lenSEC": 10,
        "con": [0, 16, 48, 48, 64, 80, 96, 112, 128, 144],
        "sec": [0, 1, 1, 3, 4, 5, 6, 7, 8, 9],
        "tile1": [11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8],
        "tile2": [19, 20, 0, 0, 0, 0, 0, 0, 9, 10, 0, 0, 0, 0, 0, 0],
    },
    "noc_th2": {
        "tileN": 2, "lenAX": 64, "lenSEC": 16,
        "con": [0, 1, 6, 6, 32, 33, 34, 35, 36, 37],
        "sec": [0, 1, 1, 1, 1, 1, 3, 3, 4, 5, 6, 7, 8, 9, 0, 0],
        "tile0_brickrow0": [1, 11, 13, 15, 17, 19, 33, 35, 1, 1, 3, 5, 7, 9, 3, 5],
        "tile0_brickrow1": [0, 12, 14, 16, 18, 20, 34, 0, 0, 2, 4, 6, 8, 10, 4, 0],
    },
    "com_bs8": {
        "tileN": 3, "lenAX": 48, "lenSEC": 24, "log": 3,
        "sec": [0, 14, 0, 0, 0, 0, 0, 0, 10, 0, 0, 26, 0, 0, 32, 40, 48, 56, 64, 72, 0, 0, 0, 0],
        "y_reference_kernel": [0, 935, 0, 410, 225, 336, 469, 624, 801, 1000],
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny10.json")
    with open(out, "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote", out)
