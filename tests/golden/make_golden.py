"""Writes the hand-derived golden vectors of tests/golden/*.json.

The reference ships no tests, fixtures or data, so these vectors are worked
out by hand from the operator's definition (each derivation is spelled out
below) rather than produced by any implementation.  Run from the repo root:
python tests/golden/make_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

# --- geometry (computeWeight, FunctionOnGPU.cu:12-25) --------------------
# xc = [2.25, 0.5, 0, 0]: u = 2.25 + 0.5 X - 1 = 1.25 + 0.5 X
#   X = 3  -> u = 2.75  -> x_min = 2,  du = 0.75
#   X = -4 -> u = -0.75 -> x_min = -1, du = 0.25
# yc = [1, 0, 0.25, 2]: v = 1 + 0.25 Y + 2 Z - 1 = 0.25 Y + 2 Z
#   (X, Y, Z) = (3, 1, 0)  -> v = 0.25 -> y_min = 0, dv = 0.25
#   (X, Y, Z) = (-4, -2, 0) -> v = -0.5 -> y_min = -1, dv = 0.5
geometry = {
    "xc": [2.25, 0.5, 0.0, 0.0],
    "yc": [1.0, 0.0, 0.25, 2.0],
    "cases": [
        {"XYZ": [3, 1, 0], "x_min": 2, "y_min": 0, "du": 0.75, "dv": 0.25},
        {"XYZ": [-4, -2, 0], "x_min": -1, "y_min": -1, "du": 0.25, "dv": 0.5},
    ],
}

# --- one SIRT iteration on a 2-voxel line ---------------------------------
# Volume X = 2 (Y = Z = 1, box starts at 0), detector nu = 3, nv = 1, one
# view with u = X + 0.5 (xc = [1.5, 1, 0, 0]) and v = Y (yc = [1, 0, 1, 0];
# Y = 0 here, so v = 0).
#   voxel 0 -> u = 0.5: pixel 0 weight 0.5, pixel 1 weight 0.5
#   voxel 1 -> u = 1.5: pixel 1 weight 0.5, pixel 2 weight 0.5
# x = [2, 6]:  s = [1, 1 + 3, 3] = [1, 4, 3],  c = [0.5, 1, 0.5]
# p = [3, 4, 5]: r = p - s/c = [3 - 2, 4 - 4, 5 - 6] = [1, 0, -1]
# BP: voxel 0: S = 0.5*1 + 0.5*0 = 0.5, C = 1 -> x0 += 0.5 * lambda
#     voxel 1: S = 0.5*0 + 0.5*(-1) = -0.5, C = 1 -> x1 -= 0.5 * lambda
# lambda = 1: x = [2.5, 5.5].
# BPT of p: voxel 0: (0.5*3 + 0.5*4)/1 = 3.5, voxel 1: (0.5*4 + 0.5*5)/1 = 4.5
line = {
    "vol": [2, 1, 1],
    "prj": [3, 1, 1],
    "xc": [1.5, 1.0, 0.0, 0.0],
    "yc": [1.0, 0.0, 1.0, 0.0],
    "x": [2.0, 6.0],
    "p": [3.0, 4.0, 5.0],
    "s": [1.0, 4.0, 3.0],
    "c": [0.5, 1.0, 0.5],
    "r": [1.0, 0.0, -1.0],
    "lam": 1.0,
    "x_after_sirt1": [2.5, 5.5],
    "bpt": [3.5, 4.5],
}

# --- MACE on one voxel with a one-to-one geometry ---------------------------
# One view, u = X, v = Y on a 1x1x1 box with a 1x1 detector: s = x, c = 1,
# so one SIRT iteration maps z to z + lambda (p - z).  Two agents with
# p1 = 1, p2 = 3; initial B = BPT_1 + BPT_2 = 1 + 3 = 4 = w1 = w2 = wbar.
# rho = 0.5, S = 1.
# lambda = 1 (SIRT(z) = p):
#   step 1: z = (4, 4),  v = (1, 3), w = (0.5(2-4)+2, 0.5(6-4)+2) = (1, 3), wbar = 2
#   step 2: z = (3, 1),  v = (1, 3), w = (0.5(2-3)+0.5, 0.5(6-1)+1.5) = (0, 4), wbar = 2
#   step 3: z = (4, 0),  v = (1, 3), w = (0.5(2-4)+0, 0.5(6-0)+2) = (-1, 5), wbar = 2
# lambda = 0.5 (SIRT(z) = (z + p)/2):
#   step 1: z = (4, 4), v = (2.5, 3.5), w = (2.5, 3.5), wbar = 3
#   step 2: z = (3.5, 2.5), v = (2.25, 2.75), w = (0.5+1.25, 1.5+1.75) = (1.75, 3.25), wbar = 2.5
#   step 3: z = (3.25, 1.75), v = (2.125, 2.375), w = (0.5+0.875, 1.5+1.625) = (1.375, 3.125), wbar = 2.25
mace = {
    "p": [1.0, 3.0],
    "rho": 0.5,
    "S": 1,
    "cases": [
        {"lam": 1.0, "wbar": [4.0, 2.0, 2.0, 2.0], "w1": [4.0, 1.0, 0.0, -1.0], "w2": [4.0, 3.0, 4.0, 5.0]},
        {"lam": 0.5, "wbar": [4.0, 3.0, 2.5, 2.25], "w1": [4.0, 2.5, 1.75, 1.375], "w2": [4.0, 3.5, 3.25, 3.125]},
    ],
}

# --- divideSlice (MultiAxis.cpp:13-31) ------------------------------------
# Z = 10 over 4 parts starting at 5: 10 = 4*2 + 2 -> parts 0,1 get 3 slices,
# parts 2,3 get 2: [5,8) [8,11) [11,13) [13,15).
slices = {"Z": 10, "N": 4, "S": 5, "parts": [[5, 8], [8, 11], [11, 13], [13, 15]]}

for name, obj in (("geometry", geometry), ("line", line), ("mace", mace), ("slices", slices)):
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(obj, f, indent=1)
print("golden vectors written to", HERE)
