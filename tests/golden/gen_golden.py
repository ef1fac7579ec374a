"""Writes the small fixtures under tests/golden/.

  python tests/golden/gen_golden.py

bsplines.npz      The five boundary-modified quadratic B-splines of the
                  reference's own test script (test_quadratic_bsplines.jl:3-73:
                  nodes 0..4, 100 points of LinRange(0,4,100), types 1,2,3,4,1)
                  evaluated in exact rational arithmetic from the piecewise
                  polynomials written in that script, then rounded once; and
                  the uniform cubic B-spline (Cox-de Boor closed form
                  2/3 - r^2 + |r|^3/2, (2-|r|)^3/6) at rational points.
julia_ranges.json Lengths and end values of Julia Float64 ranges used by the
                  drivers' particle builders: Julia's `start:step:stop`
                  recovers the rationals the endpoints denote and counts in
                  exact arithmetic, so the expected length is
                  floor((stop-start)/step)+1 over those rationals and element
                  k is start + k*step rounded once.
two_disks_energy.npz  Kinetic/strain energy history of the 2D Two Disks run
                  (20x20 cells, linear basis, USL, 3500 steps, every 10
                  steps) from the CPU restatement: a regression record that the
                  GPU run is compared against.
"""
import json
import os
import sys
from fractions import Fraction as Fr

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def q1(r):
    if Fr(-3, 2) <= r <= Fr(-1, 2):
        return Fr(1, 2) * r * r + Fr(3, 2) * r + Fr(9, 8)
    if r <= 0:
        return 1 + r
    if r <= Fr(1, 2):
        return 1 - r
    if r <= Fr(3, 2):
        return Fr(1, 2) * r * r - Fr(3, 2) * r + Fr(9, 8)
    return Fr(0)


def q2(r):
    if -1 <= r <= Fr(-1, 2):
        return 1 + r
    if r <= Fr(1, 2):
        return -r * r + Fr(3, 4)
    if r <= Fr(3, 2):
        return Fr(1, 2) * r * r - Fr(3, 2) * r + Fr(9, 8)
    return Fr(0)


def q3(r):
    if Fr(-3, 2) <= r <= Fr(-1, 2):
        return Fr(1, 2) * r * r + Fr(3, 2) * r + Fr(9, 8)
    if r <= Fr(1, 2):
        return -r * r + Fr(3, 4)
    if r <= Fr(3, 2):
        return Fr(1, 2) * r * r - Fr(3, 2) * r + Fr(9, 8)
    return Fr(0)


def q4(r):
    if Fr(-3, 2) <= r <= Fr(-1, 2):
        return Fr(1, 2) * r * r + Fr(3, 2) * r + Fr(9, 8)
    if r <= Fr(1, 2):
        return -r * r + Fr(3, 4)
    if r <= 1:
        return 1 - r
    return Fr(0)


def cubic(r):
    a = abs(r)
    if a <= 1:
        return Fr(2, 3) - a * a + a * a * a / 2
    if a <= 2:
        return (2 - a) ** 3 / 6
    return Fr(0)


def bsplines():
    pts = [Fr(4 * p, 99) for p in range(100)]
    # the script evaluates at its points; the exact rational of each Float64 point is used
    xs = np.array([float(p) for p in pts])
    fx = [Fr(float(p)) for p in pts]
    funcs = (q1, q2, q3, q4, q1)
    # The script's functions have no guard left of their support (its type 3
    # gives -3.25 at r = -2); Bsplines.jl returns 0 there.  Points left of a
    # kernel's support are stored as NaN and not compared.
    left = (Fr(-3, 2), Fr(-1), Fr(-3, 2), Fr(-3, 2), Fr(-3, 2))
    vals = np.array([[float(f(x - k)) if x - k >= left[k] else np.nan for k, f in enumerate(funcs)] for x in fx])
    rc = [Fr(k, 16) for k in range(-40, 41)]
    cub = np.array([float(cubic(r)) for r in rc])
    np.savez(os.path.join(HERE, "bsplines.npz"), points=xs, quad=vals, cubic_r=np.array([float(r) for r in rc]),
             cubic=cub)


def ranges():
    out = {}
    for name, (a, s, b) in {
        "two_disks_circle": (-0.2 + 0.5 * 0.025, 0.025, 0.2 - 0.5 * 0.025),
        "impact_plate_xy": (-0.5 * (277 * 0.36e-3) + 0.5 * 0.36e-3, 0.36e-3, 0.5 * (277 * 0.36e-3) - 0.5 * 0.36e-3),
        "impact_plate_z": (-0.5 * 20e-3 + 0.5 * 0.36e-3, 0.36e-3, 0.5 * 20e-3 - 0.5 * 0.36e-3),
    }.items():
        # Julia reconstructs the rationals the Float64 endpoints stand for
        # (Base.rat) and counts elements in that arithmetic: use the shortest
        # decimal of each endpoint, which is that rational here.
        fa, fs, fb = Fr(repr(a)), Fr(repr(s)), Fr(repr(b))
        n = int((fb - fa) / fs) + 1
        out[name] = dict(start=a, step=s, stop=b, length=n, first=float(fa), last=float(fa + (n - 1) * fs))
    with open(os.path.join(HERE, "julia_ranges.json"), "w") as f:
        json.dump(out, f, indent=1)


def two_disks_energy():
    from jump_amd import scenes
    from oracle.pyoracle import OracleEngine
    sc = scenes.two_disks()
    eng = OracleEngine(sc.params)
    sc.load(eng)
    t = 0.0
    rec = []
    # reference loop: energies after the step whose counter % 10 == 0, at t before the increment
    t, _ = eng.step(1, t)
    rec.append((0.0,) + eng.energies())
    for c in range(1, 3500):
        tc = t
        t, _ = eng.step(1, t)
        if c % 10 == 0:
            rec.append((tc,) + eng.energies())
    np.savez(os.path.join(HERE, "two_disks_energy.npz"), rec=np.array(rec))


if __name__ == "__main__":
    bsplines()
    ranges()
    two_disks_energy()
    print("fixtures written to", HERE)
