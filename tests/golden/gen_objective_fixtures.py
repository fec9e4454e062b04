"""Writes objective_f64.json: objective values at seeded points, computed with
mpmath (60 digits) from the textbook definitions, independently of the oracle.

Points: each objective at D in {2, 10, 50} for random x in the box, under the
shift, rotation and diagonal transforms of a seeded task.  Run from the repo
root: python tests/golden/gen_objective_fixtures.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 60
HERE = os.path.dirname(os.path.abspath(__file__))


def z_of(task, x):
    D = len(x)
    w = [mp.mpf(float(x[d])) - mp.mpf(float(task["shift"][d])) for d in range(D)]
    if task["xform"] == 1:
        Q = task["rot"]
        return [mp.fsum(mp.mpf(float(Q[j][d])) * w[d] for d in range(D)) for j in range(D)]
    if task["xform"] == 2:
        return [mp.mpf(float(task["diag"][j])) * w[j] for j in range(D)]
    return w


def f_of(func, z):
    D = len(z)
    pi = mp.pi
    if func == 0:
        return mp.fsum(v * v for v in z)
    if func == 1:
        return mp.fsum(100 * (z[j] ** 2 - z[j + 1]) ** 2 + (z[j] - 1) ** 2 for j in range(D - 1))
    if func == 2:
        s = mp.fsum(v * v for v in z)
        c = mp.fsum(mp.cos(2 * pi * v) for v in z)
        return -20 * mp.exp(-mp.mpf("0.2") * mp.sqrt(s / D)) - mp.exp(c / D) + 20 + mp.e
    if func == 3:
        return mp.fsum(v * v - 10 * mp.cos(2 * pi * v) + 10 for v in z)
    if func == 4:
        s = mp.fsum(v * v for v in z)
        p = mp.mpf(1)
        for j, v in enumerate(z):
            p *= mp.cos(v / mp.sqrt(j + 1))
        return 1 + s / 4000 - p
    if func == 5:
        w0 = mp.fsum(mp.mpf(0.5) ** k * mp.cos(pi * 3 ** k) for k in range(21))
        return mp.fsum(mp.fsum(mp.mpf(0.5) ** k * mp.cos(2 * pi * 3 ** k * (v + mp.mpf(0.5))) for k in range(21))
                       for v in z) - D * w0
    if func == 6:
        return mp.mpf("418.9828872724338") * D - mp.fsum(v * mp.sin(mp.sqrt(abs(v))) for v in z)
    if func == 7:
        return mp.fsum((mp.mpf(10) ** (6 * mp.mpf(j) / (D - 1)) if D > 1 else 1) * v * v for j, v in enumerate(z))
    raise ValueError(func)


def main():
    rng = np.random.Generator(np.random.PCG64(2024))
    transforms, cases = [], []
    for D in (2, 10, 50):
        for xform in (0, 1, 2):
            shift = rng.uniform(-80, 80, D).astype(np.float32)
            task = {"dim": D, "xform": xform, "shift": [float(v) for v in shift]}
            if xform == 1:
                A = rng.standard_normal((D, D))
                Q, R = np.linalg.qr(A)
                task["rot"] = [[float(v) for v in row] for row in (Q * np.sign(np.diag(R))[None, :]).astype(np.float32)]
            if xform == 2:
                task["diag"] = [float(v) for v in (1.0 - rng.random(D)).astype(np.float32)]
            transforms.append(task)
            # small-|z| points matter for the cancellation-free forms: mix box and near-optimum points
            pts = [rng.uniform(-100, 100, D).astype(np.float32), (shift + rng.normal(0, 1e-3, D)).astype(np.float32)]
            for x in pts:
                z = z_of(task, x)
                for func in range(8):
                    cases.append({"t": len(transforms) - 1, "func": func, "x": [float(v) for v in x],
                                  "f": mp.nstr(f_of(func, z), 30)})
    with open(os.path.join(HERE, "objective_f64.json"), "w") as fh:
        json.dump({"transforms": transforms, "cases": cases}, fh, separators=(",", ":"))
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
