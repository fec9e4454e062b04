"""Derives the float32 polynomial coefficients used by the FP32 evaluator.

Test infrastructure / documentation: prints Chebyshev (near-minimax) fits,
computed in 60-digit arithmetic with mpmath, rounded to float32, together with
the maximum relative error of the rounded polynomial measured in float64.
Both the device code and the CPU oracle carry these constants; the oracle's
unit tests check the resulting functions against float64 references.
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 60


def fit(name, f, a, b, n, scale=None):
    poly, err = mp.chebyfit(f, [a, b], n, error=True)
    # mpmath returns highest-degree first
    c32 = [np.float32(float(c)) for c in poly]
    xs = np.linspace(a, b, 20001)
    ref = np.array([float(f(mp.mpf(float(x)))) for x in xs])
    approx = np.zeros_like(xs)
    for c in c32:
        approx = approx * xs + np.float64(c)
    rel = np.max(np.abs(approx - ref) / np.maximum(np.abs(ref), 1e-300))
    print(f"{name}: degree {n - 1} on [{a}, {b}], max rel err (float32 coeffs) {rel:.3e}")
    print("  " + ", ".join(f"{float(c):.9e}f" for c in c32))
    return c32


def fit_rel_minimax(name, f, a, b, deg, iters=60):
    """Least squares on a 20001-point grid, reweighted toward the minimax of the relative
    error (Lawson's iteration), lowest degree first; prints the binary32 Horner error."""
    xs = np.linspace(a, b, 20001)
    ys = f(xs)
    V = np.vander(xs, deg + 1, increasing=True)
    w = np.ones_like(xs)
    for _ in range(iters):
        W = w / np.abs(ys)
        c, *_ = np.linalg.lstsq(V * W[:, None], ys * W, rcond=None)
        err = (V @ c - ys) / ys
        w = w * (np.abs(err) + 1e-30) ** 0.5
        w /= w.mean()
    c32 = c.astype(np.float32)
    x32 = xs.astype(np.float32)
    p = np.full_like(x32, c32[-1])
    for k in range(deg - 1, -1, -1):
        p = (p * x32 + c32[k]).astype(np.float32)
    rel = np.abs((p.astype(np.float64) - ys) / ys).max()
    print(f"{name}: degree {deg} on [{a}, {b}], max rel err (binary32 Horner) {rel:.3e}")
    print("  " + ", ".join(f"{float(v)!r}f" for v in c32))
    return c32


def main():
    pi = mp.pi
    # sin^2(pi t) = x * WQ(x), x = t^2, t in [0, 1/2] (the Weierstrass even terms, lowest degree
    # first; the odd terms follow by the triple-angle identity)
    def wq(x):
        t = np.sqrt(x)
        return np.where(x > 0, np.sin(np.pi * t) ** 2 / np.where(x > 0, x, 1.0), np.pi ** 2)
    fit_rel_minimax("WQ5", wq, 0.0, 0.25, 5)
    # sin(pi*a) = a * SP(x), x = a^2, a in [0, 1/2]
    fit("SP", lambda x: (mp.sin(pi * mp.sqrt(x)) / mp.sqrt(x)) if x > 0 else pi, 0.0, 0.25, 5)
    # sin(2*pi*b) = b * SN(x), x = b^2, b in [0, 1/4]
    fit("SN", lambda x: (mp.sin(2 * pi * mp.sqrt(x)) / mp.sqrt(x)) if x > 0 else 2 * pi, 0.0, 1.0 / 16, 7)
    # expm1(r) = r + r^2 * E(r), r in [-ln2/2, ln2/2]
    h = float(mp.log(2) / 2)
    fit("E", lambda r: ((mp.expm1(r) - r) / r ** 2) if r != 0 else mp.mpf(1) / 2, -h, h, 7)
    # 2^f, f in [-1/2, 1/2]
    fit("X2", lambda t: mp.power(2, t), -0.5, 0.5, 8)
    # log2((1+s)/(1-s)) = s * L(s^2), s in [0, 0.1716] -> x = s^2 in [0, 0.02944]
    smax = float((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1))
    fit("L", lambda x: (mp.log((1 + mp.sqrt(x)) / (1 - mp.sqrt(x)), 2) / mp.sqrt(x)) if x > 0 else 2 / mp.log(2),
        0.0, smax ** 2, 6)


if __name__ == "__main__":
    main()
