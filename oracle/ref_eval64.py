"""FP64 restatement of the reference's evaluation() with the reference's own constants.

TEST INFRASTRUCTURE ONLY (the product never imports it).  Follows
proposed model/SG-EMTF_DE.cu:834-1069 (and its GA copy proposed model/SG-EMTF.cu:898-1111)
formula by formula in float64, from the binary32 transformed point z = M (x + b) that the
reference forms (SG-EMTF_DE.cu:856 and every other case):
  0 Sphere       sum z^2
  1 Rosenbrock   sum_{j<D-1} 100 (z_j^2 - z_{j+1})^2 + (z_j - 1)^2
  2 Ackley       -20 exp(-0.2 sqrt(S / D)) - exp(sum cos(2 pi z) / D) + 20 + E, E = 2.718282
  3 Rastrigin    sum z^2 - 10 cos(2 pi z) + 10
  4 Griewank     1 + S / 4000 - prod cos(z_j / sqrt(j + 1)); the GA copy's product is
                 commented out (SG-EMTF.cu:1029), leaving 1 + S / 4000 - 1
  5 Weierstrass  sum_j sum_{k<=20} 0.5^k cos(pi y_jk) - D W, y_jk = tv_j 3^k rounded to
                 binary32 as the reference's cospif argument is (tv_j = 2 (z_j + 0.5) in
                 binary32, 3^k by the binary32 recurrence), W = Weierstrass_para
                 (SG-EMTF.cu:354-361): sum_k 0.5^k cos(2 PI 3^k 0.5) with PI = 3.1415926 and
                 the argument rounded to binary32 (the host cosf's argument)
  6 Schwefel     418.9829f D - sum z sin(sqrt|z|)
Cosines of binary32 arguments are reduced exactly (an integer multiple of 2 pi removed in
float64 from an exact binary32 value), so the float64 results are accurate to ~1e-15.
"""
import numpy as np

PI_REF = 3.1415926
E_REF = 2.718282
SCHWEFEL_REF = float(np.float32(418.9829))


def p3_binary32():
    """3^k, k = 0..20, by the reference's binary32 recurrence pow_val2 *= 3."""
    out, p = [], np.float32(1.0)
    for _ in range(21):
        out.append(p)
        p = np.float32(p * np.float32(3.0))
    return np.array(out, np.float32)


def cospi64(y):
    """cos(pi y) for binary32 (or float64) y, reducing y modulo 2 exactly first."""
    y = np.asarray(y, np.float64)
    r = np.fmod(y, 2.0)  # exact
    return np.cos(np.pi * r)


def weierstrass_para():
    p3 = p3_binary32()
    res = 0.0
    for k in range(21):
        arg = np.float32(2.0 * PI_REF * float(p3[k]) * 0.5)
        res += 0.5 ** k * np.cos(float(arg))
    return res


def transform32(x, diag, shift):
    """The reference's z = M (x + b) in binary32; the engine stores o = -b, x - o = x + b."""
    x = np.asarray(x, np.float32)
    return (np.asarray(diag, np.float32) * (x - np.asarray(shift, np.float32))).astype(np.float32)


def evaluate(func, z32, ga=False):
    """f of each row of z32 [n, D] (binary32 transformed points) in float64."""
    z32 = np.atleast_2d(np.asarray(z32, np.float32))
    z = z32.astype(np.float64)
    D = z.shape[1]
    if func == 0:
        return (z * z).sum(1)
    if func == 1:
        return (100.0 * (z[:, :-1] ** 2 - z[:, 1:]) ** 2 + (z[:, :-1] - 1.0) ** 2).sum(1)
    if func == 2:
        S = (z * z).sum(1)
        C = cospi64(2.0 * z).sum(1)
        return -20.0 * np.exp(-0.2 * np.sqrt(S / D)) - np.exp(C / D) + 20.0 + E_REF
    if func == 3:
        return (z * z - 10.0 * cospi64(2.0 * z) + 10.0).sum(1)
    if func == 4:
        S = (z * z).sum(1)
        if ga:
            return 1.0 + S / 4000.0 - 1.0
        j = np.arange(D, dtype=np.float64)
        return 1.0 + S / 4000.0 - np.prod(np.cos(z / np.sqrt(j + 1.0)), 1)
    if func == 5:
        tv = np.float32(2.0) * (z32 + np.float32(0.5))  # binary32, as the reference's temp_val
        p3 = p3_binary32()
        y = (tv[..., None] * p3[None, None, :]).astype(np.float32)  # binary32 cospif arguments
        w = 0.5 ** np.arange(21)
        return (cospi64(y) * w).sum((1, 2)) - D * weierstrass_para()
    if func == 6:
        return SCHWEFEL_REF * D - (z * np.sin(np.sqrt(np.abs(z)))).sum(1)
    raise ValueError(func)
