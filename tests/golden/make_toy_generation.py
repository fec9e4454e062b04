"""Writes toy_generation.json: one DE/rand/1/bin and one SBX + polynomial-mutation GA
generation of a 16 x 8 toy population, with the explicit draw tables they consume.

The operators are restated here in numpy binary32, independently of the oracle's C code,
from the reference's definitions (proposed model/SG-EMTF_DE.cu:474-593 mutant and binomial
crossover, :679-705 selection; proposed model/SG-EMTF.cu:548-606 polynomial mutation,
:711-749 SBX) and the engine's documented draw addressing (counter = {block,
purpose<<24 | sub<<16 | index, generation, task}, key = seed halves).  Only the Philox
blocks (pinned by rocRAND / Random123 known answers) and binary32 pow (pinned against
math.pow) are taken from the oracle.  Run from the repository root:
    python tests/golden/make_toy_generation.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

f32 = np.float32
P, D, SEED, GEN = 16, 8, 0x5EED0123456789AB, 1
LO, HI = f32(-5.0), f32(5.0)
P_DE_IDX, P_DE_CR, P_GA_MATE, P_GA_SBX, P_GA_PM = 2, 3, 4, 5, 6


def bits(a):
    return [int(v) for v in np.asarray(a, np.float32).view(np.uint32).ravel()]


def block(purpose, index, sub, blk):
    return oracle.draw_block(SEED, 0, GEN, purpose, index, sub, blk)


def clamp(v):
    return HI if v > HI else (LO if v < LO else v)


def sphere(x, o):
    """Sum of z_j^2 in two accumulators (rows with ((j >> 2) & 1) == 0, the rest), A + B."""
    a = b = f32(0.0)
    for j in range(D):
        z = f32(x[j] - o[j])
        v = f32(z * z)
        if ((j >> 2) & 1) == 0:
            a = f32(a + v)
        else:
            b = f32(b + v)
    return f32(a + b)


def de_generation(X, fit, o, F, CR):
    m = P
    out_x, out_f, tables = X.copy(), fit.copy(), []
    for i in range(P):
        il = i
        b0 = block(P_DE_IDX, i, 0, 0)
        words = b0[:3] + block(P_DE_IDX, i, 0, 1) + block(P_DE_IDX, i, 0, 2) + block(P_DE_IDX, i, 0, 3)
        jrand = b0[3] % D
        excl, mates, k = [il], [], 0
        while len(mates) < 3:  # curand() % n redrawn while it collides (SG-EMTF_DE.cu:559-573)
            r = words[k] % m
            k += 1
            if r not in excl:
                mates.append(r)
                excl.append(r)
        r1, r2, r3 = mates
        u16, trial = [], np.zeros(D, np.float32)
        for d in range(D):
            h, s = d & 1, d >> 1
            w = block(P_DE_CR, i, h, s >> 3)[(s >> 1) & 3]
            v16 = (w & 0xFFFF) if (s & 1) else (w >> 16)
            u16.append(v16)
            u = f32(f32(v16 + 1) * f32(2.0 ** -16))
            diff = f32(X[r2, d] - X[r3, d])
            mut = clamp(f32(X[r1, d] + f32(F * diff)))
            trial[d] = X[i, d] if (u > CR and d != jrand) else mut
        ft = sphere(trial, o)
        if fit[i] >= ft:  # ties go to the child
            out_x[i], out_f[i] = trial, ft
        tables.append({"index_words": words[:k], "jrand": jrand, "mates": [r1, r2, r3], "cr_u16": u16,
                       "trial": bits(trial), "trial_fitness": bits([ft])[0]})
    return out_x, out_f, tables


def ga_generation(X, fit, o, eta_c, eta_m, pm):
    m = P
    rng = f32(HI - LO)
    ec = f32(f32(1.0) / f32(eta_c + f32(1.0)))
    em1 = f32(eta_m + f32(1.0))
    em = f32(f32(1.0) / em1)
    out_x, out_f, tables = X.copy(), fit.copy(), []
    for i in range(P):
        il = i
        mb = block(P_GA_MATE, i & ~1, 0, 0)
        ra, rb = mb[il & 1] % m, mb[(il & 1) ^ 1] % m
        p_words, pm_words, trial = [], [], np.zeros(D, np.float32)
        for d in range(D):
            h, s = d & 1, d >> 1
            wp = block(P_GA_SBX, i, h, s >> 2)[s & 3]
            p = f32(f32(2 * (wp >> 9) + 1) * f32(2.0 ** -24))
            tv = f32(f32(2.0) * p) if p <= f32(0.5) else f32(f32(0.5) / f32(f32(1.0) - p))
            beta = f32(oracle.scalar("pow", float(tv), float(ec)))
            if il & 1:
                beta = -beta
            c = f32(f32(0.5) * f32(f32(f32(1.0) + beta) * X[ra, d] + f32(f32(1.0) - beta) * X[rb, d]))
            c = clamp(c)
            pb = block(P_GA_PM, i, h, s >> 1)
            wr, wu = pb[(s & 1) * 2], pb[(s & 1) * 2 + 1]
            p_words.append(wp)
            pm_words.append([wr, wu])
            r = f32(f32((wr >> 8) + 1) * f32(2.0 ** -24))
            if r < pm:
                u = f32(f32((wu >> 8) + 1) * f32(2.0 ** -24))
                if u > f32(0.5):  # Deb's two-sided polynomial mutation, upper side
                    delta, dl = f32(-1.0), f32(f32(HI - c) / rng)
                    iv1, iv2 = f32(f32(2.0) * f32(f32(1.0) - u)), f32(f32(2.0) * f32(u - f32(0.5)))
                else:
                    delta, dl = f32(1.0), f32(f32(c - LO) / rng)
                    iv1 = f32(f32(2.0) * u)
                    iv2 = f32(f32(1.0) - iv1)
                pw = f32(oracle.scalar("pow", float(f32(f32(1.0) - dl)), float(em1)))
                a4 = f32(oracle.scalar("pow", float(f32(iv1 + f32(iv2 * pw))), float(em)))
                delta = f32(delta * f32(a4 - f32(1.0)))
                c = clamp(f32(c + f32(delta * rng)))
            trial[d] = c
        ft = sphere(trial, o)
        if fit[i] >= ft:
            out_x[i], out_f[i] = trial, ft
        tables.append({"mates": [ra, rb], "sbx_words": p_words, "pm_words": pm_words, "trial": bits(trial),
                       "trial_fitness": bits([ft])[0]})
    return out_x, out_f, tables


def main():
    rs = np.random.default_rng(7)
    o = rs.uniform(-2, 2, D).astype(np.float32)
    X = rs.uniform(-5, 5, (P, D)).astype(np.float32)
    fit = np.array([sphere(X[i], o) for i in range(P)], np.float32)
    F, CR = f32(0.5), f32(0.9)
    dx, dfit, dtab = de_generation(X, fit, o, F, CR)
    gx, gfit, gtab = ga_generation(X, fit, o, f32(2.0), f32(5.0), f32(1.0))
    out = {"seed": SEED, "generation": GEN, "pop": P, "dim": D, "lo": float(LO), "hi": float(HI),
           "func": 0, "shift": bits(o), "x": bits(X), "fitness": bits(fit),
           "de": {"F": float(F), "CR": float(CR), "draws": dtab, "x": bits(dx), "fitness": bits(dfit)},
           "ga": {"eta_c": 2.0, "eta_m": 5.0, "pm": 1.0, "draws": gtab, "x": bits(gx), "fitness": bits(gfit)}}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "toy_generation.json"), "w") as fh:
        json.dump(out, fh)
    print("replaced (DE, GA):", int((dfit != fit).sum()), int((gfit != fit).sum()))


if __name__ == "__main__":
    main()
