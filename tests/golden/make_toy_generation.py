"""Writes toy_generation.json: one DE/rand/1/bin and one SBX + polynomial-mutation GA
generation of a 16 x 8 toy population, with the explicit draw tables they consume; and
("islands") two generations of three tasks with four islands each, different search boxes
and one smaller dimension, covering the cross-island exchange, the island migration and one
inter-task transfer event, with every intermediate state and the draws each step consumed.

The operators are restated here in numpy binary32, independently of the oracle's C code,
from the reference's definitions (proposed model/SG-EMTF_DE.cu:474-593 mutant and binomial
crossover, :679-705 selection; proposed model/SG-EMTF.cu:548-606 polynomial mutation,
:711-749 SBX; exchange SharedToGlobal + SelectFromBlocks SG-EMTF.cu:856-896, :1220-1260 with
one island offset per slot, intervalRand :1593-1600; island migration SG-EMTF_DE.cu:1093-1170;
knowledgeTransfer SG-EMTF.cu:1514-1591 with the partner draw :1696-1699; pop_init
SG-EMTF.cu:146-159) and the engine's documented draw addressing (counter = {block,
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


# ------------------------------------------------------------------ islands scenario
P_INIT, P_MIG, P_XFER, P_ISLX = 1, 7, 8, 9
ISL = dict(seed=0x15A7D5EED0000042, n_islands=4, F=0.5, CR=0.9, migrate_count=2, transfer_count=3,
           transfer_cr=0.4, generations=2)
ISL_TASKS = [(8, -5.0, 5.0), (8, -100.0, 100.0), (6, -50.0, 50.0)]  # (dim, lo, hi); Sphere, pop P


def iblock(task, gen, purpose, index, sub, blk):
    return oracle.draw_block(ISL["seed"], task, gen, purpose, index, sub, blk)


def iword(task, gen, purpose, index, sub, word):
    return iblock(task, gen, purpose, index, sub, word >> 2)[word & 3]


def u01(w):
    return f32(f32((w >> 8) + 1) * f32(2.0 ** -24))


def sphere_d(x, o):
    """Sphere of any dimension, summed like sphere() above."""
    a = b = f32(0.0)
    for j in range(len(x)):
        z = f32(x[j] - o[j])
        v = f32(z * z)
        if ((j >> 2) & 1) == 0:
            a = f32(a + v)
        else:
            b = f32(b + v)
    return f32(a + b)


def ranking(fit):
    """Indices ordered by (fitness, index), NaN last."""
    key = [(float("inf") if f != f else float(f), i) for i, f in enumerate(fit)]
    return [i for _, i in sorted(key)]


def isl_init(t, dim, lo, hi, o):
    """x = U(0,1] (hi - lo) + lo per gene (pop_init, SG-EMTF.cu:146-159)."""
    rng = f32(f32(hi) - f32(lo))
    X = np.zeros((P, dim), np.float32)
    for i in range(P):
        for d in range(dim):
            X[i, d] = f32(f32(u01(iword(t, 0, P_INIT, i, 0, d)) * rng) + f32(lo))
    return X, np.array([sphere_d(X[i], o) for i in range(P)], np.float32)


def isl_exchange(t, gen, X, fit):
    """Each island's rows are written out Fisher-Yates shuffled (SharedToGlobal: curand() %
    (k+1) for k = m-1 .. 0, staged[shuffle[p]] = row p); slot k of every island q is then
    refilled from slot k of island (q + off_k) mod n, one offset per slot shared by every
    island (SelectFromBlocks with select_interval[iter * IPB + k])."""
    n = ISL["n_islands"]
    m = P // n
    sx, sf = np.zeros_like(X), np.zeros_like(fit)
    shuffles = []
    for q in range(n):
        arr = list(range(m))
        for k in range(m - 1, -1, -1):
            tt = iword(t, gen, P_ISLX, q, 0, m - 1 - k) % (k + 1)
            arr[k], arr[tt] = arr[tt], arr[k]
        shuffles.append(arr)
        for p in range(m):
            sx[q * m + arr[p]], sf[q * m + arr[p]] = X[q * m + p], fit[q * m + p]
    offs = [iword(t, gen, P_ISLX, 0, 1, k) % n for k in range(m)]
    ox, of = np.zeros_like(X), np.zeros_like(fit)
    for q in range(n):
        for k in range(m):
            src = ((q + offs[k]) % n) * m + k
            ox[q * m + k], of[q * m + k] = sx[src], sf[src]
    return ox, of, {"shuffles": shuffles, "offsets": offs}


def isl_de(t, gen, X, fit, o, lo, hi):
    """DE/rand/1/bin inside each island (mates island-local) plus greedy selection."""
    n, F, CR = ISL["n_islands"], f32(ISL["F"]), f32(ISL["CR"])
    m, dim = P // n, X.shape[1]
    lo, hi = f32(lo), f32(hi)
    trials, tfit, mates_all = np.zeros_like(X), np.zeros_like(fit), []
    for i in range(P):
        il = i % m
        base = i - il
        b0 = iblock(t, gen, P_DE_IDX, i, 0, 0)
        words = b0[:3] + iblock(t, gen, P_DE_IDX, i, 0, 1) + iblock(t, gen, P_DE_IDX, i, 0, 2) + \
            iblock(t, gen, P_DE_IDX, i, 0, 3)
        jrand = b0[3] % dim
        excl, mates, k = [il], [], 0
        while len(mates) < 3:
            if k < len(words):
                r = words[k] % m
                k += 1
            else:  # the 15-word stream is spent: the smallest admissible index
                r = min(v for v in range(m) if v not in excl)
            if r not in excl:
                mates.append(r)
                excl.append(r)
        r1, r2, r3 = [base + r for r in mates]
        mates_all.append([r1, r2, r3, jrand])
        for d in range(dim):
            h, s = d & 1, d >> 1
            w = iblock(t, gen, P_DE_CR, i, h, s >> 3)[(s >> 1) & 3]
            v16 = (w & 0xFFFF) if (s & 1) else (w >> 16)
            u = f32(f32(v16 + 1) * f32(2.0 ** -16))
            mut = f32(X[r1, d] + f32(F * f32(X[r2, d] - X[r3, d])))
            mut = hi if mut > hi else (lo if mut < lo else mut)
            trials[i, d] = X[i, d] if (u > CR and d != jrand) else mut
        tfit[i] = sphere_d(trials[i], o)
    ox, of = X.copy(), fit.copy()
    for i in range(P):
        if fit[i] >= tfit[i]:
            ox[i], of[i] = trials[i], tfit[i]
    return ox, of, {"mates": mates_all, "trial_fitness": bits(tfit)}


def isl_migrate(X, fit):
    """On a snapshot: island q's rank-j member replaces island (q+1) mod n's rank m-k+j member
    iff strictly better (IslandMigration, SG-EMTF_DE.cu:1127-1163; the reference writes the
    fitness at rank m-k+j but the genes at rank m-1-j, this keeps genes and fitness together)."""
    n, k = ISL["n_islands"], ISL["migrate_count"]
    m = P // n
    ords = [[q * m + r for r in ranking(fit[q * m:(q + 1) * m])] for q in range(n)]
    ox, of, moves = X.copy(), fit.copy(), []
    for q in range(n):
        p = (q + 1) % n
        for j in range(k):
            src, dst = ords[q][j], ords[p][m - k + j]
            if fit[src] < fit[dst]:
                ox[dst], of[dst] = X[src], fit[src]
                moves.append([src, dst])
    return ox, of, {"moves": moves}


def isl_transfer(gen, states, shifts):
    """knowledgeTransfer (SG-EMTF.cu:1514-1591): partner = uniform other task; candidate j =
    own rank-j elite with each shared gene replaced, iff u > transfer_cr, by the partner's
    rank-j elite mapped box to box, lo1 + (v - lo2) / range2 * range1; evaluated on self and
    kept in place of own rank pop-k+j iff strictly better."""
    T, k, tcr = len(states), ISL["transfer_count"], f32(ISL["transfer_cr"])
    orders = [ranking(f) for _, f in states]
    elites = [states[t][0][orders[t][:k]].copy() for t in range(T)]
    out, info = [], []
    for t in range(T):
        X, fit = states[t]
        dim, lo, hi = ISL_TASKS[t]
        dn = iword(t, gen, P_MIG, 0, 0, 0) % (T - 1)
        dn += 1 if dn >= t else 0
        ddim, dlo, dhi = ISL_TASKS[dn]
        r_s, r_d = f32(f32(hi) - f32(lo)), f32(f32(dhi) - f32(dlo))
        same = (lo == dlo and hi == dhi)
        cands, takes = np.zeros((k, dim), np.float32), []
        for j in range(k):
            c = X[orders[t][j]].copy()
            for d in range(min(dim, ddim)):
                take = u01(iword(t, gen, P_XFER, j, d & 1, d >> 1)) > tcr
                takes.append(int(take))
                if take:
                    v = elites[dn][j, d]
                    c[d] = v if same else f32(f32(lo) + f32(f32(f32(v - f32(dlo)) / r_d) * r_s))
            cands[j] = c
        cf = np.array([sphere_d(cands[j], shifts[t]) for j in range(k)], np.float32)
        ox, of, repl = X.copy(), fit.copy(), []
        for j in range(k):
            tgt = orders[t][P - k + j]
            if cf[j] < fit[tgt]:
                ox[tgt], of[tgt] = cands[j], cf[j]
                repl.append([j, tgt])
        out.append((ox, of))
        info.append({"partner": dn, "takes": takes, "candidates": bits(cands), "candidate_fitness": bits(cf),
                     "replaced": repl})
    return out, info


def islands():
    rs = np.random.default_rng(11)
    shifts = [rs.uniform(0.4 * lo, 0.4 * hi, dim).astype(np.float32) for dim, lo, hi in ISL_TASKS]
    states = [isl_init(t, dim, lo, hi, shifts[t]) for t, (dim, lo, hi) in enumerate(ISL_TASKS)]
    rec = {"seed": ISL["seed"], "params": {k: v for k, v in ISL.items() if k != "seed"}, "pop": P,
           "tasks": [{"dim": dim, "lo": lo, "hi": hi, "shift": bits(shifts[t])} for t, (dim, lo, hi) in
                     enumerate(ISL_TASKS)],
           "init": [{"x": bits(x), "fitness": bits(f)} for x, f in states], "generations": []}
    for gen in range(1, ISL["generations"] + 1):
        g = {"generation": gen}
        if gen > 1:  # transfer events come before generations 2, 3, ... (interval 1)
            states, g["transfer"] = isl_transfer(gen - 1, states, shifts)
            g["after_transfer"] = [{"x": bits(x), "fitness": bits(f)} for x, f in states]
        g["tasks"] = []
        for t, (dim, lo, hi) in enumerate(ISL_TASKS):
            X, fit = states[t]
            X1, f1, ex = isl_exchange(t, gen, X, fit)
            X2, f2, de = isl_de(t, gen, X1, f1, shifts[t], lo, hi)
            X3, f3, mg = isl_migrate(X2, f2)
            states[t] = (X3, f3)
            g["tasks"].append({"exchange": dict(ex, x=bits(X1), fitness=bits(f1)),
                               "de": dict(de, x=bits(X2), fitness=bits(f2)),
                               "migration": dict(mg, x=bits(X3), fitness=bits(f3))})
        rec["generations"].append(g)
    rec["final"] = [{"x": bits(x), "fitness": bits(f)} for x, f in states]
    return rec


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
           "ga": {"eta_c": 2.0, "eta_m": 5.0, "pm": 1.0, "draws": gtab, "x": bits(gx), "fitness": bits(gfit)},
           "islands": islands()}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "toy_generation.json"), "w") as fh:
        json.dump(out, fh)
    print("replaced (DE, GA):", int((dfit != fit).sum()), int((gfit != fit).sum()))
    isl = out["islands"]
    print("islands: migrations", [len(t["migration"]["moves"]) for g in isl["generations"] for t in g["tasks"]],
          "transfer partners / replaced",
          [(e["partner"], len(e["replaced"])) for g in isl["generations"] for e in g.get("transfer", [])])


if __name__ == "__main__":
    main()
