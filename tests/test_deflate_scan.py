"""The divide-and-conquer deflation scan (dlaed2's rules) as dc_deflate_kernel
(imcom_amd/csrc/stedc.hip) runs it, 64 entries at a time: a chunk with no
deflating rotation and ascending deflated values is resolved at once from
per-entry flags and prefix counts, any other chunk is walked entry by entry.
Restated here in Python and checked against the scalar scan, entry for entry,
on random merges that exercise negligible z, rotations (clustered d) and
deflated values arriving out of order."""
import math

import numpy as np

EPS = 1.1102230246251565e-16


def dlapy2(x, y):
    xa, ya = abs(x), abs(y); w, z = max(xa, ya), min(xa, ya)
    if z == 0.0: return w
    q = z / w
    return w * math.sqrt(1.0 + q * q)


def rot_test(pd, pz, nd, nz, tol):
    s, c = pz, nz
    tau = dlapy2(c, s)
    t = nd - pd
    c = c / tau; s = -s / tau
    return abs(t * c * s) <= tol, c, s, tau


def scalar(ds, zs, cs, rho, n1=None):
    """One entry at a time: non-deflated (d, z, column) triples, deflated values
    (ascending, by insertion) and their columns, the rotations, and each
    non-deflated column's type (0: rows of the first n1 only, 1: mixed by a
    rotation across the halves, 2: the other rows only)."""
    n1 = len(ds) // 2 if n1 is None else n1
    typ = []
    K = len(ds); tol = 8 * EPS * max(max(abs(x) for x in ds), max(abs(x) for x in zs))
    dl, wz, nc, dfd, dfc, rots = [], [], [], [], [], []
    lastv = -math.inf
    def push(v, col):
        nonlocal lastv
        p = len(dfd)
        dfd.append(None); dfc.append(None)
        if v < lastv:
            while p > 0 and dfd[p - 1] > v:
                dfd[p] = dfd[p - 1]; dfc[p] = dfc[p - 1]; p -= 1
        else:
            lastv = v
        dfd[p] = v; dfc[p] = col
    if rho * max(abs(x) for x in zs) <= tol:
        for j in range(K): push(ds[j], cs[j])
        return dl, wz, nc, dfd, dfc, rots, typ
    j = 0
    while j < K and rho * abs(zs[j]) <= tol:
        push(ds[j], cs[j]); j += 1
    pd, pz, pcol, pmix = ds[j], zs[j], cs[j], False
    for j in range(j + 1, K):
        nd, nz, ncol = ds[j], zs[j], cs[j]
        if rho * abs(nz) <= tol:
            push(nd, ncol); continue
        r, c, s, tau = rot_test(pd, pz, nd, nz, tol)
        if r:
            rots.append((pcol, ncol, c, s))
            tp = pd * c * c + nd * s * s; tn = pd * s * s + nd * c * c
            push(tp, pcol)
            pmix = pmix or ((pcol < n1) != (ncol < n1))
            pd, pz, pcol = tn, tau, ncol
        else:
            dl.append(pd); wz.append(pz); nc.append(pcol)
            typ.append(1 if pmix else (0 if pcol < n1 else 2))
            pd, pz, pcol, pmix = nd, nz, ncol, False
    dl.append(pd); wz.append(pz); nc.append(pcol)
    typ.append(1 if pmix else (0 if pcol < n1 else 2))
    return dl, wz, nc, dfd, dfc, rots, typ


def chunked(ds, zs, cs, rho, n1=None):
    """The kernel's form: 64 entries per chunk (lanes); returns the same lists,
    the number of chunks walked entry by entry, and each non-deflated
    column's place in the type-grouped order (types 0, then 1, then 2)."""
    n1 = len(ds) // 2 if n1 is None else n1
    tcode = lambda col, mix: 1 if mix else (0 if col < n1 else 2)
    K = len(ds); tol = 8 * EPS * max(max(abs(x) for x in ds), max(abs(x) for x in zs))
    if rho * max(abs(x) for x in zs) <= tol:
        return [], [], [], list(ds), list(cs), [], 0, []
    dl = [None] * K; wz = [None] * K; nc = [None] * K; dfd = [None] * K; dfc = [None] * K; rots = []
    tp_ = [None] * K
    have, pd, pz, pcol, k, ndf, lastv, pmix = False, 0.0, 0.0, 0, 0, 0, -math.inf, False
    walked = 0
    for b in range(0, K, 64):
        lanes = range(min(64, K - b))
        nd = [ds[b + l] for l in lanes]; nz = [zs[b + l] for l in lanes]; ncol = [cs[b + l] for l in lanes]
        small = [rho * abs(z) <= tol for z in nz]; cand = [not x for x in small]
        rot = [False] * len(nd); unordered = [False] * len(nd)
        for l in lanes:
            cb = [q for q in range(l) if cand[q]]; sb = [q for q in range(l) if small[q]]
            if cand[l] and (cb or have):
                qd, qz = (nd[cb[-1]], nz[cb[-1]]) if cb else (pd, pz)
                rot[l] = rot_test(qd, qz, nd[l], nz[l], tol)[0]
            if small[l]:
                unordered[l] = nd[l] < (nd[sb[-1]] if sb else lastv)
        if not any(rot) and not any(unordered):
            cl = [l for l in lanes if cand[l]]
            if cl:
                base = k
                if have:
                    dl[k], wz[k], nc[k] = pd, pz, pcol; tp_[k] = tcode(pcol, pmix); base = k + 1
                for idx, l in enumerate(cl[:-1]):
                    dl[base + idx], wz[base + idx], nc[base + idx] = nd[l], nz[l], ncol[l]
                    tp_[base + idx] = tcode(ncol[l], False)
                k = base + len(cl) - 1
                L = cl[-1]; pd, pz, pcol = nd[L], nz[L], ncol[L]; have = True; pmix = False
            sl = [l for l in lanes if small[l]]
            for i, l in enumerate(sl):
                dfd[ndf + i], dfc[ndf + i] = nd[l], ncol[l]
            if sl: lastv = nd[sl[-1]]
            ndf += len(sl)
        else:
            walked += 1
            for l in lanes:
                vd, vz, vc = nd[l], nz[l], ncol[l]
                def push(v, col):
                    nonlocal ndf, lastv
                    p = ndf
                    if v < lastv:
                        while p > 0 and dfd[p - 1] > v:
                            dfd[p] = dfd[p - 1]; dfc[p] = dfc[p - 1]; p -= 1
                    else:
                        lastv = v
                    dfd[p] = v; dfc[p] = col; ndf += 1
                if rho * abs(vz) <= tol: push(vd, vc)
                elif not have: pd, pz, pcol, have, pmix = vd, vz, vc, True, False
                else:
                    r, c, s, tau = rot_test(pd, pz, vd, vz, tol)
                    if r:
                        rots.append((pcol, vc, c, s))
                        tp = pd * c * c + vd * s * s; tn = pd * s * s + vd * c * c
                        push(tp, pcol)
                        pmix = pmix or ((pcol < n1) != (vc < n1))
                        pd, pz, pcol = tn, tau, vc
                    else:
                        dl[k], wz[k], nc[k] = pd, pz, pcol; tp_[k] = tcode(pcol, pmix); k += 1
                        pd, pz, pcol, pmix = vd, vz, vc, False
    dl[k], wz[k], nc[k] = pd, pz, pcol; tp_[k] = tcode(pcol, pmix); k += 1
    # places: counts of types 0 and 1, then running places per type
    c0 = sum(1 for t in tp_[:k] if t == 0); c1 = sum(1 for t in tp_[:k] if t == 1)
    nxt = [0, c0, c0 + c1]; pos = []
    for t in tp_[:k]:
        pos.append(nxt[t]); nxt[t] += 1
    return dl[:k], wz[:k], nc[:k], dfd[:ndf], dfc[:ndf], rots, walked, pos



def test_chunked_scan_equals_scalar_scan():
    rng = np.random.default_rng(7)
    walked = chunks = 0
    for case in range(600):
        K = int(rng.integers(1, 400))
        kind = case % 5
        d = np.sort(rng.standard_normal(K))
        if kind in (1, 3):  # clusters of near-equal d: rotations
            idx = rng.random(K) < 0.3
            d = np.sort(np.where(idx, np.round(d, 2), d))
            d = np.sort(d + rng.standard_normal(K) * 1e-17 * (rng.random(K) < 0.5))
        z = rng.standard_normal(K)
        if kind in (2, 3):  # many negligible z
            z = np.where(rng.random(K) < 0.4, z * 1e-18, z)
        if case % 10 == 4:  # every entry negligible
            z = z * 1e-30
        cs = list(rng.permutation(K))
        rho = float(rng.uniform(0.1, 2))
        a = scalar(list(d), list(z), cs, rho)
        b = chunked(list(d), list(z), cs, rho)
        walked += b[6]
        chunks += (K + 63) // 64
        assert a[:6] == b[:6], case
        # grouped places: a permutation, types in blocks 0 | 1 | 2, scan order kept within a type
        typ, pos = a[6], b[7]
        assert sorted(pos) == list(range(len(typ)))
        order = [q for _, q in sorted(zip(pos, range(len(typ))))]
        assert [typ[q] for q in order] == sorted(typ)
        for t in (0, 1, 2):
            qs = [q for q in order if typ[q] == t]
            assert qs == sorted(qs)
    assert 0 < walked < chunks  # both ways of resolving a chunk were exercised


def test_column_types_bound_the_rows():
    """After the recorded rotations, a type-0 column of a block-diagonal Z has
    no entries in the second half's rows and a type-2 column none in the
    first half's: what the per-half eigenvector products rely on."""
    rng = np.random.default_rng(11)
    mixed = 0
    for case in range(300):
        K = int(rng.integers(2, 200))
        n1 = (K + 1) // 2
        d = np.sort(np.round(rng.standard_normal(K), 1))  # ties across the halves: rotations
        z = rng.standard_normal(K)
        z = np.where(rng.random(K) < 0.3, z * 1e-18, z)
        cs = list(rng.permutation(K))
        Z = np.zeros((K, K))
        Z[:n1, :n1] = rng.standard_normal((n1, n1))
        Z[n1:, n1:] = rng.standard_normal((K - n1, K - n1))
        dl, wz, nc, dfd, dfc, rots, typ = scalar(list(d), list(z), cs, 1.0, n1)
        for p, q, c, s in rots:  # x' = c x + s y, y' = c y - s x, in order
            x, y = Z[:, p].copy(), Z[:, q].copy()
            Z[:, p], Z[:, q] = c * x + s * y, c * y - s * x
        for col, t in zip(nc, typ):
            if t == 0:
                assert not np.any(Z[n1:, col]), case
            elif t == 2:
                assert not np.any(Z[:n1, col]), case
            else:
                mixed += 1
    assert mixed > 0
