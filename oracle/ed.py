"""Exact diagonalisation oracle for small S=1/2 XXZ lattices.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).

Restates the thermal averages of the reference's "diagonalization" worker
(diag.C:159-293 static/dynamic averages, diag.C:313-488 observables) for
H = sum_b [ Jz Sz_i Sz_j + Jxy/2 (S+_i S-_j + S-_i S+_j) ] on the chain and
square lattices with ALPS bond conventions.  The Hamiltonian is block-diagonal
in total Sz, so each sector is diagonalised densely with numpy.linalg.eigh.
Susceptibilities use the Kubo formula of diag.C:252-293,
chi_O = (1/(Z V)) sum_ij |<i|O|j>|^2 w_ij with w_ij = (e^{-b E_i} - e^{-b E_j})/(E_j - E_i),
and w_ii = b e^{-b E_i}.
"""

import itertools
import math

import numpy as np


def lattice_bonds(kind, L):
    """(src, tgt) lists and gauge with ALPS conventions (chain: bond i = (i, i+1 mod L);
    square: site x*L+y, bond 2s = (s, s+y^), 2s+1 = (s, s+x^))."""
    if kind == "chain":
        N = L
        bonds = [(i, (i + 1) % L) for i in range(L)]
        gauge = [1 if s % 2 == 0 else -1 for s in range(N)]
    elif kind == "square":
        N = L * L
        bonds = []
        for s in range(N):
            x, y = divmod(s, L)
            bonds.append((s, x * L + (y + 1) % L))
            bonds.append((s, ((x + 1) % L) * L + y))
        gauge = [1 if (s // L + s % L) % 2 == 0 else -1 for s in range(N)]
    elif kind == "site":  # one site, no bonds (the reference's "site" lattice)
        N, bonds, gauge = 1, [], [1]
    else:
        raise ValueError(kind)
    return N, bonds, np.array(gauge, dtype=float)


def _sector_states(N, nup):
    states = []
    for ups in itertools.combinations(range(N), nup):
        b = 0
        for u in ups:
            b |= 1 << u
        states.append(b)
    return np.array(sorted(states), dtype=np.int64)


def _sector_hamiltonian(states, bonds, jxy, jz):
    dim = len(states)
    H = np.zeros((dim, dim))
    cols = np.arange(dim)
    for (a, b) in bonds:
        sa = ((states >> a) & 1) - 0.5
        sb = ((states >> b) & 1) - 0.5
        H[cols, cols] += jz * sa * sb
        hop = sa != sb
        t = states[hop] ^ (1 << a) ^ (1 << b)
        rows = np.searchsorted(states, t)
        np.add.at(H, (rows, cols[hop]), 0.5 * jxy)
    return H


def _spin_sectors(N, bonds, jxy, jz, two_s):
    """Spin S = two_s/2: basis states as base-(2S+1) digits (digit q -> m = S - q),
    grouped by total S^z; H = sum_b [Jz Sz Sz + Jxy/2 (S+S- + S-S+)] with
    S+|m> = sqrt(S(S+1) - m(m+1)) |m+1>."""
    S = 0.5 * two_s
    levels = [S - q for q in range(two_s + 1)]
    ms = np.array(list(itertools.product(levels, repeat=N)), dtype=float)[:, ::-1]  # site i = column i
    tot2 = np.rint(2 * ms.sum(axis=1)).astype(int)
    out = []
    for M2 in sorted(set(tot2.tolist())):
        idx = np.nonzero(tot2 == M2)[0]
        conf = ms[idx]
        key = {tuple(np.rint(2 * c).astype(int)): n for n, c in enumerate(conf)}
        dim = len(idx)
        H = np.zeros((dim, dim))
        for n, c in enumerate(conf):
            for (a, b) in bonds:
                H[n, n] += jz * c[a] * c[b]
                for da, db in ((1, -1), (-1, 1)):
                    ma, mb = c[a] + da, c[b] + db
                    if abs(ma) > S + 1e-9 or abs(mb) > S + 1e-9:
                        continue
                    amp = np.sqrt(S * (S + 1) - c[a] * (c[a] + da)) * np.sqrt(S * (S + 1) - c[b] * (c[b] + db))
                    d = c.copy()
                    d[a], d[b] = ma, mb
                    H[key[tuple(np.rint(2 * d).astype(int))], n] += 0.5 * jxy * amp
        out.append((conf, H))
    return out


def diagonalize_spin(kind, L, T, two_s, jxy=1.0, jz=1.0):
    """Thermal averages of the spin-S XXZ model (S = two_s/2, periodic chain or
    square lattice), keyed by the reference's observable names; staggered
    quantities with the lattice's gauge (diag.C:159-293, 313-488 for spin S)"""
    beta = 1.0 / T
    N, bonds, gauge = lattice_bonds(kind, L)
    V = float(N)
    sectors = []
    for conf, H in _spin_sectors(N, bonds, jxy, jz, two_s):
        E, U = np.linalg.eigh(H)
        sectors.append((E, U, conf.sum(axis=1), conf @ gauge))
    e0 = min(sec[0].min() for sec in sectors)
    Z = sum(np.exp(-beta * (sec[0] - e0)).sum() for sec in sectors)
    res = {}
    ene = sum((np.exp(-beta * (E - e0)) * E).sum() for (E, _, _, _) in sectors) / Z
    ene2 = sum((np.exp(-beta * (E - e0)) * E * E).sum() for (E, _, _, _) in sectors) / Z
    res["Energy"] = ene
    res["Energy Density"] = ene / V
    res["Specific Heat"] = beta * beta * (ene2 - ene * ene) / V
    res["Ground State Energy"] = e0
    for name, pick in (("", 2), ("Staggered ", 3)):
        v2 = kub = 0.0
        for sec in sectors:
            E, U, d = sec[0], sec[1], sec[pick]
            w = np.exp(-beta * (E - e0))
            v2 += (w * ((U * U).T @ (d ** 2))).sum()
            A = U.T @ (d[:, None] * U)
            dE = E[:, None] - E[None, :]
            with np.errstate(divide="ignore", invalid="ignore"):
                wij = -(w[:, None] - w[None, :]) / dE
            wij = np.where(np.abs(dE) <= 1e-12, beta * w[:, None] * np.ones_like(dE), wij)
            kub += (A * A * wij).sum()
        res[name + "Magnetization^2"] = v2 / Z
        res[name + "Susceptibility"] = kub / Z / V
    res["Staggered Structure Factor"] = res["Staggered Magnetization^2"] / V
    return res


def diagonalize_transverse(kind, L, T, two_s=1, jxy=1.0, jz=1.0, hx=0.0, h=0.0, d=0.0):
    """Thermal averages of the spin-S XXZ model in a transverse field,
    H = sum_b [Jz Sz Sz + Jxy/2 (S+S- + S-S+)] - hx sum_i S^x_i - h sum_i S^z_i
    + D sum_i (S^z_i)^2 (looper/model_parameter.h:63-68, S^x = (S+ + S-)/2;
    d the single-ion D), on the whole product
    space: S^z is no longer conserved, so the Hamiltonian is diagonalised
    densely in one block and every static average and Kubo susceptibility
    (diag.C:252-293) is taken in its eigenbasis.  Sizes: (2S+1)^N states."""
    beta = 1.0 / T
    N, bonds, gauge = lattice_bonds(kind, L)
    V = float(N)
    S = 0.5 * two_s
    levels = [S - q for q in range(two_s + 1)]
    ms = np.array(list(itertools.product(levels, repeat=N)), dtype=float)[:, ::-1]
    dim = len(ms)
    key = {tuple(np.rint(2 * c).astype(int)): n for n, c in enumerate(ms)}
    H = np.zeros((dim, dim))
    dion = d  # (d names a basis state below)
    for n, c in enumerate(ms):
        H[n, n] += dion * (c * c).sum() - h * c.sum()
        for (a, b) in bonds:
            H[n, n] += jz * c[a] * c[b]
            for da, db in ((1, -1), (-1, 1)):
                ma, mb = c[a] + da, c[b] + db
                if abs(ma) > S + 1e-9 or abs(mb) > S + 1e-9:
                    continue
                amp = np.sqrt(S * (S + 1) - c[a] * (c[a] + da)) * np.sqrt(S * (S + 1) - c[b] * (c[b] + db))
                d = c.copy()
                d[a], d[b] = ma, mb
                H[key[tuple(np.rint(2 * d).astype(int))], n] += 0.5 * jxy * amp
        for a in range(N):  # -hx S^x_a = -hx/2 (S+_a + S-_a)
            for da in (1, -1):
                ma = c[a] + da
                if abs(ma) > S + 1e-9:
                    continue
                d = c.copy()
                d[a] = ma
                H[key[tuple(np.rint(2 * d).astype(int))], n] -= 0.5 * hx * np.sqrt(S * (S + 1) - c[a] * (c[a] + da))
    E, U = np.linalg.eigh(H)
    e0 = E.min()
    w = np.exp(-beta * (E - e0))
    Z = w.sum()
    res = {}
    ene = (w * E).sum() / Z
    ene2 = (w * E * E).sum() / Z
    res["Energy"] = ene
    res["Energy Density"] = ene / V
    res["Specific Heat"] = beta * beta * (ene2 - ene * ene) / V
    res["Ground State Energy"] = e0
    dE = E[:, None] - E[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        wij = -(w[:, None] - w[None, :]) / dE
    wij = np.where(np.abs(dE) <= 1e-12, beta * w[:, None] * np.ones_like(dE), wij)
    for name, d in (("", ms.sum(axis=1)), ("Staggered ", ms @ gauge)):
        A = U.T @ (d[:, None] * U)        # <i|O|j> in the eigenbasis
        A2 = U.T @ ((d ** 2)[:, None] * U)
        A4 = U.T @ ((d ** 4)[:, None] * U)
        res[name + "Magnetization^2"] = (w * np.diag(A2)).sum() / Z
        res[name + "Magnetization^4"] = (w * np.diag(A4)).sum() / Z
        res[name + "Susceptibility"] = (A * A * wij).sum() / Z / V
    res["Staggered Structure Factor"] = res["Staggered Magnetization^2"] / V
    return res


def diagonalize_spin1_chain(L, T, jxy=1.0, jz=1.0):
    """Thermal averages of the S=1 XXZ chain (periodic, L sites)"""
    return diagonalize_spin("chain", L, T, 2, jxy, jz)


def diagonalize(kind, L, T, jxy=1.0, jz=1.0, h=0.0):
    """Thermal averages keyed by the reference's observable names; h is a
    longitudinal field (site term -h S^z), diagonal in every S^z sector."""
    beta = 1.0 / T
    N, bonds, gauge = lattice_bonds(kind, L)
    V = float(N)
    sectors = []
    for nup in range(N + 1):
        states = _sector_states(N, nup)
        H = _sector_hamiltonian(states, bonds, jxy, jz)
        E, U = np.linalg.eigh(H)
        E = E - h * (nup - 0.5 * N)
        sz = np.array([[0.5 if (int(s) >> i) & 1 else -0.5 for i in range(N)] for s in states])
        mu = sz.sum(axis=1)
        ms = sz @ gauge
        sectors.append((E, U, mu, ms))
    e0 = min(sec[0].min() for sec in sectors)
    Z = sum(np.exp(-beta * (sec[0] - e0)).sum() for sec in sectors)
    res = {}
    ene = sum((np.exp(-beta * (E - e0)) * E).sum() for (E, _, _, _) in sectors) / Z
    ene2 = sum((np.exp(-beta * (E - e0)) * E * E).sum() for (E, _, _, _) in sectors) / Z
    res["Energy"] = ene
    res["Energy Density"] = ene / V
    res["Specific Heat"] = beta * beta * (ene2 - ene * ene) / V
    res["Ground State Energy"] = e0

    def static(diag_of):
        v2 = v4 = va = 0.0
        for (E, U, mu, ms) in sectors:
            d = diag_of(mu, ms)
            w = np.exp(-beta * (E - e0))
            p = U * U  # |<n|basis>|^2, columns = eigenvectors
            v2 += (w * (p.T @ (d ** 2))).sum()
            v4 += (w * (p.T @ (d ** 4))).sum()
            va += (w * (p.T @ np.abs(d))).sum()
        return v2 / Z, v4 / Z, va / Z

    def kubo(diag_of):
        tot = 0.0
        for (E, U, mu, ms) in sectors:
            d = diag_of(mu, ms)
            A = U.T @ (d[:, None] * U)
            w = np.exp(-beta * (E - e0))
            dE = E[:, None] - E[None, :]
            with np.errstate(divide="ignore", invalid="ignore"):
                wij = -(w[:, None] - w[None, :]) / dE
            same = np.abs(dE) <= 1e-12
            wij = np.where(same, beta * w[:, None] * np.ones_like(dE), wij)
            tot += (A * A * wij).sum()
        return tot / Z / V

    res["Magnetization"] = sum((np.exp(-beta * (E - e0)) * mu[0]).sum() for (E, _, mu, _) in sectors) / Z
    m2, m4, ma = static(lambda mu, ms: mu)
    res["Magnetization^2"] = m2
    res["Magnetization^4"] = m4
    res["|Magnetization|"] = ma
    res["Susceptibility"] = kubo(lambda mu, ms: mu)
    s2, s4, sa = static(lambda mu, ms: ms)
    res["Staggered Magnetization^2"] = s2
    res["Staggered Magnetization^4"] = s4
    res["|Staggered Magnetization|"] = sa
    res["Staggered Susceptibility"] = kubo(lambda mu, ms: ms)
    res["Staggered Structure Factor"] = s2 / V
    # equal-time correlations <Sz_i Sz_j> (diagonal in the basis): basis-state
    # probabilities of each sector times products of the site spins
    C = np.zeros((N, N))
    for (E, U, mu, ms), nup in zip(sectors, range(N + 1)):
        w = np.exp(-beta * (E - e0))
        pb = (U * U) @ w
        states = _sector_states(N, nup)
        sz = np.array([[0.5 if (int(st) >> i) & 1 else -0.5 for i in range(N)] for st in states]).reshape(len(states), N)
        C += sz.T @ (pb[:, None] * sz)
    C /= Z
    nx = 1 if kind == "chain" else L
    ny = L
    corr = np.zeros(N)
    scorr = np.zeros(N)
    for i in range(N):
        xi, yi = divmod(i, ny)
        for j in range(N):
            xj, yj = divmod(j, ny)
            d = ((xj - xi) % nx) * ny + (yj - yi) % ny
            corr[d] += C[i, j] / N
            scorr[d] += gauge[i] * gauge[j] * C[i, j] / N
    res["Spin Correlations"] = corr
    res["Staggered Spin Correlations"] = scorr
    res["Spin Correlations Matrix"] = C
    sfac = np.zeros(N)
    for a in range(nx):
        for b in range(ny):
            k1 = 2 * math.pi * a / nx
            k2 = 2 * math.pi * b / ny
            val = 0.0
            for i in range(N):
                xi, yi = divmod(i, ny)
                for j in range(N):
                    xj, yj = divmod(j, ny)
                    val += C[i, j] * math.cos(k1 * (xi - xj) + k2 * (yi - yj))
            sfac[a * ny + b] = val / N
    res["Spin Structure Factor"] = sfac
    return res


if __name__ == "__main__":
    r = diagonalize("chain", 4, 0.25)
    for k in sorted(r):
        print(k, r[k])
