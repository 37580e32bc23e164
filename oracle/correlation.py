"""CPU restatement of the reference's correlation estimators.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).

looper::correlation::normal_estimator::collector::commit
(correlation.h:230-296) on one tau = 0 configuration, as plain pair and
phase loops over the sites (the reference's own O(N^2) form):

* Spin Correlations[d] = 0.25 * mult_d^-1 * sum over ordered pairs (i, j) at
  distance class d of s_i s_j, s = 1 - 2 spin (correlation.h:254-267); the
  distance class of a periodic lattice is the displacement j - i, one class
  per displacement, each of multiplicity N;
* Staggered Spin Correlations[d]: the same with g_i g_j (gauge, +1 on the
  sublattice of site 0);
* with INITIAL_SITE = o: 0.25 s_o s_r per site r (correlation.h:241-252);
* Spin Structure Factor[k] = |sum_s (0.5 - spin_s) e^(i k.r_s)|^2 / V over
  the lattice momenta k = 2 pi n / L folded into (-pi, pi]
  (correlation.h:280-291).

Site s of the square lattice is (x, y) = divmod(s, L); the chain is x = 0.
"""

import math

import numpy as np


def geometry(kind, L):
    nx = 1 if kind == "chain" else L
    return nx, L, nx * L


def momentum(n, L):
    k = 2.0 * math.pi * n / L
    return k - 2.0 * math.pi if k > math.pi + 1e-9 else k


def correlations(kind, L, spins, origin=None):
    """(ucorr, scorr, sfac) of one configuration, N values each."""
    nx, ny, N = geometry(kind, L)
    spins = [int(v) & 1 for v in spins]
    sig = [1 - 2 * v for v in spins]
    gauge = [1 if (s // ny + s % ny) % 2 == 0 else -1 for s in range(N)]
    u = np.zeros(N)
    st = np.zeros(N)
    if origin is not None:
        for r in range(N):
            u[r] += sig[origin] * sig[r]
            st[r] += gauge[origin] * sig[origin] * gauge[r] * sig[r]
        u *= 0.25
        st *= 0.25
    else:
        for i in range(N):
            xi, yi = divmod(i, ny)
            for j in range(N):
                xj, yj = divmod(j, ny)
                d = ((xj - xi) % nx) * ny + (yj - yi) % ny
                u[d] += sig[i] * sig[j]
                st[d] += gauge[i] * sig[i] * gauge[j] * sig[j]
        u *= 0.25 / N
        st *= 0.25 / N
    f = np.zeros(N)
    for a in range(nx):
        for b in range(ny):
            k1, k2 = momentum(a, nx), momentum(b, ny)
            val = 0j
            for s in range(N):
                x, y = divmod(s, ny)
                val += (0.5 - spins[s]) * complex(math.cos(k1 * x + k2 * y), math.sin(k1 * x + k2 * y))
            f[a * ny + b] = abs(val) ** 2 / N
    return u, st, f


def labels(kind, L, origin=None):
    """distance and momentum labels in the reference's output format
    ("( 0 ) -- ( 1 )", "( 1.5708 )"; square: "( x,y )")."""
    nx, ny, N = geometry(kind, L)

    def coord(s):
        return "( %d )" % s if nx == 1 else "( %d,%d )" % divmod(s, ny)

    def num(v):
        return "%.6g" % v

    o = coord(0 if origin is None else origin)
    dist = [o + " -- " + coord(r) for r in range(N)]
    mom = []
    for a in range(nx):
        for b in range(ny):
            if nx == 1:
                mom.append("( %s )" % num(momentum(b, ny)))
            else:
                mom.append("( %s,%s )" % (num(momentum(a, nx)), num(momentum(b, ny))))
    return dist, mom
