"""CPU tests: pin the oracle (oracle/loop_oracle.cpp, oracle/ed.py) against
the reference's own fixtures, the reference built from its sources where that
is possible, and what the mathematics fixes."""

import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))


# ---------------------------------------------------------------------------
# Philox4x32-10 known answers (Salmon et al., SC'11, Random123 test vectors)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, out):
    assert list(O.philox(ctr, key)) == out


def test_det_log_against_libm():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(20000), [1.0, 0.5, 2.0 ** -53, 0.7071067811865476, 0.70710678118654746,
                                            1 - 2.0 ** -53, 2.0 ** -30 * 3]])
    x = x[x > 0]
    y = O.det_log(x)
    ref = np.log(x)
    ulp = np.spacing(np.abs(ref))
    assert np.all(np.abs(y - ref) <= 4 * ulp + 1e-300)
    assert O.det_log(np.array([1.0]))[0] == 0.0


def test_exponential_gaps_distribution():
    # gaps from Philox draws must be exponential with the requested rate
    lam = 37.0
    scale = 2.0 ** 52 / lam
    gaps = []
    for k in range(20000):
        x = O.philox([0, O.STAGE_DIAG, 7, k], [23125, 0])
        g = np.zeros(1, np.uint64)
        O.lib.or_exp_gap(x.ctypes.data, scale, g.ctypes.data)
        gaps.append(float(g[0]) * 2.0 ** -52)
    gaps = np.array(gaps)
    assert abs(gaps.mean() * lam - 1) < 0.03
    # Kolmogorov-Smirnov against 1 - exp(-lam t)
    s = np.sort(gaps)
    cdf = 1 - np.exp(-lam * s)
    n = len(s)
    d = max(np.max(np.arange(1, n + 1) / n - cdf), np.max(cdf - np.arange(n) / n))
    assert d < 1.63 / math.sqrt(n)  # 1% level


# ---------------------------------------------------------------------------
# lattice conventions (test/spinmodel_helper.op, test/virtual_lattice.op)
# ---------------------------------------------------------------------------
def test_chain_lattice_matches_alps_fixture():
    N, Nb, src, tgt, gauge = O.lattice(O.CHAIN, 4)
    assert [[int(a), int(b)] for a, b in zip(src, tgt)] == FX["lattices"]["chain4_bonds"]


def test_square_lattice_matches_alps_fixture():
    N, Nb, src, tgt, gauge = O.lattice(O.SQUARE, 2)
    L = 2
    coords = FX["lattices"]["square2_coords"]
    for s, (x, y) in enumerate(coords):
        assert s == x * L + y
    edges = FX["lattices"]["square2_edges"]
    assert Nb == len(edges)
    for b, (a, t, vx, vy) in enumerate(edges):
        assert (src[b], tgt[b]) == (a, t)
        # bond 2s is the (0,1) direction, bond 2s+1 the (1,0) direction
        assert (vx, vy) == ((0, 1) if b % 2 == 0 else (1, 0))
    assert list(gauge) == FX["lattices"]["square2_gauge"]


# ---------------------------------------------------------------------------
# cut rules and weights (test/compatibility.op, test/weight.op)
# ---------------------------------------------------------------------------
def test_compatibility_truth_table():
    # is_compatible(g, c0, c1) = (g & 1) ^ c0 ^ c1: candidates of graph g are
    # accepted by the philox-mode sweep exactly on these spin pairs
    for g, c0, c1, want in FX["compatibility"]:
        assert ((g & 1) ^ c0 ^ c1) == want


@pytest.mark.parametrize("row", FX["weights"], ids=lambda r: "jxy%g_jz%g_a%g" % (r["jxy"], r["jz"], r["force_scatter"]))
def test_xxz_weights_match_fixture(row):
    v, off, sign = O.xxz_weights(row["jxy"], row["jz"], row["force_scatter"])
    np.testing.assert_allclose(v, row["v"], rtol=2e-5, atol=1e-6)
    assert off == pytest.approx(row["offset"], rel=2e-5, abs=1e-6)
    assert sign == row["sign"]
    # loop equations of weight_impl.h:114-118 (sign-free bond: jxy>0 gives sign -1)
    assert -off + v[0] + v[2] == pytest.approx(row["jz"] / 4, abs=1e-12)
    assert -off + v[1] + v[3] == pytest.approx(-row["jz"] / 4, abs=1e-12)
    assert v[0] + v[1] == pytest.approx(abs(row["jxy"]) / 2, abs=1e-12)


# ---------------------------------------------------------------------------
# union-find (test/union_find.op, looper/union_find.h built in oracle/_ref)
# ---------------------------------------------------------------------------
def _uf(variant):
    f = FX["union_find"]
    n = f["n"]
    pairs = np.array(f["pairs"], np.int32).ravel()
    par = np.zeros(n, np.int32)
    root = np.zeros(n, np.int32)
    w = np.zeros(n, np.int32)
    O.lib.or_uf_pairs(n, len(f["pairs"]), pairs.ctypes.data, variant, par.ctypes.data, root.ctypes.data,
                      w.ctypes.data)
    return par, root, w


def test_union_find_matches_reference_output():
    f = FX["union_find"]
    par, root, w = _uf(0)  # weighted union with path halving, union_find.h:242-284
    assert list(par) == f["parent"]
    assert list(root) == f["root"]
    assert list(w) == f["size"]


def test_union_find_compress_variant_same_partition():
    f = FX["union_find"]
    par, root, w = _uf(1)
    assert list(root) == f["root"]
    assert list(w) == f["size"]


def test_union_find_against_reference_header_build():
    exe = os.path.join(O.ORACLE_DIR, "_ref", "looper_uf")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    f = FX["union_find"]
    inp = "%d %d\n" % (f["n"], len(f["pairs"])) + "".join("%d %d\n" % tuple(p) for p in f["pairs"])
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    ref = [tuple(map(int, l.split())) for l in out if l.strip()]
    par, root, w = _uf(0)
    assert [r[0] for r in ref] == list(par)
    assert [r[1] for r in ref] == list(root)
    assert [r[2] for r in ref] == list(w)


@pytest.mark.parametrize("kind,L,nops,graphs,grid", [
    (O.CHAIN, 2, 0, (0,), False), (O.CHAIN, 4, 9, (0, 1, 2, 3), False), (O.CHAIN, 16, 300, (0, 1), False),
    (O.CHAIN, 16, 300, (0, 1, 2, 3), True), (O.SQUARE, 2, 30, (0, 1, 2, 3), False),
    (O.SQUARE, 6, 800, (0, 1, 2, 3), True), (O.CHAIN, 64, 3000, (0,), False),
])
def test_looper_fragments_equal_leg_graph_bfs(kind, L, nops, graphs, grid):
    # looper's reconnect/unify semantics (graph_impl.h:277-295, path_integral.C
    # :455-588) give the same partition as a breadth-first search on the leg graph
    rng = np.random.default_rng(7 + nops + L)
    spins, bo, t, ty = O.random_config(kind, L, nops, rng, graphs, grid)
    a = O.looper_labels(kind, L, bo, t, ty)
    b = O.bfs_labels(kind, L, bo, t, ty)
    np.testing.assert_array_equal(a, b)


def test_bfs_labels_tiny_hand_case():
    # chain L=2, one g=0 operator on bond 0 at tau=0.5: legs l0,l1 joined,
    # u0,u1 joined; world lines: boundary 0 - l0, u0 - boundary 0, boundary 1
    # - l1, u1 - boundary 1  -> everything is one cluster
    bo = np.array([0, 1, 1], np.int32)
    t = np.array([0.5])
    ty = np.array([0], np.int32)
    lab = O.bfs_labels(O.CHAIN, 2, bo, t, ty)
    assert list(lab) == [0] * 6
    # cross graph: l0-u1, l1-u0 and the world lines close each site on itself
    # through the other: boundary0-l0-u1-boundary1-l1-u0-boundary0 -> one cluster
    lab = O.bfs_labels(O.CHAIN, 2, bo, t, np.array([1 << 2], np.int32))
    assert list(lab) == [0] * 6
    # two g=0 operators on bond 0: the segment between them on each site forms
    # one loop with the other site's segment
    bo = np.array([0, 2, 2], np.int32)
    t = np.array([0.25, 0.75])
    lab = O.bfs_labels(O.CHAIN, 2, bo, t, np.array([0, 0], np.int32))
    # elements: 0,1 boundaries; op0 legs 2..5, op1 legs 6..9
    # upper legs of op0 (4,5) and lower legs of op1 (6,7) form the inner loop
    assert lab[4] == lab[5] == lab[6] == lab[7] == 4
    assert lab[0] == lab[1] == lab[2] == lab[3] == lab[8] == lab[9] == 0


# ---------------------------------------------------------------------------
# standalone kernel: digit-for-digit (standalone/loop.op)
# ---------------------------------------------------------------------------
def test_standalone_mode_reproduces_loop_op():
    import ctypes
    m = (ctypes.c_double * 5)()
    e = (ctypes.c_double * 5)()
    O.lib.or_standalone_run(8, 0.2, 65536, 8192, m, e)
    fx = FX["standalone_loop_op"]
    names = ["Number of Clusters", "Energy Density", "Uniform Susceptibility", "Staggered Magnetization^2",
             "Staggered Susceptibility"]
    for i, nm in enumerate(names):
        assert "%.6g" % m[i] == "%.6g" % fx[nm][0], nm
        assert "%.6g" % e[i] == "%.6g" % fx[nm][1], nm


def test_standalone_mode_against_reference_build():
    exe = os.path.join(O.ORACLE_DIR, "_ref", "standalone_loop")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    import ctypes
    out = subprocess.run([exe, "-l", "16", "-t", "0.1", "-n", "4096"], capture_output=True, text=True,
                         check=True).stdout
    m = (ctypes.c_double * 5)()
    e = (ctypes.c_double * 5)()
    O.lib.or_standalone_run(16, 0.1, 4096, 512, m, e)
    vals = [l.split("=")[1].split("+-") for l in out.splitlines() if "+-" in l]
    for i in range(5):
        assert "%.6g" % m[i] == "%.6g" % float(vals[i][0])


# ---------------------------------------------------------------------------
# exact diagonalisation oracle (loop.op ED block, diag.C)
# ---------------------------------------------------------------------------
def test_ed_reproduces_reference_diagonalization():
    import ed
    ref = FX["loop_op_chain4"][0]
    assert ref["algorithm"] == "diagonalization"
    r = ed.diagonalize("chain", 4, 0.25)
    for k in ["Energy", "Energy Density", "Specific Heat", "Magnetization^2", "Magnetization^4",
              "Staggered Magnetization^2", "Staggered Magnetization^4", "Staggered Susceptibility",
              "Susceptibility", "|Magnetization|", "|Staggered Magnetization|"]:
        assert r[k] == pytest.approx(ref["results"][k][0], rel=2e-5), k


def test_ed_closed_forms():
    import ed
    # two sites joined by two bonds (chain L=2): H = 2J S1.S2, singlet -3J/2, triplet +J/2
    r = ed.diagonalize("chain", 2, 1.0)
    b = 1.0
    Z = math.exp(1.5 * b) + 3 * math.exp(-0.5 * b)
    e = (-1.5 * math.exp(1.5 * b) + 3 * 0.5 * math.exp(-0.5 * b)) / Z
    assert r["Energy"] == pytest.approx(e, rel=1e-12)
    # M^2: triplet m=+-1 contribute 1 each
    assert r["Magnetization^2"] == pytest.approx(2 * math.exp(-0.5 * b) / Z, rel=1e-12)
    # Bethe ansatz ground-state energy per site for the L->inf chain is 1/4 - ln 2;
    # L=12 at T=0.05 lies above it and within finite-size range
    r = ed.diagonalize("chain", 12, 0.05)
    assert -0.4535 < r["Energy Density"] < 0.25 - math.log(2) + 0.01


# ---------------------------------------------------------------------------
# Monte Carlo modes against ED and the reference's QMC output
# ---------------------------------------------------------------------------
def _run(chain, therm, n):
    chain.sweep(therm)
    return chain.sweep(n)


OBS = ["Energy Density", "Susceptibility", "Staggered Magnetization^2", "Staggered Susceptibility"]


@pytest.mark.parametrize("mode", ["looper", "philox"])
@pytest.mark.parametrize("kind,L,T,jz", [(O.CHAIN, 4, 0.25, 1.0), (O.CHAIN, 12, 0.1, 1.0), (O.CHAIN, 8, 0.2, 0.6),
                                         (O.CHAIN, 6, 0.3, 1.8), (O.SQUARE, 2, 0.5, 1.0)])
def test_monte_carlo_modes_match_ed(mode, kind, L, T, jz):
    import ed
    beta = 1 / T
    V = L if kind == O.CHAIN else L * L
    C = O.LooperChain if mode == "looper" else O.PhiloxChain
    ch = C(kind, L, beta, 1.0, jz, seed=31)
    rec = _run(ch, 2000, 30000)
    obs = stats.observables(rec, beta, V, 1 if kind == O.CHAIN else 2)
    ref = ed.diagonalize("chain" if kind == O.CHAIN else "square", L, T, 1.0, jz)
    refv = {"Energy Density": ref["Energy"] / V, "Susceptibility": ref["Susceptibility"],
            "Staggered Magnetization^2": ref["Staggered Magnetization^2"],
            "Staggered Susceptibility": ref["Staggered Susceptibility"]}
    for k in OBS:
        m, e = stats.mean_err(obs[k])
        assert stats.sigmas(m, e, refv[k]) < 4.0, (k, m, e, refv[k])


def test_looper_mode_matches_reference_qmc_output():
    # loop.op block 2: "loop; path integral", chain L=4, T=0.25, SWEEPS=4096
    ref = FX["loop_op_chain4"][1]
    assert ref["algorithm"] == "loop; path integral"
    ch = O.LooperChain(O.CHAIN, 4, 4.0, 1.0, 1.0, seed=23125)
    rec = _run(ch, 512, 40000)
    obs = stats.observables(rec, 4.0, 4, 1)
    for k in ["Energy Density", "Staggered Magnetization^2", "Staggered Susceptibility", "Susceptibility"]:
        m, e = stats.mean_err(obs[k])
        rm, re_ = ref["results"][k]
        assert stats.sigmas(m, e, rm, re_) < 3.0, (k, m, e, rm, re_)


def test_planted_field_mix_up_is_caught():
    # a staggered susceptibility built from the tau = 0 field (smag2) instead of
    # the tau-integral (smag) -- a plausible transposed-field mistake -- must
    # fail the ED comparison (the staggered moment does not commute with H)
    import ed
    ch = O.LooperChain(O.CHAIN, 12, 10.0, seed=5)
    rec = _run(ch, 1000, 20000)
    wrong = 10.0 * rec[:, 12] / 12
    m, e = stats.mean_err(wrong)
    ref = ed.diagonalize("chain", 12, 0.1)["Staggered Susceptibility"]
    assert stats.sigmas(m, e, ref) > 4.0
    right = 10.0 * rec[:, 16] / 12
    m, e = stats.mean_err(right)
    assert stats.sigmas(m, e, ref) < 4.0


# ---------------------------------------------------------------------------
# longitudinal field (looper's FIELD path, path_integral.C:798, 852-858)
# ---------------------------------------------------------------------------

def test_det_exp_and_tanh_against_libm():
    x = np.concatenate([np.linspace(-40, 40, 4001), [-700.0, -745.5, 0.0, 1e-300, -1e-12, 709.0]])
    e = O.det_exp(x)
    ref = np.exp(np.minimum(x, 709.0))
    ok = ref > 0
    np.testing.assert_allclose(e[ok], ref[ok], rtol=4e-16 * 4)
    t = O.det_tanh(np.linspace(-25, 25, 5001))
    np.testing.assert_allclose(t, np.tanh(np.linspace(-25, 25, 5001)), rtol=0, atol=4e-16)


def test_ed_field_two_site_closed_form():
    import ed
    # chain L=2: two bonds between the sites, singlet -3J/2, triplet J/2 - h m
    b, J, h = 1.3, 1.0, 0.45
    lev = [(-1.5 * J, 0), (0.5 * J - h, 1), (0.5 * J, 0), (0.5 * J + h, -1)]
    w = np.array([np.exp(-b * e) for e, _ in lev])
    m = np.array([mm for _, mm in lev], float)
    en = np.array([e for e, _ in lev])
    r = ed.diagonalize("chain", 2, 1 / b, 1.0, 1.0, h)
    assert r["Magnetization"] == pytest.approx((w * m).sum() / w.sum(), rel=1e-12)
    assert r["Energy"] == pytest.approx((w * en).sum() / w.sum(), rel=1e-12)
    assert r["Susceptibility"] == pytest.approx(b * (w * m * m).sum() / w.sum() / 2, rel=1e-12)


FIELD_CASES = [
    # kind, L, T, jxy, jz, h
    (O.CHAIN, 8, 0.5, 1.0, 1.0, 0.3),
    (O.CHAIN, 6, 0.3, 1.0, 0.5, 0.8),
    (O.SQUARE, 2, 0.7, 1.0, 1.0, 0.6),
]


def _field_observables(rec, beta, V):
    return {"Energy Density": rec[:, O.R_ENERGY] / V, "Magnetization": rec[:, O.R_UMAG0],
            "Susceptibility": beta * rec[:, O.R_UMAGA] ** 2 / V}


@pytest.mark.parametrize("mode", ["looper", "philox"])
@pytest.mark.parametrize("case", FIELD_CASES, ids=lambda c: "k%d_L%d_T%g_jz%g_h%g" % (c[0], c[1], c[2], c[4], c[5]))
def test_field_modes_match_exact_diagonalisation(case, mode):
    import ed
    kind, L, T, jxy, jz, h = case
    beta = 1 / T
    if mode == "looper":
        c = O.LooperChain(kind, L, beta, jxy, jz, seed=1234, field=h)
    else:
        c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=1234, chain_id=3, field=h)
    c.sweep(2000)
    rec = c.sweep(150000)  # the field lengthens autocorrelations: long runs for honest binning errors
    V = L if kind == O.CHAIN else L * L
    ref = ed.diagonalize("chain" if kind == O.CHAIN else "square", L, T, jxy, jz, h)
    ref_obs = {"Energy Density": ref["Energy"] / V, "Magnetization": ref["Magnetization"],
               "Susceptibility": ref["Susceptibility"]}
    for k, x in _field_observables(rec, beta, V).items():
        m, err = x.mean(), stats.binning_error(x)
        assert stats.sigmas(m, err, ref_obs[k]) < 4.0, (mode, case, k, m, err, ref_obs[k])
