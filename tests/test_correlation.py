"""Correlation estimators (correlation.h:59-296): the oracle restatement
pinned by closed forms, by the reference's own output (loop.op, chain L=4,
T=0.5, MEASURE[Correlations] / MEASURE[Structure Factor]) through exact
diagonalisation, and by sum rules; then the library's estimator
(lg_spin_correlations) against the oracle on random configurations, and the
worker's measurements on the GPU against exact diagonalisation."""

import json
import math
import os

import numpy as np
import pytest

import correlation as C
import ed
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))["loop_op_chain4_correlations"]


def parity(kind, L):
    nx, ny, N = C.geometry(kind, L)
    return np.array([(-1) ** (d // ny + d % ny) for d in range(N)], float)


@pytest.mark.parametrize("kind,L", [("chain", 4), ("chain", 6), ("square", 2), ("square", 4)])
def test_oracle_closed_forms(kind, L):
    nx, ny, N = C.geometry(kind, L)
    g = parity(kind, L)
    # all up: every pair agrees; only k = 0 survives, |N/2|^2 / N
    u, s, f = C.correlations(kind, L, [0] * N)
    np.testing.assert_allclose(u, 0.25)
    np.testing.assert_allclose(s, 0.25 * g)
    np.testing.assert_allclose(f, np.where(np.arange(N) == 0, N / 4.0, 0.0), atol=1e-12)
    # Neel state: s_i s_(i+d) = (-1)^(dx+dy); only k = (pi, pi) survives
    neel = [(x + y) % 2 for x in range(nx) for y in range(ny)]
    u, s, f = C.correlations(kind, L, neel)
    np.testing.assert_allclose(u, 0.25 * g)
    np.testing.assert_allclose(s, 0.25)
    kpi = (nx // 2) * ny + ny // 2 if nx > 1 else ny // 2
    np.testing.assert_allclose(f, np.where(np.arange(N) == kpi, N / 4.0, 0.0), atol=1e-12)
    # one flipped spin at site r0 in the all-up state: N - 4 agreeing pairs
    # minus disagreeing ones for every d != 0; S(k) = |N/2 delta_k0 - e^(i k r0)|^2 / N
    r0 = N - 1
    sp = [0] * N
    sp[r0] = 1
    u, s, f = C.correlations(kind, L, sp)
    exp_u = np.full(N, 0.25 * (N - 4) / N)
    exp_u[0] = 0.25
    np.testing.assert_allclose(u, exp_u)
    x0, y0 = divmod(r0, ny)
    for a in range(nx):
        for b in range(ny):
            k1, k2 = C.momentum(a, nx), C.momentum(b, ny)
            z = (N / 2.0 if (a, b) == (0, 0) else 0.0) - complex(math.cos(k1 * x0 + k2 * y0), math.sin(k1 * x0 + k2 * y0))
            assert f[a * ny + b] == pytest.approx(abs(z) ** 2 / N, abs=1e-12)
    # INITIAL_SITE: the origin's row
    u, s, f = C.correlations(kind, L, sp, origin=r0)
    exp = np.full(N, -0.25)
    exp[r0] = 0.25
    np.testing.assert_allclose(u, exp)


def test_oracle_labels_are_the_reference_output_labels():
    dist, mom = C.labels("chain", 4)
    assert dist == ["( 0 ) -- ( %d )" % d for d in range(4)]
    assert mom == ["( 0 )", "( 1.5708 )", "( 3.14159 )", "( -1.5708 )"]   # loop.op order (n = 0..L-1)
    names = {k.split("[")[0] for k in FX["results"] if "[" in k}
    assert {"Spin Correlations", "Staggered Spin Correlations", "Spin Structure Factor"} <= names
    for d in dist:
        assert "Spin Correlations[%s]" % d in FX["results"]
        assert "Staggered Spin Correlations[%s]" % d in FX["results"]
    for k in mom:
        assert "Spin Structure Factor[%s]" % k in FX["results"]
    dist, mom = C.labels("square", 2, origin=3)
    assert dist[0] == "( 1,1 ) -- ( 0,0 )" and mom[3] == "( 3.14159,3.14159 )"


def test_exact_diagonalisation_matches_reference_correlation_output():
    p = FX["params"]
    assert (p["LATTICE"], p["L"], p["T"], p["local_S"]) == ("chain lattice", "4", "0.5", "1/2")
    ref = ed.diagonalize("chain", 4, 0.5)
    dist, mom = C.labels("chain", 4)
    for name, vals, labels in [("Spin Correlations", ref["Spin Correlations"], dist),
                               ("Staggered Spin Correlations", ref["Staggered Spin Correlations"], dist),
                               ("Spin Structure Factor", ref["Spin Structure Factor"], mom)]:
        for lab, v in zip(labels, vals):
            m, e = FX["results"]["%s[%s]" % (name, lab)]
            if e == 0:
                assert m == pytest.approx(v, abs=1e-12)
            else:
                assert stats.sigmas(m, e, v) < 4.0, (name, lab, m, e, v)


@pytest.mark.parametrize("kind,L,T", [("chain", 4, 0.5), ("chain", 6, 0.3), ("square", 2, 0.7)])
def test_exact_diagonalisation_sum_rules(kind, L, T):
    nx, ny, N = C.geometry(kind, L)
    r = ed.diagonalize(kind, L, T)
    u, st, f = r["Spin Correlations"], r["Staggered Spin Correlations"], r["Spin Structure Factor"]
    assert u[0] == pytest.approx(0.25)
    assert u.sum() == pytest.approx(r["Magnetization^2"] / N)
    assert st.sum() == pytest.approx(r["Staggered Magnetization^2"] / N)
    assert f[0] == pytest.approx(r["Magnetization^2"] / N)
    kpi = (nx // 2) * ny + ny // 2 if nx > 1 else ny // 2
    assert f[kpi] == pytest.approx(r["Staggered Magnetization^2"] / N)
    # S(k) is the Fourier transform of the correlations by displacement
    for a in range(nx):
        for b in range(ny):
            k1, k2 = 2 * math.pi * a / nx, 2 * math.pi * b / ny
            v = sum(u[d] * math.cos(k1 * (d // ny) + k2 * (d % ny)) for d in range(N))
            assert f[a * ny + b] == pytest.approx(v, abs=1e-12)


@pytest.mark.parametrize("kind,L", [("chain", 2), ("chain", 4), ("chain", 10), ("chain", 70),
                                    ("square", 2), ("square", 4), ("square", 6)])
def test_library_estimator_matches_oracle(kind, L):
    import loopgpu as lg
    rng = np.random.default_rng(7 + L)
    nx, ny, N = C.geometry(kind, L)
    k = lg.CHAIN if kind == "chain" else lg.SQUARE
    for trial in range(3):
        sp = rng.integers(0, 2, N).astype(np.uint8)
        for origin in (None, int(rng.integers(0, N))):
            u, s, f = lg.spin_correlations(k, L, sp, -1 if origin is None else origin)
            eu, es, ef = C.correlations(kind, L, sp, origin)
            np.testing.assert_allclose(u, eu, rtol=0, atol=1e-15)   # exact pair counts
            np.testing.assert_allclose(s, es, rtol=0, atol=1e-15)
            np.testing.assert_allclose(f, ef, rtol=1e-12, atol=1e-12)


def test_library_estimator_argument_errors():
    import loopgpu as lg
    with pytest.raises(lg.LoopGPUError):
        lg.spin_correlations(lg.CHAIN, 5, np.zeros(5, np.uint8))
    with pytest.raises(lg.LoopGPUError):
        lg.spin_correlations(lg.SQUARE, 4, np.zeros(16, np.uint8), 16)


def _worker_vectors(w, kind, L, origin=None):
    dist, mom = C.labels(kind, L, origin)
    out = {}
    for name, labels in [("Spin Correlations", dist), ("Staggered Spin Correlations", dist),
                         ("Spin Structure Factor", mom)]:
        out[name] = [w.result("%s[%s]" % (name, lab))[:2] for lab in labels]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("kind,L,T,origin", [("chain", 4, 0.5, None), ("chain", 8, 0.4, 3), ("square", 2, 0.6, None)])
def test_worker_correlations_match_exact_diagonalisation(loopgpu, kind, L, T, origin):
    print("worker correlations", kind, L, T, origin, flush=True)
    lat = "chain lattice" if kind == "chain" else "square lattice"
    txt = ('LATTICE = "%s"\nL = %d\nT = %g\nSWEEPS = 40000\nTHERMALIZATION = 2000\nSEED = 23125\n'
           'MEASURE[Correlations] = true\nMEASURE[Structure Factor] = true\nMEASURE[Green Function] = true\n' % (lat, L, T))
    if origin is not None:
        txt += "INITIAL_SITE = %d\n" % origin
    w = loopgpu.Worker(txt)
    w.run()
    ref = ed.diagonalize(kind, L, T)
    got = _worker_vectors(w, kind, L, origin)
    nx, ny, N = C.geometry(kind, L)
    if origin is None:
        exp = {"Spin Correlations": ref["Spin Correlations"], "Staggered Spin Correlations": ref["Staggered Spin Correlations"]}
    else:
        g = np.array([(-1) ** (r // ny + r % ny) for r in range(N)], float)
        row = ref["Spin Correlations Matrix"][origin]
        exp = {"Spin Correlations": row, "Staggered Spin Correlations": g[origin] * g * row}
    exp["Spin Structure Factor"] = ref["Spin Structure Factor"]
    for name, vals in got.items():
        for i, ((m, e), v) in enumerate(zip(vals, exp[name])):
            if e == 0:
                assert m == pytest.approx(v, abs=1e-12), (name, i)
            else:
                assert stats.sigmas(m, e, v) < 4.0, (name, i, m, e, v)
    # the scalar observables are unchanged by the extra measurements
    m, e, _ = w.result("Energy Density")
    assert stats.sigmas(m, e, ref["Energy"] / N) < 4.0


@pytest.mark.gpu
def test_worker_custom_measurements_of_sz(loopgpu):
    """MEASURE_AVERAGE / MEASURE_LOCAL [name] = Sz and MEASURE_CORRELATIONS /
    MEASURE_STRUCTURE_FACTOR [name] = "Sz:Sz" (custom.h:147-250), with the
    parameters of the reference's chain L=4, T=0.5 run: against exact
    diagonalisation, against that run's output and against the built-in
    correlation estimators; other operator expressions are skipped."""
    print("worker custom measurements chain 4 0.5", flush=True)
    txt = ('LATTICE = "chain lattice"\nL = 4\nT = 0.5\nSWEEPS = 40000\nTHERMALIZATION = 2000\nSEED = 23125\n'
           'MEASURE[Correlations] = true\nMEASURE[Structure Factor] = true\n')
    txt += "".join('%s = "%s"\n' % (k, v) for k, v in FX["params"].items() if k.startswith("MEASURE_"))
    w = loopgpu.Worker(txt)
    w.run()
    ref = ed.diagonalize("chain", 4, 0.5)
    dist, mom = C.labels("chain", 4, None)
    res = FX["results"]
    m, e, _ = w.result("Sz Average")
    assert stats.sigmas(m, e, 0.0) < 4.0
    assert stats.sigmas(m, e, *res["Sz Average"]) < 4.0
    for i in range(4):
        name = "Sz Local[( %d )]" % i
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, 0.0) < 4.0, name
        assert stats.sigmas(m, e, *res[name]) < 4.0, name
    for custom, builtin, labels, exact in [("Sz Correlations", "Spin Correlations", dist, ref["Spin Correlations"]),
                                           ("Sz Structure Factor", "Spin Structure Factor", mom, ref["Spin Structure Factor"])]:
        for lab, v in zip(labels, exact):
            name = "%s[%s]" % (custom, lab)
            got = w.result(name)
            assert got == w.result("%s[%s]" % (builtin, lab)), name
            m, e, _ = got
            if e == 0:
                assert m == pytest.approx(v, abs=1e-12), name
            else:
                assert stats.sigmas(m, e, v) < 4.0, (name, m, e, v)
                assert stats.sigmas(m, e, *res[name]) < 4.0, name
    assert "Exchange" not in res
    with pytest.raises(loopgpu.LoopGPUError):
        w.result("Exchange")
