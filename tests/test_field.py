"""Longitudinal field (looper's FIELD path, path_integral.C:798, 852-858;
cluster.h:50-63): device trajectories bit-exact with the oracle's philox mode,
mixed batches, and the worker's normal-estimator observables against exact
diagonalisation."""

import numpy as np
import pytest

import oracle_lib as O
import stats

pytestmark = pytest.mark.gpu

RTOL = 1e-9
EXACT = [0, 1]
FLOAT = list(range(2, 18)) + [20, 21, 22]

TRAJ = [
    # kind, L, beta, jxy, jz, h, sweeps
    (O.CHAIN, 16, 4.0, 1.0, 1.0, 0.4, 60),
    (O.CHAIN, 12, 10.0, 1.0, 0.5, 0.9, 60),     # XY regime, strong field
    (O.SQUARE, 8, 3.0, 1.0, 1.0, 0.7, 40),
    (O.SQUARE, 6, 2.0, 1.0, 1.4, -0.5, 40),     # Ising regime (frozen graphs), negative field
]


@pytest.mark.parametrize("cfg", TRAJ, ids=lambda c: "k%d_L%d_b%g_jz%g_h%g" % (c[0], c[1], c[2], c[4], c[5]))
def test_field_trajectories_match_oracle_bit_exact(loopgpu, cfg):
    kind, L, beta, jxy, jz, h, n = cfg
    print("field trajectory", cfg, flush=True)
    g = loopgpu.Chain(loopgpu.params(kind, L, beta, jxy, jz, seed=91, chain_id=6, h=h))
    c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=91, chain_id=6, field=h)
    rg = g.sweep(n)
    rc = c.sweep(n)
    np.testing.assert_array_equal(rg[:, EXACT], rc[:, EXACT])
    np.testing.assert_allclose(rg[:, FLOAT], rc[:, FLOAT], rtol=RTOL, atol=1e-9)
    assert np.any(rg[:, O.R_EFIELD] != 0)
    gs, gbo, gt, gty = g.get_config()
    cs, cbo, ct, cty = c.state()
    np.testing.assert_array_equal(gs, cs)
    np.testing.assert_array_equal(gbo, cbo)
    np.testing.assert_array_equal(gt, ct)
    np.testing.assert_array_equal(gty, cty)


def test_mixed_field_batch_equals_single_chains(loopgpu):
    plist = [loopgpu.params(O.CHAIN, 16, 3.0, 1.0, 1.0, seed=4, chain_id=0, h=0.5),
             loopgpu.params(O.SQUARE, 6, 2.0, 1.0, 0.8, seed=4, chain_id=1),
             loopgpu.params(O.SQUARE, 4, 4.0, 1.0, 1.0, seed=4, chain_id=2, h=-0.3),
             loopgpu.params(O.CHAIN, 10, 1.0, 1.0, 1.2, seed=4, chain_id=3)]
    n = 30
    rb = loopgpu.Batch(plist).run(n)
    for i, p in enumerate(plist):
        rs = loopgpu.Chain(p).sweep(n)
        np.testing.assert_array_equal(rb[:, i, EXACT], rs[:, EXACT])
        np.testing.assert_allclose(rb[:, i, FLOAT], rs[:, FLOAT], rtol=RTOL, atol=1e-9)


def test_worker_with_field_matches_exact_diagonalisation(loopgpu):
    import ed
    L, T, h = 8, 0.5, 0.3
    w = loopgpu.Worker("LATTICE = \"chain lattice\"\nL = %d\nT = %g\nh = %g\nSWEEPS = 150000\n"
                       "THERMALIZATION = 2000\nSEED = 17" % (L, T, h))
    w.run()
    with pytest.raises(loopgpu.LoopGPUError):
        w.result("Generalized Susceptibility")   # normal estimators with a field
    ref = ed.diagonalize("chain", L, T, 1.0, 1.0, h)
    for name, rk in [("Energy Density", ref["Energy"] / L), ("Magnetization", ref["Magnetization"]),
                     ("Magnetization^2", ref["Magnetization^2"]), ("Susceptibility", ref["Susceptibility"])]:
        m, err, _ = w.result(name)
        assert stats.sigmas(m, err, rk) < 4.0, (name, m, err, rk)


def test_worker_disabled_improved_estimator_without_field(loopgpu):
    """DISABLE_IMPROVED_ESTIMATOR without a field (path_integral.C:217): the
    normal estimators of the same observables (susceptibility.h:258-339), no
    generalized or stiffness observables, against exact diagonalisation; the
    improved run of the same chain measures the same energies"""
    import ed
    L, T = 8, 0.4
    base = "LATTICE = \"chain lattice\"\nL = %d\nT = %g\nSWEEPS = 100000\nTHERMALIZATION = 2000\nSEED = 29" % (L, T)
    w = loopgpu.Worker(base + "\nDISABLE_IMPROVED_ESTIMATOR = 1")
    w.run()
    rep = w.report()
    assert "Generalized" not in rep and "Stiffness" not in rep
    ref = ed.diagonalize("chain", L, T, 1.0, 1.0)
    for name in ("Energy", "Magnetization^2", "Susceptibility", "Staggered Magnetization^2",
                 "Staggered Susceptibility", "|Staggered Magnetization|", "Specific Heat"):
        rk = ref[name]
        m, err, _ = w.result(name)
        assert stats.sigmas(m, err, rk) < 4.0, (name, m, err, rk)
    m, _, _ = w.result("Magnetization")
    assert abs(m) < 0.1
    # the estimator choice does not change the Markov chain: identical energies
    v = loopgpu.Worker(base)
    v.run()
    assert v.result("Energy") == w.result("Energy")
    assert v.result("Generalized Susceptibility")[0] > 0
