"""Annealing (looper/temperature.h:43-113, path_integral.C:212-213, 325).

CPU: the parameter parser's schedules and T(step) against the reference's
own test/temperature.ip / test/temperature.op (fixture), the four parameter
forms, the annealing <= thermalization check.  GPU: a chain under a schedule
bit-exact with the oracle's philox mode (beta of every sweep, operator
strings, spins), through a capacity rebuild and a checkpoint in the middle
of the schedule; the worker's annealed run against exact diagonalisation.
"""

import json
import os

import numpy as np
import pytest

import oracle_lib as O
import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))


def _text(params):
    return "; ".join("%s = %s" % kv for kv in params.items())


@pytest.mark.parametrize("blk", FX["temperature"], ids=lambda b: "_".join(sorted(b["params"]))[:40])
def test_schedule_matches_reference_temperature_op(blk):
    import loopgpu
    steps, temps = loopgpu.temperature_schedule(_text(blk["params"]))
    assert temps[0] == pytest.approx(blk["initial"], rel=1e-5)
    assert temps[-1] == pytest.approx(blk["final"], rel=1e-5)
    assert steps[-1] - 1 == blk["annealing_steps"]
    for step, T in blk["T"]:
        assert O.schedule_temperature(steps, temps, step) == pytest.approx(T, rel=1e-5), step


def test_schedule_closed_form_and_step_zero():
    """T(s) is linear between the pairs (temperature.h:106-108): endpoints
    exact, midpoints the mean; step 0 takes the first temperature"""
    steps, temps = [1, 11, 31], [2.0, 1.0, 0.25]
    assert O.schedule_temperature(steps, temps, 0) == 2.0
    assert O.schedule_temperature(steps, temps, 1) == 2.0
    assert O.schedule_temperature(steps, temps, 6) == 1.5
    assert O.schedule_temperature(steps, temps, 11) == 1.0
    assert O.schedule_temperature(steps, temps, 21) == pytest.approx(0.625, rel=1e-15)
    assert O.schedule_temperature(steps, temps, 31) == 0.25
    assert O.schedule_temperature(steps, temps, 10 ** 6) == 0.25
    assert O.schedule_temperature([1], [0.3], 5) == 0.3


def test_schedule_parameter_forms_agree():
    import loopgpu
    base = loopgpu.temperature_schedule("T = 0.25; T_START_0 = 1.5; T_DURATION_0 = 40; T_START_1 = 0.5; "
                                        "T_DURATION_1 = 20")
    assert list(base[0]) == [1, 41, 61]
    assert list(base[1]) == [1.5, 0.5, 0.25]
    for txt in ["TEMPERATURE = 0.25; TEMPERATURE_START_0 = 1.5; TEMPERATURE_DURATION_0 = 40; "
                "TEMPERATURE_START_1 = 0.5; TEMPERATURE_DURATION_1 = 20",
                "BETA = 4; BETA_START_0 = 1/1.5; BETA_DURATION_0 = 40; BETA_START_1 = 2; BETA_DURATION_1 = 20",
                "INVERSE_TEMPERATURE = 4; INVERSE_TEMPERATURE_START_0 = 1/1.5; INVERSE_TEMPERATURE_DURATION_0 = 40; "
                "INVERSE_TEMPERATURE_START_1 = 2; INVERSE_TEMPERATURE_DURATION_1 = 20"]:
        st, tp = loopgpu.temperature_schedule(txt)
        assert list(st) == list(base[0])
        np.testing.assert_allclose(tp, base[1], rtol=1e-15)
    # a missing or non-positive pair ends the schedule (temperature.h:79)
    st, tp = loopgpu.temperature_schedule("T = 0.5; T_START_0 = 1; T_DURATION_0 = 5; T_START_2 = 3; T_DURATION_2 = 5")
    assert list(st) == [1, 6] and list(tp) == [1.0, 0.5]
    st, tp = loopgpu.temperature_schedule("T = 0.5; T_START_0 = 1; T_DURATION_0 = 0")
    assert list(st) == [1] and list(tp) == [0.5]
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.temperature_schedule("L = 4")
    assert e.value.code == -1 and "temperature is not initialized" in str(e.value)


def test_annealing_longer_than_thermalization_is_rejected():
    """path_integral.C:212-213"""
    import loopgpu
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Worker("L = 8; T = 0.5; T_START_0 = 1; T_DURATION_0 = 101; THERMALIZATION = 100; SWEEPS = 1000")
    assert e.value.code == -1 and "longer annealing steps than thermalization" in str(e.value)
    with pytest.raises(loopgpu.LoopGPUError) as e:   # default thermalization SWEEPS/8 = 125
        loopgpu.Worker("L = 8; T = 0.5; T_START_0 = 1; T_DURATION_0 = 200; SWEEPS = 1000")
    assert "longer annealing steps" in str(e.value)
    if loopgpu.device_count() == 0:   # a schedule that fits passes the check (then needs the device)
        with pytest.raises(loopgpu.LoopGPUError) as e:
            loopgpu.Worker("L = 8; T = 0.5; T_START_0 = 1; T_DURATION_0 = 100; THERMALIZATION = 100; SWEEPS = 1000")
        assert e.value.code == -2


def test_oracle_flat_schedule_leaves_trajectory_unchanged():
    """a schedule whose temperatures are all the final one is the fixed-T chain"""
    a = O.PhiloxChain(O.SQUARE, 4, 3.0, 1.0, 0.8, seed=3, chain_id=1)
    b = O.PhiloxChain(O.SQUARE, 4, 3.0, 1.0, 0.8, seed=3, chain_id=1)
    b.set_schedule([1, 7, 20], [1 / 3.0, 1 / 3.0, 1 / 3.0])
    np.testing.assert_array_equal(a.sweep(30), b.sweep(30))


EXACT = [0, 1, 18, 19]                       # n_op, n_clusters, beta, sweep
FLOAT = list(range(2, 18)) + [20, 21, 22]


def _cmp(g, c, where):
    np.testing.assert_array_equal(g[:, EXACT], c[:, EXACT], err_msg=where)
    np.testing.assert_allclose(g[:, FLOAT], c[:, FLOAT], rtol=1e-9, atol=1e-9, err_msg=where)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,L,final_T,steps,temps", [
    (O.CHAIN, 16, 0.5, [1, 11, 26], [2.0, 0.1, 0.5]),          # colder than the final T: capacity grows
    (O.SQUARE, 6, 0.25, [1, 9, 30], [1.5, 0.8, 0.25]),
])
def test_annealed_chain_bit_exact_with_oracle(loopgpu, kind, L, final_T, steps, temps):
    beta = 1.0 / final_T
    p = loopgpu.params(kind, L, beta, 1.0, 0.9, seed=41, chain_id=6)
    print("annealed chain", kind, L, steps, temps, flush=True)
    g = loopgpu.Chain(p)
    g.set_schedule(steps, temps)
    c = O.PhiloxChain(kind, L, beta, 1.0, 0.9, seed=41, chain_id=6)
    c.set_schedule(steps, temps)
    rg, rc = g.sweep(20), c.sweep(20)
    _cmp(rg, rc, "first 20")
    np.testing.assert_array_equal(rg[:, 18], [1.0 / O.schedule_temperature(steps, temps, m) for m in range(20)])
    blob = g.save()
    rg2, rc2 = g.sweep(20), c.sweep(20)
    _cmp(rg2, rc2, "next 20")
    assert rg2[-1, 18] == beta
    for a, b in zip(g.get_config(), c.state()):
        np.testing.assert_array_equal(a, b)
    # resume from the checkpoint taken in the middle of the schedule
    h = loopgpu.Chain(p)
    h.set_schedule(steps, temps)
    h.load(blob)
    _cmp(h.sweep(20), rc2, "resumed")


@pytest.mark.gpu
def test_schedule_set_after_sweeps_keeps_configuration(loopgpu):
    """a schedule set on a running chain (capacity rebuilt for its largest
    beta) continues from the same configuration; removing it restores the
    fixed temperature"""
    p = loopgpu.params(O.CHAIN, 12, 2.0, 1.0, 1.0, seed=5, chain_id=0)
    steps, temps = [1, 30, 50], [1.0, 0.05, 0.5]
    g = loopgpu.Chain(p)
    c = O.PhiloxChain(O.CHAIN, 12, 2.0, 1.0, 1.0, seed=5, chain_id=0)
    _cmp(g.sweep(10), c.sweep(10), "fixed")
    g.set_schedule(steps, temps)
    c.set_schedule(steps, temps)
    _cmp(g.sweep(30), c.sweep(30), "annealed")
    g.set_schedule([], [])
    c.set_schedule([1], [0.5])
    _cmp(g.sweep(10), c.sweep(10), "fixed again")
    with pytest.raises(loopgpu.LoopGPUError) as e:
        g.set_schedule([1, 5], [1.0, 0.7])          # does not end at the chain's temperature
    assert e.value.code == -1


@pytest.mark.gpu
def test_worker_annealed_run_matches_exact_diagonalisation(loopgpu):
    import ed
    txt = ('LATTICE = "chain lattice"; L = 8; T = 0.25; T_START_0 = 2.0; T_DURATION_0 = 300; '
           'T_START_1 = 0.1; T_DURATION_1 = 200; THERMALIZATION = 1000; SWEEPS = 40000; SEED = 17')
    w = loopgpu.Worker(txt)
    print("annealed worker", flush=True)
    w.run()
    ref = ed.diagonalize("chain", 8, 0.25, 1.0, 1.0)
    for name in ("Energy", "Staggered Magnetization^2", "Susceptibility", "Staggered Susceptibility"):
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ref[name]) < 3.5, (name, m, e, ref[name])
    m, e, _ = w.result("Temperature")
    assert m == 0.25 and e == 0.0                    # every measured sweep at the final temperature
    # the looper-mode oracle under the same schedule agrees within the combined errors
    c = O.LooperChain(O.CHAIN, 8, 4.0, 1.0, 1.0, seed=17)
    c.set_schedule([1, 301, 501], [2.0, 0.1, 0.25])
    rec = c.sweep(41000)[1000:]
    assert np.all(rec[:, 18] == 4.0)
    mo, eo = stats.mean_err(rec[:, 2])
    m, e, _ = w.result("Energy")
    assert abs(m - mo) < 3.5 * np.hypot(e, eo), (m, e, mo, eo)
