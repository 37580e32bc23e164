"""The worker boundary (include/loopgpu_worker.h): the reference worker's
parameters, observable names and update/measure interface."""

import json
import os

import numpy as np
import pytest

import stats

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))


@pytest.mark.parametrize("txt,frag", [
    ("L=7; T=1", "even"),
    ('LATTICE="chain lattice"; T=0.1', "L not defined"),
    ('LATTICE="triangular lattice"; L=4; T=1', "unsupported LATTICE"),
    ('L=4; T=1; Jxy=-1; Jz=1; Gamma=0.3; MEASURE[Correlations]=true', "transverse field"),
    ("L=4; local_S=9/2; T=1", "local_S"),
    ('LATTICE="square lattice"; L=4; local_S=0; T=1', "local_S"),
    ("L=4; T=-1", "negative temperature"),
    ("L=4", "temperature is not initialized"),
    ("L=4; T=1; SWEEPS=[100:x]", "SWEEPS"),
    ("L=4; T=1/Q", "unknown name"),
])
def test_worker_parameter_errors(txt, frag):
    import loopgpu
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Worker(txt)
    assert e.value.code == -1
    assert frag in str(e.value)


@pytest.mark.gpu
def test_worker_matches_reference_run_and_ed(loopgpu):
    # loop.ip block: chain L=4, J=1, T=0.25 (loop.op: ED and "loop; path integral")
    w = loopgpu.Worker('LATTICE = "chain lattice"\nL = 4\nJ = 1\nT = 0.25\nSWEEPS = 65536\nSEED = 23125')
    print("worker run", flush=True)
    w.run()
    st = w.status()
    assert st["mcs"] == 65536 + 8192 and not st["can_work"]
    ed, qmc = FX["loop_op_chain4"][0]["results"], FX["loop_op_chain4"][1]["results"]
    for name in ["Energy", "Energy Density", "Staggered Magnetization^2", "Staggered Susceptibility",
                 "Susceptibility", "Magnetization^2", "Specific Heat", "Binder Ratio of Staggered Magnetization"]:
        m, e, _ = w.result(name)
        assert stats.sigmas(m, e, ed[name][0]) < 3.5, (name, m, e, ed[name][0])
        if name in qmc:
            assert stats.sigmas(m, e, qmc[name][0], qmc[name][1]) < 3.5, (name, m, e, qmc[name])
    rep = w.report()
    assert rep.startswith("[results]")
    for name in ["Number of Clusters", "Generalized Staggered Susceptibility", "|Staggered Magnetization|",
                 "Stiffness", "Temperature", "Volume"]:
        assert name + ": " in rep


@pytest.mark.gpu
def test_worker_checkpoint_at_any_step_resumes_exactly(loopgpu):
    """saved at a step count off the fetch chunk (and again at the end of
    thermalization), a worker loaded from the blob ends with every result of
    an uninterrupted run, bit for bit"""
    p = 'LATTICE = "square lattice"\nL = 6\nT = 0.3\nSWEEPS = 3000\nTHERMALIZATION = 200\nSEED = 9'
    c = loopgpu.Worker(p)
    c.set_chunk(100)
    c.run()
    for first in (1037, 200, 7):
        print("worker checkpoint at step", first, flush=True)
        a = loopgpu.Worker(p)
        a.set_chunk(100)
        assert a.run(first) == first
        assert a.status()["mcs"] == first
        blob = a.save()
        b = loopgpu.Worker(p)
        b.load(blob)
        assert b.status()["mcs"] == first
        b.set_chunk(64)
        b.run()
        assert b.status()["mcs"] == c.status()["mcs"] == 3200
        # functions of integer counts bit for bit; FP64 estimator sums (device
        # atomics, accumulated in varying order) within 1e-9
        for name in ("Energy", "Energy^2", "Number of Clusters", "Specific Heat", "Inverse Temperature"):
            assert b.result(name) == c.result(name), (first, name)
        for name in ("Staggered Magnetization^2", "Susceptibility", "Stiffness",
                     "Binder Ratio of Staggered Magnetization"):
            for x, y in zip(b.result(name), c.result(name)):
                assert x == pytest.approx(y, rel=1e-9, abs=1e-12), (first, name)
        assert [l.split(":")[0] for l in b.report().splitlines()] == [l.split(":")[0] for l in c.report().splitlines()]
    with pytest.raises(loopgpu.LoopGPUError) as e:
        b.load(blob[:100])
    assert e.value.code == -1


@pytest.mark.gpu
@pytest.mark.parametrize("par", [
    # sign problem: the sign-weighted normal estimators and "Sign"
    'L = 4\nJ = 1\nGamma = 0.8\nT = 0.5',
    # the SSE worker with S=1, a single-ion term and a (sign-free) transverse field
    'ALGORITHM = "loop; sse"\nL = 4\nlocal_S = 1\nJxy = -1\nJz = 1\nD = 0.3\nGamma = 0.4\nT = 0.5'])
def test_worker_checkpoint_resumes_signed_and_sse_runs(loopgpu, par):
    """the same resume as above for the signed and SSE workers, whose results
    come from other accumulators"""
    p = par + "\nSWEEPS = 2000\nTHERMALIZATION = 200\nSEED = 3"
    c = loopgpu.Worker(p)
    c.set_chunk(100)
    c.run()
    print("worker checkpoint", par.replace("\n", "; "), flush=True)
    a = loopgpu.Worker(p)
    a.set_chunk(100)
    assert a.run(1237) == 1237
    b = loopgpu.Worker(p)
    b.load(a.save())
    b.set_chunk(64)
    b.run()
    assert b.status()["mcs"] == c.status()["mcs"] == 2200
    names = [l.split(":")[0] for l in c.report().splitlines()]
    assert [l.split(":")[0] for l in b.report().splitlines()] == names
    for name in ("Energy", "Sign", "Number of Clusters") if "Gamma = 0.8" in par else ("Energy", "Number of Clusters"):
        assert b.result(name) == c.result(name), name
    for name in ("Magnetization^2", "Susceptibility", "Staggered Magnetization^2", "Staggered Susceptibility",
                 "Specific Heat"):
        for x, y in zip(b.result(name), c.result(name)):
            assert x == pytest.approx(y, rel=1e-9, abs=1e-12), name
