"""Serving form (include/loopgpu_server.h): continuous batching of jobs."""

import time

import numpy as np
import pytest

import stats

EXACT = ["Energy Density", "Number of Clusters", "Operators"]  # functions of integer counts only
FLOAT = ["Susceptibility", "Staggered Structure Factor", "Staggered Susceptibility", "Stiffness"]


def test_server_needs_device_or_reports_it():
    import loopgpu
    if loopgpu.device_count() > 0:
        pytest.skip("host has a device")
    with pytest.raises(loopgpu.LoopGPUError) as e:
        loopgpu.Server(max_jobs=4)
    assert e.value.code == -2


def _job(lg, kind, L, beta, jz, cid, seed=77):
    return lg.params(kind, L, beta, 1.0, jz, seed=seed, chain_id=cid)


@pytest.mark.gpu
def test_job_result_same_alone_and_in_the_mix(loopgpu):
    lg = loopgpu
    J = _job(lg, lg.SQUARE, 16, 4.0, 0.8, 1000)
    a = lg.Server(max_jobs=8)
    ja = a.submit(J, sweeps=2000, thermalization=250)
    ra = a.wait(ja)
    a.stop()
    assert ra["status"] == lg.JOB_DONE and ra["sweeps_done"] == 2250
    b = lg.Server(max_jobs=8, max_chunk=16)
    # the first job leaves while J runs, so J's chain moves down in the batch
    others = [b.submit(_job(lg, lg.CHAIN, 64, 30.0, 1.0, 1 + i), sweeps=(300, 3500, 4000)[i]) for i in range(3)]
    others.append(b.submit(_job(lg, lg.SQUARE, 32, 8.0, 0.6, 7), sweeps=2500))
    time.sleep(0.05)  # J arrives while the others are on the device
    jb = b.submit(J, sweeps=2000, thermalization=250)
    others.append(b.submit(_job(lg, lg.SQUARE, 8, 2.0, 1.0, 9), sweeps=1000))
    rb = b.wait(jb)
    for o in others:
        assert b.wait(o)["status"] == lg.JOB_DONE
    st = b.stats()
    b.stop()
    assert rb["status"] == lg.JOB_DONE
    for k in EXACT:
        assert ra["mean"][k] == rb["mean"][k], k
        assert ra["error"][k] == rb["error"][k], k
    for k in FLOAT:
        assert rb["mean"][k] == pytest.approx(ra["mean"][k], rel=1e-9, abs=1e-12), k
    assert st["done"] == 6 and st["p99_latency_s"] >= st["p50_latency_s"] > 0
    # and the same as the plain one-chain path
    ch = lg.Chain(J)
    rec = ch.sweep(2250)[250:]
    assert np.mean(rec[:, 2] / 256) == pytest.approx(ra["mean"]["Energy Density"], rel=1e-14)
    assert np.mean(rec[:, 1]) == ra["mean"]["Number of Clusters"]


@pytest.mark.gpu
def test_admission_by_footprint_queues_jobs(loopgpu):
    lg = loopgpu
    p0 = _job(lg, lg.SQUARE, 16, 8.0, 1.0, 0)
    # room for about two such jobs at a time
    import ctypes
    s = lg.Server(max_jobs=16, max_ops=2 * 5000 + 3000, max_chunk=32)
    ids = [s.submit(_job(lg, lg.SQUARE, 16, 8.0, 1.0, i), sweeps=400, thermalization=50) for i in range(6)]
    res = [s.wait(i) for i in ids]
    st = s.stats()
    s.stop()
    assert all(r["status"] == lg.JOB_DONE for r in res)
    assert max(r["queue_s"] for r in res) > 0
    assert st["done"] == 6 and st["total_updates"] > 0 and st["updates_per_s"] > 0
    # every admission that changes the layout re-captures the sweep graph;
    # the host time of those captures is counted
    assert st["graph_captures"] >= 2 and st["graph_capture_s"] > 0


@pytest.mark.gpu
def test_cancel_and_capacity_errors(loopgpu):
    lg = loopgpu
    s = lg.Server(max_jobs=4, max_ops=200000)
    with pytest.raises(lg.LoopGPUError) as e:
        s.submit(_job(lg, lg.SQUARE, 96, 384.0, 1.0, 0), sweeps=10)
    assert e.value.code == -3
    long = s.submit(_job(lg, lg.CHAIN, 256, 100.0, 1.0, 1), sweeps=10 ** 6)
    short = s.submit(_job(lg, lg.CHAIN, 16, 4.0, 1.0, 2), sweeps=500)
    assert s.wait(short)["status"] == lg.JOB_DONE
    s.cancel(long)
    r = s.wait(long, 10)
    s.stop()
    assert r["status"] == lg.JOB_CANCELLED and r["sweeps_done"] < 10 ** 6


@pytest.mark.gpu
def test_served_jobs_match_exact_diagonalisation(loopgpu):
    import ed
    lg = loopgpu
    s = lg.Server(max_jobs=8)
    cases = [(8, 0.2, 1.0), (8, 0.25, 0.6), (10, 0.5, 0.8)]
    ids = [s.submit(_job(lg, lg.CHAIN, L, 1 / T, jz, 10 + i), sweeps=40000, thermalization=2000)
           for i, (L, T, jz) in enumerate(cases)]
    for (L, T, jz), i in zip(cases, ids):
        r = s.wait(i)
        ref = ed.diagonalize("chain", L, T, 1.0, jz)
        for k, rk in [("Energy Density", ref["Energy"] / L), ("Susceptibility", ref["Susceptibility"]),
                      ("Staggered Structure Factor", ref["Staggered Magnetization^2"] / L),
                      ("Staggered Susceptibility", ref["Staggered Susceptibility"])]:
            assert stats.sigmas(r["mean"][k], r["error"][k], rk) < 3.5, (L, T, jz, k, r["mean"][k], r["error"][k], rk)
    s.stop()


@pytest.mark.gpu
def test_preempted_job_resumes_with_the_uninterrupted_result(loopgpu):
    """a job preempted twice (its configuration saved, its device memory given
    to a job that was waiting) ends with the result of the same job run
    without interruption; the waiting job runs meanwhile"""
    lg = loopgpu
    A = _job(lg, lg.SQUARE, 12, 6.0, 0.8, 300)
    B = _job(lg, lg.CHAIN, 128, 40.0, 1.0, 301)
    ref = lg.Server(max_jobs=8)
    ra = ref.wait(ref.submit(A, sweeps=20000, thermalization=2000))
    rb_ref = ref.wait(ref.submit(B, sweeps=1500, thermalization=100))
    ref.stop()
    # room for one of the two jobs at a time (operator footprints about 3100
    # and 6900)
    s = lg.Server(max_jobs=8, max_ops=8000, max_chunk=16)
    ja = s.submit(A, sweeps=20000, thermalization=2000)
    time.sleep(0.3)
    jb = s.submit(B, sweeps=1500, thermalization=100)
    assert s.poll(jb)["status"] == lg.JOB_QUEUED        # B waits for A's memory
    done_before = s.poll(ja)["sweeps_done"]
    s.preempt(ja)
    rb = s.wait(jb)
    assert rb["status"] == lg.JOB_DONE
    mid = s.poll(ja)
    assert mid["status"] in (lg.JOB_QUEUED, lg.JOB_RUNNING) and mid["sweeps_done"] < 22000
    s.preempt(ja)                                       # again, whatever its state
    r = s.wait(ja)
    st = s.stats()
    s.stop()
    assert r["status"] == lg.JOB_DONE and r["sweeps_done"] == 22000
    assert st["preemptions"] >= 1 and done_before < 22000
    for res, want in ((r, ra), (rb, rb_ref)):
        for k in EXACT:
            assert res["mean"][k] == want["mean"][k], k
            assert res["error"][k] == want["error"][k], k
        for k in FLOAT:
            assert res["mean"][k] == pytest.approx(want["mean"][k], rel=1e-9, abs=1e-12), k
    with pytest.raises(lg.LoopGPUError) as e:
        s2 = lg.Server(max_jobs=2)
        j = s2.submit(_job(lg, lg.CHAIN, 8, 1.0, 1.0, 5), sweeps=10)
        s2.wait(j)
        try:
            s2.preempt(j)
        finally:
            s2.stop()
    assert e.value.code == -4
