"""GPU parity: the HIP path against the CPU oracle, through the C-ABI.

* cluster stage on given operator/cut configurations: canonical cluster labels
  bit-exact with looper's union-find (oracle restatement of
  looper/union_find.h + graph_impl.h reconnect, pinned by test/union_find.op),
  flipped spins and operator types bit-exact;
* whole sweeps: operator strings, spins, operator counts and cluster counts
  bit-exact with the oracle's counter-based (philox) mode; FP64 estimator sums
  within rtol 1e-9 (atomic accumulation order differs);
* batched launches: every chain identical to the same chain run alone.
"""

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

RTOL = 1e-9
EXACT = [0, 1]          # n_op, n_clusters
FLOAT = list(range(2, 18)) + [20, 21, 22]   # + normal-estimator sums, Zeeman energy


def _cmp_records(gpu, cpu, where=""):
    assert gpu.shape[0] == cpu.shape[0]
    np.testing.assert_array_equal(gpu[:, EXACT], cpu[:, EXACT], err_msg="n_op/n_clusters " + where)
    np.testing.assert_allclose(gpu[:, FLOAT], cpu[:, FLOAT], rtol=RTOL, atol=1e-9, err_msg=where)


CONFIGS = [
    # kind, L, nops, graphs, time_grid
    (O.CHAIN, 2, 0, (0,), False),
    (O.CHAIN, 4, 7, (0, 1, 2, 3), False),
    (O.CHAIN, 8, 40, (0,), False),
    (O.CHAIN, 16, 300, (0, 1), False),
    (O.CHAIN, 16, 300, (0, 1, 2, 3), True),     # many equal times
    (O.SQUARE, 2, 30, (0, 1, 2, 3), False),
    (O.SQUARE, 4, 200, (0, 1), False),
    (O.SQUARE, 8, 2000, (0, 1, 2, 3), False),
    (O.SQUARE, 16, 20000, (0,), False),
    (O.CHAIN, 64, 5000, (0, 1, 2, 3), True),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "k%d_L%d_n%d" % (c[0], c[1], c[2]))
def test_cluster_labels_match_looper_union_find(loopgpu, cfg):
    kind, L, nops, graphs, grid = cfg
    rng = np.random.default_rng(1000 + L * 7 + nops)
    spins, bo, t, ty = O.random_config(kind, L, nops, rng, graphs, grid)
    seed, chain_id, sweep = 4242, 3, 17
    p = loopgpu.params(kind, L, 1.0, 1.0, 1.0, seed=seed, chain_id=chain_id)
    print("cluster parity", cfg, flush=True)
    labels, spo, tyo, nc = loopgpu.cluster_config(p, spins, bo, t, ty, sweep)
    ref = O.looper_labels(kind, L, bo, t, ty)
    np.testing.assert_array_equal(labels, ref)
    exp_sp, exp_ty, exp_nc = O.expected_flip(kind, L, seed, chain_id, sweep, spins, bo, t, ty, ref)
    np.testing.assert_array_equal(spo, exp_sp)
    np.testing.assert_array_equal(tyo, exp_ty)
    assert nc == exp_nc


TRAJ = [
    # kind, L, beta, jxy, jz, sweeps
    (O.CHAIN, 16, 4.0, 1.0, 1.0, 60),
    (O.CHAIN, 12, 10.0, 1.0, 0.6, 60),       # XY regime: cross graphs
    (O.CHAIN, 10, 3.0, 1.0, 1.7, 60),        # Ising-like: frozen graphs
    (O.SQUARE, 4, 2.0, 1.0, 1.0, 60),
    (O.SQUARE, 8, 6.0, 1.0, 0.8, 40),
    (O.CHAIN, 64, 10.0, 1.0, 1.0, 40),       # C1 shape
]


@pytest.mark.parametrize("cfg", TRAJ, ids=lambda c: "k%d_L%d_b%g_jxy%g_jz%g" % c[:5])
def test_sweeps_match_oracle_bit_exact(loopgpu, cfg):
    kind, L, beta, jxy, jz, n = cfg
    print("trajectory", cfg, flush=True)
    g = loopgpu.Chain(loopgpu.params(kind, L, beta, jxy, jz, seed=777, chain_id=5))
    c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=777, chain_id=5)
    rg = g.sweep(n)
    rc = c.sweep(n)
    _cmp_records(rg, rc, str(cfg))
    gs, gbo, gt, gty = g.get_config()
    cs, cbo, ct, cty = c.state()
    np.testing.assert_array_equal(gs, cs)
    np.testing.assert_array_equal(gbo, cbo)
    np.testing.assert_array_equal(gt, ct)
    np.testing.assert_array_equal(gty, cty)
    assert np.all(rg[:, 19] == np.arange(n))


LAUNCH_TRAJ = [
    # kind, L, beta, jxy, jz, sweeps: a few events per site, tens, and more
    # than the 512-event staging (the linking kernel's global-memory path)
    (O.SQUARE, 8, 6.0, 1.0, 0.8, 20),
    (O.CHAIN, 32, 60.0, 1.0, 1.0, 20),
    (O.CHAIN, 8, 600.0, 1.0, 1.0, 12),     # about 375 operators per bond: the fill pass's 512-time stage
    (O.CHAIN, 6, 1200.0, 1.0, 1.0, 8),     # about 750 per bond: its 1024-time stage (64 KB of LDS)
    (O.CHAIN, 4, 2000.0, 1.0, 0.9, 6),     # about 2500 events per site: beyond the 1536-event staging
]
LAUNCHES = [dict(link=0), dict(link=1), dict(link=128), dict(link=512), dict(link=1536),
            dict(link=1, group=8, blocks=3), dict(link=1, compress=1), dict(link=512, group=4, blocks=7),
            dict(group=1), dict(group=2, link=1), dict(group=1, link=0, blocks=5), dict(group=2, compress=1),
            dict(link=128, group=8, blocks=5),
            dict(link=0, group=16, blocks=64), dict(link=1536, group=32),
            dict(link=512, compress=1), dict(link=0, compress=0, blocks=9), dict(link=1536, compress=1, group=8),
            dict(flush=8), dict(flush=4, compress=0), dict(flush=4, link=128, blocks=5),
            dict(group=64, link=0), dict(group=64, link=512, compress=0)]


@pytest.mark.parametrize("cfg", LAUNCH_TRAJ, ids=lambda c: "k%d_L%d_b%g" % c[:3])
def test_launch_settings_leave_trajectories_bit_exact(loopgpu, cfg):
    """every linking kernel, lane-group size and grid gives the oracle's trajectory"""
    kind, L, beta, jxy, jz, n = cfg
    c = O.PhiloxChain(kind, L, beta, jxy, jz, seed=31, chain_id=2)
    rc = c.sweep(n)
    cs, cbo, ct, cty = c.state()
    for lc in LAUNCHES:
        print("launch", cfg, lc, flush=True)
        g = loopgpu.Chain(loopgpu.params(kind, L, beta, jxy, jz, seed=31, chain_id=2))
        g.set_launch(**lc)
        got = g.launch()
        for k, v in lc.items():
            assert got[k] == v
        rg = g.sweep(n)
        _cmp_records(rg, rc, "%s %s" % (cfg, lc))
        gs, gbo, gt, gty = g.get_config()
        np.testing.assert_array_equal(gbo, cbo)
        np.testing.assert_array_equal(gt, ct)
        np.testing.assert_array_equal(gty, cty)
        np.testing.assert_array_equal(gs, cs)


LARGE = [
    # square L, beta, sweeps, default linking kernel: the last sweeps hold
    # 2.6M, 4.4-5.2M and 7M segments, where the estimator and the collector
    # take 8 and 16 items per thread; about 640 events per site at beta = 256
    (128, 96.0, 3, 512),
    (128, 160.0, 4, 512),
    (128, 256.0, 3, 1536),
]


@pytest.mark.parametrize("cfg", LARGE, ids=lambda c: "L%d_b%g" % c[:2])
def test_large_configurations_match_oracle(loopgpu, cfg):
    """sizes of the benchmark configuration (larger estimator and collector
    chunks than the small cases above), with and without the compress pass"""
    L, beta, n, link = cfg
    print("large", cfg, flush=True)
    c = O.PhiloxChain(O.SQUARE, L, beta, 1.0, 1.0, seed=5, chain_id=2)
    rc = c.sweep(n)
    cs, cbo, ct, cty = c.state()
    assert 2 * rc[-1, 0] > 2.2e6          # past the 4-item chunks
    for comp in (1, 0):
        g = loopgpu.Chain(loopgpu.params(O.SQUARE, L, beta, 1.0, 1.0, seed=5, chain_id=2))
        assert g.launch()["link"] == link
        g.set_launch(compress=comp)
        _cmp_records(g.sweep(n), rc, "large %s compress %d" % (cfg, comp))
        gs, gbo, gt, gty = g.get_config()
        np.testing.assert_array_equal(gbo, cbo)
        np.testing.assert_array_equal(gt, ct)
        np.testing.assert_array_equal(gty, cty)
        np.testing.assert_array_equal(gs, cs)
        g.close()


def test_compress_mode_switches_keep_trajectory(loopgpu):
    """switching between the compress pass and the estimator's own root walk
    between sweeps (accumulators cleared on the switch) keeps the oracle's
    trajectory, for a chain and for a batch with a field chain"""
    n, modes = 24, [1, 0, 0, 1, 0, 1]
    c = O.PhiloxChain(O.SQUARE, 8, 4.0, 1.0, 1.0, seed=8, chain_id=5)
    rc = c.sweep(n)
    g = loopgpu.Chain(loopgpu.params(O.SQUARE, 8, 4.0, 1.0, 1.0, seed=8, chain_id=5))
    assert g.launch()["compress"] == 0          # by default the estimator walks to the roots itself
    parts = []
    for m in modes:
        g.set_launch(compress=m)
        parts.append(g.sweep(n // len(modes)))
    _cmp_records(np.concatenate(parts), rc, "chain mode switches")
    plist = [loopgpu.params(O.CHAIN, 16, 3.0, 1.0, 1.0, seed=4, chain_id=0, h=0.5),
             loopgpu.params(O.SQUARE, 6, 2.0, 1.0, 0.8, seed=4, chain_id=1),
             loopgpu.params(O.CHAIN, 10, 1.0, 1.0, 1.2, seed=4, chain_id=3)]
    b = loopgpu.Batch(plist)
    assert b.launch()["compress"] == 0          # for batches too
    parts = []
    for m in modes:
        b.set_launch(compress=m)
        parts.append(b.run(n // len(modes)))
    rb = np.concatenate(parts)
    for i, p in enumerate(plist):
        rs = loopgpu.Chain(p).sweep(n)
        np.testing.assert_array_equal(rb[:, i, [0, 1]], rs[:, [0, 1]])
        np.testing.assert_allclose(rb[:, i, FLOAT], rs[:, FLOAT], rtol=1e-9, atol=1e-9)


def test_launch_setting_errors(loopgpu):
    g = loopgpu.Chain(loopgpu.params(O.CHAIN, 8, 1.0))
    for bad in (dict(group=3), dict(link=100), dict(blocks=0), dict(compress=2), dict(flush=2), dict(flush=12)):
        with pytest.raises(loopgpu.LoopGPUError) as e:
            g.set_launch(**bad)
        assert e.value.code == -1


def test_batch_equals_single_chains(loopgpu):
    plist = [
        loopgpu.params(O.CHAIN, 8, 2.0, 1.0, 1.0, seed=11, chain_id=0),
        loopgpu.params(O.SQUARE, 4, 3.0, 1.0, 0.7, seed=11, chain_id=1),
        loopgpu.params(O.CHAIN, 32, 8.0, 1.0, 1.0, seed=11, chain_id=2),
        loopgpu.params(O.SQUARE, 6, 1.0, 1.0, 1.4, seed=11, chain_id=3),
    ]
    n = 30
    singles = [loopgpu.Chain(p).sweep(n) for p in plist]
    for link in (None, 0, 1, 128, 512, 1536):  # the default and every linking kernel (sites of several chains per block)
        b = loopgpu.Batch(plist)
        if link is not None:
            b.set_launch(link=link)
        rb = b.run(n)
        for i, rs in enumerate(singles):
            print("batch chain", i, "link", link, flush=True)
            np.testing.assert_array_equal(rb[:, i, EXACT], rs[:, EXACT])
            np.testing.assert_allclose(rb[:, i, FLOAT], rs[:, FLOAT], rtol=RTOL, atol=1e-9)


def test_graph_replay_matches_plain_launches(loopgpu):
    plist = [loopgpu.params(O.SQUARE, 8, 4.0, 1.0, 1.0, seed=5, chain_id=i) for i in range(3)]
    a = loopgpu.Batch(plist)
    b = loopgpu.Batch(plist)
    b.set_graph(True)
    ra = a.run(25)
    rb = b.run(25)
    np.testing.assert_array_equal(ra[:, :, EXACT], rb[:, :, EXACT])
    np.testing.assert_allclose(ra[:, :, FLOAT], rb[:, :, FLOAT], rtol=RTOL, atol=1e-9)


def test_set_config_roundtrip_and_resume(loopgpu):
    p = loopgpu.params(O.CHAIN, 16, 5.0, 1.0, 1.0, seed=9)
    a = loopgpu.Chain(p)
    a.sweep(20)
    cfg = a.get_config()
    _, _, _, sw = a.info()
    ra = a.sweep(15)
    b = loopgpu.Chain(p)
    b.set_config(*cfg, sweep=sw)
    rb = b.sweep(15)
    np.testing.assert_array_equal(ra[:, EXACT], rb[:, EXACT])
    np.testing.assert_allclose(ra[:, FLOAT], rb[:, FLOAT], rtol=RTOL, atol=1e-9)


def test_chain_checkpoint_blob_continues_the_trajectory(loopgpu):
    p = loopgpu.params(O.SQUARE, 8, 4.0, 1.0, 0.9, seed=21, chain_id=4)
    a = loopgpu.Chain(p)
    a.sweep(20)
    blob = a.save()
    ra = a.sweep(15)
    b = loopgpu.Chain(p)
    b.load(blob)
    rb = b.sweep(15)
    np.testing.assert_array_equal(ra[:, EXACT], rb[:, EXACT])
    np.testing.assert_allclose(ra[:, FLOAT], rb[:, FLOAT], rtol=RTOL, atol=1e-9)
    np.testing.assert_array_equal(ra[:, 19], rb[:, 19])
    other = loopgpu.Chain(loopgpu.params(O.SQUARE, 8, 4.0, 1.0, 0.9, seed=22, chain_id=4))
    for bad in (blob[:40], b"X" + blob[1:]):
        with pytest.raises(loopgpu.LoopGPUError) as e:
            b.load(bad)
        assert e.value.code == -1
    with pytest.raises(loopgpu.LoopGPUError) as e:
        other.load(blob)
    assert e.value.code == -1


def test_batch_add_keeps_every_trajectory(loopgpu):
    first = [loopgpu.params(O.CHAIN, 16, 3.0, 1.0, 1.0, seed=8, chain_id=0),
             loopgpu.params(O.SQUARE, 6, 2.0, 1.0, 0.7, seed=8, chain_id=1)]
    later = [loopgpu.params(O.SQUARE, 4, 5.0, 1.0, 1.0, seed=8, chain_id=2),
             loopgpu.params(O.CHAIN, 8, 1.0, 1.0, 1.3, seed=8, chain_id=3)]
    b = loopgpu.Batch(first)
    r1 = b.run(10)
    b.add(later)
    r2 = b.run(20)
    for i, p in enumerate(first):
        rs = loopgpu.Chain(p).sweep(30)
        np.testing.assert_array_equal(np.concatenate([r1[:, i, EXACT], r2[:, i, EXACT]]), rs[:, EXACT])
        np.testing.assert_allclose(r2[:, i, FLOAT], rs[10:, FLOAT], rtol=RTOL, atol=1e-9)
    for j, p in enumerate(later):
        rs = loopgpu.Chain(p).sweep(20)
        np.testing.assert_array_equal(r2[:, 2 + j, EXACT], rs[:, EXACT])
        np.testing.assert_allclose(r2[:, 2 + j, FLOAT], rs[:, FLOAT], rtol=RTOL, atol=1e-9)


def _decode_chain_blob(blob, n_sites):
    """the chain checkpoint's fields (include/loopgpu.h lg_chain_save)"""
    import struct
    assert blob[:8] == b"LGCHAIN2"
    sweep, ns = struct.unpack_from("<QQ", blob, 8)
    assert ns == n_sites
    off = 24
    spins = np.frombuffer(blob, np.int32, ns, off)
    off += 4 * ns
    (nops,) = struct.unpack_from("<Q", blob, off)
    off += 8
    rec = np.frombuffer(blob, np.dtype([("type", "<i4"), ("loc", "<i4"), ("time", "<f8")]), nops, off)
    return sweep, spins, rec


def test_kill_and_resume_matches_uninterrupted_oracle(loopgpu):
    """a chain saved, destroyed and resumed in a new handle continues the
    oracle's uninterrupted trajectory bit for bit; the blob holds mcs, spins
    and the operators in imaginary-time order (path_integral.C:111-124)"""
    for kind, L, beta, jz in ((O.SQUARE, 8, 5.0, 0.8), (O.CHAIN, 32, 12.0, 1.0)):
        print("kill and resume", kind, L, beta, flush=True)
        p = loopgpu.params(kind, L, beta, 1.0, jz, seed=12, chain_id=9)
        c = O.PhiloxChain(kind, L, beta, 1.0, jz, seed=12, chain_id=9)
        rc = c.sweep(50)
        a = loopgpu.Chain(p)
        ra = a.sweep(23)
        blob = a.save()
        cfg = a.get_config()
        a.close()
        del a
        sweep, spins, rec = _decode_chain_blob(blob, len(cfg[0]))
        assert sweep == 23
        np.testing.assert_array_equal(spins, cfg[0])
        bond = np.repeat(np.arange(len(cfg[1]) - 1), np.diff(cfg[1]))
        order = np.lexsort((np.arange(len(bond)), bond, cfg[2]))      # time, then bond, then position
        np.testing.assert_array_equal(rec["time"], cfg[2][order])
        np.testing.assert_array_equal(rec["type"], cfg[3][order])
        np.testing.assert_array_equal(rec["loc"], 2 * bond[order] + 1)
        assert np.all(np.diff(rec["time"]) >= 0)
        b = loopgpu.Chain(p)
        b.load(blob)
        rb = b.sweep(27)
        _cmp_records(np.concatenate([ra, rb]), rc, "kill and resume %s" % ((kind, L),))
        for x, y in zip(b.get_config(), c.state()):
            np.testing.assert_array_equal(x, y)
