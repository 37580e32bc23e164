"""lg_cluster_config_noalloc (include/loopgpu_noalloc.h) through the C-ABI:
the same labels, flipped spins, operator types and cluster count as
lg_cluster_config on the same configurations, on consecutive calls (the
second smaller than the first, so stale contents of the shared array would
show), and labels equal to looper's union-find (the oracle).  Device buffers
are the caller's, allocated through the HIP runtime the library itself uses."""

import ctypes

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        import loopgpu  # noqa: F401  (loads the library and its HIP runtime first)
        _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


class Dev:
    """a device buffer holding a copy of a host array"""

    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), max(self.arr.nbytes, 8)) == 0
        if self.arr.nbytes:
            assert hip().hipMemcpy(self.ptr, self.arr.ctypes.data, self.arr.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.arr)
        if out.nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        if self.ptr:
            hip().hipFree(self.ptr)


def _bind(lg):
    f = lg.lib.lg_cluster_config_noalloc
    vp = ctypes.c_void_p
    f.argtypes = [ctypes.POINTER(lg.LGParams), vp, vp, vp, vp, ctypes.c_int64, ctypes.c_uint64, vp, vp, vp,
                  ctypes.POINTER(ctypes.c_int64)]
    lg.lib.lg_noalloc_scratch_bytes.restype = ctypes.c_size_t
    return f


def noalloc(lg, p, spins, bo, t, ty, sweep):
    f = _bind(lg)
    n = len(t)
    ds, db = Dev(np.asarray(spins, np.int32)), Dev(np.asarray(bo, np.int32))
    dt, dy = Dev(np.asarray(t, np.float64)), Dev(np.asarray(ty, np.int32))
    lab = Dev(np.full(len(spins) + 4 * n, -1, np.int32))
    spo = Dev(np.full(len(spins), -1, np.int32))
    tyo = Dev(np.full(max(n, 1), -1, np.int32))
    nc = ctypes.c_int64(-1)
    rc = f(ctypes.byref(p), ds.ptr, db.ptr, dt.ptr, dy.ptr, n, sweep, lab.ptr, spo.ptr, tyo.ptr, ctypes.byref(nc))
    if rc != 0:
        lg._check(rc)
    return lab.get(), spo.get(), tyo.get()[:n], nc.value


CASES = [
    # kind, L, nops, graphs, time_grid: the second of each pair is smaller
    (O.SQUARE, 16, 20000, (0, 1, 2, 3), False),
    (O.CHAIN, 16, 300, (0, 1), True),
    (O.SQUARE, 128, 65536, (0, 1), False),        # the maximum operator count, 16384 sites
    (O.SQUARE, 8, 2000, (0, 1, 2, 3), False),
]


def test_noalloc_equals_cluster_config_on_consecutive_calls(loopgpu):
    lg = loopgpu
    _bind(lg)
    assert 8 << 20 < lg.lib.lg_noalloc_scratch_bytes() <= 17 << 20
    for i, (kind, L, nops, graphs, grid) in enumerate(CASES):
        print("noalloc case", i, kind, L, nops, flush=True)
        rng = np.random.default_rng(500 + i)
        spins, bo, t, ty = O.random_config(kind, L, nops, rng, graphs, grid)
        p = lg.params(kind, L, 1.0, 1.0, 1.0, seed=77, chain_id=i)
        ref = lg.cluster_config(p, spins, bo, t, ty, 31 + i)
        got = noalloc(lg, p, spins, bo, t, ty, 31 + i)
        for a, b, name in zip(got, ref, ("labels", "spins", "types", "clusters")):
            np.testing.assert_array_equal(a, b, err_msg="%s case %d" % (name, i))
        np.testing.assert_array_equal(got[0], O.looper_labels(kind, L, bo, t, ty))


def test_noalloc_equilibrated_configuration_with_field(loopgpu):
    """a configuration the device produced, and the field path (k_field_flip)"""
    lg = loopgpu
    for h in (0.0, 0.6):
        p = lg.params(O.SQUARE, 16, 6.0, 1.0, 0.8, seed=3, chain_id=4, h=h)
        g = lg.Chain(p)
        g.sweep(40)
        cfg = g.get_config()
        print("noalloc equilibrated h", h, len(cfg[2]), flush=True)
        ref = lg.cluster_config(p, *cfg, 40)
        got = noalloc(lg, p, *cfg, 40)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)


def test_noalloc_virtual_sites(loopgpu):
    """S > 1/2 (virtual sites, split world lines and the boundary
    permutation): S = 1, 3/2 chains and the S = 1 square lattice"""
    lg = loopgpu
    for kind, L, two_s, beta in ((O.CHAIN, 16, 2, 4.0), (O.CHAIN, 8, 3, 3.0), (O.SQUARE, 4, 2, 2.0)):
        p = lg.params(kind, L, beta, 1.0, 1.0, seed=5, chain_id=9, two_s=two_s)
        g = lg.Chain(p)
        g.sweep(30)
        cfg = g.get_config()
        print("noalloc virtual sites", kind, L, two_s, len(cfg[2]), flush=True)
        for sweep in (30, 31):
            ref = lg.cluster_config(p, *cfg, sweep)
            got = noalloc(lg, p, *cfg, sweep)
            for a, b, name in zip(got, ref, ("labels", "spins", "types", "clusters")):
                np.testing.assert_array_equal(a, b, err_msg="%s %s" % (name, (kind, L, two_s, sweep)))


def test_noalloc_rejects_oversized_and_invalid_input(loopgpu):
    lg = loopgpu
    rng = np.random.default_rng(9)
    spins, bo, t, ty = O.random_config(O.CHAIN, 16, 100, rng, (0, 1), False)
    p = lg.params(O.CHAIN, 16, 1.0)
    big = lg.params(O.SQUARE, 130, 1.0)                       # 16900 sites
    with pytest.raises(lg.LoopGPUError) as e:
        noalloc(lg, big, np.zeros(130 * 130, np.int32), np.zeros(2 * 130 * 130 + 1, np.int32), [], [], 0)
    assert e.value.code == -3
    bad_t = t.copy()
    bad_t[1], bad_t[0] = 0.9, 0.95                           # descending within bond 0
    if bo[1] >= 2:
        with pytest.raises(lg.LoopGPUError) as e:
            noalloc(lg, p, spins, bo, bad_t, ty, 0)
        assert e.value.code == -1
    bad_bo = bo.copy()
    bad_bo[-1] += 1                                          # n_ops != bond_offsets[n_bonds]
    with pytest.raises(lg.LoopGPUError) as e:
        noalloc(lg, p, spins, bad_bo, t, ty, 0)
    assert e.value.code == -1
    # and the next valid call is right
    for a, b in zip(noalloc(lg, p, spins, bo, t, ty, 5), lg.cluster_config(p, spins, bo, t, ty, 5)):
        np.testing.assert_array_equal(a, b)
