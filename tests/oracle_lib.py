"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test infrastructure."""

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


if not os.path.exists(LIB):
    build_oracle()

lib = ctypes.CDLL(LIB)
vp = ctypes.c_void_p
D = ctypes.c_double
I = ctypes.c_int
U = ctypes.c_uint32
lib.or_record_size.restype = I
lib.or_philox4x32_10.argtypes = [vp, vp, vp]
lib.or_det_log.argtypes = [vp, vp, I]
lib.or_det_tanh.argtypes = [vp, vp, I]
lib.or_det_exp.argtypes = [vp, vp, I]
lib.or_exp_gap.argtypes = [vp, D, vp]
lib.or_xxz_weights.argtypes = [D, D, D, vp, vp, vp]
lib.or_thresholds.argtypes = [D, D, D, vp, vp]
lib.or_lattice_info.argtypes = [I, I, I, vp, vp, vp]
lib.or_lattice_bonds.argtypes = [I, I, I, vp, vp, vp]
lib.or_uf_pairs.argtypes = [I, I, vp, I, vp, vp, vp]
lib.or_looper_create.restype = vp
lib.or_looper_create.argtypes = [I, I, D, D, D, D, U, D, I, D, D]
lib.or_looper_sweep.argtypes = [vp, I, vp]
lib.or_looper_sweep_sse.argtypes = [vp, I, vp]
lib.or_looper_destroy.argtypes = [vp]
lib.or_standalone_run.argtypes = [U, D, U, U, vp, vp]
lib.or_bfs_labels.argtypes = [I, I, I, vp, vp, vp, vp]
lib.or_looper_labels.argtypes = [I, I, I, vp, vp, vp, I, vp]
lib.or_segment_of_elements.argtypes = [I, I, I, vp, vp, vp]
lib.or_segment_of_elements.restype = I
lib.or_philox_create.restype = vp
lib.or_philox_create.argtypes = [I, I, D, D, D, D, D, U, U, I, D, D]
lib.or_philox_sweep.argtypes = [vp, I, vp]
lib.or_philox_sweep_sse.argtypes = [vp, I, vp]
lib.or_philox_set_beta.argtypes = [vp, D]
lib.or_philox_num_ops.argtypes = [vp]
lib.or_philox_num_nodes.argtypes = [vp]
lib.or_philox_state.argtypes = [vp, vp, vp, vp, vp]
lib.or_philox_roots.argtypes = [vp, vp]
lib.or_philox_load.argtypes = [vp, vp, vp, vp, vp, U]
lib.or_philox_destroy.argtypes = [vp]
lib.or_permutation_test.argtypes = [U, I, I, I, vp, vp, vp, vp]
lib.or_schedule_temperature.argtypes = [I, vp, vp, I]
lib.or_schedule_temperature.restype = D
lib.or_philox_set_schedule.argtypes = [vp, I, vp, vp]
lib.or_looper_set_schedule.argtypes = [vp, I, vp, vp]

R = lib.or_record_size()
# record fields (the device library's lg_sweep_record layout)
R_NOP, R_NC, R_ENERGY, R_UMAG0 = 0, 1, 2, 3
R_BETA, R_SWEEP, R_UMAGA, R_SMAGA, R_EFIELD = 18, 19, 20, 21, 22
CHAIN, SQUARE = 0, 1
STAGE_DIAG, STAGE_REGRAPH, STAGE_FLIP = 1, 2, 3


def _p(a):
    return a.ctypes.data


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib.or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def det_tanh(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    lib.or_det_tanh(_p(x), _p(y), len(x))
    return y


def det_exp(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    lib.or_det_exp(_p(x), _p(y), len(x))
    return y


def det_log(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    lib.or_det_log(_p(x), _p(y), len(x))
    return y


def _schedule_arrays(steps, temps):
    st = np.ascontiguousarray(steps, np.int32)
    tp = np.ascontiguousarray(temps, np.float64)
    assert st.shape == tp.shape
    return st, tp


def permutation_test(seed, n, trial1, trial2):
    """the oracle's restatement of looper::random_shuffle and
    partitioned_random_shuffle driven as test/permutation.C drives them:
    (shuffles [trial1, n], conf0, conf1, perm [trial2, n])"""
    o1 = np.zeros(trial1 * n, np.int32)
    c0, c1, pm = (np.zeros(trial2 * n, np.int32) for _ in range(3))
    lib.or_permutation_test(seed, n, trial1, trial2, _p(o1), _p(c0), _p(c1), _p(pm))
    return o1.reshape(trial1, n), c0.reshape(trial2, n), c1.reshape(trial2, n), pm.reshape(trial2, n)


def schedule_temperature(steps, temps, step):
    """T(step) of an annealing schedule (looper/temperature.h:93-113)"""
    st, tp = _schedule_arrays(steps, temps)
    return lib.or_schedule_temperature(len(st), _p(st), _p(tp), int(step))


def lattice(kind, L, nv=1):
    """(sites, bonds, source, target, gauge); nv > 1: the virtual lattice of
    spin S = nv/2 (nv virtual sites per real site)"""
    N, Nb, dim = I(), I(), I()
    assert lib.or_lattice_info(kind, L, nv, ctypes.byref(N), ctypes.byref(Nb), ctypes.byref(dim)) == 0
    src = np.zeros(Nb.value, np.int32)
    tgt = np.zeros(Nb.value, np.int32)
    gauge = np.zeros(N.value, np.int32)
    lib.or_lattice_bonds(kind, L, nv, _p(src), _p(tgt), _p(gauge))
    return N.value, Nb.value, src, tgt, gauge


def xxz_weights(jxy, jz, fs=0.0):
    v = np.zeros(4)
    off = D()
    sign = I()
    lib.or_xxz_weights(jxy, jz, fs, _p(v), ctypes.byref(off), ctypes.byref(sign))
    return v, off.value, sign.value


class LooperChain:
    def __init__(self, kind, L, beta, jxy=1.0, jz=None, seed=23125, fs=0.0, field=0.0, two_s=1, hx=0.0, d=0.0):
        jz = jxy if jz is None else jz
        self.h = lib.or_looper_create(kind, L, beta, jxy, jz, fs, seed, field, two_s if two_s >= 2 else 1, hx, d)

    def sweep(self, n):
        rec = np.zeros((n, R))
        lib.or_looper_sweep(self.h, n, _p(rec))
        return rec

    def sweep_sse(self, n):
        """n sweeps of the SSE worker (sse.C); tau-integrated sums in units of the string"""
        rec = np.zeros((n, R))
        lib.or_looper_sweep_sse(self.h, n, _p(rec))
        return rec

    def set_schedule(self, steps, temps):
        st, tp = _schedule_arrays(steps, temps)
        lib.or_looper_set_schedule(self.h, len(st), _p(st), _p(tp))

    def __del__(self):
        lib.or_looper_destroy(self.h)


class PhiloxChain:
    def __init__(self, kind, L, beta, jxy=1.0, jz=None, seed=23125, chain_id=0, fs=0.0, field=0.0, two_s=1,
                 hx=0.0, d=0.0):
        jz = jxy if jz is None else jz
        self.kind, self.L = kind, L
        nv = two_s if two_s >= 2 else 1
        self.N, self.Nb, _, _, _ = lattice(kind, L, nv)
        if d != 0:
            self.Nb += (self.N // nv) * nv * (nv - 1) // 2  # in-site bonds of the single-ion term
        if hx != 0:
            self.Nb += self.N  # then one bond per site: its operators (transverse field)
        self.h = lib.or_philox_create(kind, L, beta, jxy, jz, fs, field, seed, chain_id, nv, hx, d)

    def sweep(self, n):
        rec = np.zeros((n, R))
        lib.or_philox_sweep(self.h, n, _p(rec))
        return rec

    def sweep_sse(self, n):
        """n sweeps of the SSE worker in the device's formulation (times p / n)"""
        rec = np.zeros((n, R))
        lib.or_philox_sweep_sse(self.h, n, _p(rec))
        return rec

    def set_beta(self, beta):
        lib.or_philox_set_beta(self.h, ctypes.c_double(beta))

    def set_schedule(self, steps, temps):
        st, tp = _schedule_arrays(steps, temps)
        lib.or_philox_set_schedule(self.h, len(st), _p(st), _p(tp))

    def state(self):
        n = lib.or_philox_num_ops(self.h)
        spins = np.zeros(self.N, np.int32)
        bo = np.zeros(self.Nb + 1, np.int32)
        t = np.zeros(max(n, 1))
        ty = np.zeros(max(n, 1), np.int32)
        lib.or_philox_state(self.h, _p(spins), _p(bo), _p(t), _p(ty))
        return spins, bo, t[:n], ty[:n]

    def roots(self):
        m = lib.or_philox_num_nodes(self.h)
        r = np.zeros(max(m, 1), np.int32)
        lib.or_philox_roots(self.h, _p(r))
        return r[:m]

    def load(self, spins, bo, times, types, sweep):
        spins = np.ascontiguousarray(spins, np.int32)
        bo = np.ascontiguousarray(bo, np.int32)
        times = np.ascontiguousarray(times, np.float64)
        types = np.ascontiguousarray(types, np.int32)
        lib.or_philox_load(self.h, _p(spins), _p(bo), _p(times), _p(types), sweep)

    def __del__(self):
        lib.or_philox_destroy(self.h)


def bond_of_ops(bo):
    return np.repeat(np.arange(len(bo) - 1, dtype=np.int32), np.diff(bo))


def bfs_labels(kind, L, bo, times, types):
    bond = bond_of_ops(bo)
    N = lattice(kind, L)[0]
    n = len(times)
    lab = np.zeros(N + 4 * n, np.int32)
    t = np.ascontiguousarray(times, np.float64)
    ty = np.ascontiguousarray(types, np.int32)
    lib.or_bfs_labels(kind, L, n, _p(bond), _p(t), _p(ty), _p(lab))
    return lab


def looper_labels(kind, L, bo, times, types, compress=0):
    bond = bond_of_ops(bo)
    N = lattice(kind, L)[0]
    n = len(times)
    lab = np.zeros(N + 4 * n, np.int32)
    t = np.ascontiguousarray(times, np.float64)
    ty = np.ascontiguousarray(types, np.int32)
    lib.or_looper_labels(kind, L, n, _p(bond), _p(t), _p(ty), compress, _p(lab))
    return lab


def segment_of_elements(kind, L, bo, times):
    bond = bond_of_ops(bo)
    N = lattice(kind, L)[0]
    n = len(times)
    seg = np.zeros(N + 4 * n, np.int32)
    t = np.ascontiguousarray(times, np.float64)
    M = lib.or_segment_of_elements(kind, L, n, _p(bond), _p(t), _p(seg))
    return seg, M


def expected_flip(kind, L, seed, chain_id, sweep, spins, bo, times, types, labels):
    """Flipped spins/types for canonical labels: flip bit of a cluster =
    Philox(sweep, FLIP, min segment node of the cluster, 0).x & 1."""
    N = len(spins)
    n = len(times)
    seg, M = segment_of_elements(kind, L, bo, times)
    # min segment node per cluster (cluster keyed by canonical label)
    minseg = {}
    for e in range(N + 4 * n):
        c = labels[e]
        s = seg[e]
        if c not in minseg or s < minseg[c]:
            minseg[c] = s
    fbit = {c: int(philox([sweep, STAGE_FLIP, s, 0], [seed, chain_id])[0] & 1) for c, s in minseg.items()}
    spins_out = np.array([spins[s] ^ fbit[labels[s]] for s in range(N)], np.int32)
    types_out = np.array([types[k] ^ (fbit[labels[N + 4 * k]] ^ fbit[labels[N + 4 * k + 2]]) for k in range(n)],
                         np.int32)
    return spins_out, types_out, len(minseg)


def random_config(kind, L, nops, rng, graphs=(0, 1, 2, 3), time_grid=False):
    """Random operator/cut configuration in canonical (bond, time) order.
    Any assignment of graphs defines a valid leg graph for the partition."""
    N, Nb, _, _, _ = lattice(kind, L)
    bond = np.sort(rng.integers(0, Nb, size=nops)).astype(np.int32)
    if time_grid:
        t = rng.integers(0, 16, size=nops) / 16.0  # forces equal times on a site
    else:
        t = rng.integers(0, 1 << 52, size=nops, dtype=np.int64).astype(np.float64) * 2.0 ** -52
    order = np.lexsort((t, bond))
    bond = bond[order]
    t = t[order]
    g = rng.choice(np.array(graphs), size=nops)
    off = rng.integers(0, 2, size=nops)
    types = (off | (g << 2)).astype(np.int32)
    bo = np.zeros(Nb + 1, np.int32)
    np.add.at(bo, bond + 1, 1)
    bo = np.cumsum(bo).astype(np.int32)
    spins = rng.integers(0, 2, size=N).astype(np.int32)
    return spins, bo, t, types
