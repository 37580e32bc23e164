"""ctypes binding of libloopgpu.so (the C-ABI declared in include/loopgpu.h).

The library is the product: everything here only marshals arguments.  If the
shared object has not been built, importing this module raises -- there is no
fallback path.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libloopgpu.so")

CHAIN = 0
SQUARE = 1

REC_FIELDS = [
    "n_op", "n_clusters", "energy", "umag0", "usize2", "umag2", "usize4", "umag4", "usize", "umag",
    "smag0", "ssize2", "smag2", "ssize4", "smag4", "ssize", "smag", "w2", "beta", "sweep",
    "umag_a", "smag_a", "e_field", "sign",
]
REC_SIZE = len(REC_FIELDS)

EXPORTED = [
    "lg_last_error", "lg_device_count",
    "lg_chain_create", "lg_chain_sweep", "lg_chain_info", "lg_chain_get_config",
    "lg_chain_set_config", "lg_chain_destroy", "lg_chain_set_schedule", "lg_cluster_config",
    "lg_batch_create", "lg_batch_run", "lg_batch_tune", "lg_batch_set_graph",
    "lg_batch_chain_info", "lg_batch_destroy",
    "lg_chain_get_launch", "lg_chain_set_launch", "lg_batch_get_launch", "lg_batch_set_launch",
    "lg_exchange_create", "lg_exchange_state",
    "lg_chain_state_size", "lg_chain_save", "lg_chain_load", "lg_batch_add",
    "lg_profile_num_kernels", "lg_profile_kernel_name", "lg_chain_profile", "lg_batch_profile",
    "lg_copy_bandwidth",
]


class LGLaunch(ctypes.Structure):
    _fields_ = [("group", ctypes.c_int32), ("blocks", ctypes.c_int32), ("link", ctypes.c_int32),
                ("compress", ctypes.c_int32), ("flush", ctypes.c_int32)]


def _launch_get(fn, h):
    out = LGLaunch()
    _check(fn(h, ctypes.byref(out)))
    return {"group": out.group, "blocks": out.blocks, "link": out.link, "compress": out.compress, "flush": out.flush}


def _launch_set(fn, h, group, blocks, link, compress, flush):
    v = [-1 if x is None else x for x in (group, blocks, link, compress, flush)]
    _check(fn(h, ctypes.byref(LGLaunch(*v))))


class LGParams(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int32), ("L", ctypes.c_int32), ("beta", ctypes.c_double),
        ("jxy", ctypes.c_double), ("jz", ctypes.c_double), ("force_scatter", ctypes.c_double),
        ("seed", ctypes.c_uint32), ("chain_id", ctypes.c_uint32), ("h", ctypes.c_double),
        ("two_s", ctypes.c_int32), ("hx", ctypes.c_double), ("d", ctypes.c_double),
        ("algorithm", ctypes.c_int32),
    ]


class LoopGPUError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("loopgpu error %d: %s" % (code, msg))
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libloopgpu.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.lg_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.lg_device_count.argtypes = [P(ctypes.c_int)]
    lib.lg_chain_create.argtypes = [P(LGParams), P(vp)]
    lib.lg_chain_sweep.argtypes = [vp, ctypes.c_uint32, vp]
    lib.lg_chain_info.argtypes = [vp, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_uint64)]
    lib.lg_chain_get_config.argtypes = [vp, vp, vp, vp, vp]
    lib.lg_chain_set_config.argtypes = [vp, vp, vp, vp, vp, ctypes.c_uint64]
    lib.lg_chain_destroy.argtypes = [vp]
    lib.lg_chain_set_schedule.argtypes = [vp, ctypes.c_int32, vp, vp]
    lib.lg_cluster_config.argtypes = [P(LGParams), vp, vp, vp, vp, ctypes.c_uint64, vp, vp, vp, P(ctypes.c_int64)]
    lib.lg_batch_create.argtypes = [P(LGParams), ctypes.c_int32, P(vp)]
    lib.lg_batch_run.argtypes = [vp, ctypes.c_uint32, vp]
    lib.lg_batch_tune.argtypes = [vp, ctypes.c_uint32]
    lib.lg_batch_set_graph.argtypes = [vp, ctypes.c_int]
    lib.lg_batch_chain_info.argtypes = [vp, ctypes.c_int32, P(ctypes.c_int64), P(ctypes.c_uint64)]
    lib.lg_batch_destroy.argtypes = [vp]
    for fn in (lib.lg_chain_get_launch, lib.lg_chain_set_launch, lib.lg_batch_get_launch, lib.lg_batch_set_launch):
        fn.argtypes = [vp, vp]
    lib.lg_exchange_create.argtypes = [P(LGParams), vp, ctypes.c_int32, ctypes.c_uint32, P(vp)]
    lib.lg_chain_state_size.argtypes = [vp, P(ctypes.c_size_t)]
    lib.lg_chain_save.argtypes = [vp, vp, ctypes.c_size_t]
    lib.lg_chain_load.argtypes = [vp, vp, ctypes.c_size_t]
    lib.lg_batch_add.argtypes = [vp, P(LGParams), ctypes.c_int32]
    lib.lg_exchange_state.argtypes = [vp, vp, vp, vp]
    lib.lg_profile_num_kernels.argtypes = []
    lib.lg_profile_kernel_name.argtypes = [ctypes.c_int]
    lib.lg_profile_kernel_name.restype = ctypes.c_char_p
    lib.lg_chain_profile.argtypes = [vp, ctypes.c_uint32, vp]
    lib.lg_batch_profile.argtypes = [vp, ctypes.c_uint32, vp]
    lib.lg_copy_bandwidth.argtypes = [ctypes.c_uint64, ctypes.c_int, P(ctypes.c_double)]
    lib.lg_worker_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.lg_worker_create.argtypes = [ctypes.c_char_p, P(vp)]
    lib.lg_worker_run.argtypes = [vp, ctypes.c_uint64, P(ctypes.c_uint64)]
    lib.lg_worker_status.argtypes = [vp, P(ctypes.c_uint64), P(ctypes.c_int), P(ctypes.c_double), P(ctypes.c_int)]
    lib.lg_worker_result.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double)]
    lib.lg_worker_report.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, P(ctypes.c_size_t)]
    lib.lg_worker_save.argtypes = [vp, vp, P(ctypes.c_size_t)]
    lib.lg_worker_load.argtypes = [vp, vp, ctypes.c_size_t]
    lib.lg_worker_set_chunk.argtypes = [vp, ctypes.c_uint32]
    lib.lg_worker_destroy.argtypes = [vp]
    lib.lg_spin_correlations.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, ctypes.c_int32, vp, vp, vp]
    lib.lg_temperature_schedule.argtypes = [ctypes.c_char_p, ctypes.c_int32, P(ctypes.c_int32), vp, vp]
    return lib


lib = _load()


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        lib.lg_last_error(buf, 1024)
        raise LoopGPUError(rc, buf.value.decode(errors="replace"))


def device_count():
    n = ctypes.c_int(0)
    _check(lib.lg_device_count(ctypes.byref(n)))
    return n.value


def copy_bandwidth(nbytes, iters=20):
    """achieved GB/s (read + write) of a plain device copy of nbytes"""
    g = ctypes.c_double()
    _check(lib.lg_copy_bandwidth(int(nbytes), int(iters), ctypes.byref(g)))
    return g.value


def kernel_names():
    return [lib.lg_profile_kernel_name(i).decode() for i in range(lib.lg_profile_num_kernels())]


PATH_INTEGRAL, SSE = 0, 1   # lg_params.algorithm: "loop; path integral", "loop; sse"


def params(lattice, L, beta, jxy=1.0, jz=None, seed=23125, chain_id=0, force_scatter=0.0, h=0.0, two_s=1,
           hx=0.0, d=0.0, algorithm=PATH_INTEGRAL):
    if jz is None:
        jz = jxy
    return LGParams(int(lattice), int(L), float(beta), float(jxy), float(jz), float(force_scatter),
                    int(seed), int(chain_id), float(h), int(two_s), float(hx), float(d), int(algorithm))


def lattice_size(lattice, L, two_s=1, hx=0.0, d=0.0):
    """(sites, bonds) the device sweeps: virtual ones for S > 1/2 (2S per real site); a
    single-ion term adds the in-site bonds of each real site's virtual sites, and a
    transverse field one bond per site, after the lattice's, for the site operators"""
    nv = two_s if two_s >= 2 else 1
    n, nb = (nv * L, nv * nv * L) if lattice == CHAIN else (nv * L * L, nv * nv * 2 * L * L)
    if d != 0:
        nb += (n // nv) * nv * (nv - 1) // 2
    return n, nb + (n if hx != 0 else 0)


def _ptr(a):
    return None if a is None else a.ctypes.data


class Chain:
    """One chain on the plain one-chain-per-launch path."""

    def __init__(self, p):
        self.p = p
        self.h = ctypes.c_void_p()
        _check(lib.lg_chain_create(ctypes.byref(p), ctypes.byref(self.h)))
        self.n_sites, self.n_bonds = lattice_size(p.lattice, p.L, p.two_s, p.hx, p.d)

    def sweep(self, n, records=True):
        rec = np.zeros((n, REC_SIZE)) if records else None
        _check(lib.lg_chain_sweep(self.h, n, _ptr(rec)))
        return rec

    def profile(self, n):
        """average per-sweep duration (ms) of each sweep kernel over n sweeps"""
        k = lib.lg_profile_num_kernels()
        ms = np.zeros(k)
        _check(lib.lg_chain_profile(self.h, n, _ptr(ms)))
        return dict(zip(kernel_names(), ms))

    def save(self):
        """checkpoint blob (include/loopgpu.h lg_chain_save)"""
        n = ctypes.c_size_t()
        _check(lib.lg_chain_state_size(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib.lg_chain_save(self.h, buf, n.value))
        return buf.raw

    def load(self, blob):
        _check(lib.lg_chain_load(self.h, blob, len(blob)))

    def launch(self):
        """current launch settings (include/loopgpu.h lg_launch)"""
        return _launch_get(lib.lg_chain_get_launch, self.h)

    def set_launch(self, group=None, blocks=None, link=None, compress=None, flush=None):
        _launch_set(lib.lg_chain_set_launch, self.h, group, blocks, link, compress, flush)

    def info(self):
        ns, nb, no, sw = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64(), ctypes.c_uint64()
        _check(lib.lg_chain_info(self.h, ctypes.byref(ns), ctypes.byref(nb), ctypes.byref(no), ctypes.byref(sw)))
        return ns.value, nb.value, no.value, sw.value

    def get_config(self):
        _, nb, nops, _ = self.info()
        spins = np.zeros(self.n_sites, np.int32)
        bo = np.zeros(nb + 1, np.int32)
        t = np.zeros(max(nops, 1), np.float64)
        ty = np.zeros(max(nops, 1), np.int32)
        _check(lib.lg_chain_get_config(self.h, _ptr(spins), _ptr(bo), _ptr(t), _ptr(ty)))
        return spins, bo, t[:nops], ty[:nops]

    def set_config(self, spins, bo, times, types, sweep):
        spins = np.ascontiguousarray(spins, np.int32)
        bo = np.ascontiguousarray(bo, np.int32)
        times = np.ascontiguousarray(times, np.float64)
        types = np.ascontiguousarray(types, np.int32)
        _check(lib.lg_chain_set_config(self.h, _ptr(spins), _ptr(bo), _ptr(times), _ptr(types), int(sweep)))

    def set_schedule(self, steps, temperatures):
        """annealing schedule (include/loopgpu.h lg_chain_set_schedule)"""
        st = np.ascontiguousarray(steps, np.int32)
        tp = np.ascontiguousarray(temperatures, np.float64)
        if st.shape != tp.shape:
            raise LoopGPUError(-1, "steps and temperatures differ in length")
        _check(lib.lg_chain_set_schedule(self.h, len(st), _ptr(st) if len(st) else None, _ptr(tp) if len(tp) else None))

    def close(self):
        if self.h:
            lib.lg_chain_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many chains packed into one launch per sweep stage."""

    def __init__(self, plist):
        arr = (LGParams * len(plist))(*plist)
        self.n = len(plist)
        self.h = ctypes.c_void_p()
        _check(lib.lg_batch_create(arr, self.n, ctypes.byref(self.h)))

    def run(self, n, records=True):
        rec = np.zeros((n, self.n, REC_SIZE)) if records else None
        _check(lib.lg_batch_run(self.h, n, _ptr(rec)))
        return rec

    def tune(self, sweeps_per_trial=8):
        _check(lib.lg_batch_tune(self.h, sweeps_per_trial))

    def add(self, plist):
        """append chains (starting from all-up spins, no operators)"""
        arr = (LGParams * len(plist))(*plist)
        _check(lib.lg_batch_add(self.h, arr, len(plist)))
        self.n += len(plist)

    def launch(self):
        return _launch_get(lib.lg_batch_get_launch, self.h)

    def set_launch(self, group=None, blocks=None, link=None, compress=None, flush=None):
        _launch_set(lib.lg_batch_set_launch, self.h, group, blocks, link, compress, flush)

    def set_graph(self, enable):
        _check(lib.lg_batch_set_graph(self.h, 1 if enable else 0))

    def profile(self, n):
        k = lib.lg_profile_num_kernels()
        ms = np.zeros(k)
        _check(lib.lg_batch_profile(self.h, n, _ptr(ms)))
        return dict(zip(kernel_names(), ms))

    def chain_info(self, c):
        no, sw = ctypes.c_int64(), ctypes.c_uint64()
        _check(lib.lg_batch_chain_info(self.h, c, ctypes.byref(no), ctypes.byref(sw)))
        return no.value, sw.value

    def close(self):
        if self.h:
            lib.lg_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



class Exchange(Batch):
    """Replica exchange over inverse temperatures `betas` (include/loopgpu.h
    lg_exchange_create): records of run() are per chain; their `beta` field
    is the temperature the chain had in that sweep."""

    def __init__(self, p, betas, seed=0):
        self.n = len(betas)
        self.betas = np.asarray(betas, dtype=np.float64)
        self.h = ctypes.c_void_p()
        _check(lib.lg_exchange_create(ctypes.byref(p), _ptr(self.betas), self.n, seed, ctypes.byref(self.h)))

    def state(self):
        at = np.zeros(self.n, np.int32)
        acc = np.zeros(self.n - 1, np.uint64)
        steps = ctypes.c_uint64()
        _check(lib.lg_exchange_state(self.h, _ptr(at), _ptr(acc), ctypes.byref(steps)))
        return at, acc, steps.value

def cluster_config(p, spins, bo, times, types, sweep):
    """Cluster stage on a given configuration -> (labels, spins_out, types_out, n_clusters)."""
    spins = np.ascontiguousarray(spins, np.int32)
    bo = np.ascontiguousarray(bo, np.int32)
    times = np.ascontiguousarray(times, np.float64)
    types = np.ascontiguousarray(types, np.int32)
    n_sites = len(spins)
    nops = int(bo[-1])
    labels = np.zeros(n_sites + 4 * nops, np.int32)
    spo = np.zeros(n_sites, np.int32)
    tyo = np.zeros(max(nops, 1), np.int32)
    nc = ctypes.c_int64()
    _check(lib.lg_cluster_config(ctypes.byref(p), _ptr(spins), _ptr(bo), _ptr(times), _ptr(types),
                                 int(sweep), _ptr(labels), _ptr(spo), _ptr(tyo), ctypes.byref(nc)))
    return labels, spo, tyo[:nops], nc.value


def _wcheck(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        lib.lg_worker_last_error(buf, 1024)
        raise LoopGPUError(rc, buf.value.decode(errors="replace"))


class Worker:
    """The reference worker's parameter / run / measure interface
    (include/loopgpu_worker.h)."""

    def __init__(self, params):
        if isinstance(params, dict):
            params = "\n".join("%s = %s" % (k, ('"%s"' % v) if isinstance(v, str) else v) for k, v in params.items())
        self.h = ctypes.c_void_p()
        _wcheck(lib.lg_worker_create(params.encode(), ctypes.byref(self.h)))

    def run(self, max_steps=1 << 62):
        done = ctypes.c_uint64()
        _wcheck(lib.lg_worker_run(self.h, max_steps, ctypes.byref(done)))
        return done.value

    def status(self):
        mcs, th, pr, cw = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        _wcheck(lib.lg_worker_status(self.h, ctypes.byref(mcs), ctypes.byref(th), ctypes.byref(pr), ctypes.byref(cw)))
        return {"mcs": mcs.value, "thermalized": bool(th.value), "progress": pr.value, "can_work": bool(cw.value)}

    def result(self, name):
        m, e, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _wcheck(lib.lg_worker_result(self.h, name.encode(), ctypes.byref(m), ctypes.byref(e), ctypes.byref(t)))
        return m.value, e.value, t.value

    def report(self):
        need = ctypes.c_size_t()
        _wcheck(lib.lg_worker_report(self.h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _wcheck(lib.lg_worker_report(self.h, buf, need.value, None))
        return buf.value.decode()

    def save(self):
        n = ctypes.c_size_t(0)
        _wcheck(lib.lg_worker_save(self.h, None, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _wcheck(lib.lg_worker_save(self.h, buf, ctypes.byref(n)))
        return buf.raw[:n.value]

    def load(self, blob):
        _wcheck(lib.lg_worker_load(self.h, blob, len(blob)))

    def set_chunk(self, n):
        _wcheck(lib.lg_worker_set_chunk(self.h, n))

    def close(self):
        if self.h:
            lib.lg_worker_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_jobs", ctypes.c_int32), ("max_sites", ctypes.c_int64), ("max_bonds", ctypes.c_int64),
                ("max_ops", ctypes.c_int64), ("max_cand", ctypes.c_int64), ("max_chunk", ctypes.c_uint32)]


class Job(ctypes.Structure):
    _fields_ = [("params", LGParams), ("thermalization", ctypes.c_uint32), ("sweeps", ctypes.c_uint32)]


JOB_NOBS = 7
JOB_OBS = ["Energy Density", "Susceptibility", "Staggered Structure Factor", "Staggered Susceptibility",
           "Number of Clusters", "Stiffness", "Operators"]
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_CANCELLED = 0, 1, 2, 3


class JobResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("sweeps_done", ctypes.c_uint64), ("latency_s", ctypes.c_double),
                ("queue_s", ctypes.c_double), ("mean", ctypes.c_double * JOB_NOBS),
                ("error", ctypes.c_double * JOB_NOBS)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("done", ctypes.c_uint64), ("running", ctypes.c_uint64),
                ("queued", ctypes.c_uint64), ("p50_latency_s", ctypes.c_double), ("p99_latency_s", ctypes.c_double),
                ("mean_latency_s", ctypes.c_double), ("updates_per_s", ctypes.c_double), ("busy_s", ctypes.c_double),
                ("total_updates", ctypes.c_uint64), ("sweep_launches", ctypes.c_uint64),
                ("preemptions", ctypes.c_uint64), ("graph_captures", ctypes.c_uint64),
                ("graph_capture_s", ctypes.c_double)]


def _server_bind():
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.lg_server_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.lg_server_start.argtypes = [P(ServerConfig), P(vp)]
    lib.lg_server_submit.argtypes = [vp, P(Job), P(ctypes.c_uint64)]
    lib.lg_server_poll.argtypes = [vp, ctypes.c_uint64, P(JobResult)]
    lib.lg_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int64, P(JobResult)]
    lib.lg_server_cancel.argtypes = [vp, ctypes.c_uint64]
    lib.lg_server_preempt.argtypes = [vp, ctypes.c_uint64]
    lib.lg_server_stats_get.argtypes = [vp, P(ServerStats)]
    lib.lg_server_stop.argtypes = [vp]


_server_bind()
EXPORTED += ["lg_server_last_error", "lg_server_start", "lg_server_submit", "lg_server_poll", "lg_server_wait",
             "lg_server_cancel", "lg_server_preempt", "lg_server_stats_get", "lg_server_stop",
             "lg_worker_last_error", "lg_worker_create", "lg_worker_run", "lg_worker_status", "lg_worker_result",
             "lg_worker_report", "lg_worker_save", "lg_worker_load", "lg_worker_set_chunk", "lg_worker_destroy",
             "lg_spin_correlations", "lg_temperature_schedule"]


def spin_correlations(kind, L, spins, origin=-1):
    """Correlation estimators of one tau = 0 configuration (include/loopgpu_worker.h
    lg_spin_correlations): (Spin Correlations, Staggered Spin Correlations,
    Spin Structure Factor), N values each."""
    sp = np.ascontiguousarray(spins, dtype=np.uint8)
    n = L if kind == CHAIN else L * L
    if sp.size != n:
        raise LoopGPUError(-1, "spins must hold %d sites" % n)
    u, s, f = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = lib.lg_spin_correlations(kind, L, sp.ctypes.data, origin, u.ctypes.data, s.ctypes.data, f.ctypes.data)
    if rc != 0:
        raise LoopGPUError(rc, "lg_spin_correlations: invalid arguments")
    return u, s, f


def temperature_schedule(params):
    """(steps, temperatures) of the schedule a parameter text defines
    (include/loopgpu_worker.h lg_temperature_schedule)"""
    n = ctypes.c_int32()
    cap = 64
    st = np.zeros(cap, np.int32)
    tp = np.zeros(cap, np.float64)
    _wcheck(lib.lg_temperature_schedule(params.encode(), cap, ctypes.byref(n), st.ctypes.data, tp.ctypes.data))
    return st[:n.value].copy(), tp[:n.value].copy()


def _scheck(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        lib.lg_server_last_error(buf, 1024)
        raise LoopGPUError(rc, buf.value.decode(errors="replace"))


class Server:
    """Continuous-batching job server (include/loopgpu_server.h)."""

    def __init__(self, max_jobs=64, max_sites=1 << 20, max_bonds=1 << 21, max_ops=64 << 20, max_cand=64 << 20,
                 max_chunk=64):
        cfg = ServerConfig(max_jobs, max_sites, max_bonds, max_ops, max_cand, max_chunk)
        self.h = ctypes.c_void_p()
        _scheck(lib.lg_server_start(ctypes.byref(cfg), ctypes.byref(self.h)))

    def submit(self, p, sweeps, thermalization=None):
        if thermalization is None:
            thermalization = sweeps // 8
        j = Job(p, thermalization, sweeps)
        jid = ctypes.c_uint64()
        _scheck(lib.lg_server_submit(self.h, ctypes.byref(j), ctypes.byref(jid)))
        return jid.value

    @staticmethod
    def _res(r):
        return {"status": r.status, "sweeps_done": r.sweeps_done, "latency_s": r.latency_s, "queue_s": r.queue_s,
                "mean": dict(zip(JOB_OBS, list(r.mean))), "error": dict(zip(JOB_OBS, list(r.error)))}

    def poll(self, jid):
        r = JobResult()
        _scheck(lib.lg_server_poll(self.h, jid, ctypes.byref(r)))
        return self._res(r)

    def wait(self, jid, timeout_s=600.0):
        r = JobResult()
        _scheck(lib.lg_server_wait(self.h, jid, int(timeout_s * 1000), ctypes.byref(r)))
        return self._res(r)

    def cancel(self, jid):
        _scheck(lib.lg_server_cancel(self.h, jid))

    def preempt(self, jid):
        """save the running job at the next chunk boundary and queue it again"""
        _scheck(lib.lg_server_preempt(self.h, jid))

    def stats(self):
        s = ServerStats()
        _scheck(lib.lg_server_stats_get(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def stop(self):
        if self.h:
            lib.lg_server_stop(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
