"""ctypes binding of libcosims_engine.so (the C ABI in include/cosims_engine.h).

There is no fallback: if the HIP library is missing or no gfx950 device is
present the calls raise.  Arrays are numpy float64 / int32.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcosims_engine.so")
_LIB = None

PREC_FP64 = 0
PREC_MIXED = 1
RNG_PHILOX = 0
RNG_REFERENCE = 1
MAX_ROUNDS = 4
NOALLOC_MAX_TRAJ = 262144  # cosims_run_noalloc's limits (COSIMS_NOALLOC_MAX_*)
NOALLOC_MAX_ITER = 256

EXPORTS = [
    "cosims_params_default", "cosims_params_read", "cosims_molfile_read", "cosims_molfile_free",
    "cosims_mol_create", "cosims_mol_create_prepared", "cosims_mol_destroy", "cosims_mol_upload", "cosims_mol_info_get",
    "cosims_mol_atoms", "cosims_mol_clusters", "cosims_accel", "cosims_potential", "cosims_sample",
    "cosims_sample_reference",
    "cosims_verlet", "cosims_run", "cosims_run_noalloc", "cosims_run_batch", "cosims_tune_batch", "cosims_strerror",
    "cosims_last_error", "cosims_device_info", "cosims_device_sync", "cosims_device_fma_rate", "cosims_force_rate",
    "cosims_server_create", "cosims_server_submit", "cosims_server_wait", "cosims_server_stats_get",
    "cosims_server_destroy",
]
E_PENDING = -8
SERVE_SHORTEST_FIRST = 0
SERVE_EQUAL_SHARE = 1


class CosimsError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("temp", C.c_double), ("dt", C.c_double), ("traj", C.c_int64), ("iter", C.c_int32),
                ("charge", C.c_int32), ("seed", C.c_uint64), ("multipole", C.c_int32),
                ("dispersion_cutoff", C.c_int32), ("dispersion_radius", C.c_double),
                ("multipole_radius", C.c_double), ("multipole_order", C.c_int32),
                ("max_cluster_size", C.c_int32), ("projection", C.c_int32), ("print_rate", C.c_int32),
                ("total_charge", C.c_double), ("threads", C.c_int32), ("name", C.c_char * 128),
                ("setup_gpu", C.c_int32)]


class AtomType(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("symbol", C.c_char * 8), ("mass", C.c_double),
                ("sigma", C.c_double), ("epsilon", C.c_double), ("int_mass", C.c_int32), ("count", C.c_int32)]


class MolFile(C.Structure):
    _fields_ = [("n", C.c_int32), ("xyz", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_double)),
                ("sigma", C.POINTER(C.c_double)), ("epsilon", C.POINTER(C.c_double)),
                ("charge", C.POINTER(C.c_double)), ("charged", C.c_int32), ("total_charge", C.c_double),
                ("sigma_max", C.c_double), ("name", C.c_char * 128), ("ntypes", C.c_int32),
                ("types", C.POINTER(AtomType))]


class Atoms(C.Structure):
    _fields_ = [("n", C.c_int32), ("xyz", C.c_void_p), ("mass", C.c_void_p), ("sigma", C.c_void_p),
                ("epsilon", C.c_void_p), ("charge", C.c_void_p), ("sigma_max", C.c_double)]


class Prepared(C.Structure):
    _fields_ = [("n", C.c_int32), ("xyz", C.c_void_p), ("charge", C.c_void_p), ("sig6", C.c_void_p),
                ("eps24m", C.c_void_p), ("eps4", C.c_void_p), ("mu", C.c_double), ("a", C.c_double),
                ("b", C.c_double), ("c", C.c_double), ("ncl", C.c_int32), ("cl_start", C.c_void_p),
                ("cl_atoms", C.c_void_p), ("cl_center", C.c_void_p), ("cl_maxdist", C.c_void_p),
                ("dispersion_radius", C.c_double), ("multipole_radius", C.c_double)]


class MolInfo(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("n", "ncl", "force_type", "charged", "projection",
                                          "multipole_order")] + \
               [(k, C.c_double) for k in ("mu", "total_mass", "dipole_scalar", "a", "b", "c", "bmax",
                                           "enlarge", "gmin", "gmax", "gmean", "gmode", "gsigma", "vd_a",
                                           "vd_norm", "max_val", "cdf_frac", "ccs_const", "min_kin",
                                           "dispersion_radius", "multipole_radius")] + \
               [("rot", C.c_double * 9), ("max_xyz", C.c_double * 3), ("fit_abc", C.c_double * 3),
                ("fit_rmsd", C.c_double), ("n_boundary", C.c_int32), ("setup_ms", C.c_double),
                ("setup_boundary_ms", C.c_double), ("setup_cluster_ms", C.c_double),
                ("setup_moments_ms", C.c_double)]


class RunCfg(C.Structure):
    _fields_ = [("prec", C.c_int32), ("job", C.c_uint32), ("iter", C.c_int32), ("traj", C.c_int64),
                ("lanes", C.c_int32), ("wg_size", C.c_int32), ("blocks_per_cu", C.c_int32), ("rng", C.c_int32),
                ("chi", C.c_void_p), ("status", C.c_void_p), ("g", C.c_void_p)]


class Ccs(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("mean", "sd", "wmean", "werr", "max_ergy")] + \
               [(k, C.c_int64) for k in ("n_traj", "n_hit", "n_fail", "n_retry", "total_steps",
                                          "pair_evals", "pot_evals")] + \
               [("run_ms", C.c_double), ("n_rounds", C.c_int32),
                ("traj_kernel_ms", C.c_double * MAX_ROUNDS), ("round_steps", C.c_int64 * MAX_ROUNDS),
                ("round_evals", C.c_int64 * MAX_ROUNDS), ("round_items", C.c_int64 * MAX_ROUNDS), ("check_kernel_ms", C.c_double),
                ("sample_kernel_ms", C.c_double), ("lanes", C.c_int32), ("wg_size", C.c_int32),
                ("cl_counts", C.c_int64 * 6), ("ref_flops", C.c_double)]


class BatchJob(C.Structure):
    _fields_ = [("mol", C.c_void_p), ("seed", C.c_uint64), ("job", C.c_uint32), ("iter", C.c_int32),
                ("traj", C.c_int64), ("omega", C.c_void_p), ("omega_err", C.c_void_p), ("out", C.POINTER(Ccs))]


SIZE_CLASSES = 5


class BatchCfg(C.Structure):
    _fields_ = [("prec", C.c_int32), ("lanes", C.c_int32), ("wg_size", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("class_lanes", C.c_int32 * SIZE_CLASSES), ("class_wg", C.c_int32 * SIZE_CLASSES),
                ("own_launches", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("wg_size", C.c_int32), ("launches", C.c_int32),
                ("class_lanes", C.c_int32 * SIZE_CLASSES), ("class_wg", C.c_int32 * SIZE_CLASSES),
                ("kernel_ms", C.c_double), ("sample_ms", C.c_double), ("run_ms", C.c_double)] + \
               [(k, C.c_int64) for k in ("n_traj", "n_hit", "total_steps", "force_evals", "pot_evals",
                                          "pair_evals")] + [("ref_flops", C.c_double)]


class ServerCfg(C.Structure):
    _fields_ = [("prec", C.c_int32), ("lanes", C.c_int32), ("wg_size", C.c_int32), ("hbm_budget", C.c_int64),
                ("max_active", C.c_int32), ("min_blocks", C.c_int32), ("max_blocks", C.c_int32),
                ("setup_threads", C.c_int32), ("block_traj", C.c_int64), ("slice_traj", C.c_int64),
                ("slice_ms", C.c_double), ("policy", C.c_int32), ("min_slice_traj", C.c_int64),
                ("pipeline_depth", C.c_int32), ("issue_fraction", C.c_double),
                ("prior_rel_sd", C.c_double), ("class_lanes", C.c_int32 * SIZE_CLASSES),
                ("class_wg", C.c_int32 * SIZE_CLASSES), ("own_launches", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("params", Params), ("atoms", Atoms), ("job", C.c_uint32), ("target_rel_err", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("blocks", C.c_int32), ("ccs", Ccs), ("rel_err", C.c_double),
                ("t_submit_ms", C.c_double), ("t_ready_ms", C.c_double), ("t_admit_ms", C.c_double),
                ("t_done_ms", C.c_double), ("setup_ms", C.c_double), ("block_traj", C.c_int64),
                ("setup_boundary_ms", C.c_double), ("setup_cluster_ms", C.c_double),
                ("setup_moments_ms", C.c_double), ("setup_upload_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "done", "failed", "queued", "active", "slices", "launches",
                                          "traj", "pair_evals", "force_evals", "total_steps")] + \
               [(k, C.c_double) for k in ("kernel_ms", "busy_ms", "now_ms")] + \
               [(k, C.c_int64) for k in ("bytes_in_use", "bytes_peak", "budget")] + [("ref_flops", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CosimsError(f"{LIB_PATH} is not built; run `python -m cosims_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.cosims_params_default.argtypes = [C.POINTER(Params)]
        L.cosims_params_read.argtypes = [C.c_char_p, C.POINTER(Params)]
        L.cosims_molfile_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.POINTER(MolFile))]
        L.cosims_molfile_free.argtypes = [C.POINTER(MolFile)]
        L.cosims_mol_create.argtypes = [C.POINTER(Params), C.POINTER(Atoms), C.POINTER(vp)]
        L.cosims_mol_create_prepared.argtypes = [C.POINTER(Params), C.POINTER(Prepared), C.POINTER(vp)]
        L.cosims_mol_destroy.argtypes = [vp]
        L.cosims_mol_destroy.restype = None
        L.cosims_mol_upload.argtypes = [vp]
        L.cosims_mol_info_get.argtypes = [vp, C.POINTER(MolInfo)]
        L.cosims_mol_atoms.argtypes = [vp] + [vp] * 5
        L.cosims_mol_clusters.argtypes = [vp] + [vp] * 9
        L.cosims_accel.argtypes = [vp, C.c_int32, C.c_int64, vp, vp]
        L.cosims_potential.argtypes = [vp, C.c_int64, vp, vp]
        L.cosims_sample.argtypes = [vp, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.c_int64,
                                    vp, vp, vp, vp]
        L.cosims_sample_reference.argtypes = [vp, C.c_uint64, C.c_int32, C.c_int64, vp, vp, vp, vp]
        L.cosims_verlet.argtypes = [vp, C.c_int32, C.c_int64, vp, vp, vp, vp, vp, vp, vp]
        L.cosims_run.argtypes = [vp, C.POINTER(Params), C.POINTER(RunCfg), vp, vp, C.POINTER(Ccs)]
        L.cosims_run_noalloc.argtypes = [vp, C.POINTER(Params), C.POINTER(RunCfg), vp, vp, C.POINTER(Ccs)]
        L.cosims_run_batch.argtypes = [C.POINTER(BatchJob), C.c_int32, C.POINTER(BatchCfg), C.POINTER(BatchStats)]
        L.cosims_server_create.argtypes = [C.POINTER(ServerCfg), C.POINTER(vp)]
        L.cosims_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
        L.cosims_server_wait.argtypes = [vp, C.c_int64, C.c_double, C.POINTER(Result), vp, vp, C.c_int32]
        L.cosims_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
        L.cosims_server_destroy.argtypes = [vp]
        L.cosims_server_destroy.restype = None
        L.cosims_tune_batch.argtypes = [C.POINTER(BatchJob), C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]
        L.cosims_strerror.restype = C.c_char_p
        L.cosims_strerror.argtypes = [C.c_int]
        L.cosims_last_error.restype = C.c_char_p
        L.cosims_device_info.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_int32)]
        L.cosims_device_fma_rate.argtypes = [C.POINTER(C.c_double)]
        L.cosims_force_rate.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        L = lib()
        raise CosimsError(f"{L.cosims_strerror(rc).decode()} ({rc}): {L.cosims_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_params(**kw) -> Params:
    p = Params()
    lib().cosims_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "name":
            v = v.encode()
        setattr(p, k, v)
    return p


def read_params(path, base: Params | None = None) -> Params:
    p = base if base is not None else default_params()
    _check(lib().cosims_params_read(path.encode(), C.byref(p)))
    return p


def read_molfile(mol_path, atoms_path=None) -> dict:
    ptr = C.POINTER(MolFile)()
    _check(lib().cosims_molfile_read(mol_path.encode(), atoms_path.encode() if atoms_path else None,
                                     C.byref(ptr)))
    f = ptr.contents
    n = f.n
    out = dict(n=n,
               xyz=np.ctypeslib.as_array(f.xyz, (3 * n,)).copy().reshape(n, 3),
               mass=np.ctypeslib.as_array(f.mass, (n,)).copy(),
               sigma=np.ctypeslib.as_array(f.sigma, (n,)).copy(),
               epsilon=np.ctypeslib.as_array(f.epsilon, (n,)).copy(),
               charge=np.ctypeslib.as_array(f.charge, (n,)).copy(),
               charged=bool(f.charged), total_charge=f.total_charge, sigma_max=f.sigma_max,
               name=f.name.decode(),
               types=[dict(name=t.name.decode(), symbol=t.symbol.decode(), mass=t.mass, sigma=t.sigma,
                           epsilon=t.epsilon, int_mass=t.int_mass, count=t.count)
                      for t in (f.types[k] for k in range(f.ntypes))])
    lib().cosims_molfile_free(ptr)
    return out


def device_sync():
    _check(lib().cosims_device_sync())


def device_fma_rate() -> float:
    """Measured FP32 FMA rate (TFLOP/s) of the device for code built like the engine."""
    t = C.c_double()
    _check(lib().cosims_device_fma_rate(C.byref(t)))
    return t.value


def device_info():
    buf = C.create_string_buffer(256)
    cus = C.c_int32()
    _check(lib().cosims_device_info(buf, 256, C.byref(cus)))
    return buf.value.decode(), cus.value


class Molecule:
    """A prepared molecule resident on the GPU (one cosims_mol handle)."""

    def __init__(self, handle, params: Params):
        self._h = handle
        self.params = params

    @classmethod
    def create(cls, params: Params, xyz, mass, sigma, epsilon, charge=None, sigma_max=3.5):
        xyz, mass, sigma, epsilon = (_f64(a) for a in (xyz, mass, sigma, epsilon))
        charge = None if charge is None else _f64(charge)
        a = Atoms(n=mass.size, xyz=xyz.ctypes.data, mass=mass.ctypes.data, sigma=sigma.ctypes.data,
                  epsilon=epsilon.ctypes.data, charge=_p(charge), sigma_max=sigma_max)
        h = C.c_void_p()
        _check(lib().cosims_mol_create(C.byref(params), C.byref(a), C.byref(h)))
        return cls(h, params)

    @classmethod
    def from_file(cls, mol_path, atoms_path=None, params: Params | None = None, **kw):
        f = read_molfile(mol_path, atoms_path)
        if params is None:
            params = default_params()
        params.charge = int(f["charged"])
        params.total_charge = f["total_charge"]
        for k, v in kw.items():
            setattr(params, k, v)
        return cls.create(params, f["xyz"], f["mass"], f["sigma"], f["epsilon"],
                          f["charge"] if f["charged"] else None, f["sigma_max"])

    @classmethod
    def from_prepared(cls, params: Params, xyz, charge, sig6, eps24m, eps4, mu, abc, clusters=None,
                      dispersion_radius=0.0, multipole_radius=0.0):
        arrs = [_f64(a) for a in (xyz, charge, sig6, eps24m, eps4)]
        s = Prepared(n=arrs[1].size, xyz=arrs[0].ctypes.data, charge=arrs[1].ctypes.data,
                     sig6=arrs[2].ctypes.data, eps24m=arrs[3].ctypes.data, eps4=arrs[4].ctypes.data,
                     mu=mu, a=abc[0], b=abc[1], c=abc[2], dispersion_radius=dispersion_radius,
                     multipole_radius=multipole_radius)
        keep = [arrs]
        if clusters is not None:
            st = np.ascontiguousarray(clusters["start"], dtype=np.int32)
            at = np.ascontiguousarray(clusters["atoms"], dtype=np.int32)
            ce = _f64(clusters["center"]); md = _f64(clusters["maxdist"])
            keep.append((st, at, ce, md))
            s.ncl = md.size
            s.cl_start, s.cl_atoms, s.cl_center, s.cl_maxdist = (st.ctypes.data, at.ctypes.data,
                                                                 ce.ctypes.data, md.ctypes.data)
        h = C.c_void_p()
        _check(lib().cosims_mol_create_prepared(C.byref(params), C.byref(s), C.byref(h)))
        return cls(h, params)

    def upload(self):
        _check(lib().cosims_mol_upload(self._h))

    def close(self):
        if self._h:
            lib().cosims_mol_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self) -> dict:
        i = MolInfo()
        _check(lib().cosims_mol_info_get(self._h, C.byref(i)))
        out = {}
        for k, _ in MolInfo._fields_:
            v = getattr(i, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out

    def atoms(self) -> dict:
        n = self.info()["n"]
        xyz = np.empty((n, 3)); q, s6, e24, e4 = (np.empty(n) for _ in range(4))
        _check(lib().cosims_mol_atoms(self._h, _p(xyz), _p(q), _p(s6), _p(e24), _p(e4)))
        return dict(xyz=xyz, q=q, sig6=s6, eps24m=e24, eps4=e4)

    def clusters(self) -> dict | None:
        inf = self.info()
        n, c = inf["n"], inf["ncl"]
        if c == 0:
            return None
        st = np.empty(c + 1, np.int32); at = np.empty(n, np.int32); ce = np.empty(3 * c)
        md = np.empty(c); pos = np.empty(3 * c); q = np.empty(c); dip = np.empty(3 * c)
        quad = np.empty(9 * c); rep = np.empty(c, np.int32)
        _check(lib().cosims_mol_clusters(self._h, *[_p(a) for a in (st, at, ce, md, pos, q, dip, quad, rep)]))
        return dict(start=st, atoms=at, center=ce, maxdist=md, pos=pos, q=q, dip=dip, quad=quad,
                    replace=rep)

    def force_rate(self, prec=PREC_MIXED, lanes=16, wg_size=256):
        """Measured rate of the force sum alone (the trajectory kernels'
        ceiling): dict(tflops=reference TFLOP/s, pairs_per_s, kernel_ms)."""
        t, p, ms = C.c_double(), C.c_double(), C.c_double()
        _check(lib().cosims_force_rate(self._h, prec, lanes, wg_size, C.byref(t), C.byref(p), C.byref(ms)))
        return dict(tflops=t.value, pairs_per_s=p.value, kernel_ms=ms.value)

    def accel(self, pos, prec=PREC_FP64):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty_like(pos)
        _check(lib().cosims_accel(self._h, prec, pos.shape[0], _p(pos), _p(out)))
        return out

    def potential(self, pos):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty(pos.shape[0])
        _check(lib().cosims_potential(self._h, pos.shape[0], _p(pos), _p(out)))
        return out

    def sample(self, seed, job, iteration, first, n):
        p0 = np.empty((n, 3)); v0 = np.empty((n, 3)); g = np.empty(n); hit = np.empty(n, np.int32)
        _check(lib().cosims_sample(self._h, seed, job, iteration, first, n, _p(p0), _p(v0), _p(g), _p(hit)))
        return dict(p0=p0, v0=v0, g=g, hit=hit)

    def sample_reference(self, seed, iters, traj):
        n = iters * traj
        p0 = np.empty((n, 3)); v0 = np.empty((n, 3)); g = np.empty(n); hit = np.empty(n, np.int32)
        _check(lib().cosims_sample_reference(self._h, seed, iters, traj, _p(p0), _p(v0), _p(g), _p(hit)))
        return dict(p0=p0, v0=v0, g=g, hit=hit)

    def verlet(self, p0, v0, prec=PREC_FP64):
        p0 = _f64(p0).reshape(-1, 3); v0 = _f64(v0).reshape(-1, 3)
        n = p0.shape[0]
        chi = np.empty(n); ergy = np.empty(n); ok = np.empty(n, np.int32); steps = np.empty(n, np.int32)
        tot = np.empty(n, np.int64)
        _check(lib().cosims_verlet(self._h, prec, n, _p(p0), _p(v0), _p(chi), _p(ok), _p(steps), _p(ergy),
                                   _p(tot)))
        return dict(chi=chi, ok=ok, steps=steps, ergy=ergy, total_steps=tot)

    def run(self, prec=PREC_MIXED, job=0, iters=0, traj=0, keep=False, params: Params | None = None,
            lanes=0, wg_size=0, blocks_per_cu=0, rng=RNG_PHILOX, noalloc=False):
        """monteTraj::runTraj: iters CCS integrals of traj trajectories.
        noalloc=True calls cosims_run_noalloc (no device allocation: a static
        scratch of at most NOALLOC_MAX_TRAJ trajectories in NOALLOC_MAX_ITER
        integrals, group kernels, molecule uploaded first)."""
        p = params if params is not None else self.params
        it = iters or p.iter
        T = traj or p.traj
        cfg = RunCfg(prec=prec, job=job, iter=it, traj=T, lanes=lanes, wg_size=wg_size,
                     blocks_per_cu=blocks_per_cu, rng=rng)
        chi = status = g = None
        if keep:
            chi = np.empty(it * T); status = np.empty(it * T, np.int32); g = np.empty(it * T)
            cfg.chi, cfg.status, cfg.g = _p(chi), _p(status), _p(g)
        om = np.empty(it); err = np.empty(it)
        res = Ccs()
        fn = lib().cosims_run_noalloc if noalloc else lib().cosims_run
        _check(fn(self._h, C.byref(p), C.byref(cfg), _p(om), _p(err), C.byref(res)))
        out = {k: getattr(res, k) for k, _ in Ccs._fields_}
        for k in ("traj_kernel_ms", "round_steps", "round_evals", "round_items"):
            out[k] = list(out[k])[: res.n_rounds]
        out["cl_counts"] = list(res.cl_counts)
        out.update(omega=om, omega_err=err)
        if keep:
            out.update(chi=chi, status=status, g=g)
        return out


def _batch_jobs(jobs):
    """jobs: list of (Molecule, dict(seed=, job=, iters=, traj=))."""
    arr = (BatchJob * len(jobs))()
    keep = []
    for k, (m, spec) in enumerate(jobs):
        it = spec.get("iters") or m.params.iter
        T = spec.get("traj") or m.params.traj
        om = np.empty(it); err = np.empty(it); res = Ccs()
        keep.append((om, err, res))
        arr[k] = BatchJob(mol=m._h, seed=spec.get("seed", m.params.seed), job=spec.get("job", 0), iter=it, traj=T,
                          omega=_p(om), omega_err=_p(err), out=C.pointer(res))
    return arr, keep


def size_class(units):
    """Size class of an exact-type molecule whose force loop sweeps `units` atoms."""
    L, k = 4, 0
    while L < 64 and units >= 48 * L:
        L, k = 2 * L, k + 1
    return k


def job_class(info):
    """Size class of a molecule (Molecule.info()): the clustered force types
    take the 64-lane class."""
    return SIZE_CLASSES - 1 if info["force_type"] in (1, 2, 3, 5) else size_class(info["n"])


def job_lanes(info, class_lanes=None):
    """Lanes a job runs without a fixed lane count: 64 for the clustered
    types, else its class's entry of class_lanes, else the size rule."""
    if info["force_type"] in (1, 2, 3, 5):
        return 64
    k = size_class(info["n"])
    return (class_lanes or [0] * SIZE_CLASSES)[k] or 4 << k


def tune_batch(jobs, prec=PREC_MIXED):
    """(class_lanes, class_wg): the launch shape per size class tuned on a batch."""
    arr, _ = _batch_jobs(jobs)
    L = (C.c_int32 * SIZE_CLASSES)(); W = (C.c_int32 * SIZE_CLASSES)()
    _check(lib().cosims_tune_batch(arr, len(jobs), prec, L, W))
    return list(L), list(W)


def run_batch(jobs, prec=PREC_MIXED, lanes=0, wg_size=0, blocks_per_cu=0, class_lanes=None, class_wg=None,
              own_launches=False):
    """All jobs in one call.  By default the jobs of a launch shape whose
    molecule fits LDS share one persistent launch (each workgroup stages the
    molecule of the job it works on), larger exact-type jobs run tiled
    launches of their own and clustered jobs share table-kernel launches;
    own_launches=True gives each exact-type job with enough work a launch of
    its own (LDS-staged or tiled) instead.  lanes > 0 fixes the lanes per
    trajectory, 0 applies the size rule, -1 uses class_lanes / class_wg
    (tuned on the batch when not given).
    Returns (per-job result dicts like Molecule.run, batch statistics)."""
    arr, keep = _batch_jobs(jobs)
    cfg = BatchCfg(prec=prec, lanes=lanes, wg_size=wg_size, blocks_per_cu=blocks_per_cu,
                   own_launches=int(own_launches))
    for c in range(SIZE_CLASSES):
        cfg.class_lanes[c] = (class_lanes or [0] * SIZE_CLASSES)[c]
        cfg.class_wg[c] = (class_wg or [0] * SIZE_CLASSES)[c]
    st = BatchStats()
    _check(lib().cosims_run_batch(arr, len(jobs), C.byref(cfg), C.byref(st)))
    out = []
    for om, err, res in keep:
        d = {k: getattr(res, k) for k, _ in Ccs._fields_ if k not in ("traj_kernel_ms", "round_steps",
                                                                     "round_evals", "round_items", "cl_counts")}
        d.update(omega=om, omega_err=err)
        out.append(d)
    stats = {k: getattr(st, k) for k, _ in BatchStats._fields_}
    stats["class_lanes"] = list(st.class_lanes)
    stats["class_wg"] = list(st.class_wg)
    return out, stats


class Server:
    """CCS request server (cosims_server_*): continuous batching of requests
    over persistent launches, admission by slots and device-memory budget."""

    def __init__(self, prec=PREC_MIXED, lanes=0, wg_size=512, hbm_budget=0, max_active=64, min_blocks=3,
                 max_blocks=1000, setup_threads=0, block_traj=10000, slice_traj=0, slice_ms=0.0,
                 policy=SERVE_SHORTEST_FIRST, min_slice_traj=0, pipeline_depth=0, issue_fraction=0.0,
                 prior_rel_sd=0.0, class_lanes=None, class_wg=None, own_launches=False):
        self.cfg = ServerCfg(prec=prec, lanes=lanes, wg_size=wg_size, hbm_budget=hbm_budget, max_active=max_active,
                             min_blocks=min_blocks, max_blocks=max_blocks, setup_threads=setup_threads,
                             block_traj=block_traj, slice_traj=slice_traj, slice_ms=slice_ms, policy=policy,
                             min_slice_traj=min_slice_traj, pipeline_depth=pipeline_depth,
                             issue_fraction=issue_fraction, prior_rel_sd=prior_rel_sd,
                             own_launches=int(own_launches))
        for c in range(SIZE_CLASSES):
            self.cfg.class_lanes[c] = (class_lanes or [0] * SIZE_CLASSES)[c]
            self.cfg.class_wg[c] = (class_wg or [0] * SIZE_CLASSES)[c]
        h = C.c_void_p()
        _check(lib().cosims_server_create(C.byref(self.cfg), C.byref(h)))
        self._h = h
        self.max_blocks = max_blocks

    def submit(self, mol: dict, params: Params, job=0, target=0.005) -> int:
        """mol: a read_molfile() dict (xyz, mass, sigma, epsilon, charge, charged, sigma_max)."""
        xyz = _f64(mol["xyz"]); mass = _f64(mol["mass"]); sig = _f64(mol["sigma"]); eps = _f64(mol["epsilon"])
        q = _f64(mol["charge"]) if mol.get("charged") else None
        a = Atoms(n=mass.size, xyz=xyz.ctypes.data, mass=mass.ctypes.data, sigma=sig.ctypes.data,
                  epsilon=eps.ctypes.data, charge=None if q is None else q.ctypes.data,
                  sigma_max=mol["sigma_max"])
        p = Params()
        C.pointer(p)[0] = params
        p.charge = int(bool(mol.get("charged")))
        r = Request(params=p, atoms=a, job=job, target_rel_err=target)
        t = C.c_int64()
        _check(lib().cosims_server_submit(self._h, C.byref(r), C.byref(t)))
        return t.value

    def wait(self, ticket, timeout_ms=-1.0):
        res = Result()
        om = np.empty(self.max_blocks); err = np.empty(self.max_blocks)
        _check(lib().cosims_server_wait(self._h, ticket, timeout_ms, C.byref(res), _p(om), _p(err), self.max_blocks))
        out = {k: getattr(res, k) for k, _ in Result._fields_ if k != "ccs"}
        out.update({k: getattr(res.ccs, k) for k in ("mean", "sd", "wmean", "werr", "max_ergy", "n_traj", "n_hit",
                                                     "n_fail", "total_steps", "lanes", "wg_size")})
        out.update(omega=om[: res.blocks].copy(), omega_err=err[: res.blocks].copy())
        return out

    def stats(self):
        st = ServerStats()
        _check(lib().cosims_server_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().cosims_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
