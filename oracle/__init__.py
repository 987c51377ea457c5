"""ctypes binding of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this package;
cosims_amd never does.  The C sources cite the reference lines they restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class OrMol(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("q", C.c_void_p), ("sig6", C.c_void_p), ("eps24m", C.c_void_p), ("eps4", C.c_void_p),
        ("charged", C.c_int), ("dipole_scalar", C.c_double), ("force_type", C.c_int),
        ("ncl", C.c_int), ("cl_start", C.c_void_p), ("cl_atoms", C.c_void_p),
        ("cl_pos", C.c_void_p), ("cl_rad", C.c_void_p), ("cl_q", C.c_void_p),
        ("cl_dip", C.c_void_p), ("cl_quad", C.c_void_p), ("cl_replace", C.c_void_p),
        ("multipole_radius", C.c_double), ("dispersion_radius", C.c_double),
        ("multipole_order", C.c_int),
    ]


class OrVeldist(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("a", "mean", "mode", "sigma", "gmin", "gmax", "norm", "max_val",
                 "cdf_frac", "ccs_scale", "min_kin")]


class OrEllipsoid(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("a_unc", C.c_double), ("b_unc", C.c_double), ("c_unc", C.c_double),
                ("rmsd", C.c_double), ("cov", C.c_double * 6), ("avg", C.c_double),
                ("avg_unc", C.c_double), ("area", C.c_double), ("area_unc", C.c_double),
                ("enlarge", C.c_double)]


class OrTrajOut(C.Structure):
    _fields_ = [("chi", C.c_double), ("ok", C.c_int), ("steps", C.c_int), ("ergy", C.c_double),
                ("total_steps", C.c_int64), ("attempts", C.c_int)]


class OrIcCfg(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("job", C.c_uint32), ("iteration", C.c_uint32),
                ("gmin", C.c_double), ("gmax", C.c_double), ("max_val", C.c_double),
                ("vd_a", C.c_double), ("vd_norm", C.c_double), ("bmax", C.c_double),
                ("ea", C.c_double), ("eb", C.c_double), ("ec", C.c_double), ("projection", C.c_int)]


class OrCcs(C.Structure):
    _fields_ = [("omega", C.c_double), ("err", C.c_double), ("total", C.c_double),
                ("total2", C.c_double), ("denom", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_mol_properties.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_default_dispersion_radius.restype = C.c_double
        L.or_default_dispersion_radius.argtypes = [C.c_double]
        L.or_inertia.restype = C.c_int
        L.or_inertia.argtypes = [C.c_int, dp, dp, dp, dp]
        L.or_veldist_init.argtypes = [C.c_double, C.c_double, C.POINTER(OrVeldist)]
        L.or_veldist_pdf.restype = C.c_double
        L.or_veldist_pdf.argtypes = [C.POINTER(OrVeldist), C.c_double]
        L.or_boundary.restype = C.c_int
        L.or_boundary.argtypes = [C.POINTER(OrMol), dp, C.c_double, dp, C.c_int]
        L.or_ellipsoid_fit.argtypes = [dp, C.c_int, C.POINTER(OrEllipsoid)]
        L.or_enlarge.argtypes = [dp, C.c_int, C.POINTER(OrEllipsoid)]
        L.or_pcm_clusters.restype = C.c_int
        L.or_pcm_clusters.argtypes = [C.c_int, dp, C.c_int, ip, ip, dp, dp]
        L.or_cluster_moments.argtypes = [C.c_int, dp, dp, C.c_int, ip, ip, dp, dp, dp, dp, dp, ip]
        L.or_force_type.restype = C.c_int
        L.or_force_type.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_accel.argtypes = [C.POINTER(OrMol), dp, dp]
        L.or_potential.restype = C.c_double
        L.or_potential.argtypes = [C.POINTER(OrMol), dp]
        L.or_verlet.argtypes = [C.POINTER(OrMol), C.c_double, dp, dp, C.c_double, C.c_double,
                                C.c_double, C.c_int, C.c_double, C.POINTER(OrTrajOut)]
        L.or_verlet_batch.argtypes = [C.POINTER(OrMol), C.c_double, C.c_int64, dp, dp, C.c_double,
                                      C.c_double, C.c_double, C.c_int, C.c_double, dp, ip, ip, dp,
                                      lp, C.c_int]
        L.or_accel_batch.argtypes = [C.POINTER(OrMol), C.c_int64, dp, dp, C.c_int]
        L.or_potential_batch.argtypes = [C.POINTER(OrMol), C.c_int64, dp, dp, C.c_int]
        L.or_philox4x32.argtypes = [C.POINTER(C.c_uint32 * 4), C.POINTER(C.c_uint32 * 2),
                                    C.POINTER(C.c_uint32 * 4)]
        L.or_sample_ic.restype = C.c_int
        L.or_sample_ic.argtypes = [C.POINTER(OrIcCfg), C.c_uint32, dp, dp, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int)]
        L.or_reduce_iteration.argtypes = [C.c_int64, dp, ip, ip, C.c_double, C.c_int64,
                                          C.POINTER(OrCcs)]
        L.or_reduce_final.argtypes = [C.c_int, dp, dp] + [C.POINTER(C.c_double)] * 4
        _LIB = L
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Mol:
    """A prepared molecule (the hot path's read-only input) in oracle form."""

    def __init__(self, xyz, q, sig6, eps24m, eps4, charged, mu, force_type, clusters=None,
                 multipole_radius=20.0, dispersion_radius=0.0, multipole_order=1):
        xyz = _f64(xyz).reshape(-1, 3)
        self.n = xyz.shape[0]
        self.x, self.y, self.z = (_f64(xyz[:, k]) for k in range(3))
        self.q, self.sig6, self.eps24m, self.eps4 = (_f64(a) for a in (q, sig6, eps24m, eps4))
        self.mu = float(mu)
        s = OrMol()
        s.n = self.n
        for k in ("x", "y", "z", "q", "sig6", "eps24m", "eps4"):
            setattr(s, k, getattr(self, k).ctypes.data)
        s.charged = int(bool(charged))
        s.dipole_scalar = 28276.01363 / self.mu
        s.force_type = int(force_type)
        self.clusters = clusters
        if clusters is not None:
            self._cl = {k: (_i32(v) if k in ("start", "atoms", "replace") else _f64(v))
                        for k, v in clusters.items()}
            s.ncl = len(self._cl["rad"])
            s.cl_start = self._cl["start"].ctypes.data
            s.cl_atoms = self._cl["atoms"].ctypes.data
            s.cl_pos = self._cl["pos"].ctypes.data
            s.cl_rad = self._cl["rad"].ctypes.data
            s.cl_q = self._cl["q"].ctypes.data
            s.cl_dip = self._cl["dip"].ctypes.data
            s.cl_quad = self._cl["quad"].ctypes.data
            s.cl_replace = self._cl["replace"].ctypes.data
        s.multipole_radius = multipole_radius
        s.dispersion_radius = dispersion_radius
        s.multipole_order = multipole_order
        self.s = s

    def with_force_type(self, ft, charged=None):
        m = Mol(np.stack([self.x, self.y, self.z], 1), self.q, self.sig6, self.eps24m, self.eps4,
                self.s.charged if charged is None else charged, self.mu, ft, self.clusters,
                self.s.multipole_radius, self.s.dispersion_radius, self.s.multipole_order)
        return m

    def accel(self, pos, nthreads=0):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty_like(pos)
        lib().or_accel_batch(C.byref(self.s), pos.shape[0], pos, out, nthreads)
        return out

    def potential(self, pos, nthreads=0):
        pos = _f64(pos).reshape(-1, 3)
        out = np.empty(pos.shape[0])
        lib().or_potential_batch(C.byref(self.s), pos.shape[0], pos, out, nthreads)
        return out

    def verlet(self, dt, p0, v0, abc, projection=True, nthreads=0):
        p0 = _f64(p0).reshape(-1, 3)
        v0 = _f64(v0).reshape(-1, 3)
        n = p0.shape[0]
        chi = np.empty(n); ergy = np.empty(n)
        ok = np.empty(n, np.int32); steps = np.empty(n, np.int32)
        tot = np.empty(n, np.int64)
        lib().or_verlet_batch(C.byref(self.s), dt, n, p0, v0, abc[0], abc[1], abc[2],
                              int(projection), self.mu, chi, ok, steps, ergy, tot, nthreads)
        return dict(chi=chi, ok=ok, steps=steps, ergy=ergy, total_steps=tot)


def veldist(mu, temp):
    v = OrVeldist()
    lib().or_veldist_init(mu, temp, C.byref(v))
    return {k: getattr(v, k) for k, _ in OrVeldist._fields_}


def prepare(xyz, mass, sigma, eps, charge, charged, temp=298.0, multipole=False,
            dispersion_cutoff=False, max_cluster_size=4, multipole_order=1,
            dispersion_radius=0.0, multipole_radius=0.0, sigma_max=3.5):
    """Oracle setup pipeline: setMolProperties -> setInertia -> velocity
    distribution -> impactParam boundary/ellipsoid/enlarge -> PCM clusters ->
    multipoles (src/main.cpp:119-138, src/monteTraj.cpp:26-95)."""
    L = lib()
    xyz = _f64(xyz).reshape(-1, 3).copy()
    n = xyz.shape[0]
    mass, sigma, eps = _f64(mass), _f64(sigma), _f64(eps)
    q = _f64(charge) if (charged and charge is not None) else np.zeros(n)
    sig6 = np.empty(n); e24 = np.empty(n); e4 = np.empty(n)
    mu = C.c_double(); tm = C.c_double()
    flat = xyz.reshape(-1)
    L.or_mol_properties(n, flat, mass, sigma, eps, sig6, e24, e4, C.byref(mu), C.byref(tm))
    if dispersion_cutoff and dispersion_radius == 0:
        dispersion_radius = L.or_default_dispersion_radius(sigma_max)
    if multipole and multipole_radius == 0:
        multipole_radius = 20.0
    rot = np.empty(9); mx = np.empty(3)
    rc = L.or_inertia(n, flat, mass, rot, mx)
    if rc != 0:
        raise RuntimeError("inertia eigenvectors not orthogonal (reference fallback branch)")
    xyz = flat.reshape(-1, 3)
    vd = veldist(mu.value, temp)
    ft = L.or_force_type(int(bool(charged)), int(bool(multipole)), int(bool(dispersion_cutoff)))
    base = Mol(xyz, q, sig6, e24, e4, charged, mu.value, 6)
    cap = 400000
    pts = np.empty(3 * cap)
    npts = L.or_boundary(C.byref(base.s), mx, vd["min_kin"], pts, cap)
    pts = pts[: 3 * npts].copy()
    ell = OrEllipsoid()
    L.or_ellipsoid_fit(pts, npts, C.byref(ell))
    fit = {k: getattr(ell, k) for k in ("a", "b", "c", "a_unc", "b_unc", "c_unc", "rmsd",
                                         "avg", "avg_unc", "area", "area_unc")}
    fit["cov"] = list(ell.cov)
    L.or_enlarge(pts, npts, C.byref(ell))
    clusters = None
    if multipole or dispersion_cutoff:
        clusters = pcm_clusters(xyz, q, max_cluster_size)
    mol = Mol(xyz, q, sig6, e24, e4, charged, mu.value, ft, clusters,
              multipole_radius, dispersion_radius, multipole_order)
    bmax = max(ell.a, max(ell.b, ell.c))
    return dict(mol=mol, mu=mu.value, total_mass=tm.value, rot=rot, maxxyz=mx, vd=vd,
                boundary=pts.reshape(-1, 3), fit=fit, a=ell.a, b=ell.b, c=ell.c,
                enlarge=ell.enlarge, bmax=bmax, force_type=ft,
                ccs_const=vd["ccs_scale"] * (bmax * bmax),
                dispersion_radius=dispersion_radius, multipole_radius=multipole_radius)


def pcm_clusters(xyz, q, max_cluster_size):
    L = lib()
    xyz = _f64(xyz).reshape(-1)
    n = xyz.size // 3
    start = np.empty(n + 2, np.int32); atoms = np.empty(max(n, 1), np.int32)
    center = np.empty(3 * (n + 1)); md = np.empty(n + 1)
    ncl = L.or_pcm_clusters(n, xyz, int(max_cluster_size), start, atoms, center, md)
    if ncl < 0:
        raise RuntimeError("PCM clustering produced an empty cluster")
    start = start[: ncl + 1].copy(); center = center[: 3 * ncl].copy(); md = md[:ncl].copy()
    clq = np.empty(ncl); dip = np.empty(3 * ncl); quad = np.empty(9 * ncl)
    pos = np.empty(3 * ncl); rep = np.empty(ncl, np.int32)
    L.or_cluster_moments(n, xyz, _f64(q), ncl, start, atoms, center, clq, dip, quad, pos, rep)
    return dict(start=start, atoms=atoms[:n].copy(), center=center, rad=md, q=clq, dip=dip,
                quad=quad, pos=pos, replace=rep)


def sample_ic(cfg: dict, trajs):
    L = lib()
    c = OrIcCfg(**cfg)
    trajs = np.asarray(trajs, dtype=np.int64)
    n = trajs.size
    p0 = np.empty((n, 3)); v0 = np.empty((n, 3)); g = np.empty(n)
    hit = np.empty(n, np.int32); tries = np.empty(n, np.int32)
    pp = np.empty(3); vv = np.empty(3); gg = C.c_double(); tt = C.c_int()
    for i, t in enumerate(trajs):
        hit[i] = L.or_sample_ic(C.byref(c), int(t), pp, vv, C.byref(gg), C.byref(tt))
        p0[i] = pp; v0[i] = vv; g[i] = gg.value; tries[i] = tt.value
    return dict(p0=p0, v0=v0, g=g, hit=hit, tries=tries)


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr); k = (C.c_uint32 * 2)(*key); o = (C.c_uint32 * 4)()
    L.or_philox4x32(C.byref(c), C.byref(k), C.byref(o))
    return list(o)


def reduce_iteration(chi, ok, hit, ccs_const, update_cond):
    r = OrCcs()
    chi = _f64(chi)
    lib().or_reduce_iteration(chi.size, chi, _i32(ok), _i32(hit), ccs_const, int(update_cond),
                              C.byref(r))
    return dict(omega=r.omega, err=r.err, total=r.total, total2=r.total2, denom=r.denom)


def reduce_final(omega, err):
    o = [C.c_double() for _ in range(4)]
    omega, err = _f64(omega), _f64(err)
    lib().or_reduce_final(omega.size, omega, err, *[C.byref(x) for x in o])
    return dict(mean=o[0].value, sd=o[1].value, wmean=o[2].value, werr=o[3].value)
