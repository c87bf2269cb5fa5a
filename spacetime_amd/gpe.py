"""ctypes binding of libgpe.so, the C ABI in include/gpe.h.

This mirrors, call for call, what the reference's host loop does around its kernels
(``Spin1GpeBCC/Spin1GpeBCC.cu:1105-1743``): create buffers, upload a state, make the second
time level (forward Euler), step, run imaginary time with renormalisation, and read
observables.  States are numpy complex128 arrays shaped (ncomp, nz, ny, nx, V), which is
exactly the ABI memory layout.

There is no CPU fallback: if the HIP library is missing or no device is present, the
constructor raises.
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import Sequence, Tuple

import numpy as np

from . import lattice as _lattice

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgpe.so")

FP32, FP64 = 0, 1
LEAPFROG3, STAGGERED = 0, 1
START_EULER, START_PHASE, START_COPY = 0, 1, 2
OBJ_OVERLAP, OBJ_ENERGY, OBJ_MAGNETIZATION_Z = 0, 1, 2
FIELD_DENSITY, FIELD_SPIN, FIELD_SPIN_NORM, FIELD_UV_THETA = 0, 1, 2, 3
PHASES = {"polar": 0, "ferro": 1, "mixed": 2, "un": 3, "bn_hori": 4, "bn_vert": 5, "cyclic": 6}
BASIS_Z, BASIS_Y, BASIS_X = 0, 1, 2
PICTURE_DENSITY, PICTURE_RI, PICTURE_SLICE = 0, 1, 2
ERRORS = {-22: "EINVAL", -12: "ENOMEM", -5: "EDEVICE", -71: "ESTATE", -34: "ERANGE"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Config(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int), ("ncomp", ctypes.c_int), ("precision", ctypes.c_int), ("scheme", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("block_scale", ctypes.c_double), ("p0", ctypes.c_double * 3), ("dt", ctypes.c_double),
        ("trap", ctypes.c_double * 3), ("trap_offset", ctypes.c_double), ("trap_off_time", ctypes.c_double),
        ("c0", ctypes.c_double), ("c1", ctypes.c_double), ("c2", ctypes.c_double), ("alpha", ctypes.c_double),
        ("quad_sign", ctypes.c_double), ("q", ctypes.c_double), ("center", ctypes.c_double * 3),
        ("expansion_start", ctypes.c_double), ("expansion_rate", ctypes.c_double),
    ]


class _Ramp(ctypes.Structure):
    _fields_ = [
        ("nq", ctypes.c_int), ("q_val", _dp), ("q_dur", _dp), ("q_type", _ip),
        ("nb", ctypes.c_int), ("b_val", _dp), ("b_dur", _dp), ("b_type", _ip),
        ("tau", ctypes.c_double), ("time_scale", ctypes.c_double),
        ("scale_q", ctypes.c_double), ("scale_b", ctypes.c_double),
        ("qscale_q", ctypes.c_double), ("qscale_b", ctypes.c_double),
    ]


class _Obs(ctypes.Structure):
    _fields_ = [
        ("norm", ctypes.c_double), ("energy", ctypes.c_double), ("expect_h", ctypes.c_double),
        ("magnetization", ctypes.c_double * 3), ("max_density", ctypes.c_double), ("time", ctypes.c_double),
        ("finite", ctypes.c_int), ("spin", ctypes.c_double), ("moment", ctypes.c_double * 3),
        ("ref_spin", ctypes.c_double), ("ref_magnetization_z", ctypes.c_double),
    ]


class _AdjSpec(ctypes.Structure):
    _fields_ = [("nsweeps", ctypes.c_int), ("checkpoint_every", ctypes.c_int), ("objective", ctypes.c_int),
                ("target", ctypes.c_void_p), ("fp64_trajectory", ctypes.c_int)]


class _AdjRes(ctypes.Structure):
    _fields_ = [
        ("objective", ctypes.c_double), ("grad_trap", ctypes.c_double * 3), ("grad_static_bq", ctypes.c_double),
        ("grad_static_bb", ctypes.c_double * 3), ("grad_ramp_q", _dp), ("grad_ramp_b", _dp),
        ("forward_ms", ctypes.c_double), ("backward_ms", ctypes.c_double),
        ("checkpoint_bytes", ctypes.c_double), ("segment_bytes", ctypes.c_double),
    ]


class _Plan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nx", "ny", "nz", "V", "C", "x_per_lane", "rows", "pairs", "nseg",
                                             "tile_rows", "row_groups", "ntile", "xcd_remap")] + [
        ("lx", ctypes.c_longlong), ("waves", ctypes.c_longlong), ("blocks", ctypes.c_int)] + [
        (n, ctypes.c_uint * 3) for n in ("div_nseg", "div_V", "div_rows", "div_nz", "div_pairs")] + [
        ("slabs", ctypes.c_int)] + [(n, ctypes.c_longlong) for n in ("ys", "vs", "cs", "zs")]


class GpeError(RuntimeError):
    pass


_LIB = None


def load_library(path=LIB_PATH):
    """Load libgpe.so once; raises GpeError if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise GpeError("libgpe.so not built at %s: run `python -m spacetime_amd.build`" % path)
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.gpe_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(vp)]
    lib.gpe_destroy.argtypes = [vp]
    lib.gpe_last_error.argtypes = [vp]
    lib.gpe_last_error.restype = ctypes.c_char_p
    lib.gpe_lattice_values.argtypes = [ctypes.c_int]
    for name in ("gpe_set_state", "gpe_set_state_device"):
        getattr(lib, name).argtypes = [vp, ctypes.c_int, vp, ctypes.c_double]
    for name in ("gpe_get_state", "gpe_get_state_device"):
        getattr(lib, name).argtypes = [vp, ctypes.c_int, vp]
    lib.gpe_set_field.argtypes = [vp, ctypes.c_double, _dp, ctypes.c_double, _dp]
    lib.gpe_set_ramp.argtypes = [vp, ctypes.POINTER(_Ramp)]
    lib.gpe_start.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.gpe_step.argtypes = [vp, ctypes.c_int]
    lib.gpe_itp.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.gpe_observables.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Obs)]
    lib.gpe_scale.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.gpe_sync.argtypes = [vp]
    lib.gpe_time.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    lib.gpe_last_step_timing.argtypes = [vp, _dp, _dp]
    lib.gpe_adjoint.argtypes = [vp, ctypes.POINTER(_AdjSpec), ctypes.POINTER(_AdjRes)]
    lib.gpe_ensemble_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    lib.gpe_ensemble_destroy.argtypes = [vp]
    lib.gpe_ensemble_last_error.argtypes = [vp]
    lib.gpe_ensemble_last_error.restype = ctypes.c_char_p
    lib.gpe_ensemble_step.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.gpe_ensemble_timing.argtypes = [vp, _dp, _dp]
    lib.gpe_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int), _dp]
    lib.gpe_set_trap.argtypes = [vp, _dp]
    lib.gpe_stream_bandwidth.argtypes = [ctypes.c_longlong, ctypes.c_int, _dp, _dp]
    lib.gpe_plan_traversal.argtypes = [ctypes.POINTER(_Config), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(_Plan)]
    lib.gpe_set_traversal.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.gpe_traversal.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int)]
    lib.gpe_field_width.argtypes = [ctypes.c_int]
    lib.gpe_local_fields.argtypes = [vp, ctypes.c_int, ctypes.c_int, _dp]
    lib.gpe_fill_thomas_fermi.argtypes = [vp, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_double,
                                          ctypes.c_ulonglong, ctypes.c_double]
    lib.gpe_get_config.argtypes = [vp, ctypes.POINTER(_Config)]
    lib.gpe_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_longlong]
    lib.gpe_clock.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    lib.gpe_project_spinor.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    lib.gpe_add_noise.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong]
    lib.gpe_set_knot_ramp.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double]
    lib.gpe_write_dat.argtypes = [vp, ctypes.c_int, ctypes.c_char_p]
    lib.gpe_read_dat.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_double]
    lib.gpe_max_hamiltonian.argtypes = [vp, ctypes.c_int, _dp]
    for name in ("gpe_weighted_diff", "gpe_weighted_diff_device"):
        getattr(lib, name).argtypes = [vp, ctypes.c_int, vp, ctypes.c_double, ctypes.c_double, _dp]
    lib.gpe_save_checkpoint.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.gpe_load_checkpoint.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_int)]
    lib.gpe_write_vtk_spinor.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.gpe_write_vtk_preimage.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_double]
    lib.gpe_write_vtk_volume.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    cp = ctypes.POINTER(_Config)
    lib.gpe_host_write_dat.argtypes = [cp, vp, ctypes.c_char_p]
    lib.gpe_host_read_dat.argtypes = [cp, ctypes.c_char_p, vp]
    lib.gpe_host_write_vtk_spinor.argtypes = [cp, vp, ctypes.c_char_p, ctypes.c_int]
    lib.gpe_host_write_vtk_preimage.argtypes = [cp, vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_double]
    lib.gpe_host_write_vtk_volume.argtypes = [cp, vp, ctypes.c_char_p, ctypes.c_int]
    lib.gpe_write_bmp.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.gpe_series_create.argtypes = [ctypes.POINTER(vp)]
    lib.gpe_series_destroy.argtypes = [vp]
    lib.gpe_series_append.argtypes = [vp, vp, ctypes.c_int]
    lib.gpe_series_save.argtypes = [vp, ctypes.c_char_p]
    lib.gpe_host_write_bmp.argtypes = [cp, vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    _LIB = lib
    return lib


@dataclass
class Config:
    """Mirror of gpe_config (include/gpe.h)."""
    lattice: str = "bcc"
    ncomp: int = 1
    precision: int = FP64
    scheme: int = LEAPFROG3
    dims: Tuple[int, int, int] = (8, 8, 8)
    block_scale: float = 0.1
    p0: Sequence[float] = (0.0, 0.0, 0.0)
    dt: float = 1e-4
    trap: Sequence[float] = (1.0, 1.0, 1.0)
    trap_offset: float = 0.0
    trap_off_time: float = 1e300
    c0: float = 0.0
    c1: float = 0.0
    c2: float = 0.0
    alpha: float = 0.0
    quad_sign: float = None
    q: float = 0.0
    center: Sequence[float] = (0.0, 0.0, 0.0)
    expansion_start: float = 1e300
    expansion_rate: float = 0.0

    @property
    def V(self):
        return _lattice.get(self.lattice).V

    @property
    def shape(self):
        nx, ny, nz = self.dims
        return (self.ncomp, nz, ny, nx, self.V)

    def to_c(self):
        c = _Config()
        c.lattice = _lattice.LATTICE_IDS[self.lattice]
        c.ncomp, c.precision, c.scheme = self.ncomp, self.precision, self.scheme
        c.nx, c.ny, c.nz = self.dims
        c.block_scale, c.dt = self.block_scale, self.dt
        c.p0[:] = list(self.p0)
        c.trap[:] = list(self.trap)
        c.trap_offset, c.trap_off_time = self.trap_offset, self.trap_off_time
        c.c0, c.c1, c.c2, c.alpha, c.q = self.c0, self.c1, self.c2, self.alpha, self.q
        c.quad_sign = (1.0 if self.ncomp == 3 else -1.0) if self.quad_sign is None else self.quad_sign
        c.center[:] = list(self.center)
        c.expansion_start, c.expansion_rate = self.expansion_start, self.expansion_rate
        return c

    @classmethod
    def centred_box(cls, lattice, dims, cell_edge, **kw):
        """Box of dims blocks centred on the origin; cell_edge is the physical x edge of a block
        (the blocks are cubes except Z's, whose y edge is sqrt(3) times longer)."""
        lat = _lattice.get(lattice)
        s = cell_edge / lat.block_width
        p0 = tuple(-0.5 * s * e * d for e, d in zip(lat.block_dims, dims))
        return cls(lattice=lattice, dims=tuple(dims), block_scale=s, p0=p0, **kw)


@dataclass
class Observables:
    norm: float
    energy: float
    expect_h: float
    magnetization: np.ndarray
    max_density: float
    time: float
    finite: bool
    spin: float = 0.0
    moment: np.ndarray = None
    ref_spin: float = 0.0
    ref_magnetization_z: float = 0.0


class Stepper:
    """One simulation context on the current HIP device (gpe_create ... gpe_destroy)."""

    def __init__(self, config: Config):
        self.lib = load_library()
        self.config = config
        self._cfg = config.to_c()
        h = ctypes.c_void_p()
        rc = self.lib.gpe_create(ctypes.byref(self._cfg), ctypes.byref(h))
        if rc != 0:
            raise GpeError("gpe_create failed: %s (%d)" % (ERRORS.get(rc, "?"), rc))
        self._h = h
        self._keep = None
        self._ramp_sizes = (0, 0)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gpe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.gpe_last_error(self._h)
            raise GpeError("%s failed: %s (%d): %s" % (what, ERRORS.get(rc, "?"), rc, msg.decode() if msg else ""))

    # -- state ---------------------------------------------------------------------------------
    def set_state(self, psi, level=0, t=0.0):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        if psi.shape != self.config.shape:
            raise ValueError("state shape %s != %s" % (psi.shape, self.config.shape))
        self._check(self.lib.gpe_set_state(self._h, level, psi.ctypes.data, t), "gpe_set_state")

    def get_state(self, level=0):
        out = np.empty(self.config.shape, dtype=np.complex128)
        self._check(self.lib.gpe_get_state(self._h, level, out.ctypes.data), "gpe_get_state")
        return out

    def set_state_device(self, ptr, level=0, t=0.0):
        self._check(self.lib.gpe_set_state_device(self._h, level, ctypes.c_void_p(ptr), t), "gpe_set_state_device")

    def get_state_device(self, ptr, level=0):
        self._check(self.lib.gpe_get_state_device(self._h, level, ctypes.c_void_p(ptr)), "gpe_get_state_device")

    def fill_thomas_fermi(self, g, mu, weights=None, noise=0.0, seed=0, t=0.0):
        """Level 0 := sqrt(max(mu - V, 0) / g) (w_c + noise eps_c) on the device
        (gpe_fill_thomas_fermi); weights default to all weight in component 0."""
        C = self.config.ncomp
        w = np.zeros(C, dtype=np.complex128)
        if weights is None:
            w[0] = 1.0
        else:
            w[:] = weights
        wd = np.ascontiguousarray(np.stack([w.real, w.imag], axis=1).reshape(-1))
        self._check(self.lib.gpe_fill_thomas_fermi(self._h, g, mu, wd.ctypes.data_as(_dp), noise, seed, t),
                    "gpe_fill_thomas_fermi")

    # -- fields --------------------------------------------------------------------------------
    def set_field(self, bq=0.0, bb=(0.0, 0.0, 0.0), bq_quad=0.0, bb_quad=(0.0, 0.0, 0.0)):
        b = (ctypes.c_double * 3)(*bb)
        bq3 = (ctypes.c_double * 3)(*bb_quad)
        self._check(self.lib.gpe_set_field(self._h, bq, b, bq_quad, bq3), "gpe_set_field")

    def set_ramp(self, q_segments=(), b_segments=(), tau=1.0, time_scale=1.0, scale_q=1.0, scale_b=1.0,
                 qscale_q=0.0, qscale_b=0.0):
        """Segments are (value, duration, type); bias values are 3-vectors (gpe_ramp)."""
        qv = np.array([s[0] for s in q_segments] or [0.0], dtype=np.float64)
        qd = np.array([s[1] for s in q_segments] or [0.0], dtype=np.float64)
        qt = np.array([s[2] for s in q_segments] or [0], dtype=np.int32)
        bv = np.array([list(s[0]) for s in b_segments] or [[0.0] * 3], dtype=np.float64).reshape(-1)
        bd = np.array([s[1] for s in b_segments] or [0.0], dtype=np.float64)
        bt = np.array([s[2] for s in b_segments] or [0], dtype=np.int32)
        r = _Ramp()
        r.nq, r.q_val, r.q_dur, r.q_type = len(q_segments), qv.ctypes.data_as(_dp), qd.ctypes.data_as(_dp), \
            qt.ctypes.data_as(_ip)
        r.nb, r.b_val, r.b_dur, r.b_type = len(b_segments), bv.ctypes.data_as(_dp), bd.ctypes.data_as(_dp), \
            bt.ctypes.data_as(_ip)
        r.tau, r.time_scale = tau, time_scale
        r.scale_q, r.scale_b, r.qscale_q, r.qscale_b = scale_q, scale_b, qscale_q, qscale_b
        self._check(self.lib.gpe_set_ramp(self._h, ctypes.byref(r)), "gpe_set_ramp")
        self._ramp_sizes = (len(q_segments), len(b_segments))

    def clear_ramp(self):
        self._check(self.lib.gpe_set_ramp(self._h, None), "gpe_set_ramp")
        self._ramp_sizes = (0, 0)

    def set_trap(self, trap):
        w = (ctypes.c_double * 3)(*trap)
        self._check(self.lib.gpe_set_trap(self._h, w), "gpe_set_trap")

    def adjoint(self, nsweeps, checkpoint_every, objective=OBJ_OVERLAP, target=None, fp64_trajectory=False):
        """Objective and its gradients by the discrete adjoint (gpe_adjoint).  Returns a dict
        with J, grad_trap (3), grad_static_bq, grad_static_bb (3), grad_ramp_q (nq),
        grad_ramp_b (nb, 3), timings and memory.  fp64_trajectory: FP64 psi and mu on an FP32
        context (gradients that are small residues of cancelling per-sweep terms need it)."""
        spec = _AdjSpec()
        spec.nsweeps, spec.checkpoint_every, spec.objective = int(nsweeps), int(checkpoint_every), objective
        spec.fp64_trajectory = 1 if fp64_trajectory else 0
        tgt = None
        if target is not None:
            tgt = np.ascontiguousarray(target, dtype=np.complex128)
            if tgt.shape != self.config.shape:
                raise ValueError("target shape %s != %s" % (tgt.shape, self.config.shape))
            spec.target = tgt.ctypes.data
        nq, nb = self._ramp_sizes
        gq = np.zeros(max(nq, 1))
        gb = np.zeros((max(nb, 1), 3))
        res = _AdjRes()
        res.grad_ramp_q = gq.ctypes.data_as(_dp)
        res.grad_ramp_b = gb.ctypes.data_as(_dp)
        self._check(self.lib.gpe_adjoint(self._h, ctypes.byref(spec), ctypes.byref(res)), "gpe_adjoint")
        return dict(J=res.objective, grad_trap=np.array(res.grad_trap[:]), grad_static_bq=res.grad_static_bq,
                    grad_static_bb=np.array(res.grad_static_bb[:]), grad_ramp_q=gq[:nq], grad_ramp_b=gb[:nb],
                    forward_ms=res.forward_ms, backward_ms=res.backward_ms,
                    checkpoint_bytes=res.checkpoint_bytes, segment_bytes=res.segment_bytes)

    # -- time stepping -------------------------------------------------------------------------
    def clock(self):
        """(t0, sweeps) of the run clock (gpe_clock)."""
        t0, n = ctypes.c_double(), ctypes.c_longlong()
        self._check(self.lib.gpe_clock(self._h, ctypes.byref(t0), ctypes.byref(n)), "gpe_clock")
        return t0.value, n.value

    def set_time(self, t0, sweeps=0):
        self._check(self.lib.gpe_set_time(self._h, float(t0), int(sweeps)), "gpe_set_time")

    def project(self, phase, param=0.0, level=0):
        """Spinor phase projection on the device (gpe_project_spinor): 'polar', 'ferro',
        'mixed' (param = ferro fraction), 'un', 'bn_hori', 'bn_vert', 'cyclic' (param = the
        relative phase)."""
        self._check(self.lib.gpe_project_spinor(self._h, level, PHASES[phase], float(param)), "gpe_project_spinor")

    def add_noise(self, amplitude, seed, level=0):
        self._check(self.lib.gpe_add_noise(self._h, level, float(amplitude), int(seed)), "gpe_add_noise")

    def set_knot_ramp(self, knots, duration, bq=0.0):
        k = np.ascontiguousarray(knots, dtype=np.float64)
        self._check(self.lib.gpe_set_knot_ramp(self._h, len(k), k.ctypes.data_as(_dp), float(duration), float(bq)),
                    "gpe_set_knot_ramp")

    def write_dat(self, path, level=0):
        self._check(self.lib.gpe_write_dat(self._h, level, os.fsencode(path)), "gpe_write_dat")

    def read_dat(self, path, level=0, t=0.0):
        self._check(self.lib.gpe_read_dat(self._h, level, os.fsencode(path), float(t)), "gpe_read_dat")

    def save_checkpoint(self, directory, prefix=""):
        """Write the reference's checkpoint pair; returns the time tag T of the file names."""
        tag = ctypes.create_string_buffer(512)
        self._check(self.lib.gpe_save_checkpoint(self._h, os.fsencode(directory), prefix.encode(), tag, 512),
                    "gpe_save_checkpoint")
        return tag.value.decode()

    def load_checkpoint(self, directory, tag, prefix=""):
        """Resume from a checkpoint; returns the number of levels loaded (2: step on, 1: start)."""
        n = ctypes.c_int()
        self._check(self.lib.gpe_load_checkpoint(self._h, os.fsencode(directory), prefix.encode(), tag.encode(),
                                                 ctypes.byref(n)), "gpe_load_checkpoint")
        return n.value

    def write_vtk_spinor(self, path, level=0, basis=BASIS_Z):
        self._check(self.lib.gpe_write_vtk_spinor(self._h, level, os.fsencode(path), basis), "gpe_write_vtk_spinor")

    def write_vtk_preimage(self, path, level=0, axis=1, threshold=0.9996):
        self._check(self.lib.gpe_write_vtk_preimage(self._h, level, os.fsencode(path), axis, float(threshold)),
                    "gpe_write_vtk_preimage")

    def write_vtk_volume(self, path, level=0, basis=BASIS_Z):
        self._check(self.lib.gpe_write_vtk_volume(self._h, level, os.fsencode(path), basis), "gpe_write_vtk_volume")

    def write_bmp(self, path, level=0, style=PICTURE_DENSITY, basis=BASIS_Z):
        self._check(self.lib.gpe_write_bmp(self._h, level, os.fsencode(path), style, basis), "gpe_write_bmp")

    def start(self, mode=START_EULER, mu=0.0):
        self._check(self.lib.gpe_start(self._h, mode, mu), "gpe_start")

    def step(self, n):
        self._check(self.lib.gpe_step(self._h, int(n)), "gpe_step")

    def itp(self, n, dt):
        self._check(self.lib.gpe_itp(self._h, int(n), dt), "gpe_itp")

    def scale(self, factor, level=0):
        self._check(self.lib.gpe_scale(self._h, level, factor), "gpe_scale")

    def sync(self):
        self._check(self.lib.gpe_sync(self._h), "gpe_sync")

    def observables(self, level=0, check=True):
        o = _Obs()
        rc = self.lib.gpe_observables(self._h, level, ctypes.byref(o))
        if check or rc not in (0, -34):
            self._check(rc, "gpe_observables")
        return Observables(o.norm, o.energy, o.expect_h, np.array(o.magnetization[:]), o.max_density, o.time,
                           bool(o.finite), o.spin, np.array(o.moment[:]), o.ref_spin, o.ref_magnetization_z)

    def max_hamiltonian(self, level=0):
        """Maximum local Hamiltonian bound of a level (gpe_max_hamiltonian; scalar and spin-1)."""
        out = ctypes.c_double()
        self._check(self.lib.gpe_max_hamiltonian(self._h, level, ctypes.byref(out)), "gpe_max_hamiltonian")
        return out.value

    def weighted_diff(self, ref, phase=1.0, level=0):
        """sum dv |a|^2 |psi - a|^2, a = phase * ref (gpe_weighted_diff).  ref: a host array in
        the ABI layout, or a device pointer (int) to one in complex128."""
        out = ctypes.c_double()
        phase = complex(phase)
        if isinstance(ref, int):
            rc = self.lib.gpe_weighted_diff_device(self._h, level, ctypes.c_void_p(ref), phase.real, phase.imag,
                                                   ctypes.byref(out))
            self._check(rc, "gpe_weighted_diff_device")
        else:
            a = np.ascontiguousarray(ref, dtype=np.complex128)
            if a.shape != self.config.shape:
                raise ValueError("reference shape %s, expected %s" % (a.shape, self.config.shape))
            rc = self.lib.gpe_weighted_diff(self._h, level, a.ctypes.data, phase.real, phase.imag, ctypes.byref(out))
            self._check(rc, "gpe_weighted_diff")
        return out.value

    def local_fields(self, kind, level=0):
        """Per-unknown field map (gpe_local_fields), shaped (nz, ny, nx, V, width)."""
        w = self.lib.gpe_field_width(kind)
        if w < 0:
            raise ValueError("unknown field kind %r" % kind)
        nx, ny, nz = self.config.dims
        out = np.empty((nz, ny, nx, self.config.V, w), dtype=np.float64)
        self._check(self.lib.gpe_local_fields(self._h, level, kind, out.ctypes.data_as(_dp)), "gpe_local_fields")
        return out

    def time(self):
        t = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(self.lib.gpe_time(self._h, ctypes.byref(t), ctypes.byref(n)), "gpe_time")
        return t.value, n.value

    def autotune(self, trial_sweeps=20):
        """Pick the fastest traversal (y-tile rows, XCD remap, x values per lane) for this context
        (gpe_autotune)."""
        ty, rm, xp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        self._check(self.lib.gpe_autotune(self._h, int(trial_sweeps), ctypes.byref(ty), ctypes.byref(rm),
                                          ctypes.byref(xp), ctypes.byref(ms)), "gpe_autotune")
        return dict(tile_rows=ty.value, xcd_remap=rm.value, x_per_lane=xp.value, ms_per_sweep=ms.value)

    def set_traversal(self, tile_rows, xcd_remap=1, x_per_lane=1):
        self._check(self.lib.gpe_set_traversal(self._h, int(tile_rows), int(xcd_remap), int(x_per_lane)),
                    "gpe_set_traversal")

    def traversal(self):
        ty, rm, xp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.gpe_traversal(self._h, ctypes.byref(ty), ctypes.byref(rm), ctypes.byref(xp)),
                    "gpe_traversal")
        return dict(tile_rows=ty.value, xcd_remap=rm.value, x_per_lane=xp.value)

    def last_step_timing(self):
        ms = ctypes.c_double()
        by = ctypes.c_double()
        self._check(self.lib.gpe_last_step_timing(self._h, ctypes.byref(ms), ctypes.byref(by)), "timing")
        return ms.value, by.value


def level_layout(cfg):
    """The device layout of a configuration's levels, computed on the host (gpe_plan_traversal):
    'slabs' or 'lines', elements per x-line and the element strides of rows, unknowns, components
    and z planes."""
    lib = load_library()
    p = _Plan()
    rc = lib.gpe_plan_traversal(cfg.to_c(), 1, 1, 1, ctypes.byref(p))
    if rc != 0:
        raise GpeError("gpe_plan_traversal failed: %s (%d)" % (ERRORS.get(rc, "?"), rc))
    return dict(layout="slabs" if p.slabs else "lines", lx=p.lx, ys=p.ys, vs=p.vs, cs=p.cs, zs=p.zs)


def stream_bandwidth(nbytes=1 << 30, reps=20):
    """(copy GB/s, in-place read-2-write-1 GB/s) of 16-byte streaming kernels over two buffers of
    nbytes each on the current device (gpe_stream_bandwidth)."""
    lib = load_library()
    c, u = ctypes.c_double(), ctypes.c_double()
    rc = lib.gpe_stream_bandwidth(int(nbytes), int(reps), ctypes.byref(c), ctypes.byref(u))
    if rc != 0:
        raise GpeError("gpe_stream_bandwidth failed: %s (%d)" % (ERRORS.get(rc, "?"), rc))
    return c.value, u.value


class Ensemble:
    """Independent started Steppers advanced together, one launch per sweep for all of them
    (gpe_ensemble_*).  The member Steppers stay usable for set/get state and observables."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = ctypes.c_void_p()
        rc = self.lib.gpe_ensemble_create(arr, len(self.members), ctypes.byref(h))
        if rc != 0:
            raise GpeError("gpe_ensemble_create failed: %s (%d)" % (ERRORS.get(rc, "?"), rc))
        self._h = h

    def step(self, nsteps, graph_sweeps=0):
        rc = self.lib.gpe_ensemble_step(self._h, int(nsteps), int(graph_sweeps))
        if rc != 0:
            msg = self.lib.gpe_ensemble_last_error(self._h)
            raise GpeError("gpe_ensemble_step failed: %s (%d): %s" % (ERRORS.get(rc, "?"), rc,
                                                                     msg.decode() if msg else ""))

    def timing(self):
        ms, n = ctypes.c_double(), ctypes.c_double()
        self.lib.gpe_ensemble_timing(self._h, ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gpe_ensemble_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Series:
    """The reference's time-series (.m) file of a run (gpe_series_*): append(stepper, level) adds the
    level's time, field and spin integrals as one column; save(path) writes the eight MATLAB
    vectors and starts a new series."""

    def __init__(self):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.gpe_series_create(ctypes.byref(h))
        if rc != 0:
            raise GpeError("gpe_series_create failed: %s" % ERRORS.get(rc, rc))
        self._h = h

    def append(self, stepper, level=0):
        rc = self.lib.gpe_series_append(self._h, stepper._h, level)
        if rc != 0:
            raise GpeError("gpe_series_append failed: %s" % ERRORS.get(rc, rc))

    def save(self, path):
        rc = self.lib.gpe_series_save(self._h, os.fsencode(str(path)))
        if rc != 0:
            raise GpeError("gpe_series_save failed: %s" % ERRORS.get(rc, rc))

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gpe_series_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
