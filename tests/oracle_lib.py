"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker; this module is test infrastructure and is imported only by
tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcLattice(ctypes.Structure):
    _fields_ = [
        ("V", ctypes.c_int), ("nfaces", ctypes.c_int),
        ("face_start", _ip), ("face_d", _ip), ("face_tgt", _ip), ("face_hodge", _dp),
        ("pos", _dp), ("vol", _dp), ("block_width", ctypes.c_double * 3),
    ]


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("ncomp", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("block_scale", ctypes.c_double), ("p0", ctypes.c_double * 3), ("trap", ctypes.c_double * 3),
        ("trap_offset", ctypes.c_double), ("trap_off_time", ctypes.c_double),
        ("c0", ctypes.c_double), ("c1", ctypes.c_double), ("c2", ctypes.c_double), ("alpha", ctypes.c_double),
        ("bq", ctypes.c_double), ("bb", ctypes.c_double * 3),
        ("bq_quad", ctypes.c_double), ("bb_quad", ctypes.c_double * 3),
        ("quad_sign", ctypes.c_double), ("q", ctypes.c_double), ("center", ctypes.c_double * 3),
    ]


class OrcRamp(ctypes.Structure):
    _fields_ = [
        ("nq", ctypes.c_int), ("q_val", _dp), ("q_dur", _dp), ("q_type", _ip),
        ("nb", ctypes.c_int), ("b_val", _dp), ("b_dur", _dp), ("b_type", _ip),
        ("tau", ctypes.c_double), ("time_scale", ctypes.c_double),
        ("scale_q", ctypes.c_double), ("scale_b", ctypes.c_double),
        ("qscale_q", ctypes.c_double), ("qscale_b", ctypes.c_double),
    ]


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """Thin wrapper over liboracle.so working on complex128 numpy arrays in the ABI layout
    psi[c, z, y, x, v]."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: %s (run make -C oracle)" % path)
        self.lib = ctypes.CDLL(path)
        L = self.lib
        P, LP, RP = ctypes.POINTER(OrcParams), ctypes.POINTER(OrcLattice), ctypes.POINTER(OrcRamp)
        L.orc_apply_h.argtypes = [LP, P, ctypes.c_double, _dp, _dp]
        L.orc_update.argtypes = [LP, P, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_forward_euler.argtypes = [LP, P, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_itp_step.argtypes = [LP, P, ctypes.c_double, _dp, _dp]
        L.orc_itp_step.restype = ctypes.c_double
        L.orc_norm.argtypes = [LP, P, _dp]
        L.orc_norm.restype = ctypes.c_double
        L.orc_normalize.argtypes = [LP, P, _dp, ctypes.c_double]
        L.orc_energy.argtypes = [LP, P, ctypes.c_double, _dp]
        L.orc_energy.restype = ctypes.c_double
        L.orc_expect_h.argtypes = [LP, P, ctypes.c_double, _dp, _dp]
        L.orc_expect_h.restype = ctypes.c_double
        L.orc_magnetization.argtypes = [LP, P, _dp, _dp]
        L.orc_max_hamiltonian.argtypes = [LP, P, ctypes.c_double, _dp]
        L.orc_max_hamiltonian.restype = ctypes.c_double
        L.orc_weighted_diff.argtypes = [LP, P, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_weighted_diff.restype = ctypes.c_double
        L.orc_local_h.argtypes = [P, _dp, ctypes.c_double, _dp, _dp]
        L.orc_position.argtypes = [LP, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_run_leapfrog3.argtypes = [LP, P, RP, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.orc_run_staggered.argtypes = [LP, P, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.orc_ramp_signal.argtypes = [RP, ctypes.c_double, _dp, _dp]
        L.orc_vortex_init.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_double, _dp, _dp]
        L.orc_vortex_iterate.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.orc_vortex_iterate.restype = ctypes.c_double
        L.orc_vortex_function.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_vortex_function.restype = ctypes.c_double
        L.orc_mt_seed.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.c_uint]
        L.orc_mt_uniform_circle.argtypes = [ctypes.POINTER(ctypes.c_uint), _dp]
        L.orc_ramp_apply.argtypes = [RP, ctypes.c_double, P]
        L.orc_adjoint_sweep.argtypes = [LP, P, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_gpblocks_integrate.argtypes = [LP, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, _dp, _dp, ctypes.c_uint, ctypes.c_uint, _dp, _dp,
                                             _dp, _ip]
        L.orc_gpblocks_integrate.restype = ctypes.c_int
        self._keep = []

    # -- structures ---------------------------------------------------------------------------
    def lattice(self, lat):
        arrs = dict(
            face_start=np.ascontiguousarray(lat.face_start, dtype=np.int32),
            face_d=np.ascontiguousarray(lat.face_d.reshape(-1), dtype=np.int32),
            face_tgt=np.ascontiguousarray(lat.face_tgt, dtype=np.int32),
            face_hodge=np.ascontiguousarray(lat.face_hodge, dtype=np.float64),
            pos=np.ascontiguousarray(lat.pos.reshape(-1), dtype=np.float64),
            vol=np.ascontiguousarray(lat.vol, dtype=np.float64),
        )
        self._keep.append(arrs)
        s = OrcLattice()
        s.V = lat.V
        s.nfaces = lat.nfaces
        s.face_start = _ptr(arrs["face_start"], _ip)
        s.face_d = _ptr(arrs["face_d"], _ip)
        s.face_tgt = _ptr(arrs["face_tgt"], _ip)
        s.face_hodge = _ptr(arrs["face_hodge"], _dp)
        s.pos = _ptr(arrs["pos"], _dp)
        s.vol = _ptr(arrs["vol"], _dp)
        s.block_width[:] = [float(x) for x in lat.block_dims]
        return s

    @staticmethod
    def params(ncomp, dims, block_scale, p0, trap=(1.0, 1.0, 1.0), trap_offset=0.0, trap_off_time=1e300,
               c0=0.0, c1=0.0, c2=0.0, alpha=0.0, bq=0.0, bb=(0.0, 0.0, 0.0), bq_quad=0.0,
               bb_quad=(0.0, 0.0, 0.0), quad_sign=None, q=0.0, center=(0.0, 0.0, 0.0)):
        p = OrcParams()
        p.ncomp = ncomp
        p.nx, p.ny, p.nz = dims
        p.block_scale = block_scale
        p.p0[:] = list(p0)
        p.trap[:] = list(trap)
        p.trap_offset = trap_offset
        p.trap_off_time = trap_off_time
        p.c0, p.c1, p.c2, p.alpha = c0, c1, c2, alpha
        p.bq = bq
        p.bb[:] = list(bb)
        p.bq_quad = bq_quad
        p.bb_quad[:] = list(bb_quad)
        p.quad_sign = (1.0 if ncomp == 3 else -1.0) if quad_sign is None else quad_sign
        p.q = q
        p.center[:] = list(center)
        return p

    def ramp(self, q_segments=(), b_segments=(), tau=1.0, time_scale=1.0, scale_q=1.0, scale_b=1.0,
             qscale_q=0.0, qscale_b=0.0):
        """Segments: sequences of (value, duration, type); b values are 3-vectors."""
        qv = np.array([s[0] for s in q_segments] or [0.0], dtype=np.float64)
        qd = np.array([s[1] for s in q_segments] or [0.0], dtype=np.float64)
        qt = np.array([s[2] for s in q_segments] or [0], dtype=np.int32)
        bv = np.array([list(s[0]) for s in b_segments] or [[0.0, 0.0, 0.0]], dtype=np.float64).reshape(-1)
        bd = np.array([s[1] for s in b_segments] or [0.0], dtype=np.float64)
        bt = np.array([s[2] for s in b_segments] or [0], dtype=np.int32)
        self._keep.append((qv, qd, qt, bv, bd, bt))
        r = OrcRamp()
        r.nq, r.q_val, r.q_dur, r.q_type = len(q_segments), _ptr(qv, _dp), _ptr(qd, _dp), _ptr(qt, _ip)
        r.nb, r.b_val, r.b_dur, r.b_type = len(b_segments), _ptr(bv, _dp), _ptr(bd, _dp), _ptr(bt, _ip)
        r.tau, r.time_scale = tau, time_scale
        r.scale_q, r.scale_b, r.qscale_q, r.qscale_b = scale_q, scale_b, qscale_q, qscale_b
        return r

    # -- operations ----------------------------------------------------------------------------
    @staticmethod
    def _c(a):
        assert a.dtype == np.complex128 and a.flags.c_contiguous
        return a.ctypes.data_as(_dp)

    def apply_h(self, L, p, psi, t=0.0):
        out = np.zeros_like(psi)
        self.lib.orc_apply_h(ctypes.byref(L), ctypes.byref(p), t, self._c(psi), self._c(out))
        return out

    def update(self, L, p, nxt, cur, coef, t=0.0):
        self.lib.orc_update(ctypes.byref(L), ctypes.byref(p), t, coef, self._c(nxt), self._c(cur))

    def forward_euler(self, L, p, cur, dt, t=0.0):
        out = np.zeros_like(cur)
        self.lib.orc_forward_euler(ctypes.byref(L), ctypes.byref(p), t, dt, self._c(out), self._c(cur))
        return out

    def leapfrog3(self, L, p, a, b, dt, n, t0=0.0, ramp=None):
        self.lib.orc_run_leapfrog3(ctypes.byref(L), ctypes.byref(p), ctypes.byref(ramp) if ramp else None,
                                   t0, dt, n, self._c(a), self._c(b))

    def staggered(self, L, p, even, odd, dt, n):
        self.lib.orc_run_staggered(ctypes.byref(L), ctypes.byref(p), dt, n, self._c(even), self._c(odd))

    def itp_step(self, L, p, psi, dt):
        scratch = np.zeros_like(psi)
        return self.lib.orc_itp_step(ctypes.byref(L), ctypes.byref(p), dt, self._c(psi), self._c(scratch))

    def norm(self, L, p, psi):
        return self.lib.orc_norm(ctypes.byref(L), ctypes.byref(p), self._c(psi))

    def normalize(self, L, p, psi, target=1.0):
        self.lib.orc_normalize(ctypes.byref(L), ctypes.byref(p), self._c(psi), target)

    def energy(self, L, p, psi, t=0.0):
        return self.lib.orc_energy(ctypes.byref(L), ctypes.byref(p), t, self._c(psi))

    def expect_h(self, L, p, psi, t=0.0):
        scratch = np.zeros_like(psi)
        return self.lib.orc_expect_h(ctypes.byref(L), ctypes.byref(p), t, self._c(psi), self._c(scratch))

    def magnetization(self, L, p, psi):
        out = np.zeros(3)
        self.lib.orc_magnetization(ctypes.byref(L), ctypes.byref(p), self._c(psi), out.ctypes.data_as(_dp))
        return out

    def max_hamiltonian(self, L, p, psi, t=0.0):
        return self.lib.orc_max_hamiltonian(ctypes.byref(L), ctypes.byref(p), t, self._c(psi))

    def weighted_diff(self, L, p, psi, ref, phase=1.0):
        phase = complex(phase)
        return self.lib.orc_weighted_diff(ctypes.byref(L), ctypes.byref(p), self._c(psi), self._c(ref), phase.real,
                                          phase.imag)

    def local_h(self, p, r, psi_node, t=0.0):
        psi_node = np.ascontiguousarray(psi_node, dtype=np.complex128)
        out = np.zeros_like(psi_node)
        r = np.ascontiguousarray(r, dtype=np.float64)
        self.lib.orc_local_h(ctypes.byref(p), r.ctypes.data_as(_dp), t, self._c(psi_node), self._c(out))
        return out

    def positions(self, L, p, V):
        out = np.zeros((p.nz, p.ny, p.nx, V, 3))
        r = np.zeros(3)
        for z in range(p.nz):
            for y in range(p.ny):
                for x in range(p.nx):
                    for v in range(V):
                        self.lib.orc_position(ctypes.byref(L), ctypes.byref(p), x, y, z, v, r.ctypes.data_as(_dp))
                        out[z, y, x, v] = r
        return out

    def ramp_signal(self, ramp, t):
        bq = np.zeros(1)
        bb = np.zeros(3)
        self.lib.orc_ramp_signal(ctypes.byref(ramp), t, bq.ctypes.data_as(_dp), bb.ctypes.data_as(_dp))
        return bq[0], bb

    def vortex(self, kappa, g, rmax, zmax, rstep, zstep, ratio=1.0, maxiters=20000, eps=1e-29, rmin=0.0):
        """VortexState::setRange + iterateSolution (VortexState.cpp:42-47,109-196)."""
        rsize = int((rmax - rmin) / rstep + 0.5)
        zsize = int(zmax / zstep + 0.5)
        rs, zs = (rmax - rmin) / rsize, zmax / zsize
        f = np.zeros(rsize * zsize)
        mu = np.zeros(1)
        self.lib.orc_vortex_init(rsize, rs, rmin, zsize, zs, f.ctypes.data_as(_dp), mu.ctypes.data_as(_dp))
        self.lib.orc_vortex_iterate(kappa, g, rsize, rs, rmin, zsize, zs, ratio, f.ctypes.data_as(_dp),
                                    mu.ctypes.data_as(_dp), maxiters, eps)
        return dict(f=f, mu=float(mu[0]), rsize=rsize, rstep=rs, rplug=rmin, zsize=zsize, zstep=zs)

    def vortex_function(self, vs, r, z):
        return self.lib.orc_vortex_function(vs["f"].ctypes.data_as(_dp), vs["rsize"], vs["rstep"], vs["rplug"],
                                            vs["zsize"], vs["zstep"], r, z)

    def uniform_circle(self, seed, n):
        st = (ctypes.c_uint * 625)()
        self.lib.orc_mt_seed(st, seed)
        out = np.zeros((n, 2))
        buf = np.zeros(2)
        for i in range(n):
            self.lib.orc_mt_uniform_circle(st, buf.ctypes.data_as(_dp))
            out[i] = buf
        return out

    def gpblocks_integrate(self, L, vs, kappa, g, block_scale, minp, maxp, period, iterations):
        ns = np.zeros(iterations + 1)
        er = np.zeros(iterations + 1)
        dt = np.zeros(1)
        steps = np.zeros(1, dtype=np.int32)
        minp = np.ascontiguousarray(minp, dtype=np.float64)
        maxp = np.ascontiguousarray(maxp, dtype=np.float64)
        rc = self.lib.orc_gpblocks_integrate(
            ctypes.byref(L), kappa, g, vs["f"].ctypes.data_as(_dp), vs["rsize"], vs["rstep"], vs["rplug"],
            vs["zsize"], vs["zstep"], vs["mu"], block_scale, minp.ctypes.data_as(_dp), maxp.ctypes.data_as(_dp),
            period, iterations, ns.ctypes.data_as(_dp), er.ctypes.data_as(_dp), dt.ctypes.data_as(_dp),
            steps.ctypes.data_as(_ip))
        assert rc == 0
        return ns, er, float(dt[0]), int(steps[0])

    def adjoint(self, L, p, lat, dt, nsweeps, psi0, objective, target=None, ramp=None, ramp_bases=None):
        """FP64 oracle gradient by the discrete adjoint, storing every forward state.

        objective: 0 overlap |<target|psi_N>|^2, 1 GP energy at t_N, 2 magnetisation M_z.
        ramp_bases: optional (phi_q(t) -> [nq], phi_b(t) -> [nb]) callables giving the ramp
        sensitivities; with a ramp, p's field is taken from it at every sweep.
        Returns dict(J, grad_trap, grad_bias (3), grad_bq, grad_ramp_q, grad_ramp_b (nb,3), sens)."""
        import copy
        N = nsweeps
        dv = p.block_scale ** 3 * lat.vol
        dvb = dv[None, None, None, None, :]

        def params_at(t):
            q = copy.copy(p)
            if ramp is not None:
                self.lib.orc_ramp_apply(ctypes.byref(ramp), t, ctypes.byref(q))
            return q

        states = [psi0.copy()]
        p0 = params_at(0.0)
        states.append(self.forward_euler(L, p0, psi0, dt))
        for k in range(1, N):
            nxt = states[k - 1].copy()
            self.update(L, params_at(k * dt), nxt, states[k], 2 * dt, t=k * dt)
            states.append(nxt)
        psiN = states[N]
        pN = params_at(N * dt)
        F = (p.ncomp - 1) // 2
        if objective == 0:
            c = np.sum(dvb * np.conj(target) * psiN)
            J = abs(c) ** 2
            g = 2 * c * dvb * target
        elif objective == 1:
            J = self.energy(L, pN, psiN, t=N * dt)
            g = 2 * dvb * self.apply_h(L, pN, psiN, t=N * dt)
        else:
            J = self.magnetization(L, pN, psiN)[2]
            m = np.arange(F, -F - 1, -1, dtype=float)[:, None, None, None, None]
            g = 2 * dvb * m * psiN
        g = np.ascontiguousarray(g.astype(np.complex128))
        A = np.zeros_like(psiN)
        B = g
        sens = np.zeros((N, 7))
        s7 = np.zeros(7)
        for k in range(N - 1, -1, -1):
            self.lib.orc_adjoint_sweep(ctypes.byref(L), ctypes.byref(params_at(k * dt)), k * dt, 2 * dt,
                                       self._c(states[k]), self._c(B), self._c(A), s7.ctypes.data_as(_dp))
            sens[k] = s7
            A, B = B, A
        w = np.full(N, 2 * dt)
        w[0] = dt
        res = dict(J=J, sens=sens, grad_trap=(w[:, None] * sens[:, 4:7]).sum(0),
                   grad_bias=(w[:, None] * sens[:, 0:3]).sum(0), grad_bq=float((w * sens[:, 3]).sum()))
        if objective == 1:
            # explicit dependence of E on the parameters at t_N
            pos = self.positions(L, pN, lat.V)
            n = np.sum(np.abs(psiN) ** 2, axis=0)
            mom = [float(np.sum(dvb[0] * pos[..., a] ** 2 * n)) for a in range(3)]
            res["grad_trap"] = res["grad_trap"] + np.array([pN.trap[a] * mom[a] for a in range(3)])
            if F > 0:
                s = 1.0 if F == 1 else -1.0
                Mv = self.magnetization(L, pN, psiN)
                res["grad_bias"] = res["grad_bias"] + s * Mv
                res["explicit_mag"] = s * Mv
        if ramp_bases is not None:
            phq, phb = ramp_bases
            gq = 0.0
            gb = 0.0
            for k in range(N):
                gq = gq + w[k] * phq(k * dt) * sens[k, 3]
                gb = gb + w[k] * np.outer(phb(k * dt), sens[k, 0:3])
            if objective == 1 and F > 0:
                gb = gb + np.outer(phb(N * dt), res["explicit_mag"])
            res["grad_ramp_q"] = gq
            res["grad_ramp_b"] = gb
        return res

    def set_threads(self, n):
        return self.lib.orc_set_threads(n)


def box_dims(lat, cells, edge):
    """p0 centring used throughout: box of `cells` blocks of physical x edge `edge` centred at 0."""
    s = edge / lat.block_width
    p0 = [-0.5 * s * e * c for e, c in zip(lat.block_dims, cells)]
    return s, p0
