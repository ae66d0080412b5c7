"""ctypes front end of the parity oracle (test infrastructure only).

Loads ``oracle/liboracle_mpm.so`` (the CPU restatement in ``mpm_oracle.cpp``)
and, when it has been built, ``oracle/_ref/libref_mpm.so`` (the reference's
own headers behind ``ref_driver.cpp``).  Only ``tests/``, the smoke check and
the CPU-baseline leg of ``bench.py`` import this module, and only as the
checker or the timed CPU baseline.

Both back ends share one particle/collider/grid packing so a test can run the
same input through either and compare bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle_mpm.so")
REF_SO = os.path.join(_HERE, "_ref", "libref_mpm.so")

MASS_FLOAT = 0
MASS_FIXED = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


def _p(a, ct=_dp):
    return a.ctypes.data_as(ct)


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


_lib_cache: dict = {}


def _load(path):
    if path not in _lib_cache:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        _lib_cache[path] = C.CDLL(path)
    return _lib_cache[path]


def oracle_lib():
    return _load(ORACLE_SO)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    return _load(REF_SO)


# ---------------------------------------------------------------- packing
@dataclass
class Collider:
    """RigidColliderState (mpm_rigid_coupler.h:72-84); quat is (w, x, y, z)."""

    center: tuple = (0.0, 0.0, 0.0)
    quat: tuple = (1.0, 0.0, 0.0, 0.0)
    half_extents: tuple = (0.5, 0.5, 0.5)
    lin_vel: tuple = (0.0, 0.0, 0.0)
    ang_vel: tuple = (0.0, 0.0, 0.0)
    friction: float = 0.5
    mass: float = 1.0
    shape: int = 0  # 0 box, 1 cylinder

    def pack(self) -> np.ndarray:
        a = np.zeros(20, np.float64)
        a[0:3] = self.center
        a[3:7] = self.quat
        a[7:10] = self.half_extents
        a[10:13] = self.lin_vel
        a[13:16] = self.ang_vel
        a[16] = self.friction
        a[17] = self.mass
        a[18] = self.shape
        return a


@dataclass
class Params:
    """MPMSolver material/integration settings; defaults are mpm_solver.h:100-109."""

    E: float = 1.0e5
    nu: float = 0.3
    rho: float = 1500.0
    alpha: float = 0.40
    mu_floor: float = 0.4
    gravity: tuple = (0.0, -9.81, 0.0)

    def pack(self) -> np.ndarray:
        return np.array([self.E, self.nu, self.rho, self.alpha, self.mu_floor, *self.gravity], np.float64)


@dataclass
class Body:
    """The fields of RigidBody that TimeIntegrator::integrate reads and writes."""

    x: tuple = (0.0, 0.0, 0.0)
    quat: tuple = (1.0, 0.0, 0.0, 0.0)
    v: tuple = (0.0, 0.0, 0.0)
    w: tuple = (0.0, 0.0, 0.0)
    inv_mass: float = 1.0
    inv_I: np.ndarray = field(default_factory=lambda: np.eye(3))
    lin_damp: float = 0.01
    ang_damp: float = 0.01
    force: tuple = (0.0, 0.0, 0.0)
    torque: tuple = (0.0, 0.0, 0.0)

    def pack(self) -> np.ndarray:
        a = np.zeros(31, np.float64)
        a[0:3] = self.x
        a[3:7] = self.quat
        a[7:10] = self.v
        a[10:13] = self.w
        a[13] = self.inv_mass
        a[14:23] = np.asarray(self.inv_I, np.float64).reshape(9)
        a[23] = self.lin_damp
        a[24] = self.ang_damp
        a[25:28] = self.force
        a[28:31] = self.torque
        return a

    @staticmethod
    def unpack(a, proto: "Body") -> "Body":
        return Body(tuple(a[0:3]), tuple(a[3:7]), tuple(a[7:10]), tuple(a[10:13]), proto.inv_mass, proto.inv_I,
                    proto.lin_damp, proto.ang_damp, tuple(a[25:28]), tuple(a[28:31]))


def make_particles(x, v=None, mass=None, rho=1500.0, F=None, Cm=None):
    """Particle SoA dict the way MPMSolver::add_particle builds it (mpm_solver.h:23-34)."""
    x = _f64(x).reshape(-1, 3)
    n = x.shape[0]
    v = np.zeros((n, 3)) if v is None else _f64(v).reshape(n, 3)
    m = np.ones(n) if mass is None else np.broadcast_to(_f64(mass), (n,)).copy()
    m32 = m.astype(np.float32)
    V0 = np.where(np.float32(rho) > np.float32(1e-12), (m32 / np.float32(rho)).astype(np.float32),
                  np.float32(1.0)).astype(np.float64)
    Fm = np.tile(np.eye(3), (n, 1, 1)) if F is None else _f64(F).reshape(n, 3, 3)
    Cmat = np.zeros((n, 3, 3)) if Cm is None else _f64(Cm).reshape(n, 3, 3)
    return {"x": x, "v": v, "F": Fm, "C": Cmat, "m": m, "V0": V0}


def round32(P):
    """Round every field to float32 (what any FP32 solver stores)."""
    return {k: np.asarray(v, np.float32).astype(np.float64) for k, v in P.items()}


# ---------------------------------------------------------------- back ends
class _Base:
    def grid_dense(self, org, dims):
        raise NotImplementedError


class OracleSim(_Base):
    """The restatement; precision 'f32' or 'f64', mass_mode MASS_FLOAT or MASS_FIXED."""

    def __init__(self, dx, precision="f32", mass_mode=MASS_FLOAT, q_exp=0, params: Params | None = None):
        self.lib = oracle_lib()
        self.s = precision
        L = self.lib
        for nm, res, args in [
            ("create", C.c_void_p, [C.c_double, C.c_int, C.c_int]),
            ("destroy", None, [C.c_void_p]),
            ("params", None, [C.c_void_p, _dp]),
            ("set_particles", None, [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]),
            ("get_particles", None, [C.c_void_p, _dp, _dp, _dp, _dp]),
            ("p2g", None, [C.c_void_p]),
            ("update", None, [C.c_void_p, C.c_double]),
            ("g2p", None, [C.c_void_p, C.c_double]),
            ("step", None, [C.c_void_p, C.c_double, C.c_int]),
            ("couple", None, [C.c_void_p, C.c_int, _dp, C.c_double, _dp]),
            ("num_blocks", C.c_int64, [C.c_void_p]),
            ("grid_dense", None, [C.c_void_p, _ip, _ip, _i64p, _dp, _dp, _dp, _u8p, _u8p]),
            ("grid_set_dense", None, [C.c_void_p, _ip, _ip, _dp, _dp, _u8p]),
            ("grid_mass_q_total", C.c_int64, [C.c_void_p]),
            ("svd", None, [_dp, _dp, _dp, _dp]),
            ("hencky_tau", None, [C.c_void_p, _dp, _dp]),
            ("plasticity", None, [C.c_void_p, _dp, _dp, C.c_double, _dp]),
            ("weights", None, [C.c_void_p, _dp, _dp, _ip, _dp]),
            ("integrate_body", None, [_dp, C.c_double]),
            ("step_coupled", None, [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, _dp, C.c_int]),
        ]:
            f = getattr(L, f"orc_{precision}_{nm}")
            f.restype = res
            f.argtypes = args
        self._f = lambda nm: getattr(L, f"orc_{precision}_{nm}")
        self.h = self._f("create")(float(dx), int(mass_mode), int(q_exp))
        self.dx = dx
        self.n = 0
        self.set_params(params or Params())

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self._f("destroy")(self.h)
                self.h = None
        except Exception:
            pass

    def set_params(self, prm: Params):
        self.prm = prm
        a = prm.pack()
        self._f("params")(self.h, _p(a))

    def set_particles(self, P):
        self.n = int(P["x"].shape[0])
        arrs = [_f64(P[k]) for k in ("x", "v", "F", "C", "m", "V0")]
        self._f("set_particles")(self.h, self.n, *[_p(a) for a in arrs])

    def get_particles(self):
        n = self.n
        x, v, F, Cm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3))
        self._f("get_particles")(self.h, _p(x), _p(v), _p(F), _p(Cm))
        return {"x": x, "v": v, "F": F, "C": Cm}

    def p2g(self, dt=0.0):
        self._f("p2g")(self.h)

    def update(self, dt):
        self._f("update")(self.h, float(dt))

    def g2p(self, dt):
        self._f("g2p")(self.h, float(dt))

    def step(self, dt, n=1):
        self._f("step")(self.h, float(dt), int(n))

    def couple(self, method, col: Collider, dt):
        out = np.zeros(7)
        c = col.pack()
        self._f("couple")(self.h, int(method), _p(c), float(dt), _p(out))
        return out[0:3].copy(), out[3:6].copy(), int(out[6])

    def num_blocks(self):
        return int(self._f("num_blocks")(self.h))

    def grid_dense(self, org, dims):
        org = np.ascontiguousarray(org, np.int32)
        dims = np.ascontiguousarray(dims, np.int32)
        nn = int(np.prod(dims))
        mq = np.zeros(nn, np.int64)
        mass = np.zeros(nn)
        vel = np.zeros((nn, 3))
        force = np.zeros((nn, 3))
        act = np.zeros(nn, np.uint8)
        alloc = np.zeros(nn, np.uint8)
        self._f("grid_dense")(self.h, _p(org, _ip), _p(dims, _ip), _p(mq, _i64p), _p(mass), _p(vel), _p(force),
                              _p(act, _u8p), _p(alloc, _u8p))
        sh = tuple(int(d) for d in dims)
        return {"mass_q": mq.reshape(sh), "mass": mass.reshape(sh), "vel": vel.reshape(sh + (3,)),
                "force": force.reshape(sh + (3,)), "active": act.reshape(sh).astype(bool),
                "alloc": alloc.reshape(sh).astype(bool)}

    def grid_set_dense(self, org, mass, vel, active):
        org = np.ascontiguousarray(org, np.int32)
        dims = np.ascontiguousarray(mass.shape, np.int32)
        m = _f64(mass)
        v = _f64(vel)
        a = np.ascontiguousarray(active, np.uint8)
        self._f("grid_set_dense")(self.h, _p(org, _ip), _p(dims, _ip), _p(m), _p(v), _p(a, _u8p))

    def grid_mass_q_total(self):
        return int(self._f("grid_mass_q_total")(self.h))

    # stateless kernels ------------------------------------------------
    def svd(self, F):
        F = _f64(F, 9)
        U, S, V = np.zeros(9), np.zeros(3), np.zeros(9)
        self._f("svd")(_p(F), _p(U), _p(S), _p(V))
        return U.reshape(3, 3), S, V.reshape(3, 3)

    def hencky_tau(self, F):
        F = _f64(F, 9)
        t = np.zeros(9)
        self._f("hencky_tau")(self.h, _p(F), _p(t))
        return t.reshape(3, 3)

    def plasticity(self, F, Cm, dt):
        F = _f64(F, 9)
        Cm = _f64(Cm, 9)
        out = np.zeros(9)
        self._f("plasticity")(self.h, _p(F), _p(Cm), float(dt), _p(out))
        return out.reshape(3, 3)

    def weights(self, x):
        x = _f64(x, 3)
        fr, w = np.zeros(3), np.zeros(9)
        b = np.zeros(3, np.int32)
        self._f("weights")(self.h, _p(x), _p(fr), _p(b, _ip), _p(w))
        return fr, b, w.reshape(3, 3)

    def integrate_body(self, body: Body, dt):
        a = body.pack()
        self._f("integrate_body")(_p(a), float(dt))
        return Body.unpack(a, body)

    def step_coupled(self, method, col: Collider, body: Body, applied_force, applied_torque, dt, n,
                     kinematic=False):
        """n coupled substeps in the contract order (mpm_rigid_coupler.h:209-211,
        bench_compare_all.cpp:642-663): force += applied; P2G; update;
        coupler on the body's collider; reaction into the body; integrate;
        G2P.  col supplies shape, half extents, friction and mass.  Returns
        the body after the last substep and an (n, 20) record per substep:
        x(3) q(4) v(3) w(3) reaction force(3) torque(3) nodes_inside.
        kinematic: the body moves at its own velocities (x += v dt, the
        quaternion turned by w), the reaction is only reported."""
        a = body.pack()
        ap = np.array([*applied_force, *applied_torque], np.float64)
        rec = np.zeros((int(n), 20))
        self._f("step_coupled")(self.h, int(method), _p(col.pack()), _p(a), _p(ap), float(dt), int(n), _p(rec),
                                2 if kinematic else 1)
        return Body.unpack(a, body), rec


class RefSim(_Base):
    """The reference's MPMSolver itself (oracle/_ref/libref_mpm.so); FP32."""

    def __init__(self, dx, params: Params | None = None):
        L = self.lib = ref_lib()
        for nm, res, args in [
            ("ref_create", C.c_void_p, [C.c_double]),
            ("ref_destroy", None, [C.c_void_p]),
            ("ref_params", None, [C.c_void_p, _dp]),
            ("ref_add_particle", None, [C.c_void_p, _dp, _dp, C.c_double]),
            ("ref_set_particles", None, [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]),
            ("ref_num_particles", C.c_int64, [C.c_void_p]),
            ("ref_get_particles", None, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]),
            ("ref_step", None, [C.c_void_p, C.c_double, C.c_int]),
            ("ref_p2g", None, [C.c_void_p, C.c_double]),
            ("ref_update", None, [C.c_void_p, C.c_double]),
            ("ref_g2p", None, [C.c_void_p, C.c_double]),
            ("ref_couple", None, [C.c_void_p, C.c_int, _dp, C.c_double, _dp]),
            ("ref_num_blocks", C.c_int64, [C.c_void_p]),
            ("ref_grid_dense", None, [C.c_void_p, _ip, _ip, _dp, _dp, _dp, _u8p, _u8p]),
            ("ref_grid_set_dense", None, [C.c_void_p, _ip, _ip, _dp, _dp, _u8p]),
            ("ref_svd", None, [_dp, _dp, _dp, _dp]),
            ("ref_integrate_body", None, [_dp, C.c_double]),
        ]:
            f = getattr(L, nm)
            f.restype = res
            f.argtypes = args
        self.h = L.ref_create(float(dx))
        self.dx = dx
        self.set_params(params or Params())

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.ref_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_params(self, prm: Params):
        self.prm = prm
        a = prm.pack()
        self.lib.ref_params(self.h, _p(a))

    def add_particle(self, x, v, m):
        x, v = _f64(x, 3), _f64(v, 3)
        self.lib.ref_add_particle(self.h, _p(x), _p(v), float(m))

    def set_particles(self, P):
        n = int(P["x"].shape[0])
        arrs = [_f64(P[k]) for k in ("x", "v", "F", "C", "m", "V0")]
        self.lib.ref_set_particles(self.h, n, *[_p(a) for a in arrs])

    def get_particles(self):
        n = int(self.lib.ref_num_particles(self.h))
        x, v, F, Cm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3))
        m, V0 = np.zeros(n), np.zeros(n)
        self.lib.ref_get_particles(self.h, _p(x), _p(v), _p(F), _p(Cm), _p(m), _p(V0))
        return {"x": x, "v": v, "F": F, "C": Cm, "m": m, "V0": V0}

    def p2g(self, dt=0.0):
        self.lib.ref_p2g(self.h, float(dt))

    def update(self, dt):
        self.lib.ref_update(self.h, float(dt))

    def g2p(self, dt):
        self.lib.ref_g2p(self.h, float(dt))

    def step(self, dt, n=1):
        self.lib.ref_step(self.h, float(dt), int(n))

    def couple(self, method, col: Collider, dt):
        out = np.zeros(7)
        c = col.pack()
        self.lib.ref_couple(self.h, int(method), _p(c), float(dt), _p(out))
        return out[0:3].copy(), out[3:6].copy(), int(out[6])

    def num_blocks(self):
        return int(self.lib.ref_num_blocks(self.h))

    def grid_dense(self, org, dims):
        org = np.ascontiguousarray(org, np.int32)
        dims = np.ascontiguousarray(dims, np.int32)
        nn = int(np.prod(dims))
        mass = np.zeros(nn)
        vel = np.zeros((nn, 3))
        force = np.zeros((nn, 3))
        act = np.zeros(nn, np.uint8)
        alloc = np.zeros(nn, np.uint8)
        self.lib.ref_grid_dense(self.h, _p(org, _ip), _p(dims, _ip), _p(mass), _p(vel), _p(force), _p(act, _u8p),
                                _p(alloc, _u8p))
        sh = tuple(int(d) for d in dims)
        return {"mass": mass.reshape(sh), "vel": vel.reshape(sh + (3,)), "force": force.reshape(sh + (3,)),
                "active": act.reshape(sh).astype(bool), "alloc": alloc.reshape(sh).astype(bool)}

    def grid_set_dense(self, org, mass, vel, active):
        org = np.ascontiguousarray(org, np.int32)
        dims = np.ascontiguousarray(mass.shape, np.int32)
        m = _f64(mass)
        v = _f64(vel)
        a = np.ascontiguousarray(active, np.uint8)
        self.lib.ref_grid_set_dense(self.h, _p(org, _ip), _p(dims, _ip), _p(m), _p(v), _p(a, _u8p))

    def svd(self, F):
        F = _f64(F, 9)
        U, S, V = np.zeros(9), np.zeros(3), np.zeros(9)
        self.lib.ref_svd(_p(F), _p(U), _p(S), _p(V))
        return U.reshape(3, 3), S, V.reshape(3, 3)

    def integrate_body(self, body: Body, dt):
        a = body.pack()
        self.lib.ref_integrate_body(_p(a), float(dt))
        return Body.unpack(a, body)
