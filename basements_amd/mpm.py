"""Python host mirror of the MLS-MPM soil substep over the libbm_mpm.so C ABI.

Two layers:

* :class:`Solver` -- one device handle (``bm_mpm_handle``) with an explicit
  grid domain; every method is one C-ABI call (include/bm_mpm.h).
* :class:`MPMSolver` -- the reference's ``basements::mpm::MPMSolver`` API
  (include/basements/physics/mpm/mpm_solver.h:15-95 in the reference) with the
  same method names and argument meaning; it derives the grid domain from the
  particle cloud the way the reference's unbounded sparse grid would allocate
  it, and mirrors get_particles()/get_grid() with host copies.

There is no CPU fallback: if the HIP library is missing or no GPU is visible
these classes raise.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass

import numpy as np

from .build import LIB as LIB_PATH

INCLUDE_H = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "bm_mpm.h")

BM_OK = 0
STATUS_NAMES = {0: "BM_OK", 1: "BM_ERR_INVALID_ARG", 2: "BM_ERR_OOM", 3: "BM_ERR_HIP", 4: "BM_ERR_NONFINITE",
                5: "BM_ERR_NOT_SUPPORTED", 6: "BM_ERR_OUT_OF_DOMAIN", 7: "BM_ERR_STATE"}
MODE_REFERENCE = 0
MODE_FAST = 1
M0 = 0
M1 = 1
M45A = 2  # the research harness's M4.5a: overlap-weighted body mass, two passes
M45B = 3  # M4.5b: + Coulomb tangent blend
M45 = 4   # M4.5: + quadratic B-spline weight
SHAPE_BOX = 0
SHAPE_CYLINDER = 1
BODY_STATIC, BODY_DYNAMIC, BODY_KINEMATIC = 0, 1, 2


class BMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


# ------------------------------------------------------------------ C structs
class Desc(C.Structure):
    _fields_ = [("dx", C.c_float), ("grid_origin", C.c_int32 * 3), ("grid_dims", C.c_int32 * 3),
                ("max_particles", C.c_int64), ("mode", C.c_int32), ("sort_every", C.c_int32),
                ("stream", C.c_void_p)]


class HostParticles(C.Structure):
    _fields_ = [("x", C.POINTER(C.c_float)), ("v", C.POINTER(C.c_float)), ("F", C.POINTER(C.c_float)),
                ("C", C.POINTER(C.c_float)), ("m", C.POINTER(C.c_float)), ("V0", C.POINTER(C.c_float))]


class DeviceParticles(C.Structure):
    _fields_ = [("x", C.c_void_p), ("v", C.c_void_p), ("F", C.c_void_p), ("C", C.c_void_p), ("m", C.c_void_p),
                ("V0", C.c_void_p), ("capacity", C.c_int64)]


class ColliderC(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("quat", C.c_float * 4), ("half_extents", C.c_float * 3),
                ("lin_vel", C.c_float * 3), ("ang_vel", C.c_float * 3), ("friction", C.c_float),
                ("mass", C.c_float), ("shape", C.c_int32)]


class ReactionC(C.Structure):
    _fields_ = [("force", C.c_float * 3), ("torque", C.c_float * 3), ("nodes_inside", C.c_int32)]


@dataclass
class Collider:
    """RigidColliderState (mpm_rigid_coupler.h:72-84); quat = (w, x, y, z)."""

    center: tuple = (0.0, 0.0, 0.0)
    quat: tuple = (1.0, 0.0, 0.0, 0.0)
    half_extents: tuple = (0.5, 0.5, 0.5)
    lin_vel: tuple = (0.0, 0.0, 0.0)
    ang_vel: tuple = (0.0, 0.0, 0.0)
    friction: float = 0.5
    mass: float = 1.0
    shape: int = SHAPE_BOX

    def to_c(self) -> ColliderC:
        c = ColliderC()
        c.center[:] = [float(v) for v in self.center]
        c.quat[:] = [float(v) for v in self.quat]
        c.half_extents[:] = [float(v) for v in self.half_extents]
        c.lin_vel[:] = [float(v) for v in self.lin_vel]
        c.ang_vel[:] = [float(v) for v in self.ang_vel]
        c.friction = float(self.friction)
        c.mass = float(self.mass)
        c.shape = int(self.shape)
        return c


class BodyC(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("orientation", C.c_float * 4), ("linear_velocity", C.c_float * 3),
                ("angular_velocity", C.c_float * 3), ("force", C.c_float * 3), ("torque", C.c_float * 3),
                ("inv_mass", C.c_float), ("inv_inertia", C.c_float * 9), ("linear_damping", C.c_float),
                ("angular_damping", C.c_float), ("integrate", C.c_int32), ("applied_force", C.c_float * 3),
                ("applied_torque", C.c_float * 3), ("half_extents", C.c_float * 3), ("friction", C.c_float),
                ("mass", C.c_float), ("shape", C.c_int32)]


@dataclass
class Body:
    """The device body of the coupled loop: RigidBody state (rigid_body.h:23-60)
    plus the collider fields of RigidColliderState (mpm_rigid_coupler.h:72-84).
    Defaults are RigidBody's (identity inertia, damping 0.01)."""

    position: tuple = (0.0, 0.0, 0.0)
    orientation: tuple = (1.0, 0.0, 0.0, 0.0)
    linear_velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    force: tuple = (0.0, 0.0, 0.0)
    torque: tuple = (0.0, 0.0, 0.0)
    inv_mass: float = 1.0
    inv_inertia: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    linear_damping: float = 0.01
    angular_damping: float = 0.01
    integrate: int = 1  # BODY_STATIC, BODY_DYNAMIC or BODY_KINEMATIC
    applied_force: tuple = (0.0, 0.0, 0.0)
    applied_torque: tuple = (0.0, 0.0, 0.0)
    half_extents: tuple = (0.5, 0.5, 0.5)
    friction: float = 0.5
    mass: float = 1.0
    shape: int = SHAPE_BOX

    def to_c(self) -> BodyC:
        b = BodyC()
        for f, _ in BodyC._fields_:
            v = getattr(self, f)
            if isinstance(v, (tuple, list, np.ndarray)):
                getattr(b, f)[:] = [float(x) for x in v]
            elif f in ("integrate", "shape"):
                setattr(b, f, int(v))
            else:
                setattr(b, f, float(v))
        return b

    @staticmethod
    def from_c(b: BodyC) -> "Body":
        kw = {}
        for f, _ in BodyC._fields_:
            v = getattr(b, f)
            kw[f] = tuple(v[:]) if hasattr(v, "__len__") else v
        kw["integrate"] = int(kw["integrate"])
        return Body(**kw)

    def collider(self) -> Collider:
        return Collider(center=self.position, quat=self.orientation, half_extents=self.half_extents,
                        lin_vel=self.linear_velocity, ang_vel=self.angular_velocity, friction=self.friction,
                        mass=self.mass, shape=self.shape)


class BatchDescC(C.Structure):
    _fields_ = [("max_particles", C.c_int64), ("max_blocks", C.c_int64), ("max_scenes", C.c_int32),
                ("stream", C.c_void_p)]


class SceneDescC(C.Structure):
    _fields_ = [("dx", C.c_float), ("grid_origin", C.c_int32 * 3), ("grid_dims", C.c_int32 * 3),
                ("young_modulus", C.c_float), ("poisson_ratio", C.c_float), ("bulk_density", C.c_float),
                ("plastic_alpha", C.c_float), ("floor_friction", C.c_float), ("gravity", C.c_float * 3),
                ("walls", C.c_int32), ("coupler", C.c_int32)]


class ServerDescC(C.Structure):
    _fields_ = [("batch", BatchDescC), ("hbm_budget", C.c_int64), ("dt", C.c_float), ("max_chunk", C.c_int32),
                ("max_bypass", C.c_int32)]


class RequestC(C.Structure):
    _fields_ = [("scene", SceneDescC), ("n", C.c_int64), ("input", C.POINTER(HostParticles)),
                ("body", C.POINTER(BodyC)), ("output", C.POINTER(HostParticles)), ("substeps", C.c_int32),
                ("arrival", C.c_double)]


class RequestInfoC(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("substeps_done", C.c_int32), ("pad", C.c_int32),
                ("arrival", C.c_double), ("admitted", C.c_double), ("finished", C.c_double), ("body", BodyC),
                ("reaction", ReactionC)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = 0, 1, 2, 3, 4


class DiffDescC(C.Structure):
    _fields_ = [("dx", C.c_float), ("grid_origin", C.c_int32 * 3), ("grid_dims", C.c_int32 * 3),
                ("max_particles", C.c_int64), ("precision", C.c_int32), ("checkpoint_every", C.c_int32),
                ("reference_svd", C.c_int32), ("rollouts", C.c_int32), ("stream", C.c_void_p)]


class DiffParamsC(C.Structure):
    _fields_ = [("young_modulus", C.c_double), ("poisson_ratio", C.c_double), ("friction_deg", C.c_double),
                ("floor_friction", C.c_double), ("gravity", C.c_double * 3), ("has_tool", C.c_int32),
                ("pad", C.c_int32), ("tool_center", C.c_double * 3), ("tool_half_extents", C.c_double * 3),
                ("tool_quat", C.c_double * 4), ("tool_velocity", C.c_double * 3), ("tool_friction", C.c_double)]


class DiffGradC(C.Structure):
    _fields_ = [("tool_velocity", C.c_double * 3), ("friction_deg", C.c_double), ("young_modulus", C.c_double),
                ("alpha", C.c_double), ("mu", C.c_double), ("lambda_", C.c_double)]

# ------------------------------------------------------------------ library
_lib = None

_SIGS = {
    "bm_abi_version": (C.c_int, []),
    "bm_last_error": (C.c_char_p, []),
    "bm_mpm_create": (C.c_int, [C.POINTER(Desc), C.POINTER(C.c_void_p)]),
    "bm_mpm_destroy": (C.c_int, [C.c_void_p]),
    "bm_mpm_synchronize": (C.c_int, [C.c_void_p]),
    "bm_mpm_set_material": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    "bm_mpm_set_gravity": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    "bm_mpm_set_plastic_alpha": (C.c_int, [C.c_void_p, C.c_float]),
    "bm_mpm_set_floor_friction": (C.c_int, [C.c_void_p, C.c_float]),
    "bm_mpm_set_walls": (C.c_int, [C.c_void_p, C.c_int32]),
    "bm_mpm_add_particles": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bm_mpm_upload_particles": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(HostParticles)]),
    "bm_mpm_download_particles": (C.c_int, [C.c_void_p, C.POINTER(HostParticles)]),
    "bm_mpm_num_particles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "bm_mpm_particles_device": (C.c_int, [C.c_void_p, C.POINTER(DeviceParticles)]),
    "bm_mpm_commit_device_particles": (C.c_int, [C.c_void_p, C.c_int64]),
    "bm_mpm_step": (C.c_int, [C.c_void_p, C.c_float]),
    "bm_mpm_step_n": (C.c_int, [C.c_void_p, C.c_float, C.c_int32]),
    "bm_mpm_p2g": (C.c_int, [C.c_void_p, C.c_float]),
    "bm_mpm_grid_update": (C.c_int, [C.c_void_p, C.c_float]),
    "bm_mpm_g2p": (C.c_int, [C.c_void_p, C.c_float]),
    "bm_mpm_couple": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ColliderC), C.c_float, C.POINTER(ReactionC)]),
    "bm_mpm_set_body": (C.c_int, [C.c_void_p, C.POINTER(BodyC)]),
    "bm_mpm_get_body": (C.c_int, [C.c_void_p, C.POINTER(BodyC), C.POINTER(ReactionC)]),
    "bm_mpm_couple_body": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.POINTER(ReactionC)]),
    "bm_mpm_step_coupled": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_int32]),
    "bm_batch_create": (C.c_int, [C.POINTER(BatchDescC), C.POINTER(C.c_void_p)]),
    "bm_batch_destroy": (C.c_int, [C.c_void_p]),
    "bm_batch_synchronize": (C.c_int, [C.c_void_p]),
    "bm_batch_add_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneDescC), C.c_int64, C.POINTER(HostParticles),
                                     C.POINTER(BodyC), C.POINTER(C.c_int32)]),
    "bm_batch_remove_scene": (C.c_int, [C.c_void_p, C.c_int32]),
    "bm_batch_step": (C.c_int, [C.c_void_p, C.c_float, C.c_int32]),
    "bm_batch_download_scene": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(HostParticles)]),
    "bm_batch_get_body": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(BodyC), C.POINTER(ReactionC)]),
    "bm_batch_usage": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bm_batch_scene_error": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "bm_batch_set_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "bm_batch_autotune": (C.c_int, [C.c_void_p, C.c_float, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32)]),
    "bm_batch_footprint": (C.c_int, [C.c_int64, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]),
    "bm_server_create": (C.c_int, [C.POINTER(ServerDescC), C.POINTER(C.c_void_p)]),
    "bm_server_destroy": (C.c_int, [C.c_void_p]),
    "bm_server_clock": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "bm_server_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "bm_server_submit": (C.c_int, [C.c_void_p, C.POINTER(RequestC), C.POINTER(C.c_int64)]),
    "bm_server_poll": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32)]),
    "bm_server_info": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(RequestInfoC)]),
    "bm_server_cancel": (C.c_int, [C.c_void_p, C.c_int64]),
    "bm_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "bm_server_set_overlap": (C.c_int, [C.c_void_p, C.c_int32]),
    "bm_server_load": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64)]),
    "bm_diff_create": (C.c_int, [C.POINTER(DiffDescC), C.POINTER(C.c_void_p)]),
    "bm_diff_destroy": (C.c_int, [C.c_void_p]),
    "bm_diff_set_params": (C.c_int, [C.c_void_p, C.POINTER(DiffParamsC)]),
    "bm_diff_set_rollout_params": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(DiffParamsC)]),
    "bm_diff_set_grid_budget": (C.c_int, [C.c_void_p, C.c_int64]),
    "bm_diff_upload": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(HostParticles), C.POINTER(C.c_float)]),
    "bm_diff_forward": (C.c_int, [C.c_void_p, C.c_float, C.c_int32, C.POINTER(C.c_double)]),
    "bm_diff_forward_batch": (C.c_int, [C.c_void_p, C.c_float, C.c_int32, C.POINTER(C.c_double)]),
    "bm_diff_backward": (C.c_int, [C.c_void_p, C.POINTER(DiffGradC)]),
    "bm_diff_backward_batch": (C.c_int, [C.c_void_p, C.POINTER(DiffGradC)]),
    "bm_diff_download": (C.c_int, [C.c_void_p, C.POINTER(HostParticles)]),
    "bm_diff_download_rollout": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(HostParticles)]),
    "bm_mpm_couple_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ColliderC), C.c_float, C.c_void_p]),
    "bm_diff_map_vjp": (C.c_int, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(DiffParamsC), C.c_int32,
                                  C.c_void_p, C.c_void_p]),
    "bm_batch_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "bm_mpm_download_grid": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p]),
    "bm_mpm_upload_grid": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bm_mpm_grid_mass_total": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "bm_mpm_grid_settle": (C.c_int, [C.c_void_p, C.c_float]),
    "bm_mpm_mass_quantum_exp": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "bm_mpm_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bm_mpm_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "bm_mpm_set_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "bm_mpm_read_timing": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
}

PHASES = ("clear", "p2g", "update", "couple", "g2p", "g2p_p2g", "sort", "other")


def header_symbols(path: str = INCLUDE_H) -> list[str]:
    """Every function the public header declares."""
    txt = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(bm_\w+)\s*\(", txt, flags=re.M)))


def lib():
    """Load libbm_mpm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run `python -m basements_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.bm_abi_version() != 1:
            raise RuntimeError("libbm_mpm.so ABI version mismatch")
        _lib = L
    return _lib


def _check(rc):
    if rc != BM_OK:
        raise BMError(rc, lib().bm_last_error().decode())


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    return a if n is None else a.reshape(n)


# ------------------------------------------------------------------ Solver
class Solver:
    """One device-resident scene behind a ``bm_mpm_handle``."""

    def __init__(self, dx, origin, dims, max_particles, mode=MODE_REFERENCE, sort_every=0, stream=None):
        L = lib()
        d = Desc()
        d.dx = float(dx)
        d.grid_origin[:] = [int(v) for v in origin]
        d.grid_dims[:] = [int(v) for v in dims]
        d.max_particles = int(max_particles)
        d.mode = int(mode)
        d.sort_every = int(sort_every)
        d.stream = stream
        h = C.c_void_p()
        _check(L.bm_mpm_create(C.byref(d), C.byref(h)))
        self.h = h
        self.dx = float(dx)
        self.origin = tuple(int(v) for v in origin)
        self.dims = tuple((int(v) + 3) // 4 * 4 for v in dims)
        self.mode = mode
        self.cap = int(max_particles)

    @classmethod
    def attach(cls, handle, dx, origin, dims, max_particles, mode=MODE_REFERENCE):
        """Wrap a handle created elsewhere (a plugin's, another library's)
        without taking it over: close() leaves it to its owner."""
        if not handle:
            raise BMError(1, "attach: null handle")
        self = cls.__new__(cls)
        self.h = C.c_void_p(int(handle) if not isinstance(handle, C.c_void_p) else handle.value)
        self.dx = float(dx)
        self.origin = tuple(int(v) for v in origin)
        self.dims = tuple((int(v) + 3) // 4 * 4 for v in dims)
        self.mode = mode
        self.cap = int(max_particles)
        self._owned = False
        return self

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                lib().bm_mpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters --------------------------------------------------------
    def set_material(self, E, nu, rho):
        _check(lib().bm_mpm_set_material(self.h, E, nu, rho))

    def set_gravity(self, g):
        _check(lib().bm_mpm_set_gravity(self.h, *[float(v) for v in g]))

    def set_plastic_alpha(self, a):
        _check(lib().bm_mpm_set_plastic_alpha(self.h, a))

    def set_floor_friction(self, mu):
        _check(lib().bm_mpm_set_floor_friction(self.h, mu))

    def set_walls(self, thickness):
        _check(lib().bm_mpm_set_walls(self.h, int(thickness)))

    # particles ---------------------------------------------------------
    def add_particles(self, pos, vel, mass):
        pos = _f32(pos, -1)
        n = pos.size // 3
        vel = _f32(vel, -1)
        mass = _f32(np.broadcast_to(np.asarray(mass, np.float32), (n,)), -1)
        _check(lib().bm_mpm_add_particles(self.h, n, _ptr(pos), _ptr(vel), _ptr(mass)))

    def upload(self, P):
        n = int(np.asarray(P["x"]).reshape(-1, 3).shape[0])
        arrs = {k: _f32(P[k], -1) for k in ("x", "v", "F", "C", "m", "V0")}
        hp = HostParticles(*[arrs[k].ctypes.data_as(C.POINTER(C.c_float)) for k in ("x", "v", "F", "C", "m", "V0")])
        _check(lib().bm_mpm_upload_particles(self.h, n, C.byref(hp)))

    def num_particles(self):
        n = C.c_int64()
        _check(lib().bm_mpm_num_particles(self.h, C.byref(n)))
        return n.value

    def download(self):
        n = self.num_particles()
        out = {"x": np.zeros((n, 3), np.float32), "v": np.zeros((n, 3), np.float32),
               "F": np.zeros((n, 3, 3), np.float32), "C": np.zeros((n, 3, 3), np.float32),
               "m": np.zeros(n, np.float32), "V0": np.zeros(n, np.float32)}
        hp = HostParticles(*[out[k].ctypes.data_as(C.POINTER(C.c_float)) for k in ("x", "v", "F", "C", "m", "V0")])
        _check(lib().bm_mpm_download_particles(self.h, C.byref(hp)))
        return out

    def device_particles(self):
        dp = DeviceParticles()
        _check(lib().bm_mpm_particles_device(self.h, C.byref(dp)))
        return dp

    def commit_device_particles(self, n):
        _check(lib().bm_mpm_commit_device_particles(self.h, int(n)))

    # substep -----------------------------------------------------------
    def step(self, dt, k=1):
        if k == 1:
            _check(lib().bm_mpm_step(self.h, dt))
        else:
            _check(lib().bm_mpm_step_n(self.h, dt, int(k)))

    def p2g(self, dt=0.0):
        _check(lib().bm_mpm_p2g(self.h, dt))

    def grid_update(self, dt):
        _check(lib().bm_mpm_grid_update(self.h, dt))

    def g2p(self, dt):
        _check(lib().bm_mpm_g2p(self.h, dt))

    def synchronize(self):
        _check(lib().bm_mpm_synchronize(self.h))

    def couple(self, method, col: Collider, dt):
        r = ReactionC()
        c = col.to_c()
        _check(lib().bm_mpm_couple(self.h, int(method), C.byref(c), float(dt), C.byref(r)))
        return np.array(r.force[:], np.float64), np.array(r.torque[:], np.float64), int(r.nodes_inside)

    def couple_noalloc(self, method, col: Collider, dt, reaction_device_ptr):
        """bm_mpm_couple_noalloc: the reaction (28 bytes, bm_reaction) goes to
        the given device address on the handle's stream; nothing waits."""
        c = col.to_c()
        _check(lib().bm_mpm_couple_noalloc(self.h, int(method), C.byref(c), float(dt),
                                           C.c_void_p(int(reaction_device_ptr))))

    # device body -------------------------------------------------------
    def set_body(self, body: Body):
        b = body.to_c()
        _check(lib().bm_mpm_set_body(self.h, C.byref(b)))

    def get_body(self):
        """(Body, (force, torque, nodes_inside) of the last coupling)."""
        b, r = BodyC(), ReactionC()
        _check(lib().bm_mpm_get_body(self.h, C.byref(b), C.byref(r)))
        return Body.from_c(b), (np.array(r.force[:], np.float64), np.array(r.torque[:], np.float64),
                                int(r.nodes_inside))

    def couple_body(self, method, dt, wait=False):
        """Couple the device body; with wait the reaction is returned."""
        if not wait:
            _check(lib().bm_mpm_couple_body(self.h, int(method), float(dt), None))
            return None
        r = ReactionC()
        _check(lib().bm_mpm_couple_body(self.h, int(method), float(dt), C.byref(r)))
        return np.array(r.force[:], np.float64), np.array(r.torque[:], np.float64), int(r.nodes_inside)

    def step_coupled(self, method, dt, k=1):
        _check(lib().bm_mpm_step_coupled(self.h, int(method), float(dt), int(k)))

    # grid --------------------------------------------------------------
    def download_grid(self, org=None, dims=None):
        org = np.ascontiguousarray(self.origin if org is None else org, np.int32)
        dims = np.ascontiguousarray(self.dims if dims is None else dims, np.int32)
        sh = tuple(int(v) for v in dims)
        nn = int(np.prod(sh))
        mq = np.zeros(nn, np.int64)
        mass = np.zeros(nn, np.float32)
        vel = np.zeros((nn, 3), np.float32)
        force = np.zeros((nn, 3), np.float32)
        act = np.zeros(nn, np.uint8)
        alloc = np.zeros(nn, np.uint8)
        _check(lib().bm_mpm_download_grid(self.h, _ptr(org), _ptr(dims), _ptr(mq), _ptr(mass), _ptr(vel),
                                          _ptr(force), _ptr(act), _ptr(alloc)))
        return {"mass_q": mq.reshape(sh), "mass": mass.reshape(sh), "vel": vel.reshape(sh + (3,)),
                "force": force.reshape(sh + (3,)), "active": act.reshape(sh).astype(bool),
                "alloc": alloc.reshape(sh).astype(bool)}

    def upload_grid(self, org, mass, vel, active):
        org = np.ascontiguousarray(org, np.int32)
        mass = _f32(mass)
        dims = np.ascontiguousarray(mass.shape, np.int32)
        vel = _f32(vel)
        act = np.ascontiguousarray(active, np.uint8)
        _check(lib().bm_mpm_upload_grid(self.h, _ptr(org), _ptr(dims), _ptr(mass), _ptr(vel), _ptr(act)))

    def grid_settle(self, dvy):
        """The research harness's grid step (bm_mpm_grid_settle): v_y += dvy on
        active nodes, then clamped at 0."""
        _check(lib().bm_mpm_grid_settle(self.h, float(dvy)))

    def grid_mass_total(self):
        t = C.c_int64()
        _check(lib().bm_mpm_grid_mass_total(self.h, C.byref(t)))
        return t.value

    def mass_quantum_exp(self):
        e = C.c_int32()
        _check(lib().bm_mpm_mass_quantum_exp(self.h, C.byref(e)))
        return e.value

    def set_tuning(self, unit_particles=0, sort_every=0, variant=-1):
        """Fast-mode launch parameters (bm_mpm_set_tuning); 0 / -1 keep the current value."""
        _check(lib().bm_mpm_set_tuning(self.h, int(unit_particles), int(sort_every), int(variant)))

    def set_timing(self, on=True, kernels_only=False, g2p_only=False):
        """on: per-phase HIP events.  kernels_only: only the fast form's G2P
        and P2G kernels, timed by events their own dispatches record (no
        events between launches to perturb the stream); g2p_only: the G2P
        kernel's alone."""
        level = (3 if g2p_only else 2 if kernels_only else 1) if on else 0
        _check(lib().bm_mpm_set_timing(self.h, level))

    def read_timing(self, reset=True):
        """{phase: (total_ms, launches)} from HIP events on the handle's stream."""
        ms = np.zeros(len(PHASES), np.float64)
        calls = np.zeros(len(PHASES), np.int64)
        _check(lib().bm_mpm_read_timing(self.h, 1 if reset else 0, _ptr(ms), _ptr(calls)))
        return {p: (float(ms[i]), int(calls[i])) for i, p in enumerate(PHASES)}

    def stats(self):
        a, f = C.c_int64(), C.c_int64()
        _check(lib().bm_mpm_stats(self.h, C.byref(a), C.byref(f)))
        return {"active_blocks": a.value, "fp64_fallbacks": f.value}


# ------------------------------------------------------------------ MPMSolver mirror
class MPMSolver:
    """``basements::mpm::MPMSolver`` (mpm_solver.h:15-95) on the GPU.

    Particles are staged on the host until the first step; the device grid is
    then sized to the particle cloud plus ``margin_cells`` on every side (the
    reference's grid is unbounded and sparse).  ``get_particles_mutable()``
    returns host arrays whose edits are uploaded before the next step.
    """

    def __init__(self, mode=MODE_REFERENCE, margin_cells=16, capacity=None):
        self.dx = 0.1  # MPMSolver() : grid(0.1f)  (mpm_solver.h:17)
        self.mode = mode
        self.margin = int(margin_cells)
        self.capacity = capacity
        self._E, self._nu, self._rho = 1.0e5, 0.3, 1500.0
        self._alpha, self._mu_floor = 0.40, 0.4
        self._g = (0.0, -9.81, 0.0)
        self._host = {k: [] for k in ("x", "v", "m")}
        self._P = None          # host mirror (dict of arrays) when valid
        self._dirty = False     # host mirror edited since the last upload
        self._dev = None

    def initialize(self, dx):
        self.dx = float(dx)
        self._dev = None

    def add_particle(self, pos, vel, mass=1.0):
        P = self._materialize()
        m = np.float32(mass)
        V0 = (m / np.float32(self._rho)) if np.float32(self._rho) > np.float32(1e-12) else np.float32(1.0)
        for k, val in (("x", pos), ("v", vel)):
            P[k] = np.concatenate([P[k], np.asarray(val, np.float32).reshape(1, 3)])
        P["F"] = np.concatenate([P["F"], np.eye(3, dtype=np.float32)[None]])
        P["C"] = np.concatenate([P["C"], np.zeros((1, 3, 3), np.float32)])
        P["m"] = np.concatenate([P["m"], [m]]).astype(np.float32)
        P["V0"] = np.concatenate([P["V0"], [V0]]).astype(np.float32)
        self._dirty = True

    def set_material(self, E, nu, rho):
        self._E, self._nu = float(E), float(nu)
        P = self._materialize()
        old, new = np.float32(self._rho), np.float32(rho)
        if new > 1e-12 and old > 1e-12 and new != old and P["m"].size:
            P["V0"] = (P["V0"] * (old / new)).astype(np.float32)
            self._dirty = True
        self._rho = float(rho)

    def set_gravity(self, g):
        self._g = tuple(float(v) for v in g)

    def set_plastic_alpha(self, a):
        self._alpha = float(a)

    def set_floor_friction(self, mu):
        self._mu_floor = float(mu)

    def get_particles(self):
        return self._materialize()

    def get_particles_mutable(self):
        P = self._materialize()
        self._dirty = True
        return P

    def get_grid(self):
        if self._dev is None:
            return None
        return self._dev.download_grid()

    def step(self, dt, tape=None):
        """MPMSolver::step(dt[, tape]) (mpm_solver.h:62-87); the reference's
        tape records nothing, so `tape` is left as it is."""
        self._sync_device()
        self._dev.step(dt)
        self._P = None

    # internals ---------------------------------------------------------
    def _materialize(self):
        if self._P is None:
            if self._dev is not None:
                self._P = self._dev.download()
            else:
                self._P = {"x": np.zeros((0, 3), np.float32), "v": np.zeros((0, 3), np.float32),
                           "F": np.zeros((0, 3, 3), np.float32), "C": np.zeros((0, 3, 3), np.float32),
                           "m": np.zeros(0, np.float32), "V0": np.zeros(0, np.float32)}
        return self._P

    def _domain_for(self, x):
        lo = np.floor(x.min(0) / self.dx - 0.5).astype(int) - self.margin
        hi = np.floor(x.max(0) / self.dx - 0.5).astype(int) + 2 + self.margin
        return lo, hi - lo + 1

    def _sync_device(self):
        P = self._materialize()
        n = P["x"].shape[0]
        if self._dev is None or n > self._dev.cap:
            org, dims = self._domain_for(P["x"]) if n else (np.zeros(3, int), np.full(3, 4))
            cap = max(n, self.capacity or 0, 1)
            self._dev = Solver(self.dx, org, dims, cap, self.mode)
            self._dirty = True
        d = self._dev
        d.set_material(self._E, self._nu, self._rho)
        d.set_gravity(self._g)
        d.set_plastic_alpha(self._alpha)
        d.set_floor_friction(self._mu_floor)
        if self._dirty:
            d.upload(P)
            self._dirty = False


# ------------------------------------------------------------------ Batch
@dataclass
class SceneDesc:
    """One scene of a batch (bm_scene_desc): its node box, material, floor,
    gravity and the coupler of its body."""

    dx: float
    origin: tuple
    dims: tuple
    E: float = 1.0e5
    nu: float = 0.3
    rho: float = 1500.0
    alpha: float = 0.40
    mu_floor: float = 0.4
    gravity: tuple = (0.0, -9.81, 0.0)
    walls: int = 0
    coupler: int = M0

    def to_c(self) -> SceneDescC:
        d = SceneDescC()
        d.dx = float(self.dx)
        d.grid_origin[:] = [int(v) for v in self.origin]
        d.grid_dims[:] = [int(v) for v in self.dims]
        d.young_modulus, d.poisson_ratio, d.bulk_density = float(self.E), float(self.nu), float(self.rho)
        d.plastic_alpha, d.floor_friction = float(self.alpha), float(self.mu_floor)
        d.gravity[:] = [float(v) for v in self.gravity]
        d.walls, d.coupler = int(self.walls), int(self.coupler)
        return d


class Batch:
    """Many scenes in shared device arrays, stepped together (bm_batch_*)."""

    def __init__(self, max_particles, max_blocks, max_scenes, stream=None):
        d = BatchDescC()
        d.max_particles, d.max_blocks, d.max_scenes, d.stream = int(max_particles), int(max_blocks), \
            int(max_scenes), stream
        h = C.c_void_p()
        _check(lib().bm_batch_create(C.byref(d), C.byref(h)))
        self.h = h
        self.n = {}

    def close(self):
        if getattr(self, "h", None):
            if not getattr(self, "_borrowed", False):
                lib().bm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, desc: SceneDesc, P, body: Body | None = None) -> int:
        """Add a scene (full particle state as Solver.upload); returns its id.
        Raises BMError(BM_ERR_OOM) when the batch has no room."""
        n = int(np.asarray(P["x"]).reshape(-1, 3).shape[0])
        arrs = {k: _f32(P[k], -1) for k in ("x", "v", "F", "C", "m", "V0")}
        hp = HostParticles(*[arrs[k].ctypes.data_as(C.POINTER(C.c_float)) for k in ("x", "v", "F", "C", "m", "V0")])
        sd = desc.to_c()
        sid = C.c_int32()
        bc = body.to_c() if body is not None else None
        _check(lib().bm_batch_add_scene(self.h, C.byref(sd), n, C.byref(hp), C.byref(bc) if bc is not None else None,
                                        C.byref(sid)))
        self.n[sid.value] = n
        return sid.value

    def remove_scene(self, sid):
        _check(lib().bm_batch_remove_scene(self.h, int(sid)))
        self.n.pop(int(sid), None)

    def step(self, dt, k=1):
        _check(lib().bm_batch_step(self.h, float(dt), int(k)))

    def synchronize(self):
        _check(lib().bm_batch_synchronize(self.h))

    def download_scene(self, sid):
        n = self.n[int(sid)]
        out = {"x": np.zeros((n, 3), np.float32), "v": np.zeros((n, 3), np.float32),
               "F": np.zeros((n, 3, 3), np.float32), "C": np.zeros((n, 3, 3), np.float32),
               "m": np.zeros(n, np.float32), "V0": np.zeros(n, np.float32)}
        hp = HostParticles(*[out[k].ctypes.data_as(C.POINTER(C.c_float)) for k in ("x", "v", "F", "C", "m", "V0")])
        _check(lib().bm_batch_download_scene(self.h, int(sid), C.byref(hp)))
        return out

    def get_body(self, sid):
        b, r = BodyC(), ReactionC()
        _check(lib().bm_batch_get_body(self.h, int(sid), C.byref(b), C.byref(r)))
        return Body.from_c(b), (np.array(r.force[:], np.float64), np.array(r.torque[:], np.float64),
                                int(r.nodes_inside))

    def usage(self):
        s, p, b = C.c_int32(), C.c_int64(), C.c_int64()
        _check(lib().bm_batch_usage(self.h, C.byref(s), C.byref(p), C.byref(b)))
        return s.value, p.value, b.value

    def scene_error(self, sid):
        """BM_OK (0) or the error code the scene's download would raise."""
        c = C.c_int32()
        _check(lib().bm_batch_scene_error(self.h, int(sid), C.byref(c)))
        return c.value

    def set_tuning(self, unit_particles=0, sort_every=0, deposit_groups=0):
        _check(lib().bm_batch_set_tuning(self.h, int(unit_particles), int(sort_every), int(deposit_groups)))

    def autotune(self, dt, substeps=32):
        """(unit_particles, sort_every, deposit_groups) of the fastest candidate."""
        u, s, g = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().bm_batch_autotune(self.h, float(dt), int(substeps), C.byref(u), C.byref(s), C.byref(g)))
        return u.value, s.value, g.value

    def set_graph(self, on=True):
        _check(lib().bm_batch_set_graph(self.h, 1 if on else 0))


def _host_particles(P):
    """HostParticles over float32 copies of P's fields; returns (struct, arrays)."""
    arrs = {k: _f32(P[k], -1) for k in ("x", "v", "F", "C", "m", "V0")}
    hp = HostParticles(*[arrs[k].ctypes.data_as(C.POINTER(C.c_float)) for k in ("x", "v", "F", "C", "m", "V0")])
    return hp, arrs


def _empty_particles(n):
    return {"x": np.zeros((n, 3), np.float32), "v": np.zeros((n, 3), np.float32),
            "F": np.zeros((n, 3, 3), np.float32), "C": np.zeros((n, 3, 3), np.float32),
            "m": np.zeros(n, np.float32), "V0": np.zeros(n, np.float32)}


@dataclass
class RequestInfo:
    state: int
    status: int
    substeps_done: int
    arrival: float
    admitted: float
    finished: float
    body: Body | None
    reaction: tuple | None

    @property
    def latency(self):
        """Seconds from arrival to retirement."""
        return self.finished - self.arrival


def batch_footprint(particles, blocks, scenes):
    """Device bytes of a batch with these capacities (bm_batch_footprint)."""
    b = C.c_int64()
    _check(lib().bm_batch_footprint(int(particles), int(blocks), int(scenes), C.byref(b)))
    return b.value


class Server:
    """Request loop over one batch (bm_server_*): submit rollout requests with
    arrival times, poll to admit / advance / retire, read results when done.
    The Python side keeps every request's input and output arrays alive."""

    def __init__(self, max_particles, max_blocks, max_scenes, dt, hbm_budget=0, max_chunk=0, max_bypass=0,
                 stream=None, overlap=True):
        d = ServerDescC()
        d.batch.max_particles, d.batch.max_blocks, d.batch.max_scenes = int(max_particles), int(max_blocks), \
            int(max_scenes)
        d.batch.stream = stream
        d.hbm_budget, d.dt, d.max_chunk, d.max_bypass = int(hbm_budget), float(dt), int(max_chunk), int(max_bypass)
        h = C.c_void_p()
        _check(lib().bm_server_create(C.byref(d), C.byref(h)))
        self.h = h
        self._keep = {}  # id -> (input structs and arrays, output arrays, has_body)
        if not overlap:
            self.set_overlap(False)

    def set_overlap(self, on):
        """Overlap host staging with stepping (bm_server_set_overlap; on by default)."""
        _check(lib().bm_server_set_overlap(self.h, int(bool(on))))

    def close(self):
        if getattr(self, "h", None):
            lib().bm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self):
        t = C.c_double()
        _check(lib().bm_server_clock(self.h, C.byref(t)))
        return t.value

    def batch(self):
        """The server's batch (for set_tuning / autotune / set_graph); not owned."""
        h = C.c_void_p()
        _check(lib().bm_server_batch(self.h, C.byref(h)))
        b = Batch.__new__(Batch)
        b.h, b.n, b._borrowed = h, {}, True
        return b

    def submit(self, desc: SceneDesc, P, substeps, body: Body | None = None, arrival=-1.0, output=True):
        """Queue a request; returns its id.  BMError(BM_ERR_OOM) when it could
        never fit.  `arrival` is on the server clock (< 0 = now)."""
        n = int(np.asarray(P["x"]).reshape(-1, 3).shape[0])
        hp, arrs = _host_particles(P)
        out = _empty_particles(n) if output else None
        op = _host_particles(out)[0] if output else None
        r = RequestC()
        r.scene = desc.to_c()
        r.n = n
        r.input = C.pointer(hp)
        bc = body.to_c() if body is not None else None
        r.body = C.pointer(bc) if bc is not None else None
        r.output = C.pointer(op) if op is not None else None
        r.substeps = int(substeps)
        r.arrival = float(arrival)
        rid = C.c_int64()
        _check(lib().bm_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = (hp, arrs, op, out, body is not None)
        return rid.value

    def poll(self, capacity=64):
        """One scheduling round; returns the ids finished since the last report."""
        ids = (C.c_int64 * capacity)()
        n = C.c_int32()
        _check(lib().bm_server_poll(self.h, ids, int(capacity), C.byref(n)))
        return [ids[i] for i in range(n.value)]

    def info(self, rid) -> RequestInfo:
        i = RequestInfoC()
        _check(lib().bm_server_info(self.h, int(rid), C.byref(i)))
        has_body = self._keep.get(int(rid), (None,) * 5)[4]
        done = i.state == REQ_DONE
        return RequestInfo(i.state, i.status, i.substeps_done, i.arrival, i.admitted, i.finished,
                           Body.from_c(i.body) if (has_body and done) else None,
                           (np.array(i.reaction.force[:]), np.array(i.reaction.torque[:]), int(i.reaction.nodes_inside))
                           if (has_body and done) else None)

    def result(self, rid):
        """The output particles of a finished request (insertion order)."""
        keep = self._keep[int(rid)]
        return keep[3]

    def release(self, rid):
        """Forget a finished request's buffers on the Python side."""
        self._keep.pop(int(rid), None)

    def cancel(self, rid):
        _check(lib().bm_server_cancel(self.h, int(rid)))

    def stats(self):
        """Seconds admitting / stepping / retiring / idle, chunks and substeps run."""
        t, c = (C.c_double * 4)(), (C.c_int64 * 2)()
        _check(lib().bm_server_stats(self.h, t, c))
        return {"admit_s": t[0], "step_s": t[1], "retire_s": t[2], "idle_s": t[3], "chunks": c[0], "substeps": c[1]}

    def load(self):
        """(queued, running, bytes in use, budget)."""
        q, r, u, b = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        _check(lib().bm_server_load(self.h, C.byref(q), C.byref(r), C.byref(u), C.byref(b)))
        return q.value, r.value, u.value, b.value


# ------------------------------------------------------------------ adjoint
@dataclass
class DiffParams:
    """Parameters of a differentiable rollout (bm_diff_params)."""

    young_modulus: float = 3.5e5
    poisson_ratio: float = 0.3
    friction_deg: float = 35.0
    floor_friction: float = 0.4
    gravity: tuple = (0.0, -9.81, 0.0)
    tool: bool = False
    tool_center: tuple = (0.0, 0.0, 0.0)
    tool_half_extents: tuple = (0.01, 0.05, 0.075)
    tool_quat: tuple = (1.0, 0.0, 0.0, 0.0)
    tool_velocity: tuple = (0.0, 0.0, 0.0)
    tool_friction: float = 0.5

    def to_c(self) -> DiffParamsC:
        p = DiffParamsC()
        p.young_modulus, p.poisson_ratio = float(self.young_modulus), float(self.poisson_ratio)
        p.friction_deg, p.floor_friction = float(self.friction_deg), float(self.floor_friction)
        p.gravity[:] = [float(v) for v in self.gravity]
        p.has_tool = 1 if self.tool else 0
        p.tool_center[:] = [float(v) for v in self.tool_center]
        p.tool_half_extents[:] = [float(v) for v in self.tool_half_extents]
        p.tool_quat[:] = [float(v) for v in self.tool_quat]
        p.tool_velocity[:] = [float(v) for v in self.tool_velocity]
        p.tool_friction = float(self.tool_friction)
        return p


def _grad_dict(g: DiffGradC) -> dict:
    return {"tool_velocity": np.array(g.tool_velocity[:]), "friction_deg": g.friction_deg,
            "young_modulus": g.young_modulus, "alpha": g.alpha, "mu": g.mu, "lambda": g.lambda_}


class Rollout:
    """Differentiable rollout with a hand-written adjoint (bm_diff_*).
    ``rollouts`` > 1 runs that many rollouts of the same particles side by
    side, each with its own parameters (set_params(p, r)); the *_all methods
    return one result per rollout, the others rollout 0's."""

    def __init__(self, dx, origin, dims, max_particles, precision=32, checkpoint_every=16, reference_svd=False,
                 stream=None, rollouts=1):
        d = DiffDescC()
        d.dx = float(dx)
        d.grid_origin[:] = [int(v) for v in origin]
        d.grid_dims[:] = [int(v) for v in dims]
        d.max_particles, d.precision, d.checkpoint_every = int(max_particles), int(precision), int(checkpoint_every)
        d.reference_svd = 1 if reference_svd else 0
        d.rollouts = int(rollouts)
        d.stream = stream
        h = C.c_void_p()
        _check(lib().bm_diff_create(C.byref(d), C.byref(h)))
        self.h = h
        self.n = 0
        self.rollouts = max(1, int(rollouts))

    def close(self):
        if getattr(self, "h", None):
            lib().bm_diff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: DiffParams, rollout=None):
        """Parameters of every rollout, or of rollout ``rollout`` alone."""
        pc = p.to_c()
        if rollout is None:
            _check(lib().bm_diff_set_params(self.h, C.byref(pc)))
        else:
            _check(lib().bm_diff_set_rollout_params(self.h, int(rollout), C.byref(pc)))

    def set_grid_budget(self, nbytes: int):
        """Device bytes the forward may keep its substeps' grids in for the
        backward (default 4 GiB; 0 = recompute every segment's grids)."""
        _check(lib().bm_diff_set_grid_budget(self.h, int(nbytes)))

    def upload(self, P, target=None):
        n = int(np.asarray(P["x"]).reshape(-1, 3).shape[0])
        hp, arrs = _host_particles(P)
        t = None if target is None else _f32(target, -1)
        _check(lib().bm_diff_upload(self.h, n, C.byref(hp),
                                    t.ctypes.data_as(C.POINTER(C.c_float)) if t is not None else None))
        self.n = n

    def forward(self, dt, substeps):
        loss = C.c_double()
        _check(lib().bm_diff_forward(self.h, float(dt), int(substeps), C.byref(loss)))
        return loss.value

    def forward_all(self, dt, substeps):
        """Every rollout's loss."""
        losses = (C.c_double * self.rollouts)()
        _check(lib().bm_diff_forward_batch(self.h, float(dt), int(substeps), losses))
        return np.array(losses[:])

    def backward(self):
        """dict of d loss / d parameter."""
        g = DiffGradC()
        _check(lib().bm_diff_backward(self.h, C.byref(g)))
        return _grad_dict(g)

    def backward_all(self):
        """Every rollout's gradient dict."""
        gs = (DiffGradC * self.rollouts)()
        _check(lib().bm_diff_backward_batch(self.h, gs))
        return [_grad_dict(g) for g in gs]

    def download(self, rollout=0):
        out = _empty_particles(self.n)
        hp = _host_particles(out)[0]
        _check(lib().bm_diff_download_rollout(self.h, int(rollout), C.byref(hp)))
        return out


MAP_RETURN = 0  # bm_diff_map_vjp: the Drucker-Prager return map F -> U h(sigma) V^T
MAP_STRESS = 1  # the first Piola-Kirchhoff stress P(F) of the Hencky model


def map_vjp(which, F, Gbar, params: DiffParams, precision=64):
    """(G, Fbar) of one of the rollout's per-particle SVD maps (bm_diff_map_vjp):
    G = map(F) and Fbar = Gbar . dG/dF for arrays of 3x3 matrices."""
    F = np.ascontiguousarray(np.asarray(F, np.float64).reshape(-1, 9))
    Gb = np.ascontiguousarray(np.asarray(Gbar, np.float64).reshape(-1, 9))
    n = F.shape[0]
    G = np.zeros((n, 9))
    Fb = np.zeros((n, 9))
    pc = params.to_c()
    _check(lib().bm_diff_map_vjp(int(which), n, _ptr(F), _ptr(Gb), C.byref(pc), int(precision), _ptr(G), _ptr(Fb)))
    return G.reshape(-1, 3, 3), Fb.reshape(-1, 3, 3)
