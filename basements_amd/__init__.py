"""MI355X-native MLS-MPM soil substep (HIP kernels behind a C ABI).

The compute path is libbm_mpm.so (basements_amd/csrc, built for gfx950 by
basements_amd.build); this package holds its Python host mirror.
"""
from .mpm import (BODY_DYNAMIC, BODY_KINEMATIC, BODY_STATIC, M0, M1, M45, M45A, M45B, MAP_RETURN, MAP_STRESS,
                  MODE_FAST, MODE_REFERENCE, REQ_CANCELLED, REQ_DONE, REQ_FAILED, REQ_QUEUED, REQ_RUNNING, SHAPE_BOX, SHAPE_CYLINDER, Batch,
                  BMError, Body, Collider, DiffParams, MPMSolver, RequestInfo, Rollout, SceneDesc, Server, Solver,
                  batch_footprint, header_symbols, lib, map_vjp)

__all__ = ["BODY_DYNAMIC", "BODY_KINEMATIC", "BODY_STATIC", "M0", "M1", "M45", "M45A", "M45B", "MAP_RETURN",
           "MAP_STRESS", "MODE_FAST",
           "MODE_REFERENCE", "REQ_CANCELLED", "REQ_DONE", "REQ_FAILED", "REQ_QUEUED", "REQ_RUNNING", "SHAPE_BOX",
           "SHAPE_CYLINDER", "Batch", "BMError", "Body", "Collider", "DiffParams", "MPMSolver", "RequestInfo",
           "Rollout", "SceneDesc", "Server", "Solver", "batch_footprint", "header_symbols", "lib", "map_vjp"]
