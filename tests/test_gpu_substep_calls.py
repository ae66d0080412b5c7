"""A call of k substeps leaves the same particle state as k calls of one.

Inside a multi-substep call the fast G2P writes the particle velocity and C
on the call's last substep only (the deposit record carries what the next
P2G needs), and a re-binning in front of a G2P does not carry them over.
These tests pin that every call still ends with the full state: a handle or
batch stepped in long calls -- crossing re-binnings and, for batches, graph
replays that end a call and replays that do not -- against the same scene
stepped one substep per call, where every G2P writes everything.  Momenta are
FP32 atomic sums, so the bar is a relative 1e-4 of each field's scale (the
batch tests' bar), not bit equality."""
from __future__ import annotations

import numpy as np
import pytest

from _util import assert_close_field, gpu_solver, params_from, random_scene
from test_gpu_batch import DT, _scenes

pytestmark = pytest.mark.gpu


def _same(got, want, what, rtol=1e-4):
    for key in ("x", "v", "F", "C"):
        assert_close_field(got[key], want[key], rtol, f"{what} {key}")


def _one_by_one(S, body, coupler, n):
    for _ in range(n):
        if body is not None:
            S.step_coupled(coupler, DT, 1)
        else:
            S.step(DT, 1)


def test_single_handle_long_calls_match_single_substep_calls():
    P, prm, sd, body = _scenes()[1]  # the rolling wheel on M1
    calls = [5, 16, 16, 1, 9]        # crosses a re-binning (every 32nd substep by default)
    L = gpu_solver(P, sd.dx, sd.origin, sd.dims, prm, mode=1)
    O = gpu_solver(P, sd.dx, sd.origin, sd.dims, prm, mode=1)
    L.set_body(body)
    O.set_body(body)
    for n in calls:
        L.step_coupled(sd.coupler, DT, n)
        _one_by_one(O, body, sd.coupler, n)
        _same(L.download(), O.download(), f"after a call of {n}")


def test_single_handle_uncoupled_call_matches():
    P = random_scene(61, dx=1 / 32)
    prm = params_from()
    L = gpu_solver(P, 1 / 32, (0, -4, 0), (24, 16, 24), prm, mode=1)
    O = gpu_solver(P, 1 / 32, (0, -4, 0), (24, 16, 24), prm, mode=1)
    L.step(DT, 33)
    _one_by_one(O, None, None, 33)
    _same(L.download(), O.download(), "33 substeps")


@pytest.mark.parametrize("graph", [False, True])
def test_batch_calls_match_scenes_stepped_one_substep_per_call(graph):
    """Re-binning every 4 substeps: a replay covers 8.  The first call brings
    the batch to a replay's starting point in 4 eager substeps and ends on a
    replay; then a call of exactly one replay, one of a replay that does not
    end the call plus one that does, and one that ends on eager substeps."""
    from basements_amd import Batch

    scenes = _scenes()
    B = Batch(max_particles=200_000, max_blocks=40_000, max_scenes=8)
    B.set_graph(graph)
    ids = [B.add_scene(sd, P, body) for (P, prm, sd, body) in scenes]
    B.set_tuning(0, 4, 0)
    alone = []
    for P, prm, sd, body in scenes:
        S = gpu_solver(P, sd.dx, sd.origin, sd.dims, prm, mode=1)
        if body is not None:
            S.set_body(body)
        alone.append(S)
    for n in (12, 8, 16, 11):
        B.step(DT, n)
        for sid, S, (P, prm, sd, body) in zip(ids, alone, scenes):
            _one_by_one(S, body, sd.coupler, n)
            _same(B.download_scene(sid), S.download(), f"graph={graph} call of {n}, scene {sid}")


def test_velocity_and_c_change_between_substeps():
    """The comparisons above would not see a stale velocity or C if the
    fields stood still: one substep changes both well beyond the bar."""
    P = random_scene(62, dx=1 / 32)
    S = gpu_solver(P, 1 / 32, (0, -4, 0), (24, 16, 24), params_from(), mode=1)
    S.step(DT, 1)
    a = S.download()
    S.step(DT, 1)
    b = S.download()
    for key in ("v", "C"):
        d = np.abs(np.asarray(a[key], np.float64) - b[key]).max()
        assert d > 1e-2 * np.abs(b[key]).max(), key
