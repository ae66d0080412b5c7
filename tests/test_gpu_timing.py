"""Device timing levels of a handle (bm_mpm_set_timing): level 1 records
event pairs around every phase, level 2 only the fast form's G2P and P2G
kernels through events that their own dispatches record, level 3 the G2P
kernel's alone."""
from __future__ import annotations

import pytest

from _util import gpu_solver, params_from, random_scene

pytestmark = pytest.mark.gpu

DT = 1e-4


def _solver():
    P = random_scene(71, dx=1 / 32)
    return gpu_solver(P, 1 / 32, (0, -4, 0), (24, 16, 24), params_from(), mode=1)


def test_kernels_only_times_g2p_and_p2g_per_substep():
    S = _solver()
    S.step(DT, 2)
    S.set_timing(True, kernels_only=True)
    S.read_timing(reset=True)
    S.step(DT, 7)
    tm = S.read_timing(reset=True)
    assert tm["g2p"][1] == 7 and tm["p2g"][1] == 7
    assert tm["g2p"][0] > 0.0 and tm["p2g"][0] > 0.0
    for ph in ("clear", "update", "couple", "sort", "other"):
        assert tm[ph][1] == 0, ph


def test_g2p_only_times_the_particle_pass_alone():
    S = _solver()
    S.step(DT, 2)
    S.set_timing(True, g2p_only=True)
    S.read_timing(reset=True)
    S.step(DT, 6)
    tm = S.read_timing(reset=True)
    assert tm["g2p"][1] == 6 and tm["g2p"][0] > 0.0
    for ph in ("clear", "p2g", "update", "couple", "sort", "other"):
        assert tm[ph][1] == 0, ph


def test_every_phase_level_and_off():
    S = _solver()
    S.step(DT, 2)
    S.set_timing(True)
    S.read_timing(reset=True)
    S.step(DT, 5)
    tm = S.read_timing(reset=True)
    # (the fast form's grid update also clears the next grid: no clear phase per substep)
    for ph in ("update", "g2p", "p2g"):
        assert tm[ph][1] >= 5, ph
    S.set_timing(False)
    S.step(DT, 3)
    tm = S.read_timing(reset=True)
    assert all(calls == 0 for _, calls in tm.values())


def test_unknown_level_is_refused():
    from basements_amd import BMError
    from basements_amd.mpm import _check, lib

    S = _solver()
    with pytest.raises(BMError, match="INVALID_ARG"):
        _check(lib().bm_mpm_set_timing(S.h, 4))
