"""The request loop (bm_server_*): every request's result against the same
scene run alone on a single-scene fast handle for its own substep count,
with requests arriving while others run, admission held back by the HBM
budget, cancellation, preemption and resumption, and a failing request
beside healthy ones.

The bar is the batch's: a relative 1e-4 of each field's scale (node momenta
are FP32 atomics whose order differs between the mix and the scene alone)."""
from __future__ import annotations

import time

import numpy as np
import pytest

from test_gpu_batch import DT, _alone, _same, _same_body, _scenes

pytestmark = pytest.mark.gpu


def _server(budget=0, max_chunk=0, max_particles=200_000, max_scenes=8):
    from basements_amd import Server

    return Server(max_particles=max_particles, max_blocks=40_000, max_scenes=max_scenes, dt=DT, hbm_budget=budget,
                  max_chunk=max_chunk)


def _drain(S, ids, limit_s=120.0):
    """Poll until every id has finished; returns the largest number of
    requests seen running at once."""
    left, most = set(ids), 0
    t_end = time.monotonic() + limit_s
    while left:
        assert time.monotonic() < t_end, f"requests {sorted(left)} did not finish"
        for rid in S.poll():
            left.discard(rid)
        most = max(most, S.load()[1])
    return most


def _check_alone(S, rid, scene, substeps, what):
    from basements_amd import REQ_DONE

    P, prm, sd, body = scene
    inf = S.info(rid)
    assert inf.state == REQ_DONE and inf.substeps_done == substeps, (what, inf)
    assert inf.arrival <= inf.admitted <= inf.finished
    want, wbody = _alone(P, prm, sd, body, substeps)
    _same(S.result(rid), want, what)
    if body is not None:
        _same_body(inf.body, wbody, what)


def test_each_request_alone_and_in_the_mix():
    """Three requests with different substep counts, one arriving while the
    others run: each result equals its scene alone."""
    scenes = _scenes()
    S = _server(max_chunk=8)
    counts = [20, 13, 27]
    ids = [S.submit(scenes[0][2], scenes[0][0], counts[0], body=scenes[0][3]),
           S.submit(scenes[2][2], scenes[2][0], counts[2])]
    S.poll()  # one chunk of 8 substeps runs before the third request arrives
    assert S.load()[1] == 2
    ids.insert(1, S.submit(scenes[1][2], scenes[1][0], counts[1], body=scenes[1][3]))
    most = _drain(S, ids)
    assert most == 3  # all three on the device together
    for rid, sc, k in zip(ids, scenes, counts):
        _check_alone(S, rid, sc, k, f"request {rid}")
    assert S.load()[:3] == (0, 0, 0)


def test_budget_holds_back_admission():
    """A budget that holds one request at a time: the others wait in the
    queue, run after it, and still match their scenes alone."""
    from basements_amd import batch_footprint

    scenes = _scenes()
    one = max(batch_footprint(sc[0]["x"].shape[0], np.prod([(d + 3) // 4 for d in sc[2].dims]), 1)
              for sc in scenes)
    S = _server(budget=int(one * 1.01), max_chunk=4)
    counts = [9, 6, 11]
    ids = [S.submit(sc[2], sc[0], k, body=sc[3]) for sc, k in zip(scenes, counts)]
    most = _drain(S, ids)
    assert most == 1
    infos = [S.info(r) for r in ids]
    for a, b in zip(infos, infos[1:]):
        assert b.admitted >= a.finished - 1e-9  # arrival order kept, one at a time
    for rid, sc, k in zip(ids, scenes, counts):
        _check_alone(S, rid, sc, k, f"budgeted request {rid}")


def test_oversized_request_is_refused():
    from basements_amd import BMError

    scenes = _scenes()
    S = _server(max_particles=1000)
    with pytest.raises(BMError, match="OOM"):
        S.submit(scenes[0][2], scenes[0][0], 5)


def test_cancel_and_resume_from_a_checkpoint():
    """Cancel one queued and one running request; preempt another after k1
    substeps and resume it from its output for k2 more: the resumed result
    equals k1 + k2 substeps alone."""
    from basements_amd import REQ_CANCELLED, Body

    scenes = _scenes()
    S = _server(max_chunk=4)
    P, prm, sd, body = scenes[1]
    k1, k2 = 8, 10
    first = S.submit(sd, P, k1, body=body)
    doomed_running = S.submit(scenes[2][2], scenes[2][0], 40)
    doomed_queued = S.submit(scenes[0][2], scenes[0][0], 10, body=scenes[0][3], arrival=S.clock() + 60.0)
    S.poll()
    S.cancel(doomed_running)
    S.cancel(doomed_queued)
    assert S.info(doomed_running).state == REQ_CANCELLED and S.info(doomed_queued).state == REQ_CANCELLED
    _drain(S, [first])
    mid = S.result(first)
    mid_body = S.info(first).body
    resumed = S.submit(sd, mid, k2, body=mid_body)
    _drain(S, [resumed])
    want, wbody = _alone(P, prm, sd, body, k1 + k2)
    _same(S.result(resumed), want, "resumed request")
    _same_body(S.info(resumed).body, wbody, "resumed request")
    assert isinstance(mid_body, Body)


def test_failing_request_leaves_the_others_running():
    """A request whose particle starts outside its grid box fails with
    BM_ERR_OUT_OF_DOMAIN; the request beside it completes and matches."""
    from basements_amd import REQ_FAILED

    scenes = _scenes()
    S = _server(max_chunk=4)
    bad = {k: np.array(v, copy=True) for k, v in scenes[2][0].items()}
    bad["x"][0] = (50.0, 50.0, 50.0)
    good = S.submit(scenes[0][2], scenes[0][0], 12, body=scenes[0][3])
    failing = S.submit(scenes[2][2], bad, 12)
    _drain(S, [good, failing])
    inf = S.info(failing)
    assert inf.state == REQ_FAILED and inf.status == 6
    _check_alone(S, good, scenes[0], 12, "request beside a failing one")


def test_request_with_non_finite_state_fails_alone():
    from basements_amd import REQ_FAILED

    scenes = _scenes()
    S = _server(max_chunk=4)
    bad = {k: np.array(v, copy=True) for k, v in scenes[2][0].items()}
    bad["v"][5] = (np.nan, 0.0, 0.0)
    good = S.submit(scenes[0][2], scenes[0][0], 8, body=scenes[0][3])
    failing = S.submit(scenes[2][2], bad, 8)
    _drain(S, [good, failing])
    inf = S.info(failing)
    assert inf.state == REQ_FAILED and inf.status == 4  # BM_ERR_NONFINITE
    _check_alone(S, good, scenes[0], 8, "request beside a non-finite one")


@pytest.mark.parametrize("overlap", [True, False], ids=["overlap", "synchronous"])
def test_queued_requests_staged_while_the_device_runs(overlap):
    """A budget that holds one request at a time keeps the others queued
    while a chunk runs: with overlap on they are staged onto the device then
    (a non-finite one is caught there) and admitted from the staging area;
    retirements go out through it.  Both forms give each request's scene
    alone, and the non-finite request fails alone with BM_ERR_NONFINITE."""
    from basements_amd import REQ_FAILED, batch_footprint

    scenes = _scenes()
    one = max(batch_footprint(sc[0]["x"].shape[0], np.prod([(d + 3) // 4 for d in sc[2].dims]), 1)
              for sc in scenes)
    S = _server(budget=int(one * 1.01), max_chunk=4)
    S.set_overlap(overlap)
    bad = {k: np.array(v, copy=True) for k, v in scenes[1][0].items()}
    bad["F"][3, 1, 1] = np.inf
    counts = [9, 6, 7, 11]
    ids = [S.submit(scenes[0][2], scenes[0][0], counts[0], body=scenes[0][3]),
           S.submit(scenes[1][2], bad, counts[1], body=scenes[1][3]),
           S.submit(scenes[1][2], scenes[1][0], counts[2], body=scenes[1][3]),
           S.submit(scenes[2][2], scenes[2][0], counts[3])]
    most = _drain(S, ids)
    assert most == 1
    inf = S.info(ids[1])
    assert inf.state == REQ_FAILED and inf.status == 4, inf
    for i in (0, 2, 3):
        sc = scenes[{0: 0, 2: 1, 3: 2}[i]]
        _check_alone(S, ids[i], sc, counts[i], f"{'staged' if overlap else 'synchronous'} request {i}")
    assert S.load()[:3] == (0, 0, 0)
