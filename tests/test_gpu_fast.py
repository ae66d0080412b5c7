"""BM_MODE_FAST (fused G2P+P2G, block-binned particles, stress from the
plasticity SVD, force folded into momentum, fast reciprocal/sqrt/log/exp)
against the oracle.

Bars:
  * fixed-point node mass of a P2G from identical particle state: bit-exact;
  * total grid mass every substep: exactly the particles' quanta;
  * folded momentum vs the oracle's momentum + dt * force: rtol 2e-5 of the
    field's largest magnitude;
  * trajectories: measured against the FP64 oracle, the fast path must stay
    within 2.5x the distance of the FP32 oracle (the reference's own
    arithmetic) from that same FP64 run -- i.e. at the FP32 noise floor,
    which over 20 substeps of these scenes is ~1.6e-4 relative in v;
  * particles come back in insertion order after any number of re-binnings.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleSim
from tests._util import gpu_solver, oracle_for, params_from, random_scene, rel_err

pytestmark = pytest.mark.gpu

DX = 1.0 / 32.0
ORG = (0, -4, 0)
DIMS = (24, 16, 24)


def _pair(seed, sort_every=16, variant=None, unit=0, **kw):
    from basements_amd import MODE_FAST, Solver

    P = random_scene(seed, dx=DX, **kw)
    prm = params_from()
    g = Solver(DX, ORG, DIMS, P["x"].shape[0], MODE_FAST, sort_every=sort_every)
    if variant is not None or unit:
        g.set_tuning(unit, 0, -1 if variant is None else variant)
    g.set_material(prm.E, prm.nu, prm.rho)
    g.set_plastic_alpha(prm.alpha)
    g.set_floor_friction(prm.mu_floor)
    g.set_gravity(prm.gravity)
    g.upload(P)
    o = oracle_for(P, DX, prm, g.mass_quantum_exp())
    return P, prm, g, o


def _fp64_pair(P, prm):
    Pd = {k: np.asarray(v, np.float64) for k, v in P.items()}
    o64 = OracleSim(DX, "f64", params=prm)
    o32 = OracleSim(DX, "f32", params=prm)
    o64.set_particles(Pd)
    o32.set_particles(Pd)
    return o32, o64


def _assert_at_noise_floor(Pg, P32, P64, fields="xvFC", factor=2.5):
    for k in fields:
        floor = rel_err(P32[k], P64[k])
        got = rel_err(Pg[k], P64[k])
        assert got <= factor * floor + 1e-6, (k, got, floor)


def test_fast_p2g_mass_bit_exact_and_folded_momentum():
    P, prm, g, o = _pair(11)
    dt = 1e-4
    g.p2g(dt)
    o.p2g()
    G = g.download_grid()
    O = o.grid_dense(g.origin, g.dims)
    assert np.array_equal(G["mass_q"], O["mass_q"])
    assert np.array_equal(G["active"], O["active"])
    assert g.grid_mass_total() == o.grid_mass_q_total()
    want = O["vel"] + dt * O["force"]
    err = np.abs(G["vel"] - want).max() / np.abs(want).max()
    assert err < 2e-5, err


@pytest.mark.parametrize("variant", [4, 12, 4 | 32, 12 | 32])
def test_fast_steps_at_fp32_noise_floor(variant):
    P, prm, g, _ = _pair(12, variant=variant, strain=0.01, vel=0.05, cgrad=0.2)
    o32, o64 = _fp64_pair(P, prm)
    e = g.mass_quantum_exp()
    q = int(np.rint(np.ldexp(P["m"].astype(np.float64), -e)).sum())
    dt = 1e-4
    for _ in range(20):
        g.p2g(dt)
        assert g.grid_mass_total() == q
        g.grid_update(dt)
        g.g2p(dt)
    o32.step(dt, 20)
    o64.step(dt, 20)
    _assert_at_noise_floor(g.download(), o32.get_particles(), o64.get_particles())


@pytest.mark.parametrize("variant", [4, 12, 4 | 32])
def test_fast_order_survives_rebinning(variant):
    P, prm, g, _ = _pair(13, sort_every=3, unit=64, variant=variant, strain=0.01, vel=0.05, cgrad=0.2)
    o32, o64 = _fp64_pair(P, prm)
    dt = 1e-4
    g.step(dt, 10)
    o32.step(dt, 10)
    o64.step(dt, 10)
    Pg = g.download()
    assert np.array_equal(Pg["m"], P["m"]) and np.array_equal(Pg["V0"], P["V0"])
    _assert_at_noise_floor(Pg, o32.get_particles(), o64.get_particles())


def test_fast_matches_reference_mode_over_100_substeps():
    from basements_amd import MODE_REFERENCE

    P, prm, g, _ = _pair(14, strain=0.0, vel=0.0, cgrad=0.0)
    r = gpu_solver(P, DX, ORG, DIMS, prm, mode=MODE_REFERENCE)
    o32, o64 = _fp64_pair(P, prm)
    g.step(1e-4, 100)
    r.step(1e-4, 100)
    o32.step(1e-4, 100)
    o64.step(1e-4, 100)
    P64 = o64.get_particles()
    _assert_at_noise_floor(g.download(), o32.get_particles(), P64, fields="xF")
    _assert_at_noise_floor(r.download(), o32.get_particles(), P64, fields="xF")


@pytest.mark.parametrize("unit", [1024, 640, 64])
def test_fast_units_up_to_the_largest_size(unit):
    """A bed of 16 particles per cell (two interleaved jittered 8-per-cell
    beds), so an interior 4^3 block holds 1,024 particles: with units of 1,024
    (UNIT_MAX) each such block is one unit until the drift between
    re-binnings (every 4 substeps) pushes it past that, at 640 it is two.
    The staged G2P (forced) and the cell deposit then run the largest units
    the kernels accept; the grid mass stays bit-exact every substep and the
    state stays at the FP32 noise floor of the FP64 oracle."""
    from basements_amd import MODE_FAST, Solver
    from basements_amd.scenes import soil_block

    lo, hi = (6, 0, 6), (14, 5, 14)  # stencil bases 0-3 in y fully populated: one block layer of 1,024
    A, B = soil_block(DX, lo, hi, seed=21), soil_block(DX, lo, hi, seed=22)
    P = {k: np.concatenate([A[k], B[k]]) for k in A}
    n = P["x"].shape[0]
    rng = np.random.default_rng(23)
    P["v"] = (0.3 * rng.normal(size=(n, 3))).astype(np.float32)
    P["F"] = (np.eye(3) + 0.01 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    P["C"] = (0.2 * rng.normal(size=(n, 3, 3))).astype(np.float32)
    cell = np.floor(P["x"].astype(np.float64) / DX - 0.5).astype(np.int64)
    blk, counts = np.unique(cell // 4, axis=0, return_counts=True)
    assert counts.max() == 1024  # full interior blocks
    prm = params_from()
    g = Solver(DX, ORG, DIMS, n, MODE_FAST, sort_every=4)
    g.set_tuning(unit, 0, 4 | 32)
    g.set_material(prm.E, prm.nu, prm.rho)
    g.set_plastic_alpha(prm.alpha)
    g.set_floor_friction(prm.mu_floor)
    g.set_gravity(prm.gravity)
    g.upload(P)
    o32, o64 = _fp64_pair(P, prm)
    e = g.mass_quantum_exp()
    q = int(np.rint(np.ldexp(P["m"].astype(np.float64), -e)).sum())
    dt = 1e-4
    for _ in range(12):
        g.p2g(dt)
        assert g.grid_mass_total() == q
        g.grid_update(dt)
        g.g2p(dt)
    o32.step(dt, 12)
    o64.step(dt, 12)
    _assert_at_noise_floor(g.download(), o32.get_particles(), o64.get_particles())


def test_unit_sizes_beyond_the_largest_are_refused():
    from basements_amd import BMError, MODE_FAST, Solver

    g = Solver(DX, ORG, DIMS, 64, MODE_FAST)
    with pytest.raises(BMError, match="INVALID_ARG"):
        g.set_tuning(1025, 0, -1)
    with pytest.raises(BMError, match="INVALID_ARG"):
        g.set_tuning(15, 0, -1)


@pytest.mark.parametrize("variant", [4 | 16 | 32, 64])
def test_unknown_variants_are_refused(variant):
    """Bits 4 and 5 name the two G2P forms; asking for both, or for a bit
    above 5, is an invalid argument."""
    from basements_amd import BMError, MODE_FAST, Solver

    g = Solver(DX, ORG, DIMS, 16, MODE_FAST)
    with pytest.raises(BMError, match="INVALID_ARG"):
        g.set_tuning(0, 0, variant)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_retired_single_kernel_variants_are_refused(variant):
    """Without bit 2 the variant named the retired single-kernel G2P+P2G form."""
    from basements_amd import BMError, MODE_FAST, Solver

    g = Solver(DX, ORG, DIMS, 16, MODE_FAST)
    with pytest.raises(BMError, match="NOT_SUPPORTED"):
        g.set_tuning(0, 0, variant)


def test_fast_dt_change_redeposits():
    P, prm, g, _ = _pair(15, strain=0.01, vel=0.05, cgrad=0.2)
    o32, o64 = _fp64_pair(P, prm)
    for dt in (1e-4, 5e-5, 5e-5, 1e-4):
        g.step(dt)
        o32.step(dt)
        o64.step(dt)
    _assert_at_noise_floor(g.download(), o32.get_particles(), o64.get_particles())


def test_fast_material_change_redeposits():
    """Changing E after a substep must not reuse the stress already deposited."""
    P, prm, g, _ = _pair(16, strain=0.01, vel=0.05, cgrad=0.2)
    o32, o64 = _fp64_pair(P, prm)
    g.step(1e-4)
    o32.step(1e-4)
    o64.step(1e-4)
    from oracle.pyoracle import Params

    prm2 = Params(E=1.0e6, nu=prm.nu, rho=prm.rho, alpha=prm.alpha, mu_floor=prm.mu_floor, gravity=prm.gravity)
    g.set_material(prm2.E, prm2.nu, prm2.rho)
    o32.set_params(prm2)
    o64.set_params(prm2)
    g.step(1e-4, 3)
    o32.step(1e-4, 3)
    o64.step(1e-4, 3)
    _assert_at_noise_floor(g.download(), o32.get_particles(), o64.get_particles())


@pytest.mark.parametrize("variant", [4 | 32, 4 | 16])
def test_fast_particles_drifting_past_the_staged_tile(variant):
    """The staged G2P (variant bit 5) stages each re-binning unit's 8^3 nodes
    in LDS (the unit's block, 2 stencil layers and 1 drift layer each side)
    and gathers a particle whose stencil base left that range straight from
    the grid.  A block translating at 1.5 / 1.2 m/s with re-binning off for
    100 substeps drifts up to 2.4 cells from its binned blocks, so both
    gathers (and the deposit's off-tile path) run; the result stays at the
    FP32 noise floor, as it does for the per-particle G2P (bit 4)."""
    from basements_amd import MODE_FAST, Solver

    dt, n = 5e-4, 100
    P = random_scene(17, dx=DX, strain=0.01, vel=0.05, cgrad=0.2)
    P["v"] = (P["v"] + np.array([1.5, 0.0, 1.2], np.float32)).astype(np.float32)
    prm = params_from(g=(0.0, 0.0, 0.0))
    g = Solver(DX, ORG, DIMS, P["x"].shape[0], MODE_FAST, sort_every=100000)
    g.set_tuning(0, 0, variant)
    g.set_material(prm.E, prm.nu, prm.rho)
    g.set_plastic_alpha(prm.alpha)
    g.set_floor_friction(prm.mu_floor)
    g.set_gravity(prm.gravity)
    g.upload(P)
    o32, o64 = _fp64_pair(P, prm)
    g.step(dt, n)
    o32.step(dt, n)
    o64.step(dt, n)
    Pg = g.download()
    # stencil base relative to the block the particle was binned in (from the
    # initial state): the staged tile covers bases -1 .. 4 on each axis
    def base(x):
        return np.floor(x.astype(np.float64) / DX - 0.5).astype(np.int64) - np.array(ORG)

    b0, b1 = base(P["x"]), base(Pg["x"])
    rel = b1 - 4 * (b0 >> 2)
    outside = ((rel < -1) | (rel > 4)).any(axis=1)
    assert outside.sum() > 50, int(outside.sum())
    _assert_at_noise_floor(Pg, o32.get_particles(), o64.get_particles())
