"""The production tile shapes against the FP64 oracle, through the C ABI.

The context picks 4x4x4 tiles for large grids, and the autotuner times the
most compact 64-cell shapes (2x4x8, 2x8x4, 4x2x8) on the uneven-load box;
small grids fall back to 2x2x2 tiles.  These cases pass the tile shape
explicitly, so the 512-thread deposit (its one-lane-per-subcell and its
crowded-tile paths), the 11-bit ranking over 1,728 slots and the LDS sizes of
64-cell tiles all run against the oracle:
  * FP32 on an obstacle box with inflow, 20 full steps, at every production
    shape: n, J and faceB within 1e-5 (global relative L2) of the FP64
    oracle, particle counts within 2 per species;
  * FP32 at 96 ppc (every tile takes the crowded-tile deposit path);
  * FP64 at 4x4x4 (LDS above 64 KiB): the oracle to 1e-9 after 8 steps,
    identical particle counts;
  * a reduced uneven-load box (device uniform fill of H+, O+ photoions
    accumulated near an absorbing sphere, so that many tiles hold no O+ and a
    few are crowded) stepped against the oracle at every autotune shape: FP64
    to 1e-9 with identical particle counts; FP32 within 1e-5 plus the deposit
    of two particles.  A particle that ends a step within FP32 rounding
    (~1e-7 dx) of a face of the exterior layer or of the obstacle can cross it
    a step earlier or later in FP32 than in FP64, and the reference deletes a
    particle after depositing it in the exterior cell: for one step its
    deposit is in one side's moments and not in the other's, whatever the
    final counts (seen here at the +z wall).  One particle of weight w and
    charge q moves rho by at most q w / dV in L2 (its eight weights have L2
    norm <= 1) and J by that times its speed."""
import numpy as np
import pytest

from rhybrid_amd import problem as P

pytestmark = pytest.mark.gpu

_ORACLE = {}


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def oracle_after(oracle_mod, key, pr, steps, prepare=None):
    """The oracle's fields and particle counts after `steps` full steps
    (computed once per key: the oracle is deterministic)."""
    if key not in _ORACLE:
        ora = oracle_mod.Oracle(pr)
        if prepare is not None:
            prepare(ora)
        ora.step(steps)
        _ORACLE[key] = ({n: ora.field(n).copy() for n in ("cellRhoQi", "cellJi", "faceB", "nodeE")},
                        [ora.num_particles(s) for s in range(ora.nspecies)])
        del ora
    return _ORACLE[key]


def run_ctx(gpu_mod, pr, prec, tile, steps, prepare=None):
    print("case %s precision=%s tile=%s np=%d steps=%d" % (pr.name, "fp64" if prec else "fp32", tile,
                                                         pr.n_particles, steps), flush=True)
    ctx = gpu_mod.Context(pr, precision=prec, tile=tile)
    lay = ctx.layout()
    assert (lay["tx"], lay["ty"], lay["tz"]) == tuple(tile)
    if prepare is not None:
        prepare(ctx)
    ctx.step(steps)
    out = ({n: ctx.field(n) for n in ("cellRhoQi", "cellJi", "faceB", "nodeE")},
           [ctx.num_particles(s) for s in range(len(pr.species))])
    ctx.close()
    return out


def obstacle_inflow(ppc, n=32, seed=41):
    pr = P.config_obstacle(nx=n, ny=n, nz=n, ppc=ppc, R_cells=n / 8.0, seed=seed, inflow=True)
    for sp in pr.species:
        sp.capacity = 2 * pr.particles[0].n
    return pr


@pytest.mark.parametrize("tile", [(4, 4, 4), (2, 4, 8), (2, 8, 4), (4, 2, 8)])
def test_production_tiles_fp32_vs_oracle(gpu_mod, oracle_mod, tile):
    pr = obstacle_inflow(ppc=8)
    steps = 20
    fo, no = oracle_after(oracle_mod, "obstacle32_ppc8", pr, steps)
    fg, ng = run_ctx(gpu_mod, pr, gpu_mod.RH_FP32, tile, steps)
    for name in ("cellRhoQi", "cellJi", "faceB"):
        r = rel(fg[name], fo[name])
        print("tile %s: %s rel L2 = %.3e" % (tile, name, r), flush=True)
        assert r < 1e-5, (name, r)
    for s in range(2):
        assert abs(ng[s] - no[s]) <= 2, (s, ng[s], no[s])


def test_crowded_tiles_fp32_vs_oracle(gpu_mod, oracle_mod):
    """96 ppc: 6,144 particles per 4x4x4 tile of a species, above the 5,120
    at which a tile takes the eight-lanes-per-subcell deposit."""
    pr = obstacle_inflow(ppc=96, n=16, seed=42)
    steps = 10
    fo, no = oracle_after(oracle_mod, "obstacle16_ppc96", pr, steps)
    fg, ng = run_ctx(gpu_mod, pr, gpu_mod.RH_FP32, (4, 4, 4), steps)
    for name in ("cellRhoQi", "cellJi", "faceB"):
        r = rel(fg[name], fo[name])
        print("crowded: %s rel L2 = %.3e" % (name, r), flush=True)
        assert r < 1e-5, (name, r)
    for s in range(2):
        assert abs(ng[s] - no[s]) <= 2, (s, ng[s], no[s])


def test_fp64_64cell_tiles_vs_oracle(gpu_mod, oracle_mod):
    pr = obstacle_inflow(ppc=8, n=16, seed=43)
    steps = 8
    fo, no = oracle_after(oracle_mod, "obstacle16_ppc8", pr, steps)
    fg, ng = run_ctx(gpu_mod, pr, gpu_mod.RH_FP64, (4, 4, 4), steps)
    for name in fo:
        r = rel(fg[name], fo[name])
        print("fp64 4x4x4: %s rel L2 = %.3e" % (name, r), flush=True)
        assert r < 1e-9, (name, r)
    assert ng == no


def exosphere_small(seed=44):
    """Reduced uneven-load box: 64x48x48 cells, H+ 24 ppc filled at timestep
    1 and entering through the +x wall, O+ photoions of a corona around an
    absorbing sphere of 6 cells (scale height 2.5 cells)."""
    pr = P.config_exosphere(nx=64, ny=48, nz=48, ppc=24, R_cells=6.0, H_cells=2.5, seed=seed, sub=2)
    pr.species[0].capacity = int(1.4 * 24 * pr.grid.nc)
    pr.species[1].capacity = 600000
    return pr


PREFILL = 150


def exo_prepare(pr):
    def prep(sim):
        fill, own = pr.meta["fill"][0], pr.injectors
        for s, inj, steps in ((0, fill, (1,)), (1, own[1], range(1, PREFILL + 1))):
            sim.set_injector(s, inj)
            for t in steps:
                sim.set_timestep(t)
                sim.inject(s)
            sim.set_injector(s, own[s])
        sim.set_timestep(0)
    return prep


@pytest.mark.parametrize("tile", [(4, 4, 4), (2, 4, 8), (2, 8, 4), (4, 2, 8)])
def test_uneven_load_box_vs_oracle(gpu_mod, oracle_mod, tile):
    pr = exosphere_small()
    steps = 10
    fo, no = oracle_after(oracle_mod, "exosphere64x48x48", pr, steps, exo_prepare(pr))
    assert no[1] > 50000  # the O+ cloud is there
    f64, n64 = run_ctx(gpu_mod, pr, gpu_mod.RH_FP64, tile, steps, exo_prepare(pr))
    fg, ng = run_ctx(gpu_mod, pr, gpu_mod.RH_FP32, tile, steps, exo_prepare(pr))
    print("exosphere tile %s: particles oracle %s fp64 %s fp32 %s" % (tile, no, n64, ng), flush=True)
    r64 = {name: rel(f64[name], fo[name]) for name in fo}
    r32 = {name: rel(fg[name], fo[name]) for name in fo}
    # two particles whose deletion falls in another step, plus any count
    # difference: per particle q w / dV (rho), times the largest speed (J)
    dV = pr.grid.dx ** 3
    vmax = np.sqrt(pr.params.max_ion_speed2)  # the ion speed clamp (and a bound for injected speeds here)
    qw = max(abs(sp.q * sp.w) for sp in pr.species)
    drho = (2 + sum(abs(ng[s] - no[s]) for s in range(2))) * qw / dV
    allow = {"cellRhoQi": 1e-5 + drho / np.linalg.norm(fo["cellRhoQi"]),
             "cellJi": 1e-5 + drho * vmax / np.linalg.norm(fo["cellJi"]),
             "faceB": 1e-5, "nodeE": None}
    for name in fo:
        print("exosphere tile %s: %s rel L2 fp64 %.3e fp32 %.3e (allowed %s)"
              % (tile, name, r64[name], r32[name], allow[name]), flush=True)
    for name in fo:
        assert r64[name] < 1e-9, ("fp64", name, r64[name])
    assert n64 == no
    for name in ("cellRhoQi", "cellJi", "faceB"):
        assert r32[name] < allow[name], ("fp32", name, r32[name], allow[name])
    for s in range(2):
        assert abs(ng[s] - no[s]) <= 2, (s, ng[s], no[s])


def test_species_streams_bitwise(gpu_mod):
    """rh_set_species_streams: each species' push, migration and deposit on
    a stream of its own changes nothing — the reduced uneven-load box (crowded
    O+ tiles beside a filled H+ box) gives bit-identical grid arrays and
    particle stores after 6 steps with and without it, including the steps'
    injection and boundary conditions on the context's stream."""
    pr = exosphere_small(seed=45)
    out = []
    for on in (False, True):
        print("case species_streams=%s" % on, flush=True)
        ctx = gpu_mod.Context(pr, precision=gpu_mod.RH_FP32, tile=(4, 4, 4))
        exo_prepare(pr)(ctx)
        ctx.set_species_streams(on)
        ctx.step(3)
        ctx.step(3)
        fields = [ctx.field(n) for n in ("cellRhoQi", "cellJi", "faceB", "nodeE", "cellUe")]
        parts = [ctx.particles(s) for s in range(2)]
        out.append((fields, parts))
        ctx.close()
    (fa, pa), (fb, pb) = out
    for a, b in zip(fa, fb):
        assert np.array_equal(a, b)
    for s in range(2):
        assert len(pa[s][0]) > 1000
        for a, b in zip(pa[s], pb[s]):
            assert np.array_equal(a, b), s
