"""Batched ensembles (capaz_batch_*) on device 0.

A batch step must give every realization the bits of capaz_step on that
realization alone, for every launch shape, with and without graph capture and
the persistent tile queue; the Corey closure also matches the oracle bit for
bit.  The autotuner must leave the state untouched, and single-domain calls
on a member must compose with batch steps.
"""
import numpy as np
import pytest

from capaz_amd import capaz, workloads
from capaz_cases import bit_equal, case, random_state
from oracle import oracle

pytestmark = pytest.mark.gpu

# uneven realizations, including flat and thin ones
DIMS = [(40, 36, 33), (17, 9, 5), (3, 3, 3), (70, 12, 1), (5, 24, 40), (33, 33, 33)]


def members(keys, dims, seed):
    out = []
    for r, (nx, ny, nz) in enumerate(dims):
        k = dict(keys, NX=nx, NY=ny, NZ=nz)
        st = random_state(k, seed + r)
        st["dims"] = (nx, ny, nz)
        out.append(st)
    return out


def alone(keys, m, nsteps):
    nx, ny, nz = m["dims"]
    d = capaz.Domain(dict(keys, NX=nx, NY=ny, NZ=nz), {f: v for f, v in m.items() if f != "dims"})
    d.run(nsteps, first_step=1)
    out = d.fields()
    d.close()
    return out


def keys_of(model):
    if model == "2ph":
        return case("2ph", 8, 8, 8, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=20, DT=1e-5, K_0=6.64e-11)
    return case("3ph", 8, 8, 8, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=1, DT=1e-4)


def assert_same(got, want, label):
    for f, w in want.items():
        assert bit_equal(got[f], w), f"{label}: field {f} differs"


@pytest.mark.parametrize("model", ["2ph", "3ph"])
def test_batch_equals_single_domain_runs(model):
    keys = keys_of(model)
    if model == "2ph":
        dims = [d for d in DIMS if d[0] >= 3]
    else:
        dims = DIMS
    ms = members(keys, dims, 100)
    b = capaz.Batch(keys, ms)
    b.run(7)
    for r, m in enumerate(ms):
        assert_same(b.domains[r].fields(), alone(keys, m, 7), f"realization {r} {m['dims']}")
    b.close()


@pytest.mark.parametrize("shape", [(64, 8), (64, 4), (32, 8)])
@pytest.mark.parametrize("persistent", [1, 0])
def test_batch_launch_shapes(shape, persistent):
    keys = keys_of("3ph")
    ms = members(keys, DIMS[:4], 200)
    want = [alone(keys, m, 4) for m in ms]
    for graph in (1, 0):
        for kchunk in (3, 16):
            b = capaz.Batch(keys, ms)
            b.tune(tile_x=shape[0], tile_y=shape[1], kchunk=kchunk, persistent=persistent, graph=graph)
            b.run(3)
            b.run(1)
            for r in range(len(ms)):
                got = {f: b.domains[r].field(f) for f in ("P_w", "S_w", "S_n", "P_n", "P_g")}
                assert_same(got, {f: want[r][f] for f in got}, f"{shape} p{persistent} g{graph} k{kchunk} r{r}")
            b.close()


@pytest.mark.parametrize("model", ["2ph", "3ph"])
def test_batch_reference_closures(model):
    # the reference's own closures (Brooks-Corey / van Genuchten, with pow):
    # still the single-domain bits, realization by realization
    if model == "2ph":
        keys = case("2ph", 8, 8, 8, SOURCE=20)
    else:
        keys = case("3ph", 8, 8, 8, SOURCE=1)
    ms = members(keys, DIMS[:4], 700)
    b = capaz.Batch(keys, ms)
    b.run(5)
    for r, m in enumerate(ms):
        assert_same(b.domains[r].fields(), alone(keys, m, 5), f"realization {r} {m['dims']}")
    b.close()


def test_batch_corey_matches_oracle():
    keys = keys_of("3ph")
    ms = members(keys, [(12, 10, 9), (7, 11, 5)], 300)
    b = capaz.Batch(keys, ms)
    b.run(5)
    for r, m in enumerate(ms):
        nx, ny, nz = m["dims"]
        o = oracle.OracleDomain(dict(keys, NX=nx, NY=ny, NZ=nz), {f: v for f, v in m.items() if f != "dims"})
        o.run(5)
        got = b.domains[r].fields()
        for f in o.fields():
            assert bit_equal(got[f], o.field(f)), f"realization {r}: field {f} differs from the oracle"
    b.close()


def test_autotune_keeps_state_and_picks_a_supported_shape():
    keys = keys_of("2ph")
    ms = members(keys, [(40, 36, 33), (17, 9, 5), (33, 33, 33)], 400)
    b = capaz.Batch(keys, ms)
    b.run(3)
    before = [d.fields() for d in b.domains]
    ms_per_step = b.autotune(4)
    assert ms_per_step > 0
    t = b.tuning()
    assert (t["tile_x"], t["tile_y"]) in [(64, 8), (64, 4), (32, 8)] and t["kchunk"] in (8, 16, 32)
    for r, d in enumerate(b.domains):
        assert_same(d.fields(), before[r], f"realization {r} after autotune")
    b.run(2)
    for r, m in enumerate(ms):
        assert_same(b.domains[r].fields(), alone(keys, m, 5), f"realization {r} after 3 + autotune + 2")
    b.close()


def test_member_calls_compose_with_batch_steps():
    keys = keys_of("3ph")
    ms = members(keys, DIMS[:3], 500)
    b = capaz.Batch(keys, ms)
    b.run(3)
    b.domains[1].run(2, first_step=4)          # one member alone
    b.domains[2].time_step_function(4 * keys["DT"])  # entry-point sequence on another
    b.run(1)
    want = [alone(keys, ms[0], 4), alone(keys, ms[1], 6), alone(keys, ms[2], 5)]
    for r in range(3):
        assert_same(b.domains[r].fields(), want[r], f"realization {r}")
    b.close()


def test_batch_rejects_zero_porosity():
    keys = keys_of("2ph")
    ms = members(keys, [(9, 9, 9)], 600)
    ms[0]["m"][4, 4, 4] = 0.0
    b = capaz.Batch(keys, ms)
    with pytest.raises(capaz.CapazError, match="m == 0"):
        b.run(1)
    b.close()


def test_ensemble_workload_runs_and_stays_bounded():
    keys, reals = workloads.ensemble(count=8, seed=1003)
    b = capaz.Batch(keys, reals)
    b.run(20)
    for r, d in enumerate(b.domains):
        Sn = d.field("S_n")
        P = d.field("P_w")
        assert np.isfinite(P).all() and (Sn >= 0).all() and (Sn <= 1).all(), f"realization {r}"
    b.close()
