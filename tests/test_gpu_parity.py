"""Parity of the HIP backend with the reference, through the C ABI on device 0.

Bars, per field: the relative error max|a-b| / max(|a|,|b|) taken over the
field (normalised by the field's magnitude), plus the element-wise ratio for
the state fields P_w, S_w, S_n, which never approach zero:
  * the Corey closure has no transcendental calls, so GPU and oracle must
    agree bit for bit after any number of steps;
  * the reference closures call pow/exp, which the GPU takes from the ROCm
    device library instead of glibc: <= 1e-12 after one step and <= 1e-9
    after the full run;
  * the fused step and the entry-point sequence run the same device
    arithmetic and must agree bit for bit, every field.
"""
import glob
import os

import numpy as np
import pytest

from capaz_amd import capaz, workloads
from capaz_cases import bit_equal, case, max_rel, norm_rel, random_state
from golden.make_golden import ENTRIES, load
from oracle import oracle

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref*.npz")))
GIDS = [os.path.basename(p)[:-4] for p in GOLDEN]
TOL_STEP1, TOL_RUN = 1e-12, 1e-9


def entries_of(model):
    return [e for e in ENTRIES if model == "3phe" or e not in ("H_E_current_calculation", "E_calculation")]


STATE = ("P_w", "S_w", "S_n")


def check_fields(got: dict, want: dict, tol: float, label: str):
    worst = {}
    for f, w in want.items():
        r = norm_rel(got[f], w)
        worst[f] = r
        assert r <= tol, f"{label}: field {f} relative error {r:.3e} > {tol:g}"
        if f in STATE:
            e = max_rel(got[f], w)
            assert e <= tol, f"{label}: field {f} element-wise relative error {e:.3e} > {tol:g}"
    return worst


@pytest.mark.parametrize("path", GOLDEN, ids=GIDS)
def test_golden_entry_points(path):
    meta, z = load(path)
    keys = meta["keys"]
    names = oracle.model_fields(keys)
    d = capaz.Domain(keys, {f: z[f"in/{f}"] for f in names})
    for e in entries_of(meta["model"]):
        d.call(e, keys["DT"])
        check_fields(d.fields(), {f: z[f"ent/{e}/{f}"] for f in names}, TOL_STEP1, e)
    n = meta["nsteps"]
    for s in range(2, n + 1):
        d.time_step_function(s * keys["DT"])
    check_fields(d.fields(), {f: z[f"out{n}/{f}"] for f in names}, TOL_RUN, f"after {n} steps")


def elementwise_outside_cancellation(got: np.ndarray, want: np.ndarray, floor: float = 1e-3) -> float:
    """Element-wise relative error max |a-b| / |b| over the cells whose
    reference value is at least `floor` times the field's largest magnitude
    (the cells a field-normalised bar could hide; values near zero come from
    cancellation and carry no relative accuracy)."""
    big = np.abs(want) >= floor * np.abs(want).max() if want.size and np.abs(want).max() > 0 else np.zeros(0, bool)
    if not big.any():
        return 0.0
    return max_rel(got[big], want[big])


@pytest.mark.parametrize("path", GOLDEN, ids=GIDS)
def test_golden_elementwise_outside_cancellation(path):
    """The per-field bars are normalised by the field's magnitude for the
    fields that cross zero (velocities); this reports the element-wise error
    that such a bar allows, after each entry point of the first step: <= 1e-12
    on every field at every cell holding at least 1e-3 of the field's largest
    magnitude.  The worst case per field is printed."""
    meta, z = load(path)
    keys = meta["keys"]
    names = oracle.model_fields(keys)
    d = capaz.Domain(keys, {f: z[f"in/{f}"] for f in names})
    worst = {}
    for e in entries_of(meta["model"]):
        d.call(e, keys["DT"])
        got = d.fields()
        for f in names:
            r = elementwise_outside_cancellation(got[f], z[f"ent/{e}/{f}"])
            worst[f] = max(worst.get(f, 0.0), r)
    print(f"{os.path.basename(path)} element-wise worst (cells >= 1e-3 of max): " +
          ", ".join(f"{f} {r:.1e}" for f, r in sorted(worst.items())), flush=True)
    bad = {f: r for f, r in worst.items() if r > TOL_STEP1}
    assert not bad, bad


@pytest.mark.parametrize("path", GOLDEN, ids=GIDS)
def test_golden_fused(path):
    meta, z = load(path)
    keys = meta["keys"]
    names = oracle.model_fields(keys)
    d = capaz.Domain(keys, {f: z[f"in/{f}"] for f in names})
    d.run(1, first_step=1)
    check_fields(d.fields(), {f: z[f"out1/{f}"] for f in names}, TOL_STEP1, "fused step 1")
    n = meta["nsteps"]
    d.run(n - 1, first_step=2)
    check_fields(d.fields(), {f: z[f"out{n}/{f}"] for f in names}, TOL_RUN, f"fused after {n} steps")


RANDOM_CASES = [
    ("3ph", (40, 36, 32), dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=1, DT=1e-4), True),
    ("2ph", (40, 36, 32), dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=3, DT=1e-4), True),
    ("3ph", (33, 30, 28), dict(DT=1e-4), False),
    ("2ph", (33, 30, 28), dict(DT=1e-4, SOURCE=4), False),
    ("3phe", (20, 18, 16), dict(DT=1e-4), False),
    ("3phe", (1, 9, 1), dict(DT=1e-4), False),
    ("3phe", (9, 8, 1), dict(DT=1e-4, SOURCE=2), False),
    ("3phe", (1, 12, 9), dict(DT=1e-4), False),
    ("3phe", (3, 3, 3), dict(DT=1e-4), False),
    ("3phe", (20, 18, 16), dict(DT=1e-4, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05), False),
    ("3ph", (3, 3, 3), dict(CLOSURE="corey"), True),
    ("3ph", (1, 15, 1), dict(), False),
    ("3ph", (7, 3, 1), dict(SOURCE=2, CLOSURE="corey"), True),
    ("2ph", (3, 5, 1), dict(SOURCE=1), False),
]


@pytest.mark.parametrize("model,dims,over,exact", RANDOM_CASES,
                         ids=[f"{m}-{'x'.join(map(str, d))}-{'corey' if o.get('CLOSURE') else 'ref'}"
                              for m, d, o, _ in RANDOM_CASES])
def test_random_state_vs_oracle(model, dims, over, exact):
    keys = case(model, *dims, **over)
    st = random_state(keys, seed=sum(dims) + 7 * len(model))
    od = oracle.OracleDomain(keys, st)
    fused = capaz.Domain(keys, st)
    seq = capaz.Domain(keys, st)
    for s, tol in ((1, TOL_STEP1), (10, TOL_RUN)):
        first = 1 if s == 1 else 2
        count = 1 if s == 1 else 9
        od.run(count, first_step=first)
        fused.run(count, first_step=first)
        for t in range(first, first + count):
            seq.time_step_function(t * keys["DT"])
        want = od.fields()
        gf, gs = fused.fields(), seq.fields()
        for f in od.names:
            assert bit_equal(gf[f], gs[f]), f"fused != sequence on {f} after {s} steps"
            if exact:
                assert bit_equal(gf[f], want[f]), f"{f} after {s} steps: max rel {max_rel(gf[f], want[f]):.3e}"
        if not exact:
            check_fields(gf, want, tol, f"after {s} steps")


LAUNCH_SHAPES = [dict(), dict(tile_y=8), dict(tile_y=4), dict(fused_kernel=2), dict(fused_kernel=2, tile_y=4),
                 dict(fused_kernel=0), dict(kchunk=3)]
SHAPE_CASES = [("2ph", dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=3, DT=1e-4)),
               ("3ph", dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=1, DT=1e-4)),
               ("2ph", dict(DT=1e-4, SOURCE=4)), ("3ph", dict(DT=1e-4))]


@pytest.mark.parametrize("model,over", SHAPE_CASES,
                         ids=[f"{m}-{'corey' if o.get('CLOSURE') else 'ref'}" for m, o in SHAPE_CASES])
def test_every_launch_shape_gives_the_same_bits(model, over):
    """The fused step's launch shapes (tile_y 4 / 8 and the automatic one,
    which is 4 for the two-phase Corey-2 step and 8 otherwise; the carried
    or recomputed centre coefficients; one thread per cell; a fixed chunk
    of planes) give the same bits after 6 steps on a grid wider and taller
    than one tile, and, with the Corey closure, the oracle's."""
    keys = case(model, 70, 13, 11, **over)
    st = random_state(keys, seed=31)
    runs = []
    for tune in LAUNCH_SHAPES:
        d = capaz.Domain(keys, st)
        d.tune(**tune)
        d.run(6)
        runs.append((tune, d.fields()))
        d.close()
    base = runs[0][1]
    for tune, got in runs[1:]:
        for f in base:
            assert bit_equal(got[f], base[f]), f"{tune}: {f} differs from the automatic shape"
    if over.get("CLOSURE") == "corey":
        od = oracle.OracleDomain(keys, st)
        od.run(6)
        want = od.fields()
        for f in od.names:
            assert bit_equal(base[f], want[f]), f"{f}: max rel {max_rel(base[f], want[f]):.3e} against the oracle"


def test_headline_physics_bitwise_at_96():
    """The benchmark physics (three-phase, Corey-2, lognormal K, top
    injection) at 96^3: one and three fused steps equal the oracle bit for bit."""
    keys, st = workloads.three_phase_injection(96, seed=7)
    od = oracle.OracleDomain(keys, st)
    d = capaz.Domain(keys, st)
    od.run(1)
    d.run(1)
    for f in ("P_w", "S_w", "S_n", "P_n", "P_g", "Xi_w", "roS_g", "S_g"):
        assert bit_equal(d.field(f), od.field(f)), (f, max_rel(d.field(f), od.field(f)))
    od.run(2, first_step=2)
    d.run(2, first_step=2)
    for f in od.names:
        assert bit_equal(d.field(f), od.field(f)), (f, max_rel(d.field(f), od.field(f)))


def pivot_search_state(n: int = 40):
    """The benchmark physics (three-phase, beta_w 4.5e-10, Corey-2) with
    S_w in [0.005, 0.02] everywhere: the 3x3 Newton matrix's first pivot
    ro0_w * beta_w * S_w is below 1e-8 in every interior cell, so the inverse
    takes gauss.cpp:39-58's signed row search (both later rows qualify, so
    row 0 is swapped twice)."""
    keys, st = workloads.three_phase_injection(n, seed=11)
    rng = np.random.default_rng(12)
    st = dict(st)
    st["S_w"] = rng.uniform(0.005, 0.02, st["S_w"].shape)
    st["S_n"] = rng.uniform(0.6, 0.8, st["S_n"].shape)
    inner = (slice(1, -1),) * 3
    dF0 = keys["RO_W"] * keys["BETA_W"] * st["S_w"][inner]
    dF3 = keys["RO_N"] * keys["BETA_N"] * st["S_n"][inner]
    dF6 = keys["RO_G"] * (1.0 - st["S_w"][inner] - st["S_n"][inner]) / keys["P_ATM"]
    assert (dF0 < 1e-8).all() and (dF3 > 1e-8).all() and (dF6 > 1e-8).all()
    return keys, st


def test_three_phase_pivot_search_path_bitwise():
    """Every interior cell's first Newton iteration takes the signed pivot
    search (the out-of-line general routine on the device): three fused
    steps and three entry-point sequences equal the oracle bit for bit,
    every field."""
    keys, st = pivot_search_state()
    od = oracle.OracleDomain(keys, st)
    fused = capaz.Domain(keys, st)
    seq = capaz.Domain(keys, st)
    od.run(3)
    fused.run(3)
    for t in range(1, 4):
        seq.time_step_function(t * keys["DT"])
    assert np.isfinite(od.field("P_w")).all()
    for f in od.names:
        w = od.field(f)
        assert bit_equal(fused.field(f), w), (f, max_rel(fused.field(f), w))
        assert bit_equal(seq.field(f), w), (f, max_rel(seq.field(f), w))
    fused.close()
    seq.close()


def test_energy_workload_vs_oracle():
    """The energy benchmark's model (reference three-phase deck with the
    energy equation) on a 20 x 18 x 14 grid: 1 step within 1e-12 and 6 steps
    within 1e-9 of the oracle, every field."""
    keys, st = workloads.energy_3ph(20, 18, 14)
    od = oracle.OracleDomain(keys, st)
    d = capaz.Domain(keys, st)
    od.run(1)
    d.run(1)
    check_fields(d.fields(), od.fields(), TOL_STEP1, "energy step 1")
    od.run(5, first_step=2)
    d.run(5, first_step=2)
    check_fields(d.fields(), od.fields(), TOL_RUN, "energy after 6 steps")


def test_energy_pivot_search_cells_vs_oracle():
    """Water and oil incompressible and four cells with S_w + S_n = 1: their
    5x5 Newton matrices have a zero pressure pivot, so the Newton pass leaves
    them to its second kernel (the general routine with the pivot search;
    here it finds none, and the update turns H_* non-finite as the oracle's
    does).  One step: the same non-finite cells, every other value within
    1e-12 of the oracle."""
    keys, st = workloads.energy_3ph(20, 18, 14)
    keys = dict(keys, BETA_W=0.0, BETA_N=0.0)
    st = {f: np.array(v, copy=True) for f, v in st.items()}
    for c in ((3, 4, 5), (4, 5, 6), (5, 2, 8), (10, 10, 10)):
        st["S_w"][c], st["S_n"][c] = 0.625, 0.375
    od = oracle.OracleDomain(keys, st)
    d = capaz.Domain(keys, st)
    od.run(1)
    d.run(1)
    bad = 0
    for f, w in od.fields().items():
        g = d.field(f)
        fin = np.isfinite(w)
        assert np.array_equal(fin, np.isfinite(g)), f
        bad += int((~fin).sum())
        assert max_rel(g[fin], w[fin]) <= TOL_STEP1, (f, max_rel(g[fin], w[fin]))
    assert bad == 16
    d.close()


@pytest.mark.parametrize("gas_free", [False, True])
def test_energy_newton_two_passes_equal_the_general_routine(gas_free):
    """On the device: the energy model's Newton in two passes (energy_newton
    1: no pivot search, then the refused cells with it) and the general
    routine in one pass (energy_newton 0) give every field the same bits,
    through the six-pass step and through the entry-point sequence; the
    gas-free case (incompressible water and oil, S_w + S_n = 1 in four cells)
    sends cells to the second pass."""
    keys, st = workloads.energy_3ph(20, 18, 14)
    if gas_free:
        keys = dict(keys, BETA_W=0.0, BETA_N=0.0)
        st = {f: np.array(v, copy=True) for f, v in st.items()}
        for c in ((3, 4, 5), (4, 5, 6), (5, 2, 8), (10, 10, 10)):
            st["S_w"][c], st["S_n"][c] = 0.625, 0.375
    for energy_kernel in (1, 0):
        got = []
        for newton in (1, 0):
            d = capaz.Domain(keys, st)
            d.tune(energy_kernel=energy_kernel, energy_newton=newton)
            assert d.tuning()["energy_newton"] == newton
            d.run(3)
            got.append(d.fields())
            d.close()
        for f, v in got[1].items():
            assert bit_equal(got[0][f], v), (energy_kernel, f)


def test_full_size_fused_equals_sequence():
    """256^3 headline workload: two fused steps and two entry-point sequences
    give identical bits in every field."""
    keys, st = workloads.headline(256)
    a = capaz.Domain(keys, st)
    a.run(2)
    pa = {f: a.field(f) for f in ("P_w", "S_w", "S_n", "P_n", "P_g", "roS_w", "Xi_g")}
    a.close()
    b = capaz.Domain(keys, st)
    b.time_step_function(keys["DT"])
    b.time_step_function(2 * keys["DT"])
    for f, v in pa.items():
        assert bit_equal(v, b.field(f)), f
    assert np.all(np.isfinite(pa["P_w"]))


@pytest.mark.parametrize("closure", ["reference", "corey"])
@pytest.mark.parametrize("dims", [(20, 18, 16), (9, 8, 1), (1, 12, 9), (300, 5, 4), (130, 3, 3)])
def test_energy_passes_equal_the_sequence(dims, closure):
    keys = case("3phe", *dims, DT=1e-4, SOURCE=1, CLOSURE=closure)
    st = random_state(keys, seed=sum(dims) + 11)
    seq = capaz.Domain(keys, st)
    seq.tune(energy_kernel=0)
    seq.run(3)
    d = capaz.Domain(keys, st)
    d.run(3)
    for f in seq.names:
        assert bit_equal(d.field(f), seq.field(f)), f
    seq.close()
    d.close()


def test_energy_full_size_six_passes_equal_the_sequence():
    """The energy workload at 256 x 256 x 128: two steps of the six-pass
    energy step and two entry-point sequences give identical bits in every
    field."""
    keys, st = workloads.energy_3ph(256, 256, 128)
    a = capaz.Domain(keys, st)
    assert a.tuning()["energy_kernel"] == 1
    a.run(2)
    pa = a.fields()
    a.close()
    b = capaz.Domain(keys, st)
    b.time_step_function(keys["DT"])
    b.time_step_function(2 * keys["DT"])
    for f, v in pa.items():
        assert bit_equal(v, b.field(f)), f
    b.tune(energy_kernel=0)
    b.run(1, first_step=3)
    a = capaz.Domain(keys, pa)
    a.run(1, first_step=3)
    for f in ("P_w", "S_w", "S_n", "S_g", "T", "E", "H_w", "ux_g", "roS_g", "E_new"):
        assert bit_equal(a.field(f), b.field(f)), f
    assert np.all(np.isfinite(pa["T"]))
    a.close()
    b.close()


def test_full_size_mass_conservation_without_gravity():
    """At 256^3 with g = 0 and boundary pressures equal to their inward
    neighbour's, the flux update conserves sum(m * roS) per phase to the
    rounding of the stored roS (the divergence telescopes)."""
    keys, st = workloads.headline(256)
    keys = dict(keys, G_CONST=0.0, DT=1e-4)
    rng = np.random.default_rng(5)
    st = dict(st)
    st["P_w"] = st["P_w"] + rng.uniform(-200, 200, st["P_w"].shape)
    st["S_w"] = rng.uniform(0.2, 0.5, st["P_w"].shape)
    st["S_n"] = rng.uniform(0.1, 0.4, st["P_w"].shape)
    d = capaz.Domain(keys, st)
    for e in ("boundary_conditions", "S_calculation", "ro_P_Xi_calculation", "u_calculation", "roS_calculation"):
        d.call(e)
    m = d.field("m")[1:-1, 1:-1, 1:-1].astype(np.longdouble)
    for ph in ("w", "n", "g"):
        new = d.field("roS_" + ph)[1:-1, 1:-1, 1:-1].astype(np.longdouble)
        old = d.field("roS_" + ph + "_old")[1:-1, 1:-1, 1:-1].astype(np.longdouble)
        # extended-precision sums: the stored roS rounding is the only loss
        total = float((m * new).sum() - (m * old).sum())
        moved = float(np.abs(m * (new - old)).sum())
        noise = float(4 * new.size * np.abs(m * new).max() * 2.0 ** -53)
        assert moved > 1e3 * noise, ph
        assert abs(total) <= noise, (ph, total, noise)


def test_zero_porosity_cell_uses_reference_rule():
    """An interior cell with m == 0 keeps the reference's skip rule
    (cpu.cpp:625-626): capaz_step falls back to the entry-point sequence and
    still matches the oracle."""
    keys = case("3ph", 9, 8, 7, CLOSURE="corey", DT=1e-4)
    st = random_state(keys, 21)
    st["m"][3, 4, 4] = 0.0
    od = oracle.OracleDomain(keys, st)
    d = capaz.Domain(keys, st)
    od.run(3)
    d.run(3)
    for f in ("P_w", "S_w", "S_n"):
        assert bit_equal(d.field(f), od.field(f)), f


def test_face_buffers_match_reference_layout():
    """capaz_load_face / capaz_save_face use the planes and buffer layout of
    load_/save_exchange_data_part_* (cpu.cpp:769-888)."""
    import ctypes
    from capaz_amd import hip
    keys = case("3ph", 6, 5, 4)
    st = random_state(keys, 3)
    d = capaz.Domain(keys, st)
    P = st["P_w"]  # (k, j, i)
    L = capaz.load()
    fid = capaz.FIELD_ID["P_w"]
    want = {(0, 0): P[:, :, 1], (0, 1): P[:, :, -2], (1, 0): P[:, 1, :], (1, 1): P[:, -2, :],
            (2, 0): P[1, :, :], (2, 1): P[-2, :, :]}
    buf = hip.DeviceBuffer(8 * 64)
    for (axis, side), plane in want.items():
        assert L.capaz_load_face(d._ctx, fid, axis, side, ctypes.c_void_p(buf.address)) == 0
        d.sync()
        # buffer index a + na*b with (a, b) = (j, k), (i, k), (i, j)
        np.testing.assert_array_equal(buf.download(plane.size), plane.reshape(-1))
    buf.upload(np.arange(6 * 5, dtype=np.float64))  # z face: Nx*Ny
    assert L.capaz_save_face(d._ctx, fid, 2, 1, ctypes.c_void_p(buf.address)) == 0
    d.sync()
    np.testing.assert_array_equal(d.field("P_w")[-1], np.arange(30.0).reshape(5, 6))
    buf.close()


@pytest.mark.parametrize("dims,offset", [((9, 8, 7), 0), ((9, 7, 6), 0), ((9, 8, 7), 8)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_batched_faces_equal_the_per_face_calls(axis, dims, offset):
    """capaz_load_faces packs every field's two exchange planes in one launch,
    [field][side][plane], as the per-field, per-side capaz_load_face calls
    lay them out; capaz_save_faces writes the first and last planes as
    capaz_save_face does.  After fused steps (ping-pong state, stale
    intermediates) both see the current state, the planes of the fields as
    read after the step.  z planes of an even size in a
    16-byte aligned buffer take the two-doubles-per-lane path; an odd plane
    (9 x 7) and a buffer 8 bytes off alignment take the one-double path."""
    import ctypes
    from capaz_amd import hip
    nx, ny, nz = dims
    keys = case("3ph", nx, ny, nz, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05)
    st = random_state(keys, 4)
    d = capaz.Domain(keys, st)
    d.run(3)
    L = capaz.load()
    names = ("P_w", "S_w", "S_n", "P_n", "P_g")
    ids = (ctypes.c_int * 5)(*[capaz.FIELD_ID[f] for f in names])
    plane = {0: ny * nz, 1: nx * nz, 2: nx * ny}[axis]
    big, one = hip.DeviceBuffer(8 * plane * 10 + offset), hip.DeviceBuffer(8 * plane)
    at = ctypes.c_void_p(big.address + offset)

    def download_big():
        raw = big.download(plane * 10 + offset // 8)
        return raw[offset // 8:].reshape(5, 2, plane)

    assert L.capaz_load_faces(d._ctx, ids, 5, axis, at) == 0
    d.sync()
    got = download_big()
    for q, f in enumerate(names):
        for s in (0, 1):
            assert L.capaz_load_face(d._ctx, ids[q], axis, s, ctypes.c_void_p(one.address)) == 0
            d.sync()
            np.testing.assert_array_equal(got[q, s], one.download(plane))
    # the loads saw the state as the reference holds it after the step (P_n,
    # P_g interior: the previous P_w), which reading the fields materializes
    n_axis = {0: nx, 1: ny, 2: nz}[axis]
    for q, f in enumerate(names):
        full = d.field(f)
        for s, pos in ((0, 1), (1, n_axis - 2)):
            want = full[:, :, pos] if axis == 0 else (full[:, pos, :] if axis == 1 else full[pos, :, :])
            np.testing.assert_array_equal(got[q, s], want.reshape(-1), err_msg=f"{f} side {s}")
    # unpack a known buffer and compare with the per-face writes on a twin
    e = capaz.Domain(keys, st)
    e.run(3)
    vals = np.arange(plane * 10, dtype=np.float64).reshape(5, 2, plane) + 0.25
    big.upload(np.concatenate([np.zeros(offset // 8), vals.ravel()]))
    assert L.capaz_save_faces(d._ctx, ids, 5, axis, at) == 0
    for q in range(5):
        for s in (0, 1):
            one.upload(vals[q, s])
            assert L.capaz_save_face(e._ctx, ids[q], axis, s, ctypes.c_void_p(one.address)) == 0
            e.sync()
    for f in names:
        assert bit_equal(d.field(f), e.field(f)), f
    big.close()
    one.close()
    d.close()
    e.close()


PACK_CASES = {
    "3ph-corey2": ("3ph", (9, 8, 7), dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05), {}),
    "3ph-corey2-nx1": ("3ph", (1, 8, 7), dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05), {}),
    "3ph-corey2.5": ("3ph", (9, 7, 6), dict(CLOSURE="corey", COREY_N=2.5, S_WR_0=0.1, S_NR_0=0.05), {}),
    "3ph-reference": ("3ph", (9, 8, 7), dict(SOURCE=1), {}),
    "3ph-corey2-cellkernel": ("3ph", (9, 8, 7), dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05),
                              dict(fused_kernel=0)),
    "3ph-corey2-nocarry": ("3ph", (9, 8, 7), dict(CLOSURE="corey"), dict(fused_kernel=2, tile_y=4)),
    "2ph-reference": ("2ph", (9, 8, 7), {}, {}),
    "2ph-corey": ("2ph", (9, 7, 6), dict(CLOSURE="corey", COREY_N=2.0), {}),
}


@pytest.mark.parametrize("name", list(PACK_CASES))
def test_halo_pack_equals_load_faces_after_each_step(name):
    """capaz_set_halo_pack: every step writes planes 1 and Nz-2 of its new
    state into the buffer -- capaz_load_faces(axis 2) of the state fields
    after that step, bit for bit, except P_n and P_g of the planes' interior
    cells, which hold the new P_w.  The state itself is unchanged by the pack
    (a twin without it agrees bit for bit), and each step rewrites it.  The
    pack copied straight into the ghost planes that capaz_ghost_planes names
    gives the next step what capaz_save_faces of the same buffer gives it."""
    import ctypes
    from capaz_amd import hip
    model, (nx, ny, nz), over, tune = PACK_CASES[name]
    keys = case(model, nx, ny, nz, **over)
    st = random_state(keys, 11)
    d, e = capaz.Domain(keys, st), capaz.Domain(keys, st)
    for dom in (d, e):
        if tune:
            dom.tune(**tune)
    names = ("P_w", "S_w", "S_n", "P_n", "P_g") if model == "3ph" else ("P_w", "S_n")
    plane = nx * ny
    pack, ref = hip.DeviceBuffer(8 * plane * 2 * len(names)), hip.DeviceBuffer(8 * plane * 2 * len(names))
    d.set_halo_pack(pack.address)
    L = capaz.load()
    ids = (ctypes.c_int * len(names))(*[capaz.FIELD_ID[f] for f in names])
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    inner = ((jj > 0) & (jj < ny - 1) & (((ii > 0) & (ii < nx - 1)) | (nx < 3))).reshape(-1)
    for step in range(1, 4):
        d.run(1, first_step=step)
        e.run(1, first_step=step)
        assert L.capaz_load_faces(e._ctx, ids, len(names), 2, ctypes.c_void_p(ref.address)) == 0
        d.sync()
        e.sync()
        got = pack.download(plane * 2 * len(names)).reshape(len(names), 2, plane)
        want = ref.download(plane * 2 * len(names)).reshape(len(names), 2, plane)
        for q, f in enumerate(names):
            w = want[q].copy()
            if f in ("P_n", "P_g"):
                w[:, inner] = want[0][:, inner]
            assert bit_equal(got[q], w), f"{name} step {step}: field {f}"
        for f in names:
            assert bit_equal(d.field(f), e.field(f)), f"{name} step {step}: state {f}"
        # the exchange's receiving end: d takes the pack straight into its
        # ghost planes (capaz_ghost_planes), e through capaz_save_faces; the
        # next step must see the same ghost values in both
        ghosts = d.ghost_planes()
        assert len(ghosts) == 2 * len(names)
        for q in range(len(names)):
            for side in (0, 1):
                hip.copy(ghosts[2 * q + side], pack.address + 8 * plane * (2 * q + side), 8 * plane)
        ref.upload(got.reshape(-1))
        assert L.capaz_save_faces(e._ctx, ids, len(names), 2, ctypes.c_void_p(ref.address)) == 0
        e.sync()
    d.run(1, first_step=4)
    e.run(1, first_step=4)
    for f in names:
        assert bit_equal(d.field(f), e.field(f)), f"{name} after the exchanges: state {f}"
    got = pack.download(plane * 2 * len(names)).reshape(len(names), 2, plane)
    d.set_halo_pack(0)
    d.run(1, first_step=5)
    d.sync()
    assert bit_equal(pack.download(plane * 2 * len(names)).reshape(len(names), 2, plane),
                     got), "the pack was written after it was turned off"
    pack.close()
    ref.close()
    d.close()
    e.close()


def test_halo_pack_at_the_shard_shape():
    """The pack at the shape it is measured on: one z-slab of the 256^3
    headline over 8 GPUs (256 x 256 x 34).  After each of three steps the
    pack equals capaz_load_faces of the state (P_n, P_g interior: the new
    P_w), and the pack taken into the ghost planes through
    capaz_ghost_planes gives the next step what capaz_save_faces gives it."""
    import ctypes
    from capaz_amd import hip
    keys, st = workloads.headline(256)
    keys = dict(keys, NZ=34)
    st = {f: np.ascontiguousarray(v[:34]) for f, v in st.items()}
    d, e = capaz.Domain(keys, st), capaz.Domain(keys, st)
    del st
    names = ("P_w", "S_w", "S_n", "P_n", "P_g")
    plane = 256 * 256
    pack, ref = hip.DeviceBuffer(8 * plane * 10), hip.DeviceBuffer(8 * plane * 10)
    d.set_halo_pack(pack.address)
    L = capaz.load()
    ids = (ctypes.c_int * 5)(*[capaz.FIELD_ID[f] for f in names])
    jj, ii = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    inner = ((jj > 0) & (jj < 255) & (ii > 0) & (ii < 255)).reshape(-1)
    for step in range(1, 4):
        d.run(1, first_step=step)
        e.run(1, first_step=step)
        assert L.capaz_load_faces(e._ctx, ids, 5, 2, ctypes.c_void_p(ref.address)) == 0
        d.sync()
        e.sync()
        got = pack.download(plane * 10).reshape(5, 2, plane)
        want = ref.download(plane * 10).reshape(5, 2, plane)
        want[3:, :, inner] = want[0][None, :, inner]
        assert bit_equal(got, want), f"step {step}"
        ghosts = d.ghost_planes()
        for q in range(10):
            hip.copy(ghosts[q], pack.address + 8 * plane * q, 8 * plane)
        ref.upload(got.reshape(-1))
        assert L.capaz_save_faces(e._ctx, ids, 5, 2, ctypes.c_void_p(ref.address)) == 0
        e.sync()
    d.run(1, first_step=4)
    e.run(1, first_step=4)
    for f in names:
        assert bit_equal(d.field(f), e.field(f)), f
    pack.close()
    ref.close()
    d.close()
    e.close()


def test_halo_pack_rejections():
    """The energy models and m == 0 interior cells are rejected (the step
    takes the entry-point sequence there), as is Nz below 3, and a batch
    whose member has a pack."""
    from capaz_amd import hip
    buf = hip.DeviceBuffer(8 * 1024)
    d = capaz.Domain(case("3phe", 5, 5, 5), random_state(case("3phe", 5, 5, 5), 1))
    with pytest.raises(RuntimeError, match="energy"):
        d.set_halo_pack(buf.address)
    with pytest.raises(RuntimeError, match="energy"):
        d.ghost_planes()
    d.close()
    d = capaz.Domain(case("3ph", 5, 5, 1), random_state(case("3ph", 5, 5, 1), 1))
    with pytest.raises(RuntimeError, match="Nz"):
        d.set_halo_pack(buf.address)
    d.close()
    keys = case("3ph", 5, 5, 5)
    st = random_state(keys, 1)
    st["m"][2, 2, 2] = 0.0
    d = capaz.Domain(keys, st)
    d.set_halo_pack(buf.address)
    with pytest.raises(RuntimeError, match="m == 0"):
        d.run(1)
    d.set_halo_pack(0)
    d.run(1)
    d.close()
    # batched steps do not write a member's pack
    keys = case("3ph", 6, 5, 5, CLOSURE="corey")
    b = capaz.Batch(keys, [dict(dims=(6, 5, 5), **random_state(keys, 2))])
    b.domains[0].set_halo_pack(buf.address)
    with pytest.raises(RuntimeError, match="halo pack"):
        b.run(1)
    b.domains[0].set_halo_pack(0)
    b.run(1)
    b.close()
    buf.close()


def test_invalid_field_reports_error():
    d = capaz.Domain(case("2ph", 5, 5, 5), None)
    with pytest.raises(capaz.CapazError, match="does not exist"):
        d.field("S_g")
    with pytest.raises(capaz.CapazError, match="not in this model"):
        d.E_calculation()
