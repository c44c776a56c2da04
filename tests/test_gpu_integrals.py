"""Batched integrals on the GPU against the oracle on the same seeded inputs and nodes.

Tolerances: FP64 (plain) relative error <= 1e-12; mixed precision (FP32 integrand, FP64
nodes/denominators/accumulation) relative error <= 1e-8.  Determinism: a result does not
depend on the batch it runs in, on the launch shape's grid, or on graph replay.
"""
import math

import numpy as np
import pytest

import diffrg_amd.workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_MIXED = 1e-8


def rel_err(a, b):
    return np.abs(a - b) / np.abs(b)


# ------------------------------------------------------------------------------ threshold (C0)

def test_threshold_fp64_and_mixed(gpu, qp):
    c = W.threshold()
    x, w = qp.get_points(64), qp.get_weights(64)
    ref = O.c0_threshold(c["m2"], x, w, c["x_extent"], c["k"], 1.0, c["d"])
    d_m2 = gpu.DeviceArray.from_host(c["m2"])
    d_x, d_w = gpu.DeviceArray.from_host(x), gpu.DeviceArray.from_host(w)
    out = gpu.DeviceArray(len(c["m2"]))
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        print(f"[case] threshold 256x64 mode={mode}", flush=True)
        gpu.threshold_batch(d_m2, d_x, d_w, out, c["x_extent"], c["k"], mode, c["d"])
        got = out.to_host()
        err = rel_err(got, ref)
        print(f"  max rel err {err.max():.3e}", flush=True)
        assert err.max() <= tol


def test_threshold_near_pole_mixed(gpu, qp):
    # m2 -> -k^2: the FP64 denominator keeps the mixed path accurate
    m2 = np.array([-0.99, -0.995, -0.999, -0.9999])
    x, w = qp.get_points(64), qp.get_weights(64)
    ref = O.c0_threshold(m2, x, w, 20.0, 1.0, 1.0, 4)
    d_m2, d_x, d_w = (gpu.DeviceArray.from_host(a) for a in (m2, x, w))
    out = gpu.DeviceArray(len(m2))
    print("[case] threshold near pole mixed", flush=True)
    gpu.threshold_batch(d_m2, d_x, d_w, out, 20.0, 1.0, gpu.MIXED, 4)
    err = rel_err(out.to_host(), ref)
    print(f"  max rel err {err.max():.3e}", flush=True)
    assert err.max() <= TOL_MIXED


# ------------------------------------------------------------------------------ self-energy (C1)

def selfenergy_batch(gpu, qp, c):
    n = len(c["p"])
    return gpu.SEBatch(qp, c["p"], c["nq"], c["na"], [c["Z"]] * n, *[[v] * n for v in c["coord"]],
                       x_extent=c["x_extent"], k=c["k"])


def selfenergy_oracle(qp, c):
    n = len(c["p"])
    xq = (qp.get_points(c["nq"]), qp.get_weights(c["nq"]))
    xa = (qp.get_points(c["na"]), qp.get_weights(c["na"]))
    coord = O.logcoord(c["n_tab"], *c["coord"])
    return O.c1_selfenergy(c["p"], [c["Z"]] * n, [coord] * n, [xq] * n, [xa] * n, c["x_extent"], c["k"])


def test_selfenergy_fp64_and_mixed(gpu, qp):
    c = W.selfenergy()
    ref = selfenergy_oracle(qp, c)
    b = selfenergy_batch(gpu, qp, c)
    assert b.evals == 64 * 64 * 32 * 32
    out = gpu.DeviceArray(len(c["p"]))
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        print(f"[case] selfenergy 64x(64,32,32) mode={mode} shape={b.shape()}", flush=True)
        b.run(out, mode)
        err = rel_err(out.to_host(), ref)
        print(f"  max rel err {err.max():.3e}", flush=True)
        assert err.max() <= tol


@pytest.mark.parametrize("block,tile,grid", [(512, 16384, 0), (256, 4096, 0), (256, 8192, 512)])
def test_shared_layout_equals_tile_records(gpu, qp, block, tile, grid):
    # Integrals that share nodes and table (C1) take their tiles from the kernel arguments and the
    # block index; one more integral with another radial order makes the batch read tile records
    # instead.  Each of the 64 shared integrals gives the same bits either way.
    c = W.selfenergy()
    n = len(c["p"])
    b = selfenergy_batch(gpu, qp, c)
    b.set_shape(block, tile, grid)
    nq = np.full(n + 1, c["nq"])
    nq[-1] = c["nq"] // 2
    rec = gpu.SEBatch(qp, list(c["p"]) + [1.0], nq, [c["na"]] * (n + 1), [c["Z"]] * (n + 1),
                      *[[v] * (n + 1) for v in c["coord"]], x_extent=c["x_extent"], k=c["k"])
    rec.set_shape(block, tile, grid)
    out, out_rec = gpu.DeviceArray(n), gpu.DeviceArray(n + 1)
    ref = selfenergy_oracle(qp, c)
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        print(f"[case] shared layout vs records block={block} tile={tile} grid={grid} mode={mode}", flush=True)
        b.run(out, mode)
        rec.run(out_rec, mode)
        shared = out.to_host()
        assert np.array_equal(shared, out_rec.to_host()[:n])
        assert rel_err(shared, ref).max() <= tol


def test_selfenergy_partials_never_stale(gpu, qp):
    # Cross-block partials are published with device-scope atomics and no cache-wide fence.
    # Alternating the modes makes every published partial differ from the one a slot held
    # before (by ~1e-10), so reading a stale partial would break the FP64 result's bitwise
    # reproducibility and its 1e-12 parity.
    c = W.selfenergy(n_p=32)
    ref = selfenergy_oracle(qp, c)
    b = selfenergy_batch(gpu, qp, c)
    out = gpu.DeviceArray(len(c["p"]))
    for block, tile, grid in ((256, 1024, 0), (512, 4096, 0), (128, 2048, 256)):
        b.set_shape(block, tile, grid)
        print(f"[case] alternating modes block={block} tile={tile} grid={grid} tiles={b.shape()['n_tiles']}",
              flush=True)
        first = None
        for rep in range(8):
            b.run(out, gpu.MIXED)
            mixed = out.to_host()
            b.run(out, gpu.FP64)
            fp64 = out.to_host()
            if first is None:
                first = fp64
            assert np.array_equal(fp64, first)
            assert rel_err(fp64, ref).max() <= TOL_FP64
            assert rel_err(mixed, ref).max() <= TOL_MIXED


@pytest.mark.parametrize("block,tile,grid", [(64, 1024, 0), (128, 4096, 0), (512, 32768, 0), (256, 8192, 512),
                                             (1024, 16384, 0)])
def test_selfenergy_shapes(gpu, qp, block, tile, grid):
    c = W.selfenergy(n_p=16)
    ref = selfenergy_oracle(qp, c)
    b = selfenergy_batch(gpu, qp, c)
    b.set_shape(block, tile, grid)
    out = gpu.DeviceArray(len(c["p"]))
    print(f"[case] selfenergy shape block={block} tile={tile} grid={grid}", flush=True)
    b.run(out, gpu.FP64)
    assert rel_err(out.to_host(), ref).max() <= TOL_FP64


# ------------------------------------------------------------------------------ finite T (C2)

def finite_T_device(gpu, c, qp):
    x, w = qp.get_points(c["nx"]), qp.get_weights(c["nx"])
    temps = sorted(c["rules"])
    mx = np.concatenate([c["rules"][t][0] for t in temps])
    mw = np.concatenate([c["rules"][t][1] for t in temps])
    starts = np.concatenate([[0], np.cumsum([len(c["rules"][t][0]) for t in temps])])
    off = {t: int(starts[i]) for i, t in enumerate(temps)}
    m_off = np.array([off[t] for t in c["T"]], np.int32)
    m_n = np.array([len(c["rules"][t][0]) for t in c["T"]], np.int32)
    D = gpu.DeviceArray.from_host
    dev = {"T": D(c["T"]), "p0": D(c["p0"]), "pabs": D(c["pabs"]), "m1": D(c["m1"]), "m2": D(c["m2"]),
           "mx": D(mx), "mw": D(mw), "m_off": D(m_off), "m_n": D(m_n), "x": D(x), "w": D(w)}
    return dev, x, w


def test_finite_T_fp64_and_mixed(gpu, qp):
    c = W.finite_T()
    dev, x, w = finite_T_device(gpu, c, qp)
    mats = [c["rules"][t] for t in c["T"]]
    ref = O.c2_finiteT(c["T"], c["p0"], c["pabs"], c["m1"], c["m2"], mats, x, w, c["q_max"], c["k"], 1.0, c["d"],
                       c["radial_mode"])
    out = gpu.DeviceArray(len(c["T"]))
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        print(f"[case] finite-T 512x48x129 mode={mode}", flush=True)
        gpu.ft_batch(dev, out, c["q_max"], c["k"], mode, c["d"], c["radial_mode"])
        got = out.to_host()
        err = rel_err(got, ref)
        for t in W.TEMPS:
            sel = c["T"] == t
            print(f"  T={t}: max rel err {err[sel].max():.3e}", flush=True)
        assert err.max() <= tol


@pytest.mark.parametrize("block", [64, 128, 256, 512])
def test_finite_T_block_sizes(gpu, qp, block):
    # 100 Matsubara nodes (a full and a partial 64-lane chunk) and 37 radial nodes (not a multiple
    # of the waves per integral)
    c = W.finite_T(n=64, nx=37, n_max=100)
    dev, x, w = finite_T_device(gpu, c, qp)
    mats = [c["rules"][t] for t in c["T"]]
    ref = O.c2_finiteT(c["T"], c["p0"], c["pabs"], c["m1"], c["m2"], mats, x, w, c["q_max"], c["k"], 1.0, c["d"],
                       c["radial_mode"])
    out = gpu.DeviceArray(len(c["T"]))
    gpu.ft_set_block(block)
    try:
        for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
            print(f"[case] finite-T 64x37x201 block={block} mode={mode}", flush=True)
            gpu.ft_batch(dev, out, c["q_max"], c["k"], mode, c["d"], c["radial_mode"])
            assert rel_err(out.to_host(), ref).max() <= tol
    finally:
        gpu.ft_set_block(0)
    with pytest.raises(gpu.DfrgInvalidArgument):
        gpu.ft_set_block(96)


def test_finite_T_reference_mapping(gpu, qp):
    # radial_mode 0: the reference's x = q^2/k^2 mapping with a Monien rule from the provider
    c = W.finite_T(n=32)
    T = c["T"]
    rules = {}
    for t in W.TEMPS:
        rules[t] = (qp.get_matsubara_points(t, 1.0), qp.get_matsubara_weights(t, 1.0))
    c["rules"] = rules
    c["radial_mode"] = 0
    dev, x, w = finite_T_device(gpu, c, qp)
    mats = [rules[t] for t in T]
    ref = O.c2_finiteT(T, c["p0"], c["pabs"], c["m1"], c["m2"], mats, x, w, 20.0, 1.0, 1.0, 4, 0)
    out = gpu.DeviceArray(len(T))
    print("[case] finite-T reference mapping", flush=True)
    gpu.ft_batch(dev, out, 20.0, 1.0, gpu.FP64, 4, 0)
    assert rel_err(out.to_host(), ref).max() <= TOL_FP64


# ------------------------------------------------------------------------------ uneven batch (C3)

def uneven_subset(n_sub):
    c = W.uneven()
    return {"p": c["p"][:n_sub], "nq": c["nq"][:n_sub], "na": c["na"][:n_sub], "tables": c["tables"][:n_sub],
            "coord": c["coord"], "x_extent": c["x_extent"], "k": c["k"]}


def uneven_oracle(qp, c):
    n = len(c["p"])
    xq = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in c["nq"]]
    xa = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in c["na"]]
    coords = [O.logcoord(len(t), *c["coord"]) for t in c["tables"]]
    return O.c1_selfenergy(c["p"], c["tables"], coords, xq, xa, c["x_extent"], c["k"])


def make_uneven(gpu, qp, c):
    n = len(c["p"])
    return gpu.SEBatch(qp, c["p"], c["nq"], c["na"], c["tables"], *[[v] * n for v in c["coord"]],
                       x_extent=c["x_extent"], k=c["k"])


def test_uneven_batch_parity(gpu, qp):
    c = uneven_subset(384)
    ref = uneven_oracle(qp, c)
    b = make_uneven(gpu, qp, c)
    out = gpu.DeviceArray(len(c["p"]))
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        for grid in (0, 1024):
            b.set_shape(256, 8192, grid)
            print(f"[case] uneven 384 mode={mode} grid={grid}", flush=True)
            b.run(out, mode)
            err = rel_err(out.to_host(), ref)
            print(f"  max rel err {err.max():.3e}", flush=True)
            assert err.max() <= tol


def test_uneven_full_size(gpu, qp):
    """All 4096 integrals of the uneven batch (347,385,856 evaluations), FP64 and mixed, against the
    oracle; a persistent launch and a one-tile-per-block launch of the same tiles agree bitwise."""
    c = uneven_subset(len(W.uneven()["p"]))
    print(f"[case] uneven full size ({len(c['p'])} integrals): oracle", flush=True)
    ref = uneven_oracle(qp, c)
    b = make_uneven(gpu, qp, c)
    out = gpu.DeviceArray(len(c["p"]))
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        res = []
        for grid in (0, 1024):
            b.set_shape(256, 65536, grid)
            print(f"[case] uneven full size mode={mode} grid={grid} tiles={b.shape()['n_tiles']}", flush=True)
            b.run(out, mode)
            res.append(out.to_host())
            err = rel_err(res[-1], ref)
            print(f"  max rel err {err.max():.3e}", flush=True)
            assert err.max() <= tol
        assert np.array_equal(res[0], res[1])


def test_uneven_alone_equals_in_mix(gpu, qp):
    c = uneven_subset(256)
    b = make_uneven(gpu, qp, c)
    out = gpu.DeviceArray(len(c["p"]))
    for mode in (gpu.FP64, gpu.MIXED):
        print(f"[case] alone vs mix mode={mode}", flush=True)
        b.run(out, mode)
        mix = out.to_host()
        b.run(out, mode)
        assert np.array_equal(out.to_host(), mix)  # run to run
        for i in (0, 7, 100, 255):
            one = {"p": c["p"][i:i + 1], "nq": c["nq"][i:i + 1], "na": c["na"][i:i + 1],
                   "tables": c["tables"][i:i + 1], "coord": c["coord"], "x_extent": 20.0, "k": 1.0}
            b1 = make_uneven(gpu, qp, one)
            o1 = gpu.DeviceArray(1)
            b1.run(o1, mode)
            assert o1.to_host()[0] == mix[i]  # bitwise


def test_mixed_forms_in_one_batch(gpu, qp):
    # The mixed path has a narrow integrand form (one-FMA exponent rounding, no guard on the table
    # sum), chosen per integral: tables starting at >= 1e-30 and |p - q|^2 64 log2(e)/k^2 < 2^30.
    # Integrals with start 0 or p = 1e5 take the general form; both meet the mixed tolerance, and
    # each integral's result is the same alone as in the batch.
    c = uneven_subset(48)
    n = len(c["p"])
    start = np.full(n, c["coord"][0])
    start[1::3] = 0.0
    p = np.array(c["p"], dtype=float)
    p[2::6] = 1e5
    stop, bias = c["coord"][1], c["coord"][2]
    xq = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in c["nq"]]
    xa = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in c["na"]]
    coords = [O.logcoord(len(t), s0, stop, bias) for t, s0 in zip(c["tables"], start)]
    ref = O.c1_selfenergy(p, c["tables"], coords, xq, xa, c["x_extent"], c["k"])

    def batch(idx):
        return gpu.SEBatch(qp, p[idx], c["nq"][idx], c["na"][idx], [c["tables"][i] for i in idx], start[idx],
                           [stop] * len(idx), [bias] * len(idx), x_extent=c["x_extent"], k=c["k"])

    b = batch(np.arange(n))
    v = b.variants(gpu.MIXED)
    narrow = (start >= 1e-30) & ((np.abs(p) + c["k"] * np.sqrt(c["x_extent"])) ** 2 * 64 / np.log(2)
                                 / c["k"] ** 2 < 2.0 ** 30)
    assert np.array_equal(v, np.where(narrow, 2, 1)) and 0 < narrow.sum() < n
    assert np.array_equal(b.variants(gpu.FP64), np.zeros(n))
    out = gpu.DeviceArray(n)
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        print(f"[case] two mixed forms, 48 integrals, mode={mode}", flush=True)
        b.run(out, mode)
        mix = out.to_host()
        err = rel_err(mix, ref)
        print(f"  max rel err narrow {err[narrow].max():.3e} general {err[~narrow].max():.3e}", flush=True)
        assert err.max() <= tol
        for i in (0, 1, 2):  # narrow, start 0, large p
            b1 = batch(np.array([i]))
            o1 = gpu.DeviceArray(1)
            b1.run(o1, mode)
            assert o1.to_host()[0] == mix[i]  # bitwise


@pytest.mark.parametrize("block,grid", [(64, 0), (256, 0), (1024, 0), (256, 64)])
def test_selfenergy_order_extremes(gpu, qp, block, grid):
    # The batch's limits: radial orders 1..1024 (odd ones too, and more than one tile's 128 radial
    # nodes), angular orders 2..64, tables of 2..256 points; p = 0 (|p - q| = q) and p beyond the
    # table's last point
    rng = np.random.default_rng(77)
    cases = [(1, 2, 2, 0.0), (1024, 2, 256, 5.0), (3, 64, 2, 0.3), (129, 64, 256, 60.0), (7, 16, 37, 1e-3),
             (300, 8, 128, 12.0)]
    nq = np.array([c[0] for c in cases])
    na = np.array([c[1] for c in cases])
    p = np.array([c[3] for c in cases])
    start, stop, bias = 1e-3, 40.0, 6.0
    tables = [1.0 + 0.5 * np.tanh(rng.uniform(-1, 1) * np.sin(np.linspace(0, 3, c[2]))) for c in cases]
    xq = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in nq]
    xa = [(qp.get_points(int(v)), qp.get_weights(int(v))) for v in na]
    coords = [O.logcoord(len(t), start, stop, bias) for t in tables]
    ref = O.c1_selfenergy(p, tables, coords, xq, xa, 20.0, 1.0)
    b = gpu.SEBatch(qp, p, nq, na, tables, start, stop, bias, x_extent=20.0, k=1.0)
    b.set_shape(block, 8192, grid)
    out = gpu.DeviceArray(len(p))
    for mode, tol in ((gpu.FP64, TOL_FP64), (gpu.MIXED, TOL_MIXED)):
        print(f"[case] order extremes block={block} grid={grid} mode={mode} tiles={b.shape()['n_tiles']}",
              flush=True)
        b.run(out, mode)
        err = rel_err(out.to_host(), ref)
        print(f"  max rel err {err.max():.3e}", flush=True)
        assert err.max() <= tol


def test_uneven_graph_replay(gpu, qp):
    c = uneven_subset(128)
    b = make_uneven(gpu, qp, c)
    out = gpu.DeviceArray(len(c["p"]))
    for grid in (0, 512):
        b.set_shape(256, 4096, grid)
        b.run(out, gpu.MIXED)
        eager = out.to_host()
        for rep in range(3):
            print(f"[case] graph replay grid={grid} rep={rep}", flush=True)
            out2 = gpu.DeviceArray(len(c["p"]))
            b.graph_launch(out, gpu.MIXED, launches=2)
            gpu.stream_sync()
            assert np.array_equal(out.to_host(), eager)
            b.run(out2, gpu.MIXED)  # eager work between replays
            assert np.array_equal(out2.to_host(), eager)


def test_uneven_autotune_keeps_results(gpu, qp, tmp_path):
    c = uneven_subset(64)
    ref = uneven_oracle(qp, c)
    b = make_uneven(gpu, qp, c)
    cache = str(tmp_path / "tune.txt")
    print("[case] autotune", flush=True)
    r1 = b.autotune(gpu.FP64, reps=1, cache_path=cache)
    assert r1["block"] in (128, 256, 512, 1024) and r1["ms"] > 0
    r2 = b.autotune(gpu.FP64, reps=1, cache_path=cache)  # cache hit
    assert (r2["block"], r2["tile_evals"], r2["grid"]) == (r1["block"], r1["tile_evals"], r1["grid"])
    out = gpu.DeviceArray(len(c["p"]))
    b.run(out, gpu.FP64)
    assert rel_err(out.to_host(), ref).max() <= TOL_FP64


def test_invalid_arguments(gpu, qp):
    with pytest.raises(gpu.DfrgInvalidArgument):
        gpu.SEBatch(qp, [1.0], [64], [24], [np.ones(8)], [1e-3], [40.0], [6.0])  # angular order not 2^k
    with pytest.raises(gpu.DfrgInvalidArgument):
        gpu.SEBatch(qp, [1.0], [64], [32], [np.ones(300)], [1e-3], [40.0], [6.0])  # table too large
    c = W.selfenergy(n_p=2)
    b = selfenergy_batch(gpu, qp, c)
    with pytest.raises(gpu.DfrgInvalidArgument):
        b.set_shape(100, 1024, 0)


def test_empty_and_negative_batch_sizes(gpu, qp):
    """An empty call set is a no-op (no launch, no pointer touched); a negative size or an empty
    self-energy batch is an argument error, not a launch."""
    L = gpu.lib()
    print("[case] empty threshold / finite-T batches", flush=True)
    assert L.dfrg_threshold_batch(4, None, 0, None, None, 1, 20.0, 1.0, gpu.FP64, None, None) == 0
    assert L.dfrg_ft_batch(4, 1, 0, *([None] * 11), 16, 20.0, 1.0, gpu.MIXED, None, None) == 0
    print("[case] negative sizes", flush=True)
    with pytest.raises(gpu.DfrgInvalidArgument, match="n_int"):
        gpu.check(L.dfrg_threshold_batch(4, None, -1, None, None, 1, 20.0, 1.0, gpu.FP64, None, None))
    with pytest.raises(gpu.DfrgInvalidArgument, match="n_int"):
        gpu.check(L.dfrg_ft_batch(4, 1, -3, *([None] * 11), 16, 20.0, 1.0, gpu.FP64, None, None))
    with pytest.raises(gpu.DfrgInvalidArgument, match="NULL"):
        gpu.check(L.dfrg_ft_batch(4, 1, 2, *([None] * 11), 16, 20.0, 1.0, gpu.FP64, None, None))
    with pytest.raises(gpu.DfrgInvalidArgument, match="n_int"):
        gpu.SEBatch(qp, [], [], [], [], [], [], [])


def test_graph_capture_replays_batches(gpu, qp):
    # Threshold, finite-T and self-energy launches captured into one hipGraph (two call sets of
    # each).  Between replays the same batches run eagerly in the other mode, which overwrites the
    # outputs with other values; every replay must restore the eager results of its own mode bitwise.
    ct = W.threshold()
    x, w = qp.get_points(64), qp.get_weights(64)
    d_m2 = gpu.DeviceArray.from_host(ct["m2"])
    d_x, d_w = gpu.DeviceArray.from_host(x), gpu.DeviceArray.from_host(w)
    out_t = gpu.DeviceArray(len(ct["m2"]))
    cf = W.finite_T(n=64, nx=37, n_max=100)
    dev, _, _ = finite_T_device(gpu, cf, qp)
    out_f = gpu.DeviceArray(len(cf["T"]))
    cs = W.selfenergy(n_p=16)
    b = selfenergy_batch(gpu, qp, cs)
    out_s = gpu.DeviceArray(len(cs["p"]))
    outs = (out_t, out_f, out_s)

    def enqueue(mode):
        gpu.threshold_batch(d_m2, d_x, d_w, out_t, ct["x_extent"], ct["k"], mode, ct["d"])
        gpu.ft_batch(dev, out_f, cf["q_max"], cf["k"], mode, cf["d"], cf["radial_mode"])
        b.run(out_s, mode)

    for mode, other in ((gpu.FP64, gpu.MIXED), (gpu.MIXED, gpu.FP64)):
        print(f"[case] graph of threshold + finite-T + self-energy, mode={mode}", flush=True)
        enqueue(mode)
        eager = [o.to_host() for o in outs]
        g = gpu.Graph(lambda: enqueue(mode), times=2)
        for rep in range(4):
            enqueue(other)
            assert not np.array_equal(out_s.to_host(), eager[2])
            g.launch()
            for o, e in zip(outs, eager):
                assert np.array_equal(o.to_host(), e)


def test_graph_capture_ends_when_the_enqueuing_code_raises(gpu, qp):
    # an exception in the enqueuing code (after one captured launch) ends the capture, discards the
    # graph and propagates; the stream then runs eager launches again
    c = W.threshold()
    x, w = qp.get_points(64), qp.get_weights(64)
    d_m2 = gpu.DeviceArray.from_host(c["m2"])
    d_x, d_w = gpu.DeviceArray.from_host(x), gpu.DeviceArray.from_host(w)
    out = gpu.DeviceArray(len(c["m2"]))

    def enqueue_then_raise():
        gpu.threshold_batch(d_m2, d_x, d_w, out, c["x_extent"], c["k"], gpu.FP64, c["d"])
        raise ValueError("caller error inside the capture")

    print("[case] capture ended by a Python exception", flush=True)
    with pytest.raises(ValueError):
        gpu.Graph(enqueue_then_raise)
    gpu.threshold_batch(d_m2, d_x, d_w, out, c["x_extent"], c["k"], gpu.FP64, c["d"])
    ref = O.c0_threshold(c["m2"], x, w, c["x_extent"], c["k"], 1.0, c["d"])
    assert rel_err(out.to_host(), ref).max() <= TOL_FP64
