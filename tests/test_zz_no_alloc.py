"""dfrg_se_batch_run_noalloc (static device scratch, no allocation) against dfrg_se_batch_run: the
same kernels and summation order, so results are bitwise equal.  Two calls in a row, the second on
a smaller batch, in FP64 and mixed mode, with one-tile-per-block and persistent launch shapes; the
size limits of the static scratch are rejected before anything is launched."""
import numpy as np
import pytest

import diffrg_amd.workloads as W

pytestmark = pytest.mark.gpu


def _batch(gpu, qp, c):
    n = len(c["p"])
    return gpu.SEBatch(qp, c["p"], c["nq"], c["na"], c["tables"], *[[v] * n for v in c["coord"]],
                       x_extent=c["x_extent"], k=c["k"])


def _uneven(n):
    c = W.uneven()
    return {"p": c["p"][:n], "nq": c["nq"][:n], "na": c["na"][:n], "tables": c["tables"][:n],
            "coord": c["coord"], "x_extent": c["x_extent"], "k": c["k"]}


def _selfenergy():
    c = W.selfenergy()
    n = len(c["p"])
    return {"p": c["p"], "nq": [c["nq"]] * n, "na": [c["na"]] * n, "tables": [c["Z"]] * n,
            "coord": c["coord"], "x_extent": c["x_extent"], "k": c["k"]}


def test_noalloc_matches_run_bitwise(gpu, qp):
    s = gpu.Stream()
    large = _batch(gpu, qp, _selfenergy())  # 64 integrals {64, 32, 32}
    small = _batch(gpu, qp, _uneven(24))    # then a smaller, uneven one
    for b in (large, small):
        out_ref = gpu.DeviceArray(b.n)
        out = gpu.DeviceArray(b.n)
        for mode in (gpu.FP64, gpu.MIXED):
            for block, tile, grid in ((512, 16384, 0), (256, 4096, 512)):
                b.set_shape(block, tile, grid)
                print(f"[case] noalloc n={b.n} mode={mode} shape={b.shape()}", flush=True)
                b.run(out_ref, mode, s.ptr)
                b.run_noalloc(out, mode, s)
                s.synchronize()
                ref, got = out_ref.to_host(), out.to_host()
                assert np.all(np.isfinite(ref))
                assert np.array_equal(got, ref)
                # a second call in a row reuses the same scratch
                b.run_noalloc(out, mode, s)
                s.synchronize()
                assert np.array_equal(out.to_host(), ref)


def test_noalloc_limits_and_arguments(gpu, qp):
    s = gpu.Stream()
    b = _batch(gpu, qp, _uneven(8))
    out = gpu.DeviceArray(b.n)
    with pytest.raises(gpu.DfrgInvalidArgument):
        b.run_noalloc(out, gpu.FP64, None)  # a stream is required
    # one radial node per tile: 512 uneven integrals (about 30,000 radial nodes) exceed the tile
    # limit of the static scratch
    big = _batch(gpu, qp, _uneven(512))
    big.set_shape(256, 1, 0)
    n_tiles = big.shape()["n_tiles"]
    assert n_tiles > 16384
    big_out = gpu.DeviceArray(big.n)
    print(f"[case] noalloc limit: {n_tiles} tiles", flush=True)
    with pytest.raises(gpu.DfrgInvalidArgument, match="static scratch"):
        big.run_noalloc(big_out, gpu.FP64, s)
    # the batch still runs normally afterwards, and the limit check launched nothing
    big.set_shape(512, 65536, 0)
    ref = gpu.DeviceArray(big.n)
    big.run(ref, gpu.FP64, s.ptr)
    big.run_noalloc(big_out, gpu.FP64, s)
    s.synchronize()
    assert np.array_equal(big_out.to_host(), ref.to_host())
