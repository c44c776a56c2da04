"""LPA flow server on the GPU: final u(rho) against the oracle's RK4 flow (<= 1e-8 absolute),
continuous batching (admission by HBM footprint, slot limits, FIFO queue), determinism alone vs
in the mix, graph chunking, and eviction of a diverging flow."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_FLOW = 1e-8


def run_all(fs, ids, max_iter=100000):
    out = {}
    it = 0
    while len(out) < len(ids):
        fs.progress()
        for fid in ids:
            if fid not in out and fs.poll(fid) in (2, 3):
                out[fid] = fs.result(fid) if fs.poll(fid) == 2 else None
        it += 1
        assert it < max_iter
    return out


def oracle_flow(qp, n, T, n_steps, k_end, m2_0=0.2, lam=1.0, nx=64):
    x, w = qp.get_points(nx), qp.get_weights(nx)
    fp = O.flow_params(n, 4.0, T, 0.0, 4.0, m2_0, lam)
    return O.flow_lpa(fp, x, w, 20.0, 1.0, k_end, n_steps)


def test_single_flow_vs_oracle(gpu, qp):
    fs = gpu.FlowServer(qp, hbm_budget=1 << 20, max_resident=4, n_steps=2000, steps_per_launch=50)
    print("[case] flow n=96 T=0.1, 2000 RK4 steps", flush=True)
    fid = fs.submit(96, 0.1, mu=0.2)
    res = run_all(fs, [fid])
    u, lat = res[fid]
    ref = oracle_flow(qp, 96, 0.1, 2000, 1e-3)
    err = np.max(np.abs(u - ref))
    print(f"  max |u - u_oracle| = {err:.3e}, latency {lat:.3f} s", flush=True)
    assert err <= TOL_FLOW
    assert u[0] > 0.2  # fluctuations raise the mass in the symmetric phase


@pytest.mark.parametrize("nx", [24, 96])
def test_flow_node_counts(gpu, qp, nx):
    # fewer nodes than lanes, and more (the kernel's loop over nodes)
    fs = gpu.FlowServer(qp, hbm_budget=1 << 20, max_resident=2, nx=nx, n_steps=300, k_end=0.2,
                        steps_per_launch=20)
    print(f"[case] flow n=40 T=0.15 with {nx} nodes, 300 RK4 steps", flush=True)
    fid = fs.submit(40, 0.15)
    u, _ = run_all(fs, [fid])[fid]
    ref = oracle_flow(qp, 40, 0.15, 300, 0.2, nx=nx)
    err = np.max(np.abs(u - ref))
    print(f"  max |u - u_oracle| = {err:.3e}", flush=True)
    assert err <= TOL_FLOW


def test_mix_equals_alone_and_admission(gpu, qp):
    specs = [(64, 0.03), (200, 0.25), (128, 0.1), (77, 0.3), (150, 0.01), (64, 0.2)]
    n_steps = 200
    # alone: one server per flow
    alone = []
    for n, T in specs:
        fs = gpu.FlowServer(qp, hbm_budget=1 << 20, max_resident=1, n_steps=n_steps, k_end=0.3, steps_per_launch=7)
        fid = fs.submit(n, T)
        alone.append(run_all(fs, [fid])[fid][0])
        fs.close()
    # mix: budget fits only ~2 flows of 200 points, at most 3 resident
    budget = gpu.flow_bytes(200) * 2 + gpu.flow_bytes(64)
    fs = gpu.FlowServer(qp, hbm_budget=budget, max_resident=3, n_steps=n_steps, k_end=0.3, steps_per_launch=13)
    ids = [fs.submit(n, T) for n, T in specs]
    print("[case] 6 flows through 3 slots", flush=True)
    max_res, out = 0, {}
    while len(out) < len(ids):
        r, q, d = fs.progress()
        max_res = max(max_res, r)
        assert r <= 3
        for fid in ids:
            if fid not in out and fs.poll(fid) == 2:
                out[fid] = fs.result(fid)[0]
    assert max_res >= 2
    for i, fid in enumerate(ids):
        assert np.array_equal(out[fid], alone[i])  # bitwise: same result alone and in the mix
    for (n, T), fid in zip(specs[:2], ids[:2]):
        ref = oracle_flow(qp, n, T, n_steps, 0.3)
        assert np.max(np.abs(out[fid] - ref)) <= TOL_FLOW
    st = fs.stats()
    assert st["finished"] == 6 and st["p99_s"] >= st["p50_s"] > 0


def test_budget_rejects_oversized_flow(gpu, qp):
    fs = gpu.FlowServer(qp, hbm_budget=gpu.flow_bytes(100), max_resident=2)
    with pytest.raises(gpu.DfrgInvalidArgument):
        fs.submit(101, 0.1)
    with pytest.raises(gpu.DfrgInvalidArgument):
        fs.submit(3, 0.1)


def test_diverging_flow_is_evicted(gpu, qp):
    # m2_0 well below the pole -k^2: E^2 < 0, the flow becomes non-finite and is evicted,
    # while the healthy flow beside it finishes unaffected
    fs = gpu.FlowServer(qp, hbm_budget=1 << 20, max_resident=2, n_steps=100, k_end=0.5, steps_per_launch=10)
    bad = fs.submit(32, 0.1, m2_0=-3.0)
    good = fs.submit(32, 0.1)
    print("[case] diverging flow", flush=True)
    for _ in range(100):
        fs.progress()
        if fs.poll(bad) in (2, 3) and fs.poll(good) in (2, 3):
            break
    assert fs.poll(bad) == 3
    assert fs.poll(good) == 2
    with pytest.raises(gpu.DfrgError):
        fs.result(bad)
    u, _ = fs.result(good)
    ref = oracle_flow(qp, 32, 0.1, 100, 0.5)
    assert np.max(np.abs(u - ref)) <= TOL_FLOW


def test_timing_counts_the_steps_taken(gpu, qp):
    # dfrg_fs_timing: each graph launch advances the resident flows by the steps they have left
    # (a flow admitted at the base step takes steps_per_launch per launch until its n_steps)
    fs = gpu.FlowServer(qp, hbm_budget=1 << 20, max_resident=4, n_steps=130, k_end=0.2, steps_per_launch=25)
    print("[case] flow timing: flows of 40 and 64 points, 130 steps, 25 per launch", flush=True)
    a = fs.submit(40, 0.15)
    b = fs.submit(64, 0.2)
    run_all(fs, [a, b])
    t = fs.timing()
    # both admitted together: 6 launches (5 x 25 + the last 5 steps), each flow takes 130 steps
    assert t["launches"] == 6, t
    assert t["point_steps"] == (40 + 64) * 130, t
    assert t["device_s"] > 0, t
