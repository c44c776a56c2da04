"""Host-side argument checks of the Python wrappers (no GPU): a device output that is too small or
of the wrong dtype is refused before any kernel could write past it."""
import numpy as np
import pytest

import diffrg_amd as D


class _Fake(D.DeviceArray):
    def __init__(self, n, dtype=np.float64):  # no allocation: only the bookkeeping fields
        self.dtype, self.n, self.ptr = np.dtype(dtype), n, 1

    def __del__(self):  # nothing to free
        pass


def test_need_accepts_large_enough_output():
    D._need(_Fake(64), 64, np.float64, "out")
    D._need(_Fake(100), 64, np.float64, "out")


@pytest.mark.parametrize("arr,msg", [(_Fake(63), "holds 63"), (_Fake(64, np.float32), "dtype"),
                                     (object(), "live DeviceArray")])
def test_need_refuses(arr, msg):
    with pytest.raises(ValueError, match=msg):
        D._need(arr, 64, np.float64, "out")


def test_wrappers_check_outputs_first():
    small = _Fake(3)
    b = object.__new__(D.SEBatch)  # no library object behind it
    b.n, b.ptr = 8, 1
    try:
        with pytest.raises(ValueError, match="holds 3"):
            b.run(small)
        with pytest.raises(ValueError, match="holds 3"):
            b.graph_launch(small)
    finally:
        b.ptr = None
    with pytest.raises(ValueError, match="holds 3"):
        D.threshold_batch(_Fake(8), _Fake(64), _Fake(64), small)
    with pytest.raises(ValueError, match="holds 3"):
        D.ft_batch({"T": _Fake(8)}, small, 6.0)


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N>1 stops before touching a GPU, with a JSON line that says why."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]
