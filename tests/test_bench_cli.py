"""bench.py's command line on the CPU: a request for more than one GPU is refused with a
message (the library measures one GPU's share; the sharding is a design), and the
storage table it reports traffic for names every storage mode of the library."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_storage_tables_cover_every_mode():
    sys.path.insert(0, REPO)
    import bench
    from mbl import abi
    modes = (abi.STORE_NATIVE, abi.STORE_F32S, abi.STORE_F16S, abi.STORE_BF16S, abi.STORE_F16C, abi.STORE_F16D)
    assert set(bench.BYTES_PER_NODE) == set(modes) == set(bench.STORAGE_NAMES)
    assert all(bench.BYTES_PER_NODE[m] == (72 if m in (abi.STORE_NATIVE, abi.STORE_F32S) else 36) for m in modes)
    # the committed PMC traffic of the FP32 and 16-bit kernels is within 5 % of the algorithmic bytes
    for m in (abi.STORE_NATIVE, abi.STORE_F16S, abi.STORE_F16D):
        t = bench.traffic_from_profile(m)
        alg = bench.BYTES_PER_NODE[m] * 16384 * 16384
        assert t is not None and alg <= t < 1.05 * alg
