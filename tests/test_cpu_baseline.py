"""The CPU baseline leg of bench.py: the oracle compiled -O3 -march=native for this host
computes the same RHS as the checker build, and the thread count follows the affinity
mask and OMP_NUM_THREADS."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_baseline_build_matches_checker():
    import oracle as O
    from bbdg import basis, mesh

    N = 5
    m = mesh.kuhn_mesh(3, 0.1, 3, periodic=False)
    ma = O.MeshArrays(m, N, [1, 1, 1, -1, -1, -1])
    rng = np.random.default_rng(0)
    q = np.ascontiguousarray(rng.uniform(-1, 1, (m.K, O.Np(N), 6)))
    a = O.rhs(1, N, q, ma, nthreads=2)
    b = O.rhs(1, N, q, ma, nthreads=2, L=O.baseline_lib())
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-13
    assert "-march=native" in O.BASELINE_FLAGS and os.path.basename(O.baseline_path()).startswith("liboracle_baseline_")


def test_thread_count_follows_affinity_and_cap(monkeypatch):
    sys.path.insert(0, ROOT)
    import bench

    avail = len(os.sched_getaffinity(0))
    monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
    assert bench.cpu_threads()[:2] == (avail, avail)
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.cpu_threads()[0] == 1
