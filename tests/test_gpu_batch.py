"""GPU parity of batches of elements of mixed degree (bb_batch_*) against the oracle's
mixed-degree restatement: RHS, LSRK45 steps, graph replays, every tiling of every degree
group, walls and periodic meshes, and a single-degree batch against the single-mesh
solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def make(ns, degs, seed, periodic=True):
    from bbdg import mesh
    from bbdg.batch import Batch

    rng = np.random.default_rng(seed)
    meshes = [mesh.kuhn_mesh(n, 0.1, seed + i, periodic=periodic) for i, n in enumerate(ns)]
    degrees = [rng.integers(degs[0], degs[1] + 1, m.K) for m in meshes]
    b = Batch(meshes, degrees, [0.5 * m.h / degs[1] ** 2 for m in meshes])
    fns = []
    for i in range(len(ns)):
        c, w = rng.uniform(0.3, 0.7, 3), rng.uniform(0.1, 0.2)
        fns.append(lambda x, y, z, c=c, w=w: np.stack(
            [np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / w ** 2), 0.2 * np.sin(2 * np.pi * y),
             0.1 * np.cos(2 * np.pi * z), 0.1 * np.sin(2 * np.pi * x)]))
    return b, b.interpolate(fns)


@pytest.mark.parametrize("ns,degs,periodic", [([2, 3], (1, 6), True), ([3], (3, 10), False), ([2, 2, 3], (2, 9), True)])
def test_batch_rhs_parity(ctx, orc, ns, degs, periodic):
    from bbdg.batch import BatchSolver

    b, q = make(ns, degs, seed=sum(ns) + degs[1], periodic=periodic)
    ba = orc.BatchArrays(b, [1, 1, 1, 1])
    ref = b.from_node_major(orc.rhs_md(0, b.to_node_major(q), ba, nthreads=4))
    s = BatchSolver(b)
    s.set_state(q)
    assert rel(s.rhs("f64"), ref) <= TOL64, rel(s.rhs("f64"), ref)
    assert rel(s.rhs("mixed"), ref) <= 2e-5, rel(s.rhs("mixed"), ref)
    s.close()


def test_batch_steps_and_graph(ctx, orc):
    from bbdg.batch import BatchSolver

    b, q = make([2, 3], (2, 8), seed=11)
    ba = orc.BatchArrays(b, [1, 1, 1, 1])
    qr = b.to_node_major(q)
    for _ in range(3):
        orc.lsrk_step_md(0, qr, ba, nthreads=4)
    ref = b.from_node_major(qr)
    s = BatchSolver(b)
    s.set_state(q)
    s.step(3, "f64")
    eager = s.get_state()
    assert rel(eager, ref) <= TOL64, rel(eager, ref)
    s.set_state(q)
    s.capture(1, "f64")
    for _ in range(3):
        s.launch()
    assert rel(s.get_state(), eager) <= TOL64
    s.set_state(q)
    s.step(3, "mixed")
    assert rel(s.get_state(), ref) <= 1e-5
    s.close()


def test_batch_every_tiling(ctx, orc):
    """each degree group on each of its compiled tilings (the others on their defaults)"""
    import bbdg
    from bbdg.batch import BatchSolver

    b, q = make([3], (1, 10), seed=5)
    ba = orc.BatchArrays(b, [1, 1, 1, 1])
    ref = b.from_node_major(orc.rhs_md(0, b.to_node_major(q), ba, nthreads=4))
    s = BatchSolver(b)
    s.set_state(q)
    tried = 0
    for N in b.Ns:
        for lift in ("gather", "pingpong", "wave", "wave_gt", "wave1"):
            for eb in (1, 2, 3, 4):
                try:
                    s.set_tuning("f64", N, bbdg.Solver.tiling_code(eb, lift))
                except bbdg.BBError:
                    continue
                tried += 1
                assert rel(s.rhs("f64"), ref) <= TOL64, (N, lift, eb)
    assert tried >= len(b.Ns)
    t = s.autotune("mixed")
    assert set(t) == set(b.Ns)
    assert rel(s.rhs("mixed"), ref) <= 2e-5
    s.close()


def test_single_degree_batch_equals_solver(ctx):
    import bbdg
    from bbdg import basis, mesh
    from bbdg.batch import Batch, BatchSolver

    N = 5
    m = mesh.kuhn_mesh(3, 0.1, 2, periodic=True)
    fn = lambda x, y, z: np.stack([np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y), 0 * x, 0.1 + 0 * x, np.sin(2 * np.pi * z)])
    b = Batch([m], [np.full(m.K, N)], [0.5 * m.h / N ** 2])
    s = BatchSolver(b)
    s.set_state(b.interpolate([fn]))
    got = s.rhs("f64")
    sol = bbdg.Solver(m, N, "wave")
    sol.set_state(basis.interpolate(m, N, fn))
    ref = sol.rhs("fused", "f64")  # (K, Np, F) host layout
    ref_dev = np.concatenate([ref[b.order[e]].T.reshape(-1) for e in range(b.K)])
    assert rel(got, ref_dev) <= 1e-13
    s.close()


def test_batch_steps_on_lanes(ctx, orc):
    """a batch of 3072 elements in four degree groups, stepped on several streams: eager steps
    and graph replays against the oracle, mixed and FP64, and the traces of a new state after
    a run in the other precision"""
    from bbdg.batch import BatchSolver

    b, q = make([8], (1, 4), seed=21)
    assert b.K == 3072
    ba = orc.BatchArrays(b, [1, 1, 1, 1])
    qr = b.to_node_major(q)
    for _ in range(2):
        orc.lsrk_step_md(0, qr, ba, nthreads=4)
    ref = b.from_node_major(qr)
    s = BatchSolver(b)
    try:
        s.set_state(q)
        s.step(2, "mixed")
        assert rel(s.get_state(), ref) <= 1e-5
        s.set_state(q)
        s.step(2, "f64")
        eager = s.get_state()
        assert rel(eager, ref) <= TOL64, rel(eager, ref)
        s.set_state(q)
        s.capture(1, "f64")
        for _ in range(2):
            s.launch()
        assert rel(s.get_state(), eager) <= TOL64
        # a mixed run, then FP64 steps from a new state: the FP64 traces are gathered again
        s.set_state(q)
        s.step(1, "mixed")
        s.set_state(q)
        s.step(2, "f64")
        assert rel(s.get_state(), eager) <= TOL64
        t = s.autotune("mixed")
        assert set(t) == set(b.Ns)
        s.set_state(q)
        s.step(2, "mixed")
        assert rel(s.get_state(), ref) <= 1e-5
    finally:
        s.close()
