"""bb_solver_step_noalloc, the stepping entry point that allocates nothing (it has no
counterpart in the reference; it serves callers that cannot allocate device memory), against
the entry point it varies, bb_solver_step on the fused path: same solver, same initial
state, same tiling, so the results must agree bit for bit.  Two calls in a row on one GPU,
the second on a smaller mesh, share the library's scratch array; a mesh beyond the scratch
is refused from the host before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PREC = {"f64": 0, "mixed": 1}


def step_both(pde, N, n, prec, nsteps, seed):
    import bbdg
    from bbdg import _lib as B
    from bbdg import mesh

    m = mesh.kuhn_mesh(n, 0.1, seed, periodic=pde == "wave")
    s = bbdg.Solver(m, N, pde)
    rng = np.random.default_rng(seed)
    q0 = rng.uniform(-1, 1, (m.K, s.Np, s.F))
    dt = 0.5 * m.h / N ** 2
    print(f"noalloc: {pde} N={N} K={m.K} prec={prec} steps={nsteps}", flush=True)
    s.set_state(q0)
    s.step(dt, nsteps, "fused", prec)
    ref = s.get_state()
    d = bbdg.DeviceArray.from_numpy(np.ascontiguousarray(q0.transpose(0, 2, 1)), ctx=s.ctx)  # [K][F][Np]
    rc = B.lib().bb_solver_step_noalloc(s.h, PREC[prec], C.c_void_p(d.ptr), float(dt), int(nsteps))
    s.ctx.check(rc)
    s.ctx.sync()
    got = d.numpy().transpose(0, 2, 1)
    d.free()
    s.close()
    return got, ref


@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_noalloc_matches_step_twice_second_smaller(ctx, prec):
    # Maxwell N=7 on 3^3 x 6 elements (odd stage count: the state ends in the scratch and is
    # copied back), then the wave system N=5 on 2^3 x 6 elements (even stage count)
    got, ref = step_both("maxwell", 7, 3, prec, 3, 11)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    got, ref = step_both("wave", 5, 2, prec, 2, 12)
    assert np.array_equal(got, ref), np.abs(got - ref).max()


def test_noalloc_refuses_a_state_beyond_the_scratch(ctx):
    import bbdg
    from bbdg import _lib as B
    from bbdg import mesh

    # Maxwell N=9 on 8^3 x 6 elements: 4,055,040 state words = 31 MiB, with the FP64
    # residual 62 MiB (fits 64 MiB); on 9^3 x 6 elements 45 + 45 MiB does not
    m = mesh.kuhn_mesh(9, 0.1, 3, periodic=False)
    s = bbdg.Solver(m, 9, "maxwell")
    d = bbdg.DeviceArray.zeros((m.K, 6, 220), ctx=s.ctx)
    rc = B.lib().bb_solver_step_noalloc(s.h, 0, C.c_void_p(d.ptr), 1e-4, 1)
    assert rc == B.BB_ERR_ARG
    assert b"scratch" in B.lib().bb_last_error(s.ctx.h)
    d.free()
    s.close()
