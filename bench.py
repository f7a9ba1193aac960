"""Headline benchmark: DOF-updates/s per LSRK45 step of 3-D Maxwell at N=9 on one MI355X.

Workload (the precision configuration): Maxwell (Hx,Hy,Hz,Ex,Ey,Ez), N=9 (Np=220,
Nfp=55), K=196608 tets (32^3 x 6 Kuhn cubes, interior vertices jittered +-0.1h, seed 3),
PEC walls, upwind flux, Ez = sin(pi x) sin(pi y), LSRK45 with dt = 0.5 h / N^2.

value      : DOF-updates/s per step of the fused mixed-precision path (FP32 derivative /
             flux / L0 / chain, FP64 lift accumulation and FP64 solution accumulator),
             each step one replay of a hipGraph of its 5 stage kernels; state resident in
             HBM.
baselines  : the plain FP64 GPU path (one kernel per operator, the reference's RHS
             structure) and the fused FP64 path, same workload.
errors     : per-field relative L2 / max error of the mixed path against the fused FP64
             path after the timed steps, and after 1, 10, 100 and 200 steps from the initial
             state (the FP64 path matches the oracle to 1e-12).
roofline   : the fused mixed stage kernel, HBM-bound; achieved = algorithmic bytes per
             stage / average stage duration (HIP events on the library's stream).
cpu_baseline: the oracle's OpenMP restatement of the reference RHS + LSRK45 (FP64) on the
             host cores, on a bounded sample of the same workload (16^3 x 6 elements); on
             that sample the GPU's fused mixed and FP64 states after the same steps are
             compared with the oracle's, per field (gpu_vs_oracle).
also       : the operator microbenchmark (N=7, K=262144), the 3-D wave at N=7, the dense
             nodal comparator, the batched mixed-degree workload and the single-face lift
             conditioning study (N = 8, 11, 14, 17, K=65536), each with its own sizes.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bernsteinbasis.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def maxwell_ic(x, y, z):
    zero = 0 * x
    return np.stack([zero, zero, zero, zero, zero, np.sin(np.pi * x) * np.sin(np.pi * y)])


def field_errors(got, ref, names):
    """per field: L2 of the difference relative to the L2 of the whole reference state
    (fields that are ~0 in the exact solution make a per-field ratio meaningless), and the
    max abs difference next to the field's max abs value"""
    out = {}
    nr = np.linalg.norm(ref)
    for f, nm in enumerate(names):
        d = got[..., f] - ref[..., f]
        out[nm] = {"l2_rel_to_state": float(np.linalg.norm(d) / nr), "max_abs": float(np.abs(d).max()),
                   "field_max_abs": float(np.abs(ref[..., f]).max())}
    tot = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    return out, tot


def error_study(s, q0, dt, marks=(1, 10, 100, 200)):
    """Precision study of the Maxwell N=9 workload: the mixed path (FP32 derivative, flux,
    L0 and chain; FP64 lift accumulation and solution) against the FP64 fused path, both
    from the same initial state, per field after 1, 10, 100 and 200 LSRK45 steps.  The FP64
    fused path is the reference: it matches the oracle's RHS to 1e-12 per evaluation."""
    names = ["Hx", "Hy", "Hz", "Ex", "Ey", "Ez"]
    saved = {}
    s.set_state(q0)
    done = 0
    for mk in marks:
        s.step(dt, mk - done, "fused", "mixed")
        done = mk
        saved[mk] = s.get_state()
    s.set_state(q0)
    done = 0
    out = {}
    for mk in marks:
        s.step(dt, mk - done, "fused", "f64")
        done = mk
        ref = s.get_state()
        got = saved.pop(mk)
        per_field, tot = field_errors(got, ref, names)
        for f, nm in enumerate(names):
            nf = float(np.linalg.norm(ref[..., f]))
            per_field[nm]["l2_rel_to_field"] = (float(np.linalg.norm(got[..., f] - ref[..., f]) / nf)
                                                if nf > 1e-3 * np.linalg.norm(ref) else None)
        out[str(mk)] = {"rel_l2_all": tot, "fields": per_field}
        log(f"mixed vs f64 after {mk} steps: rel L2 {tot:.3e}, Ez max {per_field['Ez']['max_abs']:.3e}")
        del got, ref
    return out


def time_graph(s, dt, steps, warmup, prec, path="fused"):
    s.capture(dt, 1, path, prec)
    for _ in range(warmup):
        s.launch()
    s.ctx.sync()
    s.ctx.timer_start()
    for _ in range(steps):
        s.launch()
    ms = s.ctx.timer_stop()
    s.ctx.sync()
    return ms / steps


def time_stage_kernel(s, prec, reps=20):
    """average duration of one fused stage kernel (steady stage, residual read)"""
    s.step(0.0, 1, "fused", prec)  # warm
    s.ctx.sync()
    s.ctx.timer_start()
    s.step(0.0, reps, "fused", prec)
    return s.ctx.timer_stop() / (5 * reps)


def stream_bandwidth(ctx, gib=2.0, reps=10):
    """achievable HBM bandwidth of a plain copy kernel (bytes read + written per second)"""
    import bbdg
    from bbdg import _lib as B

    n = int(gib * 2 ** 30) // 16 * 16
    a = bbdg.DeviceArray.zeros((n // 8,), np.float64, ctx=ctx)
    b = bbdg.DeviceArray.zeros((n // 8,), np.float64, ctx=ctx)
    import ctypes as C

    L = B.lib()
    ctx.check(L.bb_stream_copy(ctx.h, C.c_void_p(b.ptr), C.c_void_p(a.ptr), n))
    ctx.sync()
    ctx.timer_start()
    for _ in range(reps):
        ctx.check(L.bb_stream_copy(ctx.h, C.c_void_p(b.ptr), C.c_void_p(a.ptr), n))
    ms = ctx.timer_stop() / reps
    a.free()
    b.free()
    return 2 * n / (ms * 1e-3) * 1e-9


def operator_bench(ctx, reps=10):
    """Operator microbenchmark: K=262144 tets at N=7 (Np=120, Nfp=36), F=4 fields, volume
    coefficients and face data i.i.d. U[-1,1] (seed 2), geometric factors cycled from the
    32^3x6 jittered mesh (seed 1).  Times the three derivative matvecs with geometric
    factors (grad) and the 4-face lift, FP64 and FP32, SoA layout [K][F][n]."""
    import bbdg
    from bbdg import mesh

    N, F, K = 7, 4, 262144
    NP, NFP = 120, 36
    m = mesh.kuhn_mesh(32, 0.1, 1, periodic=False)
    geo_h = np.ascontiguousarray(m.geo[np.arange(K) % m.K])
    rng = np.random.default_rng(2)
    x_h = rng.uniform(-1, 1, (K, F, NP))
    f_h = rng.uniform(-1, 1, (K, F, 4 * NFP))
    geo = bbdg.DeviceArray.from_numpy(geo_h, ctx=ctx)
    out = {}
    for dt, name, sz in ((np.float64, "f64", 8), (np.float32, "f32", 4)):
        x = bbdg.DeviceArray.from_numpy(x_h.astype(dt), ctx=ctx)
        xv = x.view((NP, K, F), (1, F * NP, NP))
        o = [bbdg.DeviceArray.empty((K, F, NP), dt, ctx=ctx) for _ in range(3)]
        ov = [a.view(xv.shape, xv.strides) for a in o]
        bbdg.grad_(ov[0], ov[1], ov[2], N, xv, geo)
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            bbdg.grad_(ov[0], ov[1], ov[2], N, xv, geo)
        ms = ctx.timer_stop() / reps
        b = K * (F * 4 * NP * sz + 9 * 8)
        out[f"grad_{name}"] = {"us": ms * 1e3, "GBs": b / (ms * 1e-3) * 1e-9, "bytes": b,
                               "dof_per_s": K * F * NP / (ms * 1e-3)}
        for a in o:
            a.free()
        fd = bbdg.DeviceArray.from_numpy(f_h.astype(dt), ctx=ctx)
        fv = fd.view((4 * NFP, K, F), (1, F * 4 * NFP, 4 * NFP))
        lo = bbdg.DeviceArray.empty((K, F, NP), dt, ctx=ctx)
        lv = lo.view((NP, K, F), (1, F * NP, NP))
        L = bbdg.BernsteinLift(N, dt, "acc64")
        bbdg.mul_(lv, L, fv)
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            bbdg.mul_(lv, L, fv)
        ms = ctx.timer_stop() / reps
        b = K * F * (4 * NFP + NP) * sz
        out[f"lift_{name}"] = {"us": ms * 1e3, "GBs": b / (ms * 1e-3) * 1e-9, "bytes": b,
                               "dof_per_s": K * F * NP / (ms * 1e-3)}
        if dt == np.float32:
            # the same lift with L0 on the matrix cores (FP32 MFMA), timed beside the VALU one
            Lm = bbdg.BernsteinLift(N, dt, "acc64_mfma")
            bbdg.mul_(lv, Lm, fv)
            ctx.sync()
            ctx.timer_start()
            for _ in range(reps):
                bbdg.mul_(lv, Lm, fv)
            ms = ctx.timer_stop() / reps
            out["lift_f32_mfma_l0"] = {"us": ms * 1e3, "GBs": b / (ms * 1e-3) * 1e-9, "bytes": b,
                                       "dof_per_s": K * F * NP / (ms * 1e-3)}
        for a in (x, fd, lo):
            a.free()
    geo.free()
    for k, v in out.items():
        v["frac_of_8TBs"] = v["GBs"] / HBM_PEAK_GBS
    return out


def lift_study(ctx, K=65536):
    """Lift conditioning study: single-face lift, N in {8, 11, 14, 17}, K columns of face
    coefficients i.i.d. N(0,1) (seed 4), and the same scaled per column by 10^U[-3,3].
    Variants: FP64; FP32 in, FP32 chain, FP32 out ("fp32"); FP32 in, FP64 chain and
    weighting ("acc64"); FP32 in, float-float chain ("twosum").  Errors are relative L2 over
    all outputs against the FP64 GPU lift of the FP64 input (the reference algorithm in
    Float64) and, to isolate the arithmetic, against the FP64 lift of the FP32-rounded
    input.  The FP64 GPU lift itself is checked against an extended-precision restatement
    on the same inputs in tests/test_gpu_operators.py (test_face_lift_fp64_conditioning)."""
    import bbdg

    rows = []
    rng = np.random.default_rng(4)
    for N in (8, 11, 14, 17):
        nfp, np_ = (N + 1) * (N + 2) // 2, (N + 1) * (N + 2) * (N + 3) // 6
        base = rng.standard_normal((K, nfp))
        scale = 10.0 ** rng.uniform(-3, 3, (K, 1))
        for dist, x in (("normal", base), ("scaled", base * scale)):
            L = bbdg.BernsteinLift(N)

            def run(xh, mode, out_dt):
                xd = bbdg.DeviceArray.from_numpy(xh, ctx=ctx)
                xv = xd.view((nfp, K), (1, nfp))
                od = bbdg.DeviceArray.empty((K, np_), out_dt, ctx=ctx)
                ov = od.view((np_, K), (1, np_))
                bbdg.face_lift_multiply_(ov, L, xv, mode=mode)
                ctx.sync()
                ctx.timer_start()
                bbdg.face_lift_multiply_(ov, L, xv, mode=mode)
                us = ctx.timer_stop() * 1e3
                r = od.numpy().astype(np.float64)
                xd.free()
                od.free()
                return r, us

            ref, us64 = run(x, "plain", np.float64)
            x32 = x.astype(np.float32)
            ref32in, _ = run(x32.astype(np.float64), "plain", np.float64)
            rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
            row = {"N": N, "dist": dist, "fp64": {"us": us64}}
            for name, mode, odt in (("fp32", "plain", np.float32), ("acc64", "acc64", np.float64),
                                    ("twosum", "twosum", np.float64)):
                got, us = run(x32, mode, odt)
                row[name] = {"err_vs_fp64": rel(got, ref), "err_vs_fp64_of_fp32_input": rel(got, ref32in), "us": us}
            rows.append(row)
            log(json.dumps(row))
    return rows


def wave_n7_bench(ctx, steps=20, warmup=3, n=32, plain_steps=3):
    """Secondary workload at the low end of the metric's degree range: the 3-D acoustic
    wave (p,u,v,w) at N=7 (Np=120) on the same 32^3 x 6 jittered Kuhn mesh (seed 3),
    periodic, Gaussian pulse p = exp(-25 |x - 0.5|^2), LSRK45 dt = 0.5 h / N^2; fused mixed,
    fused FP64 and plain FP64 steps, each autotuned and replayed from a hipGraph."""
    import bbdg
    from bbdg import basis, mesh

    N = 7
    m = mesh.kuhn_mesh(n, 0.1, 3, periodic=True)
    q0 = basis.interpolate(m, N, lambda x, y, z: np.stack(
        [np.exp(-25 * ((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2)), 0 * x, 0 * x, 0 * x]))
    s = bbdg.Solver(m, N, "wave", ctx=ctx)
    dt = 0.5 * m.h / N ** 2
    out = {"workload": f"wave3d N=7 K={m.K} (32^3x6 Kuhn, jitter seed 3) periodic upwind LSRK45", "dofs": s.dofs}
    states = {}
    for prec in ("mixed", "f64"):
        s.set_state(q0)
        t = s.autotune(prec)
        s.set_state(q0)
        ms = time_graph(s, dt, steps, warmup, prec)
        states[prec] = s.get_state()
        out[f"fused_{prec}"] = {"ms_per_step": ms, "value": s.dofs / (ms * 1e-3), "tiling": t}
        log(f"wave N=7 fused {prec}: tiling={t} {ms:.3f} ms/step {s.dofs / ms * 1e3:.3e} DOF/s")
    s.set_state(q0)
    s.step(dt, 1, "plain", "f64")
    s.ctx.sync()
    s.ctx.timer_start()
    s.step(dt, plain_steps, "plain", "f64")
    ms = s.ctx.timer_stop() / plain_steps
    out["plain_f64"] = {"ms_per_step": ms, "value": s.dofs / (ms * 1e-3)}
    log(f"wave N=7 plain f64: {ms:.3f} ms/step")
    out["speedup_mixed_vs_plain_f64"] = out["fused_mixed"]["value"] / out["plain_f64"]["value"]
    out["rel_l2_mixed_vs_f64"] = float(np.linalg.norm(states["mixed"] - states["f64"]) / np.linalg.norm(states["f64"]))
    out["steps_compared"] = steps + warmup
    s.close()
    return out


def nodal_bench(ctx, Ns=(3, 5, 7, 9), n=32, reps=5):
    """Dense nodal comparator (extras/3d_nodal_wave.jl:57-73) against the Bernstein paths,
    3-D acoustic wave on the 32^3 x 6 periodic mesh (seed 3), FP64, per degree:
    - one RHS: dense nodal (nine Np x Np and one Np x 4Nfp library GEMMs over all K*F
      columns, no RK update) against the Bernstein plain path (one kernel per operator) and
      the fused stage kernel (FP64 and mixed), both timed per LSRK stage (RHS + RK update);
    - the README's volume / surface split: nine dense derivative GEMMs against nine
      Bernstein derivative stencils (bb_deriv_apply), one dense LIFT GEMM against the
      Bernstein 4-face lift (bb_lift_apply), on the same K*F columns."""
    import bbdg
    from bbdg import basis, mesh
    from bbdg.nodal import NodalOperators, NodalRHS
    from bbdg.operators import (BernsteinDerivativeMatrix_3D_r, BernsteinDerivativeMatrix_3D_s,
                                BernsteinDerivativeMatrix_3D_t)

    m = mesh.kuhn_mesh(n, 0.1, 3, periodic=True)
    fn = lambda x, y, z: np.stack([np.exp(-25 * ((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2)), 0 * x, 0 * x, 0 * x])

    def timed(f):
        f()
        ctx.sync()
        ctx.timer_start()
        for _ in range(reps):
            f()
        return ctx.timer_stop() / reps

    rows = []
    F, K = 4, m.K
    for N in Ns:
        NP, NFP = (N + 1) * (N + 2) * (N + 3) // 6, (N + 1) * (N + 2) // 2
        ops = NodalOperators(N, ctx)
        c = basis.interpolate(m, N, fn)
        qn = bbdg.DeviceArray.from_numpy(np.ascontiguousarray(np.einsum("ab,kbf->kfa", ops.V, c)), ctx=ctx)
        nr = NodalRHS(m, N, "wave", ops=ops, ctx=ctx)
        ms_dense = timed(lambda: nr.rhs_device(qn))
        ms_dense_vol = timed(nr.derivatives_device)
        ms_dense_surf = timed(nr.lift_device)
        flops = ops.flops_per_rhs(K, F)
        # Bernstein operators on the same columns: nine derivative stencils, one 4-face lift
        sh, st = (NP, K, F), (1, F * NP, NP)
        Ds = [cls(N) for cls in (BernsteinDerivativeMatrix_3D_r, BernsteinDerivativeMatrix_3D_s,
                                 BernsteinDerivativeMatrix_3D_t)]

        def bern_volume():
            for cc in range(3):
                for ax in range(3):
                    bbdg.mul_(nr.d9[3 * cc + ax].view(sh, st), Ds[ax], nr.fu[cc].view(sh, st))

        ms_bern_vol = timed(bern_volume)
        Lb = bbdg.BernsteinLift(N)
        fv = nr.flux.view((4 * NFP, K, F), (1, F * 4 * NFP, 4 * NFP))
        ms_bern_surf = timed(lambda: bbdg.mul_(nr.du.view(sh, st), Lb, fv))
        nr.close()
        qn.free()
        s = bbdg.Solver(m, N, "wave", ctx=ctx)
        s.set_state(c)
        dt = 0.5 * m.h / N ** 2
        s.step(dt, 1, "plain", "f64")
        ctx.sync()
        ctx.timer_start()
        s.step(dt, 1, "plain", "f64")
        ms_plain = ctx.timer_stop() / 5  # one step = 5 stages
        fused = {}
        for prec in ("f64", "mixed"):
            s.autotune(prec)
            fused[prec] = time_stage_kernel(s, prec, reps=4)
        s.close()
        row = {"N": N, "dofs": K * F * NP, "ms_dense_nodal_rhs_f64": ms_dense,
               "dense_gemm_tflops": flops / (ms_dense * 1e-3) * 1e-12,
               "ms_dense_volume": ms_dense_vol, "ms_dense_surface": ms_dense_surf,
               "ms_bernstein_volume": ms_bern_vol, "ms_bernstein_surface": ms_bern_surf,
               "speedup_volume": ms_dense_vol / ms_bern_vol, "speedup_surface": ms_dense_surf / ms_bern_surf,
               "ms_bernstein_plain_stage_f64": ms_plain, "ms_bernstein_fused_stage_f64": fused["f64"],
               "ms_bernstein_fused_stage_mixed": fused["mixed"],
               "speedup_plain_f64_vs_dense": ms_dense / ms_plain,
               "speedup_fused_f64_vs_dense": ms_dense / fused["f64"]}
        rows.append(row)
        log(f"nodal N={N}: dense RHS {ms_dense:.3f} ms ({row['dense_gemm_tflops']:.1f} TFLOP/s; volume "
            f"{ms_dense_vol:.3f}, surface {ms_dense_surf:.3f}); Bernstein volume {ms_bern_vol:.3f}, surface "
            f"{ms_bern_surf:.3f}; stages: plain {ms_plain:.3f}, fused f64 {fused['f64']:.3f}, fused mixed "
            f"{fused['mixed']:.3f} ms")
    return rows


def batch_bench(ctx, steps=100, n_inst=64, seed=5):
    """Batching workload: n_inst independent acoustic-wave instances, each a jittered
    periodic Kuhn mesh of n^3 x 6 tets with n ~ U{4..16}; every element has its own
    degree N ~ U{3..10}; Gaussian pulses p = exp(-|x - c|^2 / w^2) with c ~ U[0.3,0.7]^3,
    w ~ U[0.05,0.2], u = v = w = 0 (all from seed 5); per-instance step 0.5 h / 10^2.
    Elements are grouped by degree; every stage runs the transfer pass of the faces whose
    neighbour has another degree, then each degree group's fused stage kernel, the groups on
    four streams; a step is one hipGraph replay, tilings are autotuned per group.  Timed: steps
    LSRK45 steps in mixed precision and in FP64; errors of mixed against FP64 after them."""
    from bbdg import mesh
    from bbdg.batch import Batch, BatchSolver

    t0 = time.time()
    rng = np.random.default_rng(seed)
    ns = rng.integers(4, 17, n_inst)
    meshes = [mesh.kuhn_mesh(int(n), 0.1, seed * 1000 + i, periodic=True) for i, n in enumerate(ns)]
    degrees = [rng.integers(3, 11, m.K) for m in meshes]
    fns = []
    for i in range(n_inst):
        c, w = rng.uniform(0.3, 0.7, 3), rng.uniform(0.05, 0.2)
        fns.append(lambda x, y, z, c=c, w=w: np.stack(
            [np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / w ** 2), 0 * x, 0 * x, 0 * x]))
    b = Batch(meshes, degrees, [0.5 * m.h / 10 ** 2 for m in meshes])
    q0 = b.interpolate(fns)
    log(f"batch: {n_inst} instances, K={b.K}, DOF={b.dofs}, groups {[(N, c) for N, _, c in b.groups()]}, "
        f"set-up {time.time() - t0:.1f}s")
    s = BatchSolver(b, ctx=ctx)
    out = {"instances": n_inst, "K": b.K, "dofs": b.dofs, "steps": steps,
           "elements_per_degree": {int(N): int(c) for N, _, c in b.groups()}}
    states = {}
    for prec in ("mixed", "f64"):
        s.set_state(q0)
        tun = s.autotune(prec)
        s.set_state(q0)
        s.capture(1, prec)
        s.launch()
        ctx.sync()
        s.set_state(q0)
        ctx.timer_start()
        for _ in range(steps):
            s.launch()
        ms = ctx.timer_stop() / steps
        states[prec] = s.get_state()
        out[prec] = {"ms_per_step": ms, "dof_updates_per_s": b.dofs / (ms * 1e-3), "tilings": tun}
        log(f"batch {prec}: {ms:.3f} ms/step, {b.dofs / ms * 1e3:.3e} DOF/s, tilings {tun}")
    d = states["mixed"] - states["f64"]
    ref = states["f64"]
    per = {}
    for f, nm in enumerate(["p", "u", "v", "w"]):
        num = den = mx = 0.0
        for N, st, c in b.groups():
            dv, rv = b.field_view(d, N)[:, f], b.field_view(ref, N)[:, f]
            num += float((dv ** 2).sum())
            den += float((rv ** 2).sum())
            mx = max(mx, float(np.abs(dv).max()))
        per[nm] = {"l2_rel_to_state": num ** 0.5 / float(np.linalg.norm(ref)), "max_abs": mx}
    out["errors_mixed_vs_f64"] = {"rel_l2_all": float(np.linalg.norm(d) / np.linalg.norm(ref)), "fields": per}
    out["speedup_mixed_vs_f64"] = out["f64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    log(f"batch mixed vs f64 after {steps} steps: rel L2 {out['errors_mixed_vs_f64']['rel_l2_all']:.3e}")
    s.close()
    return out


def cpu_threads():
    """every core this process may run on (sched_getaffinity), capped by OMP_NUM_THREADS
    where the environment sets it (the host's CPU share per GPU)"""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    n = avail
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return n, avail, cap


def cpu_baseline(n_sample, steps):
    """The reference's threaded CPU path (examples/3d_advec_multithreaded.jl:31-44 applied to
    the Maxwell RHS of examples/3d_maxwell.jl:36-93, LSRK45), as the oracle's C restatement
    compiled -O3 -march=native for this host, OpenMP static schedule over elements."""
    import oracle as O
    from bbdg import basis, mesh

    threads, avail, cap = cpu_threads()
    t0 = time.perf_counter()
    L = O.baseline_lib()
    t_build = time.perf_counter() - t0
    N = 9
    m = mesh.kuhn_mesh(n_sample, 0.1, 3, periodic=False)
    ma = O.MeshArrays(m, N, [1, 1, 1, -1, -1, -1])
    q = np.ascontiguousarray(basis.interpolate(m, N, maxwell_ic))
    dt = 0.5 * m.h / N ** 2
    q0 = q.copy()
    O.lsrk_step(1, N, q, dt, ma, nthreads=threads, L=L)  # warm
    t = time.perf_counter()
    for _ in range(steps):
        O.lsrk_step(1, N, q, dt, ma, nthreads=threads, L=L)
    el = time.perf_counter() - t
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        O.rhs(1, N, q, ma, nthreads=threads, L=L)
    el_rhs = (time.perf_counter() - t) / reps
    dofs = m.K * 220 * 6
    # the oracle as the checker: the GPU's fused mixed and FP64 paths after the same steps
    # from the same initial state on this mesh, per field against the oracle's state
    import bbdg

    names = ["Hx", "Hy", "Hz", "Ex", "Ey", "Ez"]
    vs_oracle = {"steps": steps + 1, "K": m.K, "reference": "oracle C restatement (FP64) of the reference RHS + "
                                                            "LSRK45, same mesh, initial state and dt"}
    s = bbdg.Solver(m, N, "maxwell", flux="upwind")
    for prec in ("mixed", "f64"):
        s.set_state(q0)
        s.step(dt, steps + 1, "fused", prec)
        per_field, tot = field_errors(s.get_state(), q, names)
        vs_oracle[prec] = {"rel_l2_all": tot, "fields": per_field}
        log(f"gpu fused {prec} vs oracle after {steps + 1} steps (K={m.K}): rel L2 {tot:.3e}")
    s.close()
    return {"value": dofs * steps / el, "unit": "DOF-updates/s per step", "cores": threads, "kind": "port",
            "per_rhs_value": dofs / el_rhs, "ms_per_step": el / steps * 1e3,
            "nproc": os.cpu_count(), "affinity_cores": avail, "omp_num_threads_env": cap,
            "cpu_model": O.cpu_model(), "flags": " ".join(O.BASELINE_FLAGS + ["-fopenmp"]),
            "schedule": "static", "K": m.K, "N": N, "build_s": t_build, "gpu_vs_oracle": vs_oracle,
            "sample": f"{steps} LSRK45 steps (+{reps} RHS evaluations) of Maxwell N=9 FP64 on the {n_sample}^3x6 "
                      f"(K={m.K}) jittered PEC mesh: the oracle's C restatement of the reference's threaded RHS, "
                      f"gcc {' '.join(O.BASELINE_FLAGS)} on this host, OpenMP static schedule, {threads} threads "
                      f"({avail} cores in the affinity mask, OMP_NUM_THREADS={cap}), {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=32, help="cubes per edge (K = 6 n^3)")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-ops", action="store_true", help="skip the operator microbenchmark")
    ap.add_argument("--lift-study", action="store_true",
                    help="run only the lift conditioning study (single-face lift precision variants)")
    ap.add_argument("--no-lift-study", action="store_true", help="skip the lift conditioning study")
    ap.add_argument("--ops-only", action="store_true", help="run only the operator microbenchmark")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched mixed-degree workload")
    ap.add_argument("--no-wave7", action="store_true", help="skip the N=7 wave workload")
    ap.add_argument("--no-nodal", action="store_true", help="skip the dense nodal comparator")
    ap.add_argument("--nodal-only", action="store_true", help="run only the dense nodal comparator")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched mixed-degree workload")
    ap.add_argument("--batch-steps", type=int, default=100, help="timed steps of the batched workload")
    ap.add_argument("--stage-only", action="store_true",
                    help="run only a few fused mixed stages (for counter collection)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    import bbdg
    from bbdg import basis, mesh

    if args.lift_study:
        print(json.dumps({"lift_conditioning": lift_study(bbdg.default_context())}))
        return
    if args.nodal_only:
        print(json.dumps({"dense_nodal_comparator": nodal_bench(bbdg.default_context())}))
        return
    if args.batch_only:
        ctx = bbdg.default_context()
        print(json.dumps({"batch_mixed_degree": batch_bench(ctx, steps=args.batch_steps)}))
        stream_bandwidth(ctx, gib=1.0, reps=1)  # the calibration dispatch for HBM byte counters
        return
    if args.ops_only:
        ctx = bbdg.default_context()
        ops = operator_bench(ctx)
        for k, v in ops.items():
            log(f"operator {k}: {v['us']:.1f} us, {v['GBs']:.0f} GB/s ({v['frac_of_8TBs']:.2f} of 8 TB/s)")
        print(json.dumps({"operators_N7_K262144_F4": ops}))
        stream_bandwidth(ctx, gib=1.0, reps=1)  # the calibration dispatch for HBM byte counters
        return
    N, n = 9, args.n
    t0 = time.time()
    m = mesh.kuhn_mesh(n, 0.1, 3, periodic=False)
    q0 = basis.interpolate(m, N, maxwell_ic)
    log(f"mesh K={m.K} + IC in {time.time() - t0:.1f}s")
    s = bbdg.Solver(m, N, "maxwell", flux="upwind")
    info = s.ctx.info()
    dt = 0.5 * m.h / N ** 2
    dofs = s.dofs
    if args.stage_only:
        s.set_state(q0)
        t = s.autotune("mixed")
        s.step(dt, args.steps, "fused", "mixed")
        s.ctx.sync()
        log(f"stage-only: {args.steps} fused mixed steps, tiling {t}")
        # one copy of exactly 1 GiB: the calibration dispatch for HBM byte counters
        stream_bandwidth(s.ctx, gib=1.0, reps=1)
        return
    res = {}
    for prec in ("mixed", "f64"):
        s.set_state(q0)
        eb = s.autotune(prec)
        s.set_state(q0)
        ms = time_graph(s, dt, args.steps, args.warmup, prec)
        final = s.get_state()
        stage_ms = time_stage_kernel(s, prec)
        b, _ = s.stage_bytes(prec)
        res[prec] = {"ms_per_step": ms, "value": dofs / (ms * 1e-3), "tiling": eb, "stage_ms": stage_ms,
                     "stage_bytes": b, "state": final}
        log(f"fused {prec}: tiling={eb} {ms:.3f} ms/step {dofs / ms * 1e3:.3e} DOF/s, stage {stage_ms * 1e3:.1f} us "
            f"= {b / stage_ms * 1e-6:.0f} GB/s")
    study = error_study(s, q0, dt)
    # plain FP64 path (one kernel per operator)
    s.set_state(q0)
    s.step(dt, 1, "plain", "f64")
    s.ctx.sync()
    s.ctx.timer_start()
    s.step(dt, args.plain_steps, "plain", "f64")
    ms_plain = s.ctx.timer_stop() / args.plain_steps
    log(f"plain f64: {ms_plain:.3f} ms/step")
    # accuracy of the mixed path after the same number of steps (warmup + timed)
    names = ["Hx", "Hy", "Hz", "Ex", "Ey", "Ez"]
    per_field, tot = field_errors(res["mixed"]["state"], res["f64"]["state"], names)
    nsteps_err = args.steps + args.warmup
    mx = res["mixed"]
    # first stage reads no residual; average the algorithmic bytes over the five stages
    w = float(dofs)
    sres = 4.0
    geo_b = m.K * (9 * 8 + 8 + 20 * 8 + 4 * 8)
    step_bytes = 5 * (w * 16 + w * sres + geo_b) + 4 * w * sres
    achieved = step_bytes / 5 / (mx["ms_per_step"] * 1e-3 / 5) * 1e-9
    # HBM traffic of the chosen stage kernel, from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes over this workload (scaled by the 1 GiB copy calibration)
    traffic = None
    prof = os.path.join(ROOT, "profiles", "stage_maxwell_n9_k196608.json")
    kname = f"k_stage<9, 1, true, {mx['tiling'] % 16}, {mx['tiling'] // 16}, false>"
    if os.path.exists(prof):
        rec = json.load(open(prof))["kernels"].get(kname)
        if rec:
            traffic = rec["fetch_bytes"] + rec["write_bytes"]
    copy_gbs = stream_bandwidth(s.ctx)
    log(f"stream copy: {copy_gbs:.0f} GB/s")
    ops = None
    if not args.no_ops:
        s.close()
        ops = operator_bench(bbdg.default_context())
        for k, v in ops.items():
            log(f"operator {k}: {v['us']:.1f} us, {v['GBs']:.0f} GB/s ({v['frac_of_8TBs']:.2f} of 8 TB/s)")
    wave7 = wave_n7_bench(bbdg.default_context()) if not args.no_wave7 else None
    nodal = nodal_bench(bbdg.default_context()) if not args.no_nodal else None
    batch = None
    if not args.no_batch:
        batch = batch_bench(bbdg.default_context(), steps=args.batch_steps)
    lift_cond = None if args.no_lift_study else lift_study(bbdg.default_context())
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.3e} DOF/s ({cpu['sample']})")
    out = {
        "metric": "DOF-updates/s per RK step at N=7-9: FP32+FP64-lift batched vs plain FP64",
        "value": mx["value"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (analytic TM110 initial condition on a seeded jittered mesh)",
        "config": {"workload": f"maxwell3d N=9 K={m.K} ({n}^3x6 Kuhn, jitter seed 3) PEC upwind LSRK45 "
                               f"dt=0.5h/N^2, fused mixed precision, hipGraph step",
                   "N": N, "K": m.K, "fields": 6, "dofs": dofs, "tiling": mx["tiling"],
                   "device": info},
        "baselines": {
            "plain_f64": {"value": dofs / (ms_plain * 1e-3), "ms_per_step": ms_plain},
            "fused_f64": {"value": res["f64"]["value"], "ms_per_step": res["f64"]["ms_per_step"],
                          "tiling": res["f64"]["tiling"]},
        },
        "speedup_vs_plain_f64": mx["value"] / (dofs / (ms_plain * 1e-3)),
        "speedup_vs_fused_f64": mx["value"] / res["f64"]["value"],
        "per_rhs_value": mx["value"] * 5,
        "errors_vs_f64": {"steps": nsteps_err, "rel_l2_all": tot, "fields": per_field},
        "errors_vs_f64_by_step": {"reference": "fused FP64 path from the same state (matches the oracle RHS to "
                                               "1e-12 per evaluation)", "by_step": study},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/stage_maxwell_n9_k196608.json (rocprofv3 FETCH_SIZE x2.0, WRITE_SIZE x1.0)",
                     "kernel": kname, "bytes_per_launch": step_bytes / 5,
                     "launch_us_graph": mx["ms_per_step"] * 1e3 / 5, "launch_us_eager": mx["stage_ms"] * 1e3,
                     "stream_copy_GBs": copy_gbs, "frac_of_stream_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
        "operators_N7_K262144_F4": ops,
        "wave_n7": wave7,
        "dense_nodal_comparator": nodal,
        "batch_mixed_degree": batch,
        "lift_conditioning": lift_cond,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
