#!/usr/bin/env python3
"""Benchmark of the MI355X SDP bounding path.  Prints ONE JSON line.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                [--no-cpu] [--no-plain] [--cpu-seconds S] [--plain-nodes P]

Workloads (a "step" bounds the nodes listed; inputs are built before the
timed region, instance data is uploaded inside each bound call):
  c5        (default) branch and bound of the 8x8x8 toroidal grid (n = 512,
            weights round(100 N(0,1))), --c5-nodes node bounds from the root
            per step; plain and CPU legs on a shared sample of tree nodes,
            the CPU leg on one core and on worker processes; the root's
            last ADMM call replayed on identical inputs in the chosen and
            plain forms and by the FP64 oracle (bound error at 1e-6)
  c2-root   root SDPbound of G(100, 0.3), integer weights uniform on
            +-{1..10}, seed 2 (IPM + cutting-plane loop with triangle,
            pentagonal and heptagonal cuts, params file values)
  c3-batch  64 nodes of the 15x15 toroidal grid (weights +-1, seed 3) with
            0..75 random fixings each (n 150..225), bounded in one batch;
            triangle + pentagonal cuts (TriIneq 1500, PentIneq 500), no
            heptagonal cuts, no diff give-up
  c1-bab    full branch and bound of g05_60.0 (rudy), time to the proven
            optimum
Forms:
  chosen    all nodes of a step in one batch (lockstep kernels, one
            workgroup per node), sign-iteration projection on FP32 MFMA
            with FP64 finish, PCG cut solve to 1e-7 relative residual
  plain     one node at a time, FP64 Jacobi eigen-projection, PCG cut solve
            to 1e-12 (a direct solve in the reference); timed on all nodes
            of a step, or on the first --plain-nodes of c3-batch
The metric value is node bounds per second of the chosen form.
`cpu_baseline` times the CPU FP64 oracle (LAPACK ADMM, dense Cholesky cut
solve) on a bounded sample of the same nodes (c5: on worker processes, one
per available core up to 16, and on one core).  `roofline` prices the
projection kernel from its HIP-event time in a separate profiled step.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

# the CPU baseline is timed on one core: one MKL thread whatever the host sets
os.environ["MKL_NUM_THREADS"] = "1"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "time-to-proven-optimum + node bounds/sec, chosen vs plain FP64 GPU at 1e-6"
PEAK_FP32_MFMA = 157.3  # TFLOP/s, dense
PEAK_FP64_MFMA = 78.6
PHASES = ["prep", "cg", "w", "eig", "post", "safe", "ipm", "trisep", "sa", "gw", "zfac"]
CHOSEN = dict(projection=1, cg_rtol=1e-7, cg_adapt=0.01, gw_mult=4)
PLAIN = dict(projection=0, cg_rtol=1e-12)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Timer:
    """Per-phase HIP-event times of the library (madam_profile)."""

    def __init__(self, lib):
        self.lib = lib

    def read(self):
        ms = np.zeros(len(PHASES))
        cnt = np.zeros(len(PHASES), dtype=np.int64)
        self.lib.madam_profile_read(len(PHASES), ms.ctypes.data_as(C.POINTER(C.c_double)),
                                    cnt.ctypes.data_as(C.POINTER(C.c_long)), 1)
        return ms, cnt

    def __enter__(self):
        self.lib.madam_profile(1)
        self.read()
        return self

    def __exit__(self, *a):
        self.ms, self.cnt = self.read()
        self.lib.madam_profile(0)


def wl_c2(args):
    import madam_amd as G
    from madam_amd import instances as I
    SPL = I.objective_matrix(*I.config2(seed=2))
    n = SPL.shape[0]

    def step(form, sample=None):
        B = G.Bounder(SPL, G.default_params(), **(CHOSEN if form == "chosen" else PLAIN))
        b = B.sdpbound([B.node()], ranks=[0])[0]
        return {"bound": float(b), "bounds": np.array([b]), "admm_iters": int(B.ctx.admm_iters),
                "nodes": 1, "n3_iters": float(B.ctx.admm_iters) * n ** 3, "n": n}

    def cpu(seconds):
        from oracle import oracle_py as O
        t0 = time.perf_counter()
        reps, bound = 0, None
        while reps == 0 or time.perf_counter() - t0 < seconds:
            C.CDLL(None).srand(2020)
            s = O.Solver(SPL, O.default_params())
            bound, _ = s.root_bound()
            reps += 1
        dt = time.perf_counter() - t0
        return reps / dt, f"{reps} oracle root SDPbound evaluations of the same node in {dt:.1f} s", bound

    cfg = {"workload": "root SDPbound of G(100,0.3), w uniform on +-{1..10}, seed 2, n=100, "
                       "tri+pent+hept cuts, params-file values",
           "n": n, "nodes_per_step": 1}
    return cfg, step, cpu


def wl_c3(args):
    import madam_amd as G
    from madam_amd import instances as I
    SPL = I.objective_matrix(*I.torus2d(15, 15, seed=3))
    N = SPL.shape[0]
    fix = I.random_fixings(N - 1, 64, seed=33, kmin=0, kmax=75)
    params = dict(include_Hepta=0, TriIneq=1500, adjust_TriIneq=0, PentIneq=500, use_diff=0)
    # Every node is bounded as its own subproblem (its own incumbent, from
    # its own rounding heuristic), in both forms and in the oracle.

    def step(form, sample=None):
        sel = fix if sample is None else fix[:sample]
        n3 = sum((N - int(xf.sum())) ** 3 for xf, _ in sel) / len(sel)
        if form == "chosen":
            B = G.Bounder(SPL, G.default_params(**params), own_incumbent=1, verbose=1, **CHOSEN)
            bs = B.sdpbound([B.node(xf, sx) for xf, sx in sel], ranks=[1] * len(sel))
            it = int(B.ctx.admm_iters)
        else:
            bs, it = [], 0
            for q, (xf, sx) in enumerate(sel):
                B = G.Bounder(SPL, G.default_params(**params), **PLAIN)
                bs.append(B.sdpbound([B.node(xf, sx)], ranks=[1])[0])
                it += int(B.ctx.admm_iters)
                log(f"plain: node {q + 1} of {len(sel)} bounded")
            bs = np.array(bs)
        return {"bound": float(np.mean(bs)), "bounds": np.asarray(bs), "admm_iters": it,
                "nodes": len(sel), "n3_iters": float(it) * n3}

    def cpu(seconds):
        from oracle import oracle_py as O
        p = O.default_params(**params)
        t0 = time.perf_counter()
        done = 0
        for xf, sx in fix:
            C.CDLL(None).srand(2020)
            s = O.Solver(SPL, p)
            node = s.load_subproblem(xf, sx)
            s.evaluate(node, 1)
            done += 1
            log(f"cpu oracle: node {done} bounded")
            if time.perf_counter() - t0 >= seconds:
                break
        dt = time.perf_counter() - t0
        return done / dt, f"the first {done} of the 64 nodes bounded by the oracle in {dt:.1f} s", None

    cfg = {"workload": "64 nodes of the 15x15 torus (w +-1, seed 3) with 0..75 random fixings "
                       "(n 150..225) in one batch, tri+pent cuts (TriIneq 1500, PentIneq 500), "
                       "no hepta, no diff give-up, each node with its own incumbent",
           "n": "150..225", "nodes_per_step": 64}
    return cfg, step, cpu


def wl_c1(args):
    import madam_amd as G
    from madam_amd import instances as I
    SPL = I.objective_matrix(*I.read_rudy(os.path.join(ROOT, "tests", "golden", "rudy", "g05_60.0")))

    def step(form, sample=None):
        kw = dict(CHOSEN, batch=16) if form == "chosen" else dict(PLAIN, batch=1)
        B = G.Bounder(SPL, G.default_params(), **kw)
        opt = B.bab()
        return {"bound": float(opt), "bounds": np.array([opt]), "admm_iters": int(B.ctx.admm_iters),
                "nodes": int(B.ctx.nodes), "n3_iters": float(B.ctx.admm_iters) * 60 ** 3,
                "root_bound": float(B.ctx.root_bound)}

    def cpu(seconds):
        from oracle import oracle_py as O
        C.CDLL(None).srand(2020)
        t0 = time.perf_counter()
        s = O.Solver(SPL, O.default_params())
        opt = s.bab()
        dt = time.perf_counter() - t0
        return s.s.nodes / dt, f"oracle B&B of g05_60.0: optimum {opt:.0f}, {s.s.nodes} nodes in {dt:.2f} s", opt

    cfg = {"workload": "full B&B of g05_60.0 (G(60,0.5), w=1), params-file values",
           "n": 60, "nodes_per_step": "B&B"}
    return cfg, step, cpu


def wl_c4(args):
    """Learned warm start: 2000 training graphs G(100, 0.5), w in {-1,0,1}
    (targets from the GPU FP64 ADMM, untimed), then the step = warm-started
    ADMM (basic SDP, tol 1e-5) on 200 fresh graphs, batched; the plain form
    is the cold start; both report ADMM iterations."""
    import madam_amd as G
    from madam_amd import instances as I
    from madam_amd import learned as LW
    t0 = time.perf_counter()
    Ltr = [I.objective_matrix(*I.config4_graph(10000 + s)) for s in range(args.c4_train)]
    Lev = [I.objective_matrix(*I.config4_graph(20000 + s)) for s in range(200)]
    Xs = [r.X for r in LW.solve_basic(Ltr, tol=1e-6)]
    log(f"c4: targets for {len(Ltr)} graphs in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    model, losses = LW.LearnedWarmStart.train(Ltr, Xs, epochs=2, batch=4096, lr=2e-3, seed=0)
    log(f"c4: {len(losses)} training steps on the GPU in {time.perf_counter() - t1:.1f} s, "
        f"loss {losses[:20].mean():.4f} -> {losses[-20:].mean():.4f}")
    info = {"train_graphs": len(Ltr), "train_steps": int(len(losses)),
            "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-20:].mean())}

    def step(form, sample=None):
        if form == "chosen":
            starts = [model.start(L) for L in Lev]
            res = LW.solve_basic(Lev, tol=1e-5, starts=starts)
        else:
            res = LW.solve_basic(Lev, tol=1e-5)
        bs = np.array([r.f for r in res])
        it = sum(r.nbit for r in res)
        return {"bound": float(bs.mean()), "bounds": bs, "admm_iters": it, "nodes": len(Lev),
                "n3_iters": float(it) * 100 ** 3}

    def cpu(seconds):
        return None, "not timed for this workload", None

    cfg = {"workload": "learned warm start: 2000 training graphs G(100,0.5), w in {-1,0,1}; "
                       "basic-SDP ADMM (tol 1e-5) on 200 fresh graphs, warm vs cold start",
           "n": 100, "nodes_per_step": 200, **info}
    return cfg, step, cpu


def wl_bab150(args):
    """Branch and bound of pw05_150.0 (rudy, n=150; the reference recorded
    optimum 17852, root bound 18083.50, 66941 nodes in 576 s on 240 cores).
    Chosen: node pipeline of 1024 slots; a step bounds args.bab_nodes nodes
    (0: the whole tree, to the proven optimum).  Plain: one node at a time,
    FP64 Jacobi, on the first 200 nodes, and beside it the chosen form one
    node at a time on the same 200 nodes (the projection and cut-solve gain
    without batching); CPU: the oracle on the first 40.  Early nodes are
    mostly branched after the IPM bound (the diff test), deep ones run the
    whole cutting-plane loop, so nodes/s depends on how far the search got."""
    import madam_amd as G
    from madam_amd import instances as I
    path = os.path.join(ROOT, "tests", "golden", "rudy150", "pw05_150.0")
    SPL = I.objective_matrix(*I.read_rudy(path))

    def step(form, sample=None):
        if form == "chosen":
            B = G.Bounder(SPL, G.default_params(), batch=1024, verbose=1, **CHOSEN)
            opt = B.bab(max_nodes=args.bab_nodes)
        else:
            B1 = G.Bounder(SPL, G.default_params(), batch=1, **CHOSEN)
            t1 = time.perf_counter()
            B1.bab(max_nodes=200)
            t1 = time.perf_counter() - t1
            extra = {"chosen_unbatched_same_nodes": {
                "value": int(B1.ctx.nodes) / t1, "nodes": int(B1.ctx.nodes),
                "admm_calls": int(B1.ctx.admm_calls), "incumbent": float(B1.ctx.LB)}}
            B = G.Bounder(SPL, G.default_params(), batch=1, **PLAIN)
            opt = B.bab(max_nodes=200)
        r = {"bound": float(opt), "bounds": np.array([opt]), "admm_iters": int(B.ctx.admm_iters),
             "nodes": int(B.ctx.nodes), "n3_iters": float(B.ctx.admm_iters) * 140 ** 3,
             "open": int(B.ctx.open_nodes), "admm_calls": int(B.ctx.admm_calls)}
        if form != "chosen":
            r["extra"] = extra
        return r

    def cpu(seconds):
        from oracle import oracle_py as O
        C.CDLL(None).srand(2020)
        t0 = time.perf_counter()
        s = O.Solver(SPL, O.default_params())
        s.bab(40)
        dt = time.perf_counter() - t0
        return s.s.nodes / dt, f"oracle B&B of pw05_150.0, first {s.s.nodes} nodes in {dt:.1f} s", None

    cfg = {"workload": "branch and bound of pw05_150.0 (n=150, rudy), node pipeline of 1024, "
                       + ("to the proven optimum" if args.bab_nodes <= 0 else f"first {args.bab_nodes} nodes"),
           "n": 150, "nodes_per_step": "B&B"}
    return cfg, step, cpu


C5_PARAMS = dict(TriIneq=3000, adjust_TriIneq=0, PentIneq=1000, HeptaIneq=1000)


def vendor_fp64_eig(n, reps=5):
    """AMD's own FP64 symmetric eigensolver on this GPU as a reference point
    for the projection (measurement only, nothing of the product uses it):
    rocSOLVER dsyevd -- with eigenvectors (what an FP64 project_SDP needs,
    operators.c:579 calls dsyevr) and values only -- on a random symmetric
    matrix of order n, ms per call from HIP events around the call."""
    import madam_amd as G
    G.lib()  # the HIP runtime the library loaded
    hip = C.CDLL("libamdhip64.so.7", mode=getattr(os, "RTLD_NOLOAD", 4) | C.RTLD_GLOBAL)
    rb = C.CDLL("/opt/rocm/lib/librocblas.so")
    rs = C.CDLL("/opt/rocm/lib/librocsolver.so")
    V = C.c_void_p

    def ck(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: error {rc}")

    h = V()
    ck(rb.rocblas_create_handle(C.byref(h)), "rocblas_create_handle")
    bufs = [V() for _ in range(5)]
    sizes = [8 * n * n, 8 * n * n, 8 * n, 8 * n, 4]
    try:
        for b, sz in zip(bufs, sizes):
            ck(hip.hipMalloc(C.byref(b), C.c_size_t(sz)), "hipMalloc")
        A, A0, D, E, info = bufs
        rng = np.random.default_rng(n)
        M = rng.uniform(-0.5, 0.5, (n, n))
        M = np.asfortranarray(np.tril(M) + np.tril(M, -1).T)
        ck(hip.hipMemcpy(A0, M.ctypes.data_as(V), C.c_size_t(8 * n * n), 1), "hipMemcpy")
        ev = [V(), V()]
        for e in ev:
            ck(hip.hipEventCreate(C.byref(e)), "hipEventCreate")
        out = {}
        for name, evect in (("ms_with_vectors", 211), ("ms_values_only", 213)):
            tot = 0.0
            for r in range(reps + 1):
                ck(hip.hipMemcpy(A, A0, C.c_size_t(8 * n * n), 3), "hipMemcpy")
                ck(hip.hipEventRecord(ev[0], V()), "hipEventRecord")
                ck(rs.rocsolver_dsyevd(h, C.c_int(evect), C.c_int(122), C.c_int(n), A, C.c_int(n),
                                       D, E, info), "rocsolver_dsyevd")
                ck(hip.hipEventRecord(ev[1], V()), "hipEventRecord")
                ck(hip.hipEventSynchronize(ev[1]), "hipEventSynchronize")
                ms = C.c_float()
                ck(hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1]), "hipEventElapsedTime")
                if r > 0:  # the first call warms up
                    tot += ms.value
            out[name] = tot / reps
        for e in ev:
            hip.hipEventDestroy(e)
        return {"library": "rocSOLVER dsyevd (FP64)", "n": n, **out}
    finally:
        for b in bufs:
            if b.value:
                hip.hipFree(b)
        rb.rocblas_destroy_handle(h)


def _cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def _oracle_node(job):
    """One oracle SDPbound (bounding.c:25-250) of a node, in a worker
    process: (seconds, bound)."""
    SPL, params, LB, diff, xf, sx = job
    import ctypes as CC
    from oracle import oracle_py as O
    CC.CDLL(None).srand(2020)
    t0 = time.perf_counter()
    s = O.Solver(SPL, O.default_params(**params))
    s.set_lb(LB)
    s.s.diff = diff
    node = s.load_subproblem(xf, sx)
    b = s.evaluate(node, 1)
    return time.perf_counter() - t0, b


def wl_c5(args):
    """Config 5 at scale: branch and bound of the 8x8x8 toroidal grid (n =
    512, weights round(100 N(0,1)), seed 5), tri + pent + hept cuts capped at
    3000 + 1000 + 1000 per round.  A step = the search from the root with a
    budget of --c5-nodes node bounds (root included), pipeline of 64 nodes,
    chosen form; value = node bounds per second over the step.  The search
    is far from proven at that budget (the root gap is ~2%); the line
    reports the root bound, incumbent and gap.
    Comparison legs on a shared sample (--plain-nodes nodes at the top of
    the tree: 1-4 random fixings, the step's incumbent and diff, each node
    pruning against its own heuristic's incumbent): the chosen form in one
    batch, the plain FP64 form one node at a time, and the FP64 oracle on
    one core and on P worker processes.  Bound error at identical inputs:
    the root's last ADMM call (its cut set, cut duals and warm start,
    captured by madam_call_state) run again in the chosen and plain forms
    and by the FP64 oracle."""
    import madam_amd as G
    from madam_amd import instances as I
    SPL = I.objective_matrix(*I.torus3d(8, seed=5))
    N = SPL.shape[0]
    sample = I.random_fixings(N - 1, args.plain_nodes, seed=55, kmin=1, kmax=4)
    state = {}

    def step(form, sample_n=None):
        if form == "chosen":
            B = G.Bounder(SPL, G.default_params(**C5_PARAMS), batch=64, **CHOSEN)
            B.bab(max_nodes=args.c5_nodes)
            state.update(LB=float(B.ctx.LB), diff=float(B.ctx.diff),
                         root_bound=float(B.ctx.root_bound))
            return {"bound": float(B.ctx.root_bound), "bounds": np.array([B.ctx.root_bound]),
                    "admm_iters": int(B.ctx.admm_iters), "nodes": int(B.ctx.nodes),
                    "n3_iters": float(B.ctx.admm_iters) * N ** 3, "n": N,
                    "incumbent": float(B.ctx.LB), "open": int(B.ctx.open_nodes),
                    "admm_calls": int(B.ctx.admm_calls), "ipm_iters": int(B.ctx.ipm_iters),
                    "ns_fallbacks": int(B.ctx.ns_fallbacks)}
        if args.timed_form == "plain":
            # the plain FP64 form timed (and profiled) on the whole step
            B = G.Bounder(SPL, G.default_params(**C5_PARAMS), batch=64, **PLAIN)
            B.bab(max_nodes=args.c5_nodes)
            return {"bound": float(B.ctx.root_bound), "bounds": np.array([B.ctx.root_bound]),
                    "admm_iters": int(B.ctx.admm_iters), "nodes": int(B.ctx.nodes),
                    "n3_iters": float(B.ctx.admm_iters) * N ** 3, "n": N,
                    "incumbent": float(B.ctx.LB), "open": int(B.ctx.open_nodes),
                    "admm_calls": int(B.ctx.admm_calls), "ipm_iters": int(B.ctx.ipm_iters),
                    "ns_fallbacks": 0}
        # plain FP64 form, one node at a time, on the shared sample
        bs, it = [], 0
        for q, (xf, sx) in enumerate(sample):
            B = G.Bounder(SPL, G.default_params(**C5_PARAMS), LB=state["LB"], own_incumbent=1,
                          **PLAIN)
            B.ctx.diff = state["diff"]
            bs.append(B.sdpbound([B.node(xf, sx)], ranks=[1])[0])
            it += int(B.ctx.admm_iters)
            log(f"plain: sample node {q + 1} of {len(sample)} bounded")
        return {"bound": float(np.mean(bs)), "bounds": np.array(bs), "admm_iters": it,
                "nodes": len(sample), "n3_iters": float(it) * N ** 3}

    def plain_step():
        """The same step (search from the root, same budget and pipeline) in
        the plain FP64 form."""
        t0 = time.perf_counter()
        B = G.Bounder(SPL, G.default_params(**C5_PARAMS), batch=64, **PLAIN)
        B.bab(max_nodes=args.c5_nodes)
        dt = time.perf_counter() - t0
        return {"value": int(B.ctx.nodes) / dt, "seconds": dt, "nodes": int(B.ctx.nodes),
                "root_bound": float(B.ctx.root_bound), "incumbent": float(B.ctx.LB),
                "admm_iters": int(B.ctx.admm_iters), "admm_calls": int(B.ctx.admm_calls)}

    def chosen_on_sample():
        t0 = time.perf_counter()
        B = G.Bounder(SPL, G.default_params(**C5_PARAMS), LB=state["LB"], own_incumbent=1,
                      **CHOSEN)
        B.ctx.diff = state["diff"]
        bs = B.sdpbound([B.node(xf, sx) for xf, sx in sample], ranks=[1] * len(sample))
        dt = time.perf_counter() - t0
        return {"value": len(sample) / dt, "seconds": dt, "bounds": [float(b) for b in bs]}

    def cpu(seconds, pool=None):
        jobs = [(SPL, C5_PARAMS, state["LB"], state["diff"], xf, sx) for xf, sx in sample]
        t0 = time.perf_counter()
        one = [_oracle_node(jobs[0])]
        t1 = time.perf_counter() - t0
        out = {"one_core": {"value": 1.0 / t1, "seconds": t1, "nodes": 1,
                            "bound": float(one[0][1])}}
        if pool is not None:
            t0 = time.perf_counter()
            res = pool.map(_oracle_node, jobs, chunksize=1)
            dt = time.perf_counter() - t0
            out["workers"] = {"value": len(jobs) / dt, "seconds": dt, "nodes": len(jobs),
                              "bounds": [float(b) for _, b in res],
                              "node_seconds": [float(t) for t, _ in res]}
        return out

    def replay():
        """The root's last ADMM call on identical inputs in three arithmetics."""
        from oracle import oracle_py as O
        B = G.Bounder(SPL, G.default_params(**C5_PARAMS), **CHOSEN)
        t0 = time.perf_counter()
        b, prob = B.sdpbound_last_call(B.node(), rank=0)
        t_root = time.perf_counter() - t0
        calls = prob.pop("calls")
        out = {"root_seconds": t_root, "root_bound": float(b), "call": calls,
               "n": prob["L"].shape[0], "tol": prob["tol"],
               "cuts": [len(prob["tri"]), len(prob["pent"]), len(prob["hept"])]}
        for name, opts in (("chosen", dict(projection=1, cg_rtol=1e-7, cg_adapt=0.01)),
                           ("plain", dict(projection=0))):
            t0 = time.perf_counter()
            r = G.ADMM_solver(**prob, **opts)
            out[name] = {"f": r.f, "nbit": r.nbit, "seconds": time.perf_counter() - t0}
            log(f"replay: {name} form done")
        # the FP64 oracle (LAPACK dsyevr projection) on the same inputs
        p = O.default_params(**C5_PARAMS)
        s = O.Solver(SPL, p)
        s.load_subproblem(None, None)
        s.set_cuts(prob["tri"], prob["pent"], prob["hept"])
        n = prob["L"].shape[0]
        s.mat("X")[:] = prob["X"]
        s.mat("Q")[:] = prob["Q"]
        s.vec("y", n)[:] = prob["y"]
        t0 = time.perf_counter()
        f, _, nbit = s.admm(prob["sigma"], prob["tol"])
        out["oracle"] = {"f": f, "nbit": nbit, "seconds": time.perf_counter() - t0}
        log("replay: oracle done")
        for name in ("chosen", "plain"):
            out[name]["rel_err_vs_oracle"] = abs(out[name]["f"] - f) / abs(f)
        out["chosen_rel_err_vs_plain"] = abs(out["chosen"]["f"] - out["plain"]["f"]) / abs(
            out["plain"]["f"])
        return out

    cfg = {"workload": f"branch and bound of the 8x8x8 torus (n=512, weights round(100 N(0,1)), "
                       f"seed 5) from the root, {args.c5_nodes} node bounds per step, pipeline of "
                       f"64, <= 3000 + 1000 + 1000 tri/pent/hept cuts per round",
           "n": N, "nodes_per_step": args.c5_nodes}
    return cfg, step, cpu, {"chosen_on_sample": chosen_on_sample, "replay": replay,
                            "plain_step": plain_step,
                            "sample": f"{len(sample)} nodes with 1-4 random fixings (seed 55)"}


def wl_ns_batch(args):
    """Throughput of the projection kernels in a full batch: k copies of
    one node of order n (G(n,0.3) objective, triangle-free basic SDP, X = I
    start), 64 ADMM iterations each, one ADMM_solver_batch call per step.
    The bound of each node is an ordinary ADMM safe bound; the line is
    about the eig phase (roofline)."""
    import madam_amd as G
    from madam_amd import instances as I
    n, k = args.ns_n, args.ns_k
    vals = [v for v in range(-10, 11) if v != 0]
    L = I.objective_matrix(*I.gnp(n, 0.3, vals, 5))

    def step(form, sample=None):
        kk = k if form == "chosen" else min(k, 2)
        probs = [dict(L=L, X=np.eye(L.shape[0]), y=np.diag(L).copy(), Q=np.zeros_like(L),
                      sigma=1.0, tol=1e-14, itermax=64) for _ in range(kk)]
        opts = (dict(projection=1, cg_rtol=1e-7, ns_layout=args.ns_layout) if form == "chosen"
                else dict(projection=0))
        res = G.ADMM_solver_batch(probs, **opts)
        it = sum(r.nbit for r in res)
        bs = np.array([r.f for r in res])
        return {"bound": float(bs[0]), "bounds": bs, "admm_iters": it, "nodes": kk,
                "n3_iters": float(it) * L.shape[0] ** 3}

    def cpu(seconds):
        return None, "not timed for this workload", None

    cfg = {"workload": f"projection throughput: {k} nodes of order {n} in one batch, "
                       "64 ADMM iterations each", "n": n, "nodes_per_step": k}
    return cfg, step, cpu


WORKLOADS = {"c5": wl_c5, "c2-root": wl_c2, "c3-batch": wl_c3, "c1-bab": wl_c1,
             "c4-learned": wl_c4, "bab150": wl_bab150, "ns-batch": wl_ns_batch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-replay", action="store_true", help="c5: skip the identical-inputs replay")
    ap.add_argument("--no-vendor", action="store_true",
                    help="skip the rocSOLVER dsyevd reference timing of one projection")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="c5: oracle worker processes (0: the cores this process may use, "
                         "at most 16)")
    ap.add_argument("--c4-train", type=int, default=2000, help="c4-learned: training graphs")
    ap.add_argument("--bab-nodes", type=int, default=20000,
                    help="bab150: nodes per step (0: solve to the proven optimum)")
    ap.add_argument("--plain-nodes", type=int, default=None,
                    help="c3-batch: nodes bounded one at a time by the plain form (4); c5: "
                         "nodes of the shared comparison sample (8)")
    ap.add_argument("--c5-nodes", type=int, default=63,
                    help="c5: node bounds per step (root included)")
    ap.add_argument("--ns-n", type=int, default=150, help="ns-batch: node order")
    ap.add_argument("--ns-k", type=int, default=128, help="ns-batch: nodes in the batch")
    ap.add_argument("--ns-layout", type=int, default=0, choices=[0, 1, 2, 3],
                    help="chosen form's sign-iteration layout: 0 automatic, 1 one workgroup "
                         "per node, 2 tile-parallel")
    ap.add_argument("--graphs", action="store_true",
                    help="replay recurring ADMM check intervals from captured HIP graphs "
                         "(default: every kernel launched directly, measured faster)")
    ap.add_argument("--timed-form", default="chosen", choices=["chosen", "plain"],
                    help="form run in the warmup, timed and profiled steps (plain: to profile "
                         "the plain FP64 form alone; the line then reports it as value)")
    args = ap.parse_args()
    if args.plain_nodes is None:
        args.plain_nodes = 8 if args.workload == "c5" else 4
    CHOSEN["ns_layout"] = args.ns_layout
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the bounding path "
              "shards by B&B node across GPUs, which this build does not run.", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline's worker processes are forked before anything touches
    # the GPU
    cpu_model, cpu_avail = _cpu_info()
    pool = None
    if args.workload == "c5" and not args.no_cpu:
        import multiprocessing as mp
        nw = args.cpu_workers if args.cpu_workers > 0 else min(16, cpu_avail)
        pool = mp.get_context("fork").Pool(nw)
        pool_size = nw

    import madam_amd as G
    G.open_device(0)
    lib = G.lib()
    G.set_graphs(args.graphs)
    wl = WORKLOADS[args.workload](args)
    cfg, step, cpu = wl[:3]
    extra = wl[3] if len(wl) > 3 else {}

    for w in range(args.warmup):
        step(args.timed_form)
        log(f"warmup step {w + 1} of {args.warmup} done")
    t0 = time.perf_counter()
    nodes = 0
    for k in range(args.steps):
        r = step(args.timed_form)
        nodes += r["nodes"]
        log(f"timed step {k + 1} of {args.steps} done ({time.perf_counter() - t0:.2f} s)")
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = nodes / dt
    graphs = None
    if args.graphs:
        # the same step once more with every kernel launched directly
        cap, rep = G.graph_stats()
        G.set_graphs(False)
        t1 = time.perf_counter()
        step(args.timed_form)
        td = time.perf_counter() - t1
        graphs = {"captures": cap, "replays": rep, "ms_per_step_direct_launches": td * 1e3}
        log("direct-launch step done")

    roofline = None
    if not args.no_profile and args.timed_form == "chosen":
        with Timer(lib) as tm:
            rp = step(args.timed_form)
        log("profiled step done")
        eig = PHASES.index("eig")
        eig_ms = float(tm.ms[eig])
        launches = int(tm.cnt[eig])
        f32 = 2.0 * 39 * rp["n3_iters"]   # 39 FP32 products of order n per projection
        f64 = 2.0 * 3 * rp["n3_iters"]    # 3 FP64 products
        sec = eig_ms * 1e-3
        achieved = (f32 + f64) / sec / 1e12 if sec > 0 else 0.0
        t_peak = (f32 / PEAK_FP32_MFMA + f64 / PEAK_FP64_MFMA) / 1e12
        # HBM bytes per projection, measured with rocprofv3 FETCH_SIZE /
        # WRITE_SIZE (per-dispatch KiB scaled by a 64 MiB copy: fetch x2.00,
        # write x1.00), summed over the launches of one projection.  One
        # node of order 100 (tile-parallel layout): 6692 KiB -- the FP32
        # iterates pass through L2 between launches.  Algorithmic: W in, X+
        # out, 2 x 8 n^2 bytes.
        # HBM bytes per projection from rocprofv3 FETCH_SIZE / WRITE_SIZE
        # counter passes over this workload (fetch scaled x2, write x1, the
        # factors a 64 MiB copy reads), summed over the projection's launches
        traffic = {"c2-root": 6692.0 * 1024, "c5": 536144.0 * 1024}.get(args.workload)
        phase_ms = {p: round(float(v), 3) for p, v, c in zip(PHASES, tm.ms, tm.cnt) if c}
        roofline = {"bound": "mfma", "kernel": "sign-iteration projection (39 FP32-MFMA products + 3 "
                    "FP64-MFMA products; k_ns_gemm32/k_ns_gemm64 tile launches or k_ns_project)",
                    "achieved": achieved,
                    "peak": PEAK_FP32_MFMA, "unit": "TFLOP/s",
                    "frac": t_peak / sec if sec > 0 else 0.0, "traffic": traffic,
                    "frac_note": "time at the FP32/FP64 MFMA peaks for these flops / measured time",
                    "launches": launches, "avg_launch_ms": eig_ms / max(1, launches),
                    "ms_per_projection": eig_ms / max(1, rp["admm_iters"]),
                    "flops_per_projection": "84 n^3 (78 n^3 FP32 + 6 n^3 FP64)",
                    "traffic_unit": "HBM bytes per projection",
                    "algorithmic_bytes_per_projection": 16.0 * rp.get("n", 0) ** 2 or None,
                    "projections": rp["admm_iters"],
                    "phase_ms": phase_ms,
                    "phase_share": {p: round(v / max(1e-9, sum(phase_ms.values())), 4)
                                    for p, v in phase_ms.items()}}

    plain = None
    if not args.no_plain:
        sample = args.plain_nodes if args.workload == "c3-batch" else None
        t1 = time.perf_counter()
        pr = step("plain", sample)
        tp = time.perf_counter() - t1
        log("plain form done")
        nb = len(pr["bounds"])
        plain = {"value": pr["nodes"] / tp, "unit": "node bounds/s", "seconds": tp,
                 "nodes": pr["nodes"], "bound": pr["bound"], "admm_iters": pr["admm_iters"]}
        if args.workload == "c5":
            cs = extra["chosen_on_sample"]()
            log("chosen form on the sample done")
            plain["sample"] = extra["sample"] + ", one at a time"
            plain["chosen_same_sample"] = cs
            plain["chosen_over_plain"] = cs["value"] / plain["value"]
            diff = np.abs(np.array(cs["bounds"]) - pr["bounds"]) / np.abs(pr["bounds"])
            ps = extra["plain_step"]()
            log("plain form on the whole step done")
            ps["chosen_over_plain"] = value / ps["value"]
            ps["root_bound_rel_diff_vs_chosen"] = abs(ps["root_bound"] - r["bound"]) / abs(
                r["bound"])
            plain["same_workload"] = ps
        else:
            plain["chosen_over_plain"] = value / plain["value"]
            diff = np.abs(r["bounds"][:nb] - pr["bounds"]) / np.abs(pr["bounds"])
        plain["rel_bound_diff_max"] = float(diff.max())
        plain.update(pr.get("extra", {}))
        if "admm_calls" in pr:
            plain["admm_calls"] = pr["admm_calls"]
        if args.workload == "c1-bab":
            plain["time_to_optimum_s"] = tp

    cpub = None
    if not args.no_cpu and args.workload != "c4-learned":
        if args.workload == "c5":
            co = cpu(args.cpu_seconds, pool)
            w = co["workers"]
            cpub = {"value": w["value"], "unit": "node bounds/s", "cores": pool_size,
                    "kind": "port", "cpu_model": cpu_model, "cores_available": cpu_avail,
                    "sample": f"{extra['sample']} bounded by the FP64 oracle (LAPACK ADMM, "
                              f"MKL_NUM_THREADS=1) on {pool_size} worker processes, "
                              f"{w['seconds']:.1f} s",
                    "one_core": co["one_core"], "workers": w}
            if plain is not None:
                cs = plain["chosen_same_sample"]
                cpub["chosen_over_workers"] = cs["value"] / w["value"]
                cpub["rel_bound_diff_chosen_vs_oracle"] = float(np.max(
                    np.abs(np.array(cs["bounds"]) - np.array(w["bounds"])) /
                    np.abs(np.array(w["bounds"]))))
            pool.close()
        else:
            cval, sample, cbound = cpu(args.cpu_seconds)
            cpub = {"value": cval, "unit": "node bounds/s", "cores": 1, "kind": "port",
                    "cpu_model": cpu_model, "cores_available": cpu_avail,
                    "sample": sample + " (1 thread, MKL_NUM_THREADS=1)"}
            if cbound is not None:
                cpub["bound"] = float(cbound)
                cpub["rel_bound_diff_chosen_vs_oracle"] = abs(r["bound"] - cbound) / abs(cbound)

    line = {"metric": METRIC, "value": value, "unit": "node bounds/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32-mfma+f64", "data": "synthetic", "config": cfg, "roofline": roofline,
            "plain_fp64_gpu": plain, "cpu_baseline": cpub,
            "chosen": {"bound": r["bound"], "admm_iters": r["admm_iters"], **CHOSEN},
            "hip_graphs": graphs}
    if "admm_calls" in r:
        line["chosen"]["admm_calls"] = r["admm_calls"]
    if args.workload == "c5":
        line["c5"] = {"root_bound": r["bound"], "incumbent": r["incumbent"],
                      "gap": (r["bound"] - r["incumbent"]) / r["incumbent"],
                      "nodes_per_step": r["nodes"], "open_nodes": r["open"],
                      "ns_fallbacks": r["ns_fallbacks"], "ipm_iters": r["ipm_iters"],
                      "proven": r["open"] == 0}
        if not args.no_replay:
            line["bound_error_identical_inputs"] = extra["replay"]()
    if args.workload == "c1-bab" or (args.workload == "bab150" and r.get("open") == 0):
        line["time_to_optimum_s"] = ms_per_step / 1e3
        line["optimum"] = r["bound"]
    if not args.no_vendor and args.workload in ("c5", "c2-root") and roofline is not None:
        try:
            ve = vendor_fp64_eig(int(rp["n"]))
            ve["chosen_ms_per_projection"] = roofline["ms_per_projection"]
            ve["chosen_speedup_vs_vectors"] = ve["ms_with_vectors"] / roofline["ms_per_projection"]
            line["vendor_fp64_eigensolver"] = ve
            log("vendor FP64 eigensolver reference done")
        except (OSError, RuntimeError, AttributeError) as e:
            line["vendor_fp64_eigensolver"] = {"error": str(e)}
    if args.workload == "c4-learned" and plain is not None:
        line["admm_iterations"] = {"warm": r["admm_iters"], "cold": plain["admm_iters"],
                                   "saved_frac": 1.0 - r["admm_iters"] / plain["admm_iters"]}
        # wall clock per step: the warm step includes the pair features, the
        # network's inference and the warm-start construction
        line["wall_seconds"] = {"warm": ms_per_step / 1e3, "cold": plain["seconds"],
                                "saved_frac": 1.0 - (ms_per_step / 1e3) / plain["seconds"]}
    print(json.dumps(line, default=float), flush=True)


if __name__ == "__main__":
    main()
