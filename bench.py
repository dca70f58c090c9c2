"""Benchmark: Tet4 Neo-Hookean f_int + K_T global assembly on one MI355X.

Workload: jittered, node-permuted 100^3-hex Kuhn box (6.0M Tet4, 1.03M nodes,
3.09M DOF, 136.9M CSR nonzeros), per-element lognormal mu_e with
lambda_e = 5 mu_e, smooth seeded displacement with max |grad u| = 0.1
(psierra_amd.workloads.jittered_box_tet4, seed 1).  One step = one
psx_assemble call (f_int and every K_T value into the fixed CSR pattern) with
u, mu, lambda already resident in HBM.  ``value`` is elements/s of the FP32
element arithmetic / FP64 accumulation mode; the plain FP64 mode is timed
beside it.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

METRIC = "Tet4 NeoHookean f_int+K_T assembly elements/s & %HBM roofline vs plain FP64"
HBM_PEAK_GBPS = 8000.0
TRAFFIC_FILE = "tet4_assemble_rows_box100_traffic.txt"


def algorithmic_bytes(E, N, npe, nnz, coord_bytes, param_bytes):
    """Compulsory bytes of one assembly: read X and u once, connectivity
    (int32), two per-element parameters; write K values and f in FP64."""
    n = 3 * N
    return 2 * n * coord_bytes + E * npe * 4 + E * 2 * param_bytes + 8 * nnz + 8 * n


def cpu_baseline(wl, sample_elems, seed=0):
    """Time the oracle (the FP64 restatement of the reference) on a sample.

    faithful: per-element Python loop, element evaluated twice (force loop,
    tangent loop), np.add.at and coo.tocsr -- the reference's own structure
    (src/psierra/domain/solid.py:252,302; src/psierra/utility/assembly.py).
    vectorised: batched numpy over the same sample.  Single thread each.
    """
    from oracle import assembly as A

    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(wl.mesh.n_elements, sample_elems, replace=False))
    conn = wl.mesh.elements[idx]
    t0 = time.perf_counter()
    A.assemble_loop(wl.mesh.nodes, conn, wl.u, wl.mu[idx], wl.lam[idx])
    t_loop = time.perf_counter() - t0
    t0 = time.perf_counter()
    A.assemble(wl.mesh.nodes, conn, wl.u, wl.mu[idx], wl.lam[idx])
    t_vec = time.perf_counter() - t0
    return sample_elems / t_loop, sample_elems / t_vec, t_loop, t_vec


def host_info():
    """CPU model (/proc/cpuinfo), usable cores, numpy/scipy versions and the
    BLAS libraries with their thread counts."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    import scipy
    from threadpoolctl import threadpool_info

    blas = [{"api": d.get("internal_api"), "threads": d.get("num_threads")} for d in threadpool_info()
            if d.get("user_api") == "blas"]
    return {"cpu_model": model, "host_cores": len(os.sched_getaffinity(0)), "numpy": np.__version__,
            "scipy": scipy.__version__, "blas": blas}


_FORK = None   # (workload, sample indices, shared k, shared f) inherited by the forked baseline workers


def _vec_chunk(rng_):
    """Worker of the all-cores baseline: the oracle's vectorised element
    routine on sample elements [a, b), numpy/BLAS on one thread, written into
    the shared element arrays."""
    from threadpoolctl import threadpool_limits

    from oracle import fem

    wl, idx, k_all, f_all = _FORK
    a, b = rng_
    sel = idx[a:b]
    conn = wl.mesh.elements[sel]
    with threadpool_limits(1):
        k, f = fem.tet4_batch(wl.mesh.nodes[conn], wl.u[conn], wl.mu[sel], wl.lam[sel])
    k_all[a:b] = k
    f_all[a:b] = f
    return b - a


def cpu_vectorised_all_cores(wl, sample_elems, workers, seed=0):
    """The vectorised FP64 oracle on the same seeded sample with all cores:
    the element routine over `workers` forked processes (one core each, the
    workload inherited, the element results in shared anonymous memory),
    then the global scatter as the reference does it, np.add.at for f and
    coo_matrix(...).tocsr() for K, in the parent (scipy is single-threaded);
    returns (elements/s, seconds)."""
    import mmap
    import multiprocessing as mp

    from oracle import assembly as A

    global _FORK
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(wl.mesh.n_elements, sample_elems, replace=False))
    S = idx.size
    kbuf = mmap.mmap(-1, S * 144 * 8)
    fbuf = mmap.mmap(-1, S * 12 * 8)
    k_all = np.frombuffer(kbuf, dtype=np.float64).reshape(S, 12, 12)
    f_all = np.frombuffer(fbuf, dtype=np.float64).reshape(S, 12)
    bounds = np.linspace(0, S, workers + 1).astype(int)
    ranges = [(int(bounds[i]), int(bounds[i + 1])) for i in range(workers)]
    _FORK = (wl, idx, k_all, f_all)
    try:
        with mp.get_context("fork").Pool(workers) as pool:
            pool.map(_vec_chunk, [(a, min(a + 8, b)) for a, b in ranges])   # workers started and warm
            t0 = time.perf_counter()
            pool.map(_vec_chunk, ranges)
            with single_thread():
                dofs = A.compute_element_dofs(wl.mesh.elements[idx])
                rows, cols = A.create_coo_sparsity_indices(dofs)
                A.assemble_vector(f_all, dofs, 3 * wl.mesh.n_nodes)
                A.assemble_sparse_matrix(k_all, rows, cols, 3 * wl.mesh.n_nodes)
            t = time.perf_counter() - t0
    finally:
        _FORK = None
        del k_all, f_all
        kbuf.close()
        fbuf.close()
    return S / t, t


def single_thread():
    """numpy and its BLAS limited to one thread, so a CPU baseline's core
    count is 1."""
    from threadpoolctl import threadpool_limits

    return threadpool_limits(1)


def cpu_entry(value, unit, sample, **extra):
    """The ``cpu_baseline`` object of a bench line: the oracle (the FP64
    restatement of the reference, "port") timed on one host core."""
    out = {"value": value, "unit": unit, "cores": 1, "kind": "port", "sample": sample}
    out.update(host_info())
    out.update(extra)
    return out


def cpu_until(fn, items, min_s):
    """Run fn over items until at least min_s seconds have passed (one item at
    least); returns (work units done, seconds, items used)."""
    done, used = 0, 0
    t0 = time.perf_counter()
    for it in items:
        done += fn(it)
        used += 1
        if time.perf_counter() - t0 >= min_s:
            break
    return done, time.perf_counter() - t0, used


PROFILE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def profiled_traffic(name):
    """HBM bytes per dispatch recorded by the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes in profiles/<name> (``key = value`` lines, already
    scaled by the copy-kernel calibration); None when absent."""
    path = os.path.join(PROFILE_DIR, name)
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    return vals


def _event_time(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_tet10(args):
    """Tet10 40^3 box (384k elements, 531k nodes): f_int + K_T assembly."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd.engine import AssemblyPlan

    wl = W.jittered_box_tet10(n=args.n or 40, seed=2)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    f, k = plan.alloc_outputs(True)
    res = {}
    for prec in ("fp64", "fp32"):
        ms = _event_time(lambda: plan.assemble(u, mu, lam, f=f, kval=k, precision=prec, check_status=False),
                         args.steps, args.warmup)
        plan.raise_on_status(u)
        res[prec] = ms
    E, N, nnz = plan.n_elements, plan.n_nodes, plan.nnz
    b_alg = algorithmic_bytes(E, N, 10, nnz, coord_bytes=4, param_bytes=4)
    ach = b_alg / (res["fp32"] * 1e-3) / 1e9
    return {
        "metric": "Tet10 NeoHookean f_int+K_T assembly elements/s",
        "value": E / (res["fp32"] * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["fp32"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": wl.name, "elements": E, "nodes": N, "nnz": nnz, "fp64_plain_ms": res["fp64"],
                   "speedup_fp32_vs_fp64_plain": res["fp64"] / res["fp32"], "tile": plan.info["tile"]},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None},
        "cpu_baseline": None if args.no_cpu else tet10_cpu(wl),
    }


def tet10_cpu(wl, sample=3000):
    with single_thread():
        loop_eps, vec_eps, t_loop, t_vec = cpu_baseline(wl, sample)
    return cpu_entry(loop_eps, "elements/s",
                     f"{sample} seeded elements of the same Tet10 mesh; FP64 oracle per-element loop (4-point "
                     f"rule, element evaluated twice, add.at, coo.tocsr) took {t_loop:.2f}s; vectorised FP64 "
                     f"numpy on the same sample: {vec_eps:.4g} elements/s ({t_vec:.2f}s)",
                     vectorised_value=vec_eps)


def bench_fbar(args):
    """F-bar Tet4 on a jittered box (default 48^3 hexes, 663k tets): f_int and
    K_dev + K_vol on the distance-2 pattern, the whole device pipeline."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd.engine import AssemblyPlan, FbarAssembly

    n = args.n or 48
    wl = W.jittered_box_tet4(n=n, seed=1)
    t0 = time.perf_counter()
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    fb = FbarAssembly(plan)
    t_setup = time.perf_counter() - t0
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    bulk = float(np.mean(wl.lam + 2.0 / 3.0 * wl.mu))
    f = torch.empty(plan.n_dofs, dtype=torch.float64, device=dev)
    k = torch.empty(fb.nnz, dtype=torch.float64, device=dev)
    res = {}
    for prec in ("fp64", "fp32"):
        res[prec] = _event_time(
            lambda: fb.assemble(u, mu, lam, bulk, f=f, kval=k, precision=prec, check_status=False),
            args.steps, args.warmup)
    fb.assemble(u, mu, lam, bulk, f=f, kval=k, precision="fp32")   # raises on inversion
    E, N = plan.n_elements, plan.n_nodes
    b_alg = algorithmic_bytes(E, N, 4, fb.nnz, coord_bytes=4, param_bytes=4)
    ach = b_alg / (res["fp32"] * 1e-3) / 1e9
    return {
        "metric": "Tet4 F-bar NeoHookean f_int+K_T assembly elements/s",
        "value": E / (res["fp32"] * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["fp32"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": wl.name + "_fbar", "elements": E, "nodes": N, "nnz": fb.nnz,
                   "element_pattern_nnz": plan.nnz, "fp64_plain_ms": res["fp64"], "setup_s": t_setup},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None},
        "cpu_baseline": None if args.no_cpu else fbar_cpu(),
    }


def fbar_cpu(n=20):
    """Tet4Fbar's domain assembly is vectorised numpy in the reference
    (tet4_fbar.py:264-459); the oracle restates it the same way."""
    from oracle import fbar as FB
    from psierra_amd import workloads as W

    wl = W.jittered_box_tet4(n=n, seed=1)
    bulk = float(np.mean(wl.lam + 2.0 / 3.0 * wl.mu))
    m = wl.mesh
    with single_thread():
        t0 = time.perf_counter()
        FB.internal_forces(m.nodes, m.elements, wl.u, wl.mu, wl.lam)
        FB.tangent_stiffness(m.nodes, m.elements, wl.u, wl.mu, wl.lam, bulk)
        t = time.perf_counter() - t0
    return cpu_entry(m.n_elements / t, "elements/s",
                     f"the same generator at {n}^3 hexes ({m.n_elements} tets); FP64 oracle f_int and "
                     f"K_dev + K_vol (vectorised numpy, scipy sparse) took {t:.2f}s")


def bench_vjp(args):
    """Adjoint on the 48^3 Tet4 box (663k tets): g_u = K_T w and d/d(mu_e, lambda_e)."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd.engine import AssemblyPlan

    wl = W.vjp_box_tet4(n=48, seed=3)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    t = [torch.as_tensor(x, device=dev) for x in (wl.u.reshape(-1), wl.mu, wl.lam, wl.extra["w"])]
    res = {}
    for prec in ("fp64", "fp32"):
        res[prec] = _event_time(lambda: plan.vjp(*t, precision=prec, check_status=False), args.steps, args.warmup)
        plan.raise_on_status(t[0])
    E = plan.n_elements
    return {
        "metric": "Tet4 NeoHookean adjoint (K_T w, dmu, dlambda) elements/s",
        "value": E / (res["fp32"] * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["fp32"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": wl.name, "elements": E, "fp64_ms": res["fp64"]},
        "roofline": None, "cpu_baseline": None if args.no_cpu else vjp_cpu(wl),
    }


def vjp_cpu(wl, sample=20000):
    """The reference has no adjoint (its parameter gradients are finite
    differences); the CPU path to K_T w is assembling K_T, timed here."""
    with single_thread():
        loop_eps, vec_eps, t_loop, t_vec = cpu_baseline(wl, sample)
    return cpu_entry(loop_eps, "elements/s",
                     f"{sample} seeded elements of the same mesh; K_T by the FP64 oracle per-element loop "
                     f"(add.at, coo.tocsr) took {t_loop:.2f}s, vectorised {vec_eps:.4g} elements/s; the sparse "
                     f"product K_T w and the parameter gradients are not included",
                     vectorised_value=vec_eps)


def bench_ecsw(args):
    """128 ECSW instances (8..48 hexes per axis, rank 64, 2% RID) in one launch."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd.ecsw import BatchedECSW

    t0 = time.perf_counter()
    insts = W.ecsw_instances(n_instances=128, seed=4)
    t_build = time.perf_counter() - t0
    b = BatchedECSW(insts)
    ur = torch.as_tensor(np.stack([i["u_r"] for i in insts]), device=b.device)
    f_r, K_r = b.evaluate(ur)
    ms = _event_time(lambda: b.evaluate(ur, f_r, K_r, check_status=False), args.steps, args.warmup)
    n_rid = b.info["n_rid_elements"]
    r = 64
    flop_mfma = 2.0 * r * r * 12 * n_rid
    flop_total = flop_mfma + n_rid * (2.0 * 12 * 12 * r + 2 * 2.0 * 12 * r)
    tf = flop_mfma / (ms * 1e-3) / 1e12
    return {
        "metric": "ECSW batched reduced tangent sum_e w_e V_e^T K_e V_e, RID elements/s",
        "value": n_rid / (ms * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "ecsw_128_instances", "rid_elements": n_rid, "work_items": b.info["n_work_items"],
                   "total_gflop": flop_total / 1e9, "setup_s": t_build},
        "roofline": {"bound": "mfma", "achieved": tf, "peak": 157.3, "unit": "TFLOP/s", "frac": tf / 157.3,
                     "traffic": None},
        "cpu_baseline": None if args.no_cpu else ecsw_cpu(insts),
    }


def ecsw_cpu(insts, min_s=4.0):
    from oracle import ecsw as EC

    def one(ins):
        EC.evaluate_forces_and_stiffness(ins["mesh"].nodes, ins["mesh"].elements, ins["basis"], ins["rid"],
                                         ins["weights"], ins["mu"], ins["lam"], ins["u_r"])
        return len(ins["rid"])

    with single_thread():
        done, t, used = cpu_until(one, insts, min_s)
    return cpu_entry(done / t, "elements/s",
                     f"the first {used} of the same instances ({done} RID elements); FP64 oracle of "
                     f"evaluate_forces_and_stiffness (vectorised numpy over each RID) took {t:.2f}s")


def bench_ecsw_train(args):
    """ECSW training matrix (FP64) on the 48^3 Tet4 box: rank 64, 8 snapshots,
    with the linear-stiffness block (rows = 8*64 + 64*65/2 = 2592), and the
    sums of squares of both blocks."""
    import torch

    from psierra_amd import _lib
    from psierra_amd import workloads as W
    from psierra_amd.engine import AssemblyPlan, current_stream_handle

    wl = W.vjp_box_tet4(n=48, seed=3)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    r, m = 64, 8
    rng = np.random.default_rng(6)
    n = wl.mesh.n_dofs
    phi = torch.as_tensor(np.linalg.qr(rng.standard_normal((n, r)))[0], device=dev).contiguous()
    snaps = torch.as_tensor(1e-4 * rng.standard_normal((n, m)) + wl.u.reshape(-1, 1), device=dev).contiguous()
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    E = plan.n_elements
    rows = m * r + r * (r + 1) // 2
    A = torch.empty((rows, E), dtype=torch.float64, device=dev)
    nwork = int(_lib.lib().psx_ecsw_training_work_doubles(plan._h))
    work = torch.empty(nwork, dtype=torch.float64, device=dev)
    sumsq = torch.zeros(2, dtype=torch.float64, device=dev)
    status = torch.zeros(2, dtype=torch.int32, device=dev)

    def call():
        _lib.check(_lib.lib().psx_ecsw_training(plan._h, _lib.ptr(phi), r, _lib.ptr(snaps), m, _lib.ptr(mu),
                                                _lib.ptr(lam), 1, _lib.ptr(A), _lib.ptr(work), nwork,
                                                _lib.ptr(sumsq), current_stream_handle(), _lib.ptr(status)))

    ms = _event_time(call, args.steps, args.warmup)
    assert int(status[0]) == 0
    b_alg = 8.0 * rows * E + 8.0 * n * (r + m) + 16.0 * E + 4.0 * 4 * E + 8.0 * 3 * wl.mesh.n_nodes
    ach = b_alg / (ms * 1e-3) / 1e9
    return {
        "metric": "ECSW training matrix (Phi_e^T f_e per snapshot + Phi_e^T K0_e Phi_e), element-snapshots/s",
        "value": E * m / (ms * 1e-3), "unit": "element-snapshots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
        "config": {"workload": wl.name, "elements": E, "rank": r, "snapshots": m, "rows": rows,
                   "A_bytes": 8 * rows * E, "layout": "A rows written directly in runs of 16 elements; "
                   "sums of squares of both blocks from per-group partials"},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None},
        "cpu_baseline": None if args.no_cpu else ecsw_train_cpu(r, m),
    }


def ecsw_train_cpu(r, m, n=12):
    from oracle import ecsw as EC
    from psierra_amd import workloads as W

    wl = W.vjp_box_tet4(n=n, seed=3)
    rng = np.random.default_rng(6)
    nd = wl.mesh.n_dofs
    phi = np.linalg.qr(rng.standard_normal((nd, r)))[0]
    snaps = 1e-4 * rng.standard_normal((nd, m)) + wl.u.reshape(-1, 1)
    with single_thread():
        t0 = time.perf_counter()
        EC.training_matrix(wl.mesh.nodes, wl.mesh.elements, phi, snaps, wl.mu, wl.lam, include_stiffness=True)
        t = time.perf_counter() - t0
    E = wl.mesh.n_elements
    return cpu_entry(E * m / t, "element-snapshots/s",
                     f"the same recipe on a {n}^3 box ({E} tets, rank {r}, {m} snapshots, stiffness rows); "
                     f"FP64 oracle of ECSWTrainingMatrix.build took {t:.2f}s")


def _newton_problem(n):
    """Jittered n^3 Tet4 box, clamped at x = 0, a total load of -80 in z
    spread over the x = max face (about 17% tip deflection: 6 Newton
    iterations per load step in 2 steps)."""
    from psierra_amd import workloads as W

    wl = W.jittered_box_tet4(n=n, seed=9)
    X = wl.mesh.nodes
    clamp = np.nonzero(np.abs(X[:, 0] - X[:, 0].min()) < 1e-9)[0]
    tip = np.nonzero(np.abs(X[:, 0] - X[:, 0].max()) < 1e-9)[0]
    f_ext = np.zeros(3 * len(X))
    f_ext[3 * tip + 2] = -80.0 / len(tip)
    return wl, clamp, f_ext


def bench_newton(args):
    """Loaded Newton solve on the GPU path (the reference's
    NonlinearStaticAnalysis.solve, analysis/nonlinear_static.py:250-474):
    2 load steps, nr_tol 1e-8, HIP-graph assemblies, free-DOF CSC and
    residual on the GPU, the sparse solve on the host (scipy spsolve, SuperLU,
    as the reference).  A step is one whole solve from u = 0; the SuperLU
    share of it is reported beside the total."""
    import torch

    from psierra_amd.analysis import NonlinearStaticAnalysis
    from psierra_amd.boundary import DirichletConstraint, DirichletManager
    from psierra_amd.domain import SolidDomain
    from psierra_amd.element import GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField
    from psierra_amd.solver import DirectSolver

    n = args.n or 12
    wl, clamp, f_ext = _newton_problem(n)
    mgr = DirichletManager(wl.mesh.n_dofs)
    mgr.add_constraint(DirichletConstraint(clamp, (0, 1, 2), 0.0))
    mat = NeoHookeanField(wl.mu, wl.lam)
    dom = SolidDomain(wl.mesh, mat, GpuTet4TotalLagrangian(mat))

    class TimedSolver(DirectSolver):
        seconds, calls = 0.0, 0

        def solve(self, A, b):
            t0 = time.perf_counter()
            x = super().solve(A, b)
            TimedSolver.seconds += time.perf_counter() - t0
            TimedSolver.calls += 1
            return x

    ana = NonlinearStaticAnalysis(dom, linear_solver=TimedSolver())

    def one():
        dom.update_state(u=np.zeros(wl.mesh.n_dofs))
        res = ana.solve(mgr, f_ext=f_ext, n_steps=2, nr_iterations=15, nr_tol=1e-8, use_graphs=True)
        assert res.converged, "the benchmark solve did not converge"
        return res

    for _ in range(args.warmup):
        one()
    TimedSolver.seconds, TimedSolver.calls = 0.0, 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = one()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / max(args.steps, 1)
    slu = TimedSolver.seconds / max(args.steps, 1)
    tip_uz = float(res.displacement[2::3][np.abs(wl.mesh.nodes[:, 0] - wl.mesh.nodes[:, 0].max()) < 1e-9].mean())
    return {
        "metric": "loaded Newton solve (2 load steps) wall time, GPU assembly + host SuperLU",
        "value": t, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64", "data": "synthetic (seeded jittered Kuhn box, clamped at x = 0, face load)",
        "config": {"workload": f"newton_box_tet4_{n}^3", "elements": wl.mesh.n_elements,
                   "free_dofs": int(mgr.free_dofs.size), "iterations": res.iteration_history,
                   "superlu_s": slu, "superlu_calls_per_solve": TimedSolver.calls / max(args.steps, 1),
                   "rest_s": t - slu, "tip_uz": tip_uz},
        "roofline": None,
        "cpu_baseline": None if args.no_cpu else newton_cpu(n),
    }


def newton_cpu(n):
    """The oracle's Newton loop (vectorised FP64 numpy assembly, scipy CSC and
    spsolve) on the same problem, one core."""
    from oracle import newton as ON
    from psierra_amd.boundary import DirichletConstraint, DirichletManager

    wl, clamp, f_ext = _newton_problem(n)
    mgr = DirichletManager(wl.mesh.n_dofs)
    mgr.add_constraint(DirichletConstraint(clamp, (0, 1, 2), 0.0))
    with single_thread():
        t0 = time.perf_counter()
        _, ok, iters = ON.solve(wl.mesh.nodes, wl.mesh.elements, wl.mu, wl.lam, mgr.free_dofs,
                                mgr.prescribed_dofs, mgr.prescribed_values, f_ext, 2, 15, 1e-8)
        t = time.perf_counter() - t0
    assert ok
    return cpu_entry(t, "s", f"the same solve ({iters} iterations) by the FP64 oracle's Newton loop "
                             f"(vectorised numpy assembly, scipy free-DOF CSC, spsolve) took {t:.2f}s")


def bench_mass(args):
    """Consistent mass matrix (FP64, constant per mesh, used by the dynamics
    drivers) on the 100^3 Tet4 box, per-element density, into the CSR pattern."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd.engine import AssemblyPlan

    n = args.n or 100
    wl = W.jittered_box_tet4(n=n, seed=1)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    rho = torch.as_tensor(1.0e3 * np.exp(0.1 * np.random.default_rng(3).standard_normal(plan.n_elements)),
                          device=dev)
    kval = torch.empty(plan.nnz, dtype=torch.float64, device=dev)
    ms = _event_time(lambda: plan.assemble_mass(rho, kval), args.steps, args.warmup)
    E, N, nnz = plan.n_elements, plan.n_nodes, plan.nnz
    # read connectivity, coordinates and the density once; write every value once
    b_alg = 4.0 * 4 * E + 8.0 * 3 * N + 8.0 * E + 8.0 * nnz
    ach = b_alg / (ms * 1e-3) / 1e9
    return {
        "metric": "Consistent mass matrix assembly (FP64) elements/s",
        "value": E / (ms * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
        "config": {"workload": wl.name, "elements": E, "nodes": N, "nnz": nnz},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None},
        "cpu_baseline": None if args.no_cpu else mass_cpu(),
    }


def mass_cpu(n=24):
    """The oracle's consistent mass (per-element tensor times rho_e V_e,
    scattered with coo.tocsr) on the same generator at n^3 hexes."""
    from oracle import assembly as A
    from psierra_amd import workloads as W

    wl = W.jittered_box_tet4(n=n, seed=1)
    m = wl.mesh
    rho = 1.0e3 * np.exp(0.1 * np.random.default_rng(3).standard_normal(m.n_elements))
    with single_thread():
        t0 = time.perf_counter()
        A.mass_matrix(m.nodes, m.elements, rho)
        t = time.perf_counter() - t0
    return cpu_entry(m.n_elements / t, "elements/s",
                     f"the same generator at {n}^3 hexes ({m.n_elements} tets); FP64 oracle consistent mass "
                     f"(vectorised, coo.tocsr) took {t:.2f}s")


def clamped_free_dofs(mesh):
    """Free DOFs of a box clamped on its x = min face (all three components)."""
    from psierra_amd.boundary import select_nodes_by_plane

    nodes = select_nodes_by_plane(mesh, "yz", "min")
    mask = np.ones(mesh.n_dofs, dtype=bool)
    mask[(3 * nodes[:, None] + np.arange(3)).ravel()] = False
    return np.nonzero(mask)[0]


def bench_csc(args):
    """Free-DOF CSC extraction on the Tet4 box clamped at x = 0: K[free][:, free]
    with exact zeros dropped, from the FP32-mode K values, on the device."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd._lib import check, lib, ptr
    from psierra_amd.engine import AssemblyPlan, current_stream_handle
    from psierra_amd.solver import FreeDofSystem

    n = args.n or 100
    wl = W.jittered_box_tet4(n=n, seed=1)
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements)
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    f, k = plan.assemble(u, mu, lam, precision="fp32")
    free = clamped_free_dofs(wl.mesh)
    fs = FreeDofSystem(plan, free)

    def call():
        check(lib().psx_csc_extract(fs._h, ptr(k), 1, ptr(fs.colptr), ptr(fs.rowind), ptr(fs.vals),
                                    current_stream_handle()))

    ms = _event_time(call, args.steps, args.warmup)
    kept = int(fs.colptr[-1].item())
    # read the kept K values, write their values, row indices and the column pointers
    b_alg = 8.0 * kept + 8.0 * kept + 4.0 * kept + 4.0 * (fs.n_free + 1)
    ach = b_alg / (ms * 1e-3) / 1e9
    return {
        "metric": "Free-DOF CSC extraction K[free][:, free] (exact zeros dropped), nonzeros/s",
        "value": kept / (ms * 1e-3), "unit": "nonzeros/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
        "config": {"workload": wl.name + "_clamped_x0", "n_free": int(fs.n_free), "structural_nnz": int(fs.nnz_struct),
                   "kept_nnz": kept},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None},
        "cpu_baseline": None if args.no_cpu else csc_cpu(),
    }


def csc_cpu(n=48):
    """scipy's K_total = K_tan + K_contact; K_total[free][:, free].tocsc()
    (nonlinear_static.py:354-355) on the 48^3 box, K from the oracle's vectorised
    assembly."""
    import scipy.sparse as spsp

    from oracle import assembly as A
    from psierra_amd import workloads as W

    wl = W.jittered_box_tet4(n=n, seed=1)
    m = wl.mesh
    with single_thread():
        _, K = A.assemble(m.nodes, m.elements, wl.u, wl.mu, wl.lam)
        free = clamped_free_dofs(m)
        t0 = time.perf_counter()
        C = A.free_dof_csc(K, free)
        t = time.perf_counter() - t0
    return cpu_entry(C.nnz / t, "nonzeros/s",
                     f"the same generator at {n}^3 hexes ({K.nnz} nonzeros, {C.nnz} kept in K[free][:, free]); "
                     f"scipy CSR '+' and fancy slicing to CSC took {t:.2f}s")


def bench_batch(args):
    """Batched assembly: 128 Tet4 boxes of uneven size (sides 8..48 hexes per
    axis, 3k-663k tets, 25M in all), per-instance parameters, f_int + K_T of
    every instance by one launch, against a loop of one launch per mesh."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd.engine import AssemblyBatch, AssemblyPlan

    n_inst = args.n or 128
    t0 = time.perf_counter()
    wls = W.batch_boxes_tet4(n_instances=n_inst, seed=5)
    t_wl = time.perf_counter() - t0
    t0 = time.perf_counter()
    plans = [AssemblyPlan(w.mesh.nodes, w.mesh.elements) for w in wls]
    t_plan = time.perf_counter() - t0
    E = sum(p.n_elements for p in plans)
    print(f"[bench] {n_inst} instances, {E} elements (workload {t_wl:.1f}s, plans {t_plan:.1f}s)",
          file=sys.stderr, flush=True)
    dev = plans[0].device
    us = [torch.as_tensor(w.u.reshape(-1), device=dev) for w in wls]
    mus = [torch.as_tensor(w.mu, device=dev) for w in wls]
    lams = [torch.as_tensor(w.lam, device=dev) for w in wls]
    batch = AssemblyBatch(plans)
    fs, ks = batch.alloc_outputs(True)
    launch = batch.launcher(us, mus, lams, fs, ks)
    res = {}
    for prec in ("fp64", "fp32"):
        for mode in ("persistent", "tile_per_group"):
            batch.set_persistent(mode == "persistent")
            res[f"batch_{mode}_{prec}"] = _event_time(lambda: launch(prec), args.steps, args.warmup)
            batch.raise_on_status(us)
            print(f"[bench] batched {mode} {prec}: {res[f'batch_{mode}_{prec}']:.3f} ms", file=sys.stderr,
                  flush=True)

        def loop():
            for p, u, m, l, f, k in zip(plans, us, mus, lams, fs, ks):
                p.assemble(u, m, l, f=f, kval=k, precision=prec, check_status=False)

        res[f"loop_{prec}"] = _event_time(loop, args.steps, args.warmup)
        print(f"[bench] loop of {n_inst} launches {prec}: {res[f'loop_{prec}']:.3f} ms", file=sys.stderr,
              flush=True)
    best_mode = min(("persistent", "tile_per_group"), key=lambda m: res[f"batch_{m}_fp32"])
    batch.set_persistent(best_mode == "persistent")
    ms = res[f"batch_{best_mode}_fp32"]
    # the batched results are the single-plan results, bit for bit
    identical = True
    if not args.no_check:
        launch("fp32")
        for i in sorted({0, int(np.argmax([p.n_elements for p in plans])), n_inst - 1}):
            f1, k1 = plans[i].assemble(us[i], mus[i], lams[i], precision="fp32")
            identical = identical and bool(torch.equal(fs[i], f1)) and bool(torch.equal(ks[i], k1))
        if not identical:
            raise RuntimeError("batched assembly differs from the single-plan assembly")
    nnz = sum(p.nnz for p in plans)
    b_alg = sum(algorithmic_bytes(p.n_elements, p.n_nodes, 4, p.nnz, coord_bytes=4, param_bytes=4)
                for p in plans)
    ach = b_alg / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        big = wls[int(np.argmax([p.n_elements for p in plans]))]
        with single_thread():
            loop_eps, vec_eps, t_loop, t_vec = cpu_baseline(big, 20000)
        cpu = cpu_entry(loop_eps, "elements/s",
                        f"20000 seeded elements of the largest instance ({big.name}); FP64 oracle per-element "
                        f"loop (element evaluated twice, add.at, coo.tocsr) took {t_loop:.2f}s; vectorised FP64 "
                        f"numpy on the same sample: {vec_eps:.4g} elements/s ({t_vec:.2f}s)",
                        vectorised_value=vec_eps)
    return {
        "metric": "batched Tet4 NeoHookean f_int+K_T assembly, 128 uneven meshes in one launch, elements/s",
        "value": E / (ms * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded jittered Kuhn boxes, permuted node ids, one mu per instance, lam = 5 mu)",
        "config": {"workload": f"batch_boxes_tet4_{n_inst}x(8..48)^3", "instances": n_inst, "elements": E,
                   "nnz": nnz, "mode": "fp32 element arithmetic, fp64 scatter/f accumulation, per-call parameters",
                   "launch": best_mode, "ms": res, "speedup_vs_loop_fp32": res["loop_fp32"] / ms,
                   "speedup_vs_loop_fp64": res["loop_fp64"] / min(res["batch_persistent_fp64"],
                                                                    res["batch_tile_per_group_fp64"]),
                   "bit_identical_to_single_plans": identical if not args.no_check else None,
                   "setup_s": {"workload": t_wl, "plans": t_plan}},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": f"batch_permute_kernel + tet4_batch_rows<float,256,kNeoHookean> ({best_mode})"},
        "cpu_baseline": cpu,
    }


def bench_hooks(args):
    """The drop-in element hooks at the headline size, host arrays in and out
    (PCIe-inclusive): pSierra's SolidDomain calls assemble_internal_forces and
    then assemble_tangent_stiffness_matrix at each Newton iterate
    (src/psierra/analysis/nonlinear_static.py:296-297); each step here is that
    pair at a new displacement: u to the device, the f-only launch, f back,
    the fused f + K launch, f and the 1.1 GB of K values back into a scipy
    CSR matrix."""
    import torch

    from psierra_amd import workloads as W
    from psierra_amd.element import GpuTet4TotalLagrangian
    from psierra_amd.material import NeoHookeanField

    wl = W.jittered_box_tet4(n=args.n or 100, seed=1)
    el = GpuTet4TotalLagrangian(NeoHookeanField(wl.mu, wl.lam), precision="fp32")
    el.assemble_internal_forces(wl.mesh, wl.u)   # plan, material binding, tile choice
    times = []
    for step in range(args.warmup + args.steps):
        u = wl.u * (1.0 + 1e-3 * (step + 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = el.assemble_internal_forces(wl.mesh, u)
        K = el.assemble_tangent_stiffness_matrix(wl.mesh, u)
        t1 = time.perf_counter()
        if step >= args.warmup:
            times.append(t1 - t0)
    ms = 1e3 * float(np.mean(times))
    E = wl.mesh.n_elements
    return {
        "metric": "drop-in element hooks f_int then K_T (host in/out, PCIe-inclusive), elements/s",
        "value": E / (ms * 1e-3), "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded jittered Kuhn box, permuted node ids, lognormal mu_e)",
        "config": {"workload": wl.name, "elements": E, "nnz": int(K.nnz), "f_dofs": int(f.size),
                   "ms_min": 1e3 * min(times), "ms_max": 1e3 * max(times),
                   "host_bytes_per_step": int(2 * 8 * f.size + 8 * K.nnz + 8 * f.size)},
        "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=None,
                    help="hexes per box side (default: tet4 100 -> 6.0M tets, tet10 40, fbar 48); "
                         "batch: number of instances (default 128)")
    ap.add_argument("--cpu-sample", type=int, default=80000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--workload",
                    choices=("tet4", "tet10", "vjp", "ecsw", "ecsw_train", "fbar", "csc", "mass", "newton", "batch",
                             "hooks"),
                    default="tet4")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
            file=sys.stderr,
        )
        sys.exit(2)
    if args.workload != "tet4":
        fn = {"tet10": bench_tet10, "vjp": bench_vjp, "ecsw": bench_ecsw, "ecsw_train": bench_ecsw_train,
              "fbar": bench_fbar, "csc": bench_csc, "mass": bench_mass, "newton": bench_newton,
              "batch": bench_batch, "hooks": bench_hooks}[
            args.workload]
        print(json.dumps(fn(args)), flush=True)
        return

    from psierra_amd import workloads as W

    args.n = args.n or 100
    t0 = time.perf_counter()
    wl = W.jittered_box_tet4(n=args.n, seed=1)
    t_build = time.perf_counter() - t0
    print(f"[bench] workload {wl.name}: E={wl.mesh.n_elements} N={wl.mesh.n_nodes} ({t_build:.1f}s)",
          file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        with single_thread():
            loop_eps, vec_eps, t_loop, t_vec = cpu_baseline(wl, args.cpu_sample)
        # all cores of this process's CPU share (at most 16: the box's share of one GPU)
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
        par_eps, t_par = cpu_vectorised_all_cores(wl, args.cpu_sample, workers)
        cpu = {
            "value": loop_eps,
            "unit": "elements/s",
            "cores": 1,
            "kind": "port",
            "sample": (
                f"{args.cpu_sample} seeded elements of the same mesh; FP64 oracle per-element loop "
                f"(element evaluated twice, add.at, coo.tocsr) took {t_loop:.2f}s on one core; "
                f"vectorised FP64 numpy on the same sample: {vec_eps:.4g} elements/s ({t_vec:.2f}s) on one core, "
                f"{par_eps:.4g} elements/s ({t_par:.2f}s) with the element routine over {workers} processes "
                f"(one BLAS thread each) and the add.at / coo.tocsr scatter on one"
            ),
            "vectorised_value": vec_eps,
            "vectorised_all_cores": {"value": par_eps, "cores": workers, "seconds": t_par},
        }
        cpu.update(host_info())
        print(f"[bench] cpu baseline loop {loop_eps:.4g} el/s, vectorised {vec_eps:.4g} el/s, "
              f"{workers} cores {par_eps:.4g} el/s", file=sys.stderr, flush=True)

    import torch

    from psierra_amd.engine import AssemblyPlan

    torch.cuda.init()
    t0 = time.perf_counter()
    plan = AssemblyPlan(wl.mesh.nodes, wl.mesh.elements, tile=args.tile)
    t_plan = time.perf_counter() - t0
    dev = plan.device
    u = torch.as_tensor(wl.u.reshape(-1), device=dev)
    mu = torch.as_tensor(wl.mu, device=dev)
    lam = torch.as_tensor(wl.lam, device=dev)
    f, k = plan.alloc_outputs(True)
    E, N, nnz = plan.n_elements, plan.n_nodes, plan.nnz
    # the element's material is fixed, as in the reference's element (its
    # assemble_* hooks take no parameters): bound once, outside the timing
    t0 = time.perf_counter()
    plan.bind_material(mu, lam)
    torch.cuda.synchronize()
    t_bind = time.perf_counter() - t0
    print(f"[bench] plan {t_plan:.1f}s nnz={nnz} tiles={plan.info['n_tiles']} tile={plan.info['tile']}",
          file=sys.stderr, flush=True)

    def timed(precision, steps, warmup, bound=True):
        m, l = (None, None) if bound else (mu, lam)
        for _ in range(warmup):
            plan.assemble(u, m, l, f=f, kval=k, precision=precision, check_status=False)
        plan.raise_on_status(u)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in evs:
            a.record()
            plan.assemble(u, m, l, f=f, kval=k, precision=precision, check_status=False)
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        plan.raise_on_status(u)
        per = [a.elapsed_time(b) for a, b in evs]
        return wall / steps * 1e3, float(np.mean(per)), float(np.median(per))

    ms64_wall, ms64_ev, _ = timed("fp64", args.steps, args.warmup, bound=False)
    f64 = f.clone()
    k64 = k.clone()
    print(f"[bench] fp64 plain {ms64_ev:.3f} ms/assembly", file=sys.stderr, flush=True)
    _, ms64_bound_ev, _ = timed("fp64", args.steps, args.warmup)
    print(f"[bench] fp64 with the bound material {ms64_bound_ev:.3f} ms/assembly", file=sys.stderr, flush=True)
    _, ms32_call_ev, _ = timed("fp32", args.steps, args.warmup, bound=False)
    print(f"[bench] fp32 with per-call parameters {ms32_call_ev:.3f} ms/assembly", file=sys.stderr, flush=True)
    ms32_wall, ms32_ev, ms32_med = timed("fp32", args.steps, args.warmup)
    print(f"[bench] fp32/fp64-acc {ms32_ev:.3f} ms/assembly", file=sys.stderr, flush=True)

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    errors = {"fp32_vs_fp64_gpu": {"K": rel(k, k64), "f": rel(f, f64)}}
    if not args.no_check:
        from oracle import assembly as A

        indptr, indices = plan.pattern()
        rng = np.random.default_rng(2024)
        sample = rng.choice(N, 20, replace=False)
        f_ref, blocks = A.node_rows(wl.mesh.nodes, wl.mesh.elements, wl.u, wl.mu, wl.lam, sample)
        k32h, f32h, k64h, f64h = k.cpu().numpy(), f.cpu().numpy(), k64.cpu().numpy(), f64.cpu().numpy()
        num = {"K32": 0.0, "K64": 0.0, "f32": 0.0, "f64": 0.0}
        den = {"K": 0.0, "f": 0.0}
        for i, a in enumerate(sample):
            a = int(a)
            cols = indices[indptr[3 * a] : indptr[3 * a + 1]]
            ref = np.zeros((3, cols.size))
            for s, b in enumerate(cols[::3] // 3):
                ref[:, 3 * s : 3 * s + 3] = blocks[a][int(b)]
            g32 = np.stack([k32h[indptr[3 * a + r] : indptr[3 * a + r + 1]] for r in range(3)])
            g64 = np.stack([k64h[indptr[3 * a + r] : indptr[3 * a + r + 1]] for r in range(3)])
            num["K32"] += float(((g32 - ref) ** 2).sum())
            num["K64"] += float(((g64 - ref) ** 2).sum())
            den["K"] += float((ref**2).sum())
            num["f32"] += float(((f32h[3 * a : 3 * a + 3] - f_ref[i]) ** 2).sum())
            num["f64"] += float(((f64h[3 * a : 3 * a + 3] - f_ref[i]) ** 2).sum())
            den["f"] += float((f_ref[i] ** 2).sum())
        errors["vs_fp64_oracle_20_node_rows"] = {
            "fp32": {"K": (num["K32"] / den["K"]) ** 0.5, "f": (num["f32"] / den["f"]) ** 0.5},
            "fp64": {"K": (num["K64"] / den["K"]) ** 0.5, "f": (num["f64"] / den["f"]) ** 0.5},
        }

    # copy probe: achievable HBM bandwidth in this process
    src = torch.empty(1 << 28, dtype=torch.float64, device=dev)  # 2 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a_ev, b_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a_ev.record()
    for _ in range(5):
        dst.copy_(src)
    b_ev.record()
    torch.cuda.synchronize()
    copy_gbps = 5 * 2 * src.numel() * 8 / (a_ev.elapsed_time(b_ev) * 1e-3) / 1e9
    del src, dst

    traffic = profiled_traffic(TRAFFIC_FILE) if args.n == 100 else None
    b_alg32 = algorithmic_bytes(E, N, 4, nnz, coord_bytes=4, param_bytes=4)
    b_alg64 = algorithmic_bytes(E, N, 4, nnz, coord_bytes=8, param_bytes=8)
    achieved = b_alg32 / (ms32_ev * 1e-3) / 1e9
    value = E / (ms32_wall * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32_wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered Kuhn box, permuted node ids, lognormal mu_e)",
        "config": {
            "workload": wl.name,
            "elements": E,
            "nodes": N,
            "dofs": 3 * N,
            "nnz": nnz,
            "mode": "fp32 element arithmetic, fp64 scatter/f accumulation",
            "material": "per-element (mu, lam) bound to the plan once at setup (psx_plan_bind_material), "
                        "as the drop-in element hooks do; the fp64 plain form takes them per call",
            "fp32_per_call_params_ms": ms32_call_ev,
            "fp64_plain_ms": ms64_ev,
            "fp64_bound_material_ms": ms64_bound_ev,
            "fp64_plain_elements_per_s": E / (ms64_wall * 1e-3),
            "speedup_fp32_vs_fp64_plain": ms64_ev / ms32_ev,
            "speedup_fp32_vs_fp64_same_path": {"per_call_params": ms64_ev / ms32_call_ev,
                                               "bound_material": ms64_bound_ev / ms32_ev},
            "event_ms_mean": ms32_ev,
            "event_ms_median": ms32_med,
            "setup_s": {"workload": t_build, "plan": t_plan, "bind_material": t_bind},
            "tile": plan.info["tile"],
            "errors": errors,
            "copy_probe_GBps": copy_gbps,
            "B_alg_bytes": {"fp32_mode": b_alg32, "fp64_mode": b_alg64},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic.get("traffic_bytes_fp32") if traffic else None,
            "traffic_source": TRAFFIC_FILE if traffic else None,
            "kernel": f"psx_assemble = permute_u_kernel + tet4_assemble_rows<float,{plan.info['tile']},"
                      "kNeoHookeanBound>",
            "fp64_plain_frac": b_alg64 / (ms64_ev * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        },
        "roofline_fp64_plain": {
            "bound": "hbm",
            "achieved": b_alg64 / (ms64_ev * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": b_alg64 / (ms64_ev * 1e-3) / 1e9 / HBM_PEAK_GBPS,
            "traffic": traffic.get("traffic_bytes_fp64") if traffic else None,
            "traffic_source": TRAFFIC_FILE if traffic else None,
            "kernel": f"psx_assemble = permute_u_kernel + tet4_assemble_rows<double,{plan.info['tile']},"
                      "kNeoHookean>",
            "elements_per_s": E / (ms64_ev * 1e-3),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
