#!/usr/bin/env python3
"""Benchmark of the MI355X P1 matrix-free multigrid path.

Workload (one GPU): P1 constant-coefficient Laplace on the 24-tetrahedron unit cube
(meshes/cube_24el.msh) refined to level 8 (67,502,849 DoFs, 66,716,415 inner).
A "step" is one FP64 operator apply y = A x over all inner DoFs, inputs resident in HBM.
value = inner DoFs x steps / time (GDoF/s), timed with HIP events on the launch stream.

Also reported:
  roofline      algorithmic bytes of the apply kernel (16 B per inner DoF: read x, write y)
                over its average launch duration, against the 8 TB/s HBM peak; `traffic`
                from the committed rocprofv3 counter summary when present;
  solve         V(3,3) Chebyshev-smoothed V-cycles to relative residual 1e-10, plain FP64 and
                mixed precision (FP32 V-cycles inside FP64 defect correction), with the error of
                the mixed solution against the FP64 one;
  cpu_baseline  the CPU oracle (C++/OpenMP restatement of the reference's apply loops) timed on
                this host on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

# FP32 V-cycles per FP64 residual in the headline mixed-precision solve (--ir-inner); the other workloads
# take one, so their mixed solves run the same number of V-cycles as the FP64 ones
IR_INNER = 3

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0


def load_pkg():
    spec = importlib.util.spec_from_file_location("hyteg_vara_amd", os.path.join(ROOT, "hyteg-vara_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hyteg_vara_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class Hip:
    """Minimal HIP runtime access for streams and events."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def stream(self):
        s = C.c_void_p()
        self._ok(self.lib.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s

    def event(self):
        e = C.c_void_p()
        self._ok(self.lib.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e

    def record(self, e, s):
        self._ok(self.lib.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, a, b):
        self._ok(self.lib.hipEventSynchronize(b), "hipEventSynchronize")
        ms = C.c_float()
        self._ok(self.lib.hipEventElapsedTime(C.byref(ms), a, b), "hipEventElapsedTime")
        return float(ms.value)

    def sync(self, s):
        self._ok(self.lib.hipStreamSynchronize(s), "hipStreamSynchronize")


def timed(hip, stream, fn, reps):
    e0, e1 = hip.event(), hip.event()
    hip.sync(stream)
    hip.record(e0, stream)
    for _ in range(reps):
        fn()
    hip.record(e1, stream)
    return hip.elapsed_ms(e0, e1)


def profile_traffic():
    """HBM bytes per apply launch from the committed rocprofv3 counter summary, if present."""
    path = os.path.join(ROOT, "profiles", "apply_cube24_L8_f64_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            return json.load(fh).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(mesh_name, L, target_s, inner):
    """Times the CPU oracle's apply on this host (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol  # test infrastructure, used here only as the CPU baseline

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    m = ol.Mesh(mesh_name)
    n = m.ndofs(L)
    op = ol.Operator(m, L, L)
    s, d = ol.Function(m, L, L), ol.Function(m, L, L)
    s.set(L, np.random.default_rng(7).uniform(-1, 1, n))
    op.apply(s, d, L, ol.INNER)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        op.apply(s, d, L, ol.INNER)
        reps += 1
        el = time.perf_counter() - t0
        if el >= target_s or reps >= 200:
            break
    out = {"value": inner * reps / el / 1e9, "unit": "GDoF/s", "cores": threads, "kind": "port",
           "sample": f"{mesh_name} L{L} FP64 P1 constant apply (inner DoFs, incl. downward halo refresh), "
                     f"{reps} applies in {el:.1f} s on {threads} OpenMP threads"}
    # the same apply on one thread (a bounded sample of a few applies)
    try:
        gomp = C.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        reps1, t0 = 0, time.perf_counter()
        while True:
            op.apply(s, d, L, ol.INNER)
            reps1 += 1
            el1 = time.perf_counter() - t0
            if el1 >= target_s / 2 or reps1 >= 50:
                break
        gomp.omp_set_num_threads(threads)
        out["single_thread"] = {"value": inner * reps1 / el1 / 1e9, "unit": "GDoF/s", "cores": 1,
                                "sample": f"{reps1} applies in {el1:.1f} s"}
    except OSError as e:
        out["single_thread"] = {"value": None, "error": str(e)}
    out["host"] = host_info()
    out["config1_solve"] = config1_cpu(ol, threads)
    return out


def config1_problem(ndofs, bnd):
    """Right-hand side of the 6-tet Kuhn cube problem at level 5: f ~ U[-1, 1] per DoF (seed 2021), zero on the
    Dirichlet boundary."""
    b0 = np.random.default_rng(2021).uniform(-1.0, 1.0, ndofs)
    b0[bnd] = 0.0
    return b0


def config1_cpu(ol, threads):
    """The reference's CPU-runnable case timed with the oracle on this host: -Laplace u = f on the 6-tet Kuhn cube,
    level 5, zero Dirichlet, 10 V(2,2) cycles with weighted Jacobi (omega = 2/3), coarse level 0 by CG."""
    m, L = ol.Mesh("cube_kuhn_6el.msh"), 5
    _, bnd = m.coords(L)
    op = ol.Operator(m, 0, L)
    x, b = ol.Function(m, 0, L), ol.Function(m, 0, L)
    b.set(L, config1_problem(len(bnd), bnd))
    t0 = time.perf_counter()
    res = ol.gmg(op, x, b, 0, L, 10, pre=2, post=2, smoother="jacobi", relax=2.0 / 3.0)
    el = time.perf_counter() - t0
    return {"seconds": round(el, 4), "cycles": 10, "rel_residual": float(res[-1] / res[0]), "threads": threads}


def bench_config1(hyg, hip, st):
    """The same 10 V(2,2) Jacobi cycles on the 6-tet Kuhn cube at level 5 on the GPU (eager, residual norm after every
    cycle)."""
    gm, L = hyg.Mesh("cube_kuhn_6el.msh"), 5
    _, bnd = gm.coords(L)
    b0 = config1_problem(len(bnd), bnd)
    op = hyg.P1ConstantOperator(gm, 0, L)
    g = hyg.GMG(op, 0, L, pre=2, post=2, smoother="jacobi", relax=2.0 / 3.0)
    x, b = hyg.Function(gm, 0, L), hyg.Function(gm, 0, L)
    b.upload(L, b0)
    g.solve(x, b, L, 10)  # warm-up
    x.upload(L, np.zeros(len(bnd)))
    hip.sync(st)
    t0 = time.perf_counter()
    res = g.solve(x, b, L, 10)
    el = time.perf_counter() - t0
    return {"workload": "-Laplace u = f, cube_kuhn_6el.msh level 5 (35,937 DoFs), zero Dirichlet, f ~ U[-1,1] (seed 2021), "
                        "10 V(2,2) cycles, weighted Jacobi 2/3, coarse level 0 CG",
            "seconds": round(el, 4), "cycles": 10, "rel_residual": float(res[-1] / res[0])}


def host_info():
    """CPU model, NUMA layout and the oracle's compiler flags of the host the CPU baseline ran on."""
    info = {"cpu_model": None, "logical_cpus": os.cpu_count(), "numa_nodes": None,
            "compiler_flags": "g++ -O3 -march=x86-64-v3 -std=c++17 -fopenmp (oracle/Makefile)"}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    base = "/sys/devices/system/node"
    try:
        nodes = sorted(d for d in os.listdir(base) if d.startswith("node") and d[4:].isdigit())
        info["numa_nodes"] = {}
        for nd in nodes:
            with open(os.path.join(base, nd, "cpulist")) as fh:
                info["numa_nodes"][nd] = fh.read().strip()
    except OSError:
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


def graph_solves(hyg, hip, st, op, mesh, L, bv, kw, every, max_cycles=60):
    """Time to 1e-10 of one problem, FP64 against mixed precision, both as captured hipGraphs with the same
    residual-check cadence: FP64 replays `every` captured V-cycles between FP64 residual norms; the mixed solve
    replays one captured defect-correction step (FP64 residual, `every` FP32 V-cycles, FP64 update) per norm.
    Both go through single-problem batches (hyg_batch_*), whose captured cycle and defect-correction step are the
    whole operation sequence of hyg_gmg_vcycle / hyg_ir_solve."""
    g64 = hyg.GMG(op, 0, L, hyg.F64, **kw)
    g32 = hyg.GMG(op, 0, L, hyg.F32, **kw)
    xf, bf = hyg.Function(mesh, 0, L), hyg.Function(mesh, 0, L)
    x32, b32 = hyg.Function(mesh, 0, L, hyg.F32), hyg.Function(mesh, 0, L, hyg.F32)
    xt, bt = hyg.Function(mesh, L, L), hyg.Function(mesh, L, L)
    bf.upload(L, bv)
    bt.upload(L, bv)
    b64 = hyg.Batch([g64], [xf], [bf])
    b32b = hyg.Batch([g32], [x32], [b32])

    def fp64_solve():
        r0 = b64.residual_norms(st)[0]
        hist, c = [r0], 0
        while hist[-1] > 1e-10 * r0 and c < max_cycles:
            for _ in range(every):
                b64.vcycle(True, st)
            c += every
            hist.append(b64.residual_norms(st)[0])
        return hist, c

    # warm-up: captures both graphs
    fp64_solve_once = lambda: (b64.vcycle(True, st), b64.residual_norms(st))
    fp64_solve_once()
    b32b.ir_solve([xt], [bt], inner=every, max_outer=1, rel_tol=1e-10, use_graph=True, stream=st)
    xf.interpolate(0.0, L)
    xt.interpolate(0.0, L)
    hip.sync(st)
    t0 = time.perf_counter()
    h64, c64 = fp64_solve()
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    hmx = b32b.ir_solve([xt], [bt], inner=every, max_outer=max_cycles // every + 1, rel_tol=1e-10, use_graph=True, stream=st)[0]
    tmx = time.perf_counter() - t0
    u64, umx = xf.download(L), xt.download(L)
    return {"captured": "whole V-cycle (FP64) / whole defect-correction step (mixed) replayed as one hipGraph",
            "residual_check_every_cycles": every,
            "fp64": {"seconds": round(t64, 5), "cycles": c64, "rel_residual": float(h64[-1] / h64[0])},
            "mixed": {"seconds": round(tmx, 5), "outer_steps": len(hmx) - 1, "inner_cycles": every,
                      "rel_residual": float(hmx[-1] / hmx[0]),
                      "rel_error_vs_fp64": float(np.abs(umx - u64).max() / np.abs(u64).max())},
            "speedup_mixed_vs_fp64": round(t64 / tmx, 3)}


def bench_divkgrad(hyg, hip, st, steps, level, seed):
    """Variable-coefficient workload: -div(k grad u) on the 24-tet cube, k = exp(g), g a sum of 32
    seeded Fourier modes of wavenumber <= 8, k evaluated at the 5 quadrature points of every element in
    every apply. Apply throughput in FP64 and FP32; V(3,3) Chebyshev cycles to 1e-10 in FP64 and as FP32
    cycles inside FP64 defect correction. g has standard deviation 0.5 (k within about e^-1.5 .. e^1.5):
    with standard deviation 1 the multigrid of rediscretised coarse operators diverges on this mesh, in
    the oracle as on the device."""
    modes = hyg.fourier_field(seed, modes=32, kmax=8, std=0.5)
    mesh = hyg.Mesh("cube_24el.msh")
    L = level
    n = mesh.ndofs(L)
    bnd = mesh.boundary(L)
    inner = int((~bnd).sum())
    lib = hyg.load()
    op = hyg.P1DivKGradOperator(mesh, 0, L, 1.0, modes)
    out = {"workload": f"P1 -div(k grad u), cube_24el.msh level {L}, k = exp(32 Fourier modes, |kappa| <= 8, std 0.5, seed {seed})",
           "dofs_inner": inner}
    rng = np.random.default_rng(seed)
    for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
        x, y = hyg.Function(mesh, L, L, dt), hyg.Function(mesh, L, L, dt)
        x.upload(L, rng.uniform(-1, 1, n))

        def ap():
            hyg.check(lib.hyg_apply(op.h, x.h, y.h, L, hyg.INNER, hyg.REPLACE, st))

        for _ in range(2):
            ap()
        ms = timed(hip, st, ap, steps) / steps
        out[f"apply_{name}"] = {"ms": round(ms, 4), "GDoF/s": round(inner / (ms / 1e3) / 1e9, 3)}
    # coefficient evaluation: per lattice anchor 32 modes x (3 FMA phase + 30 x 2 FMA) + 30 exp + 6 x 5 FMA
    anchors = int(mesh.info()[3]) * ((1 << L) * ((1 << L) + 1) * ((1 << L) + 2) // 6)
    out["kbar_flops_per_apply"] = anchors * (32 * (6 + 120) + 60)
    # the stated field: 32 modes of amplitude 1.0 each (g has standard deviation 4); the apply does the same work
    # whatever the amplitude, only the solve depends on it (it diverges for amplitude 1: see the solve below, std 0.5)
    op1 = hyg.P1DivKGradOperator(mesh, 0, L, 1.0, hyg.fourier_field(seed, modes=32, kmax=8, std=4.0))
    for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
        x, y = hyg.Function(mesh, L, L, dt), hyg.Function(mesh, L, L, dt)
        x.upload(L, rng.uniform(-1, 1, n))

        def ap1():
            hyg.check(lib.hyg_apply(op1.h, x.h, y.h, L, hyg.INNER, hyg.REPLACE, st))

        for _ in range(2):
            ap1()
        ms = timed(hip, st, ap1, steps) / steps
        out[f"apply_amplitude1_{name}"] = {"ms": round(ms, 4), "GDoF/s": round(inner / (ms / 1e3) / 1e9, 3),
                                          "field": "32 modes, amplitude 1.0 each (std of g = 4), seed %d" % seed}
    del op1
    # a second operator with the coefficient integrals kept (computeAndStoreLocalElementMatrices)
    ops = hyg.P1DivKGradOperator(mesh, 0, L, 1.0, modes)
    ops.store_local_matrices()
    for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
        x, y = hyg.Function(mesh, L, L, dt), hyg.Function(mesh, L, L, dt)
        x.upload(L, rng.uniform(-1, 1, n))

        def aps():
            hyg.check(lib.hyg_apply(ops.h, x.h, y.h, L, hyg.INNER, hyg.REPLACE, st))

        for _ in range(2):
            aps()
        ms = timed(hip, st, aps, steps) / steps
        out[f"apply_stored_{name}"] = {"ms": round(ms, 4), "GDoF/s": round(inner / (ms / 1e3) / 1e9, 3)}
    # surrogate operator (P1SurrogateOperator::interpolateStencils): degree-4 polynomial stencils of the cell
    # interiors fitted on level min(L, 5), on a smooth field (3 modes, |kappa| <= 1, std 0.2) that such polynomials
    # approximate; the interior apply reads x and writes y only (16 B per DoF), face/edge/vertex points use the stored
    # integrals. Error: relative l-inf against the variable operator's apply of the same vector.
    smooth = hyg.fourier_field(seed, modes=3, kmax=1, std=0.2)
    opv = hyg.P1DivKGradOperator(mesh, L, L, 1.0, smooth)
    opv.store_local_matrices()
    opq = hyg.P1DivKGradOperator(mesh, L, L, 1.0, smooth)
    t0 = time.perf_counter()
    opq.interpolate_stencils(4, min(L, 5))
    sur = {"degree": 4, "interpolation_level": min(L, 5), "field": f"3 Fourier modes, |kappa| <= 1, std 0.2, seed {seed}",
           "setup_seconds": round(time.perf_counter() - t0, 3)}
    xv = rng.uniform(-1, 1, n)
    for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
        x, y = hyg.Function(mesh, L, L, dt), hyg.Function(mesh, L, L, dt)
        x.upload(L, xv)

        def apq():
            hyg.check(lib.hyg_apply(opq.h, x.h, y.h, L, hyg.INNER, hyg.REPLACE, st))

        for _ in range(2):
            apq()
        ms = timed(hip, st, apq, steps) / steps
        sur[f"apply_{name}"] = {"ms": round(ms, 4), "GDoF/s": round(inner / (ms / 1e3) / 1e9, 3),
                                "GB/s_alg": round((16.0 if dt == hyg.F64 else 8.0) * inner / (ms / 1e3) / 1e9, 1)}
        if dt == hyg.F64:
            yq = y.download(L)
            yv = hyg.Function(mesh, L, L)
            opv.apply(x, yv, L, hyg.INNER)
            yr = yv.download(L)
            sur["rel_error_vs_variable"] = float(np.abs(yq - yr).max() / np.abs(yr).max())
    out["surrogate"] = sur
    del opv, opq
    # VALU use. Executed flops: kKBar as counted above; kElemApply per point of a macro-cell 24 x 4 FMA to
    # assemble the stencil from the incident elements' integrals + 15 FMA to apply it (222 flops). The kKBar time
    # is the on-the-fly apply minus the stored one (the rest is the same launches). The reference's work per apply
    # (p1_div_k_grad_affine_q3.hpp:85-95: integrateAll 309 adds + 463 muls + 2 divs per micro-element, plus a 4 x 4
    # gather-multiply-scatter, 32 flops) is reported as an equivalent rate.
    pts = int(mesh.info()[3]) * ((1 << L) + 1) * ((1 << L) + 2) * ((1 << L) + 3) // 6
    micro = int(mesh.info()[3]) * (1 << L) ** 3
    valu = {}
    for name, peak in (("f64", 78.6e12), ("f32", 157.3e12)):
        t_on, t_st = out[f"apply_{name}"]["ms"] / 1e3, out[f"apply_stored_{name}"]["ms"] / 1e3
        t_kbar = max(t_on - t_st, 1e-9)
        valu[name] = {"kKBar_tflops": round(out["kbar_flops_per_apply"] / t_kbar / 1e12, 2),
                      "kKBar_valu_frac": round(out["kbar_flops_per_apply"] / t_kbar / peak, 4),
                      "stored_apply_tflops": round(222.0 * pts / t_st / 1e12, 2),
                      "stored_apply_valu_frac": round(222.0 * pts / t_st / peak, 4),
                      "reference_equivalent_gflops_on_the_fly": round((774 + 32) * micro / t_on / 1e9, 1)}
    out["valu"] = valu
    out["valu_note"] = ("peaks: FP64 78.6, FP32 157.3 TFLOP/s (dense vector); executed flops of this build's kernels; "
                        "reference_equivalent = the reference's per-element flop count x micro-elements / time")
    b = hyg.Function(mesh, 0, L)
    bv = rng.uniform(-1.0, 1.0, n)
    bv[bnd] = 0.0
    b.upload(L, bv)
    Lr = min(L, 5)
    pv, pt = hyg.Function(mesh, Lr, Lr), hyg.Function(mesh, Lr, Lr)
    pv.upload(Lr, np.random.default_rng(5).uniform(-1, 1, mesh.ndofs(Lr)))
    rho = op.power_iteration(pv, pt, 40, Lr)
    kw = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * rho, 1.2 * rho))
    g64 = hyg.GMG(op, 0, L, hyg.F64, **kw)
    g32 = hyg.GMG(op, 0, L, hyg.F32, **kw)
    x64, xmx = hyg.Function(mesh, 0, L), hyg.Function(mesh, 0, L)
    g64.solve_tol(x64, b, L, 1, 1e-10, st)
    hyg.ir_solve(op, g32, xmx, b, L, inner=1, max_outer=1, rel_tol=1e-10, stream=st)
    for f in (x64, xmx):
        f.interpolate(0.0, L)
    hip.sync(st)
    t0 = time.perf_counter()
    r64 = g64.solve_tol(x64, b, L, 100, 1e-10, st)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    rmx = hyg.ir_solve(op, g32, xmx, b, L, inner=1, max_outer=100, rel_tol=1e-10, stream=st)
    tmx = time.perf_counter() - t0
    u64, umx = x64.download(L), xmx.download(L)
    out["solve"] = {"cycle": f"V(3,3) Chebyshev order 2, rho(D^-1 A) = {rho:.5f}",
                    "field": "std 0.5 (the amplitude-1 field diverges with rediscretised coarse operators)",
                    "fp64": {"seconds": round(t64, 4), "cycles": len(r64) - 1, "rel_residual": float(r64[-1] / r64[0])},
                    "mixed": {"seconds": round(tmx, 4), "outer_steps": len(rmx) - 1, "inner_cycles": 1, "rel_residual": float(rmx[-1] / rmx[0]),
                              "rel_error_vs_fp64": float(np.abs(umx - u64).max() / np.abs(u64).max())},
                    "speedup_mixed_vs_fp64": round(t64 / tmx, 3)}
    # the same solves with the coefficient integrals kept
    g64s = hyg.GMG(ops, 0, L, hyg.F64, **kw)
    g32s = hyg.GMG(ops, 0, L, hyg.F32, **kw)
    for f in (x64, xmx):
        f.interpolate(0.0, L)
    hip.sync(st)
    t0 = time.perf_counter()
    r64 = g64s.solve_tol(x64, b, L, 100, 1e-10, st)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    rmx = hyg.ir_solve(ops, g32s, xmx, b, L, inner=1, max_outer=100, rel_tol=1e-10, stream=st)
    tmx = time.perf_counter() - t0
    out["solve_stored"] = {"fp64": {"seconds": round(t64, 4), "cycles": len(r64) - 1, "rel_residual": float(r64[-1] / r64[0])},
                           "mixed": {"seconds": round(tmx, 4), "outer_steps": len(rmx) - 1, "inner_cycles": 1,
                                     "rel_residual": float(rmx[-1] / rmx[0])},
                           "speedup_mixed_vs_fp64": round(t64 / tmx, 3)}
    return out


def bench_batched(hyg, hip, st, steps, nprob, seed):
    """Independent problems solved together (hyg_batch_*): nprob problems drawn from one seed, each on the 6-tet
    (Kuhn) or the 24-tet cube at a level drawn from {4, 5, 6, 7}, with the constant Laplacian or -div(k grad u), k
    the 32-mode log-normal field of the variable-coefficient workload (its own seed per problem, local matrices
    stored). Every grid operation of all problems is one persistent launch. Reported: the batched FP64 apply
    (aggregate GDoF/s, roofline), the captured batched FP32 V-cycle and the host time of one graph launch, and the
    time to 1e-10 of every problem: batched FP64 V(3,3) Chebyshev cycles against batched mixed precision (FP32
    cycles inside FP64 defect correction), both captured, the residual checked after every 3 cycles in both; and
    the same mixed solves run one problem after another without batching."""
    rng = np.random.default_rng(seed)
    meshes = {name: hyg.Mesh(name) for name in ("cube_kuhn_6el.msh", "cube_24el.msh")}
    every = 3
    probs = []
    for i in range(nprob):
        L = int(rng.integers(4, 8))
        name = ("cube_kuhn_6el.msh", "cube_24el.msh")[int(rng.integers(0, 2))]
        field = bool(rng.integers(0, 2))
        mesh = meshes[name]
        if field:
            op = hyg.P1DivKGradOperator(mesh, 0, L, 1.0, hyg.fourier_field(seed * 1000 + i, modes=32, kmax=8, std=0.5))
            op.store_local_matrices()
        else:
            op = hyg.P1ConstantOperator(mesh, 0, L)
        Lr = min(L, 5)
        pv, pt = hyg.Function(mesh, Lr, Lr), hyg.Function(mesh, Lr, Lr)
        pv.upload(Lr, np.random.default_rng(5).uniform(-1, 1, mesh.ndofs(Lr)))
        rho = op.power_iteration(pv, pt, 40, Lr)
        kw = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * rho, 1.2 * rho))
        n = mesh.ndofs(L)
        bnd = mesh.boundary(L)
        bv = rng.uniform(-1.0, 1.0, n)
        bv[bnd] = 0.0
        p = dict(L=L, mesh=name, field=field, op=op, inner=int((~bnd).sum()), bv=bv,
                 g64=hyg.GMG(op, 0, L, hyg.F64, **kw), g32=hyg.GMG(op, 0, L, hyg.F32, **kw),
                 xf=hyg.Function(mesh, 0, L), bf=hyg.Function(mesh, 0, L),
                 x32=hyg.Function(mesh, 0, L, hyg.F32), b32=hyg.Function(mesh, 0, L, hyg.F32),
                 xt=hyg.Function(mesh, L, L), bt=hyg.Function(mesh, L, L),
                 src=hyg.Function(mesh, L, L), dst=hyg.Function(mesh, L, L))
        p["bf"].upload(L, bv)
        p["bt"].upload(L, bv)
        p["src"].upload(L, rng.uniform(-1.0, 1.0, n))
        probs.append(p)
    inner = sum(p["inner"] for p in probs)
    # tiles: the cell-kernel configuration of every (mesh, level, value type, mode), tuned on the mesh; each
    # batched sweep takes the one of its largest problem
    tuned = {}
    for name, mesh in meshes.items():
        opt = hyg.P1ConstantOperator(mesh, 4, 7)
        for lev in range(4, 8):
            for dt, dn in ((hyg.F64, "f64"), (hyg.F32, "f32")):
                tuned[f"{name[:-4]}_L{lev}_{dn}"] = opt.tune(lev, dt, reps=3, stream=st)
    b64 = hyg.Batch([p["g64"] for p in probs], [p["xf"] for p in probs], [p["bf"] for p in probs])
    b32 = hyg.Batch([p["g32"] for p in probs], [p["x32"] for p in probs], [p["b32"] for p in probs])
    out = {"workload": f"{nprob} independent P1 problems (seed {seed}): levels from {{4,5,6,7}}, 6-tet Kuhn or 24-tet cube, "
                       "constant or 32-mode log-normal k (std 0.5, stored local matrices)",
           "problems": {"levels": [p["L"] for p in probs], "cube24": sum(p["mesh"] == "cube_24el.msh" for p in probs),
                        "variable_k": sum(p["field"] for p in probs)},
           "dofs_inner_total": inner, "cell_config": "per batched sweep, tuned for its largest problem", "tuned": tuned}
    # batched FP64 apply: the constant-coefficient problems' sweeps in one launch, the variable-coefficient ones'
    # elementwise applies in another, then one combine launch. Algorithmic bytes: 16 per inner DoF (read x, write y),
    # plus for the variable coefficient the stored integrals of the six micro-cells per lattice point (48 B).
    def apply_block(sel, bytes_per_dof):
        ps = [p for p in probs if sel(p)]
        if not ps:
            return None
        bt = b64 if len(ps) == len(probs) else hyg.Batch([p["g64"] for p in ps], [p["xf"] for p in ps], [p["bf"] for p in ps])
        srcs, dsts = [p["src"] for p in ps], [p["dst"] for p in ps]
        bt.apply(srcs, dsts, hyg.INNER, st)
        ms = timed(hip, st, lambda: bt.apply(srcs, dsts, hyg.INNER, st), steps) / steps
        dofs = sum(p["inner"] for p in ps)
        byts = sum(p["inner"] * bytes_per_dof(p) for p in ps)
        gbs = byts / (ms / 1e3) / 1e9
        return {"problems": len(ps), "dofs_inner": dofs, "ms": round(ms, 4), "GDoF/s": round(dofs / (ms / 1e3) / 1e9, 3),
                "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes": byts}}

    per_dof = lambda p: 64.0 if p["field"] else 16.0
    out["apply_f64"] = apply_block(lambda p: True, per_dof)
    out["apply_f64_constant_only"] = apply_block(lambda p: not p["field"], per_dof)
    out["apply_f64_variable_only"] = apply_block(lambda p: p["field"], per_dof)
    # captured batched FP32 cycle: device time per replay and host time to launch the graph
    ms_graph = timed(hip, st, lambda: b32.vcycle(True, st), 5) / 5
    ms_eager = timed(hip, st, lambda: b32.vcycle(False, st), 5) / 5
    hip.sync(st)
    t0 = time.perf_counter()
    b32.vcycle(True, st)
    launch_us = (time.perf_counter() - t0) * 1e6
    hip.sync(st)
    out["vcycle_f32"] = {"graph_ms": round(ms_graph, 3), "eager_batched_ms": round(ms_eager, 3),
                         "graph_launch_host_us": round(launch_us, 1)}
    # time to 1e-10 of every problem
    for p in probs:
        p["xf"].interpolate(0.0, p["L"])
    b64.vcycle(True, st)
    b64.residual_norms(st)
    b32.ir_solve([p["xt"] for p in probs], [p["bt"] for p in probs], inner=every, max_outer=1, rel_tol=1e-10,
                 use_graph=True, stream=st)
    for p in probs:
        p["xf"].interpolate(0.0, p["L"])
        p["xt"].interpolate(0.0, p["L"])
    hip.sync(st)
    t0 = time.perf_counter()
    r0 = b64.residual_norms(st)
    r, c64 = r0, 0
    while np.any(r > 1e-10 * r0) and c64 < 60:
        for _ in range(every):
            b64.vcycle(True, st)
        c64 += every
        r = b64.residual_norms(st)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    hmx = b32.ir_solve([p["xt"] for p in probs], [p["bt"] for p in probs], inner=every, max_outer=20, rel_tol=1e-10,
                       use_graph=True, stream=st)
    tmx = time.perf_counter() - t0
    err = max(float(np.abs(p["xt"].download(p["L"]) - p["xf"].download(p["L"])).max() /
                    np.abs(p["xf"].download(p["L"])).max()) for p in probs)
    # the same mixed solves one problem at a time (hyg_ir_solve, eager launches); a first call with no outer step
    # creates each solver's defect-correction work functions outside the timed region
    for p in probs:
        p["xt"].interpolate(0.0, p["L"])
        hyg.ir_solve(p["op"], p["g32"], p["xt"], p["bt"], p["L"], inner=every, max_outer=0, rel_tol=1e-10, stream=st)
    hip.sync(st)
    t0 = time.perf_counter()
    for p in probs:
        hyg.ir_solve(p["op"], p["g32"], p["xt"], p["bt"], p["L"], inner=every, max_outer=20, rel_tol=1e-10, stream=st)
    tseq = time.perf_counter() - t0
    out["solve"] = {"cycle": "V(3,3) Chebyshev order 2 per problem (bounds 0.3/1.2 x its power-iteration rho), coarse level 0",
                    "residual_check_every_cycles": every,
                    "fp64_batched_graph": {"seconds": round(t64, 4), "cycles": c64, "max_rel_residual": float(np.max(r / r0))},
                    "mixed_batched_graph": {"seconds": round(tmx, 4), "outer_steps": hmx.shape[1] - 1, "inner_cycles": every,
                                            "max_rel_residual": float(np.max(hmx[:, -1] / hmx[:, 0])),
                                            "max_rel_error_vs_fp64": err},
                    "mixed_one_problem_at_a_time_eager": {"seconds": round(tseq, 4)},
                    "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
                    "speedup_batched_vs_one_at_a_time": round(tseq / tmx, 3)}
    return out


def bench_p2(hyg, hip, st, steps, level, seed):
    """P2 constant-coefficient -Laplace on the 24-tet cube (P2 level 7 has the DoFs of P1 level 8): apply
    throughput in FP64 and FP32 (merged per-class stencils: 65 terms at vertex DoFs, 19-27 at edge DoFs), and
    V(3,3) Chebyshev cycles to 1e-10 in FP64 and as FP32 cycles inside FP64 defect correction."""
    mesh = hyg.Mesh("cube_24el.msh")
    L = level
    n = mesh.ndofs(L + 1)
    bnd = mesh.boundary(L + 1)
    inner = int((~bnd).sum())
    lib = hyg.load()
    op = hyg.P2ConstantOperator(mesh, 0, L)
    out = {"workload": f"P2 -Laplace, cube_24el.msh P2 level {L} ({n} DoFs)", "dofs_inner": inner,
           "algorithmic_bytes_per_dof_f64": 16, "flops_per_dof": 150}
    # row-kernel configuration (segment length x planes in flight) per level, value type and mode
    tuned = {}
    for lv in range(max(1, L - 2), L + 1):
        for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
            tuned[f"L{lv}_{name}"] = op.tune(lv, dt, reps=3, stream=st)
    out["row_kernel_config"] = {"tuned": tuned, "meaning": "config k: z-segments of (8, 16, 32, 64)[k // 2] planes, "
                                "k % 2 + 1 planes in flight; modes apply, add, residual, Jacobi, Chebyshev first, next"}
    rng = np.random.default_rng(seed)
    for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
        x, y = hyg.Function(mesh, L, L, dt, space="p2"), hyg.Function(mesh, L, L, dt, space="p2")
        x.upload(L, rng.uniform(-1, 1, n))

        def ap():
            hyg.check(lib.hyg_apply(op.h, x.h, y.h, L, hyg.INNER, hyg.REPLACE, st))

        for _ in range(2):
            ap()
        ms = timed(hip, st, ap, steps) / steps
        gbs = inner * (16 if dt == hyg.F64 else 8) / (ms / 1e3) / 1e9
        out[f"apply_{name}"] = {"ms": round(ms, 4), "GDoF/s": round(inner / (ms / 1e3) / 1e9, 3), "GB/s": round(gbs, 1),
                                "roofline_frac": round(gbs / PEAK_HBM_GBS, 4)}
        del x, y
    b = hyg.Function(mesh, 0, L, space="p2")
    bv = rng.uniform(-1.0, 1.0, n)
    bv[bnd] = 0.0
    b.upload(L, bv)
    Lr = min(L, 4)
    pv, pt = hyg.Function(mesh, Lr, Lr, space="p2"), hyg.Function(mesh, Lr, Lr, space="p2")
    pv.upload(Lr, np.random.default_rng(5).uniform(-1, 1, mesh.ndofs(Lr + 1)))
    rho = op.power_iteration(pv, pt, 40, Lr)
    kw = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * rho, 1.2 * rho))
    g64 = hyg.GMG(op, 0, L, hyg.F64, **kw)
    g32 = hyg.GMG(op, 0, L, hyg.F32, **kw)
    x64, xmx = hyg.Function(mesh, 0, L, space="p2"), hyg.Function(mesh, 0, L, space="p2")
    g64.solve_tol(x64, b, L, 1, 1e-10, st)
    hyg.ir_solve(op, g32, xmx, b, L, inner=1, max_outer=1, rel_tol=1e-10, stream=st)
    for f in (x64, xmx):
        f.interpolate(0.0, L)
    hip.sync(st)
    t0 = time.perf_counter()
    r64 = g64.solve_tol(x64, b, L, 60, 1e-10, st)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    rmx = hyg.ir_solve(op, g32, xmx, b, L, inner=1, max_outer=60, rel_tol=1e-10, stream=st)
    tmx = time.perf_counter() - t0
    u64, umx = x64.download(L), xmx.download(L)
    out["solve"] = {"cycle": f"V(3,3) Chebyshev order 2, rho(D^-1 A) = {rho:.5f}, coarse P2 level 0 CG",
                    "fp64": {"seconds": round(t64, 4), "cycles": len(r64) - 1, "rel_residual": float(r64[-1] / r64[0])},
                    "mixed": {"seconds": round(tmx, 4), "outer_steps": len(rmx) - 1, "inner_cycles": 1, "rel_residual": float(rmx[-1] / rmx[0]),
                              "rel_error_vs_fp64": float(np.abs(umx - u64).max() / np.abs(u64).max())},
                    "speedup_mixed_vs_fp64": round(t64 / tmx, 3)}
    return out


def bench_blending(hyg, hip, st, steps, level):
    """Blended P1 Laplacian (P1ElementwiseBlendingLaplaceOperator, p1_diffusion_blending_q3) on the configuration-5
    spherical shell (5, 2, [0.55, 1]) with the IcosahedralShellMap: setup (the per-point stencil weights, summed on
    the device from the blended element matrices) and apply throughput; the apply streams x, y and 15 weights per
    point and cell (136 B per inner DoF in FP64)."""
    mesh = hyg.Mesh("shell", 5, 2, 0.55, 1.0)
    mesh.set_shell_map()
    L = level
    n = mesh.ndofs(L)
    inner = int((~mesh.boundary(L)).sum())
    lib = hyg.load()
    t0 = time.perf_counter()
    op = hyg.P1BlendingLaplaceOperator(mesh, L, L)
    out = {"workload": f"P1 blended Laplacian, spherical shell (5, 2, [0.55, 1]) level {L}", "dofs_inner": inner,
           "setup_seconds": round(time.perf_counter() - t0, 3)}
    rng = np.random.default_rng(11)
    for dt, name, b in ((hyg.F64, "f64", 136.0), (hyg.F32, "f32", 68.0)):
        x, y = hyg.Function(mesh, L, L, dt), hyg.Function(mesh, L, L, dt)
        x.upload(L, rng.uniform(-1, 1, n))

        def ap():
            hyg.check(lib.hyg_apply(op.h, x.h, y.h, L, hyg.INNER, hyg.REPLACE, st))

        for _ in range(2):
            ap()
        ms = timed(hip, st, ap, steps) / steps
        gbs = b * inner / (ms / 1e3) / 1e9
        out[f"apply_{name}"] = {"ms": round(ms, 4), "GDoF/s": round(inner / (ms / 1e3) / 1e9, 3), "GB/s_alg": round(gbs, 1),
                                "roofline_frac": round(gbs / PEAK_HBM_GBS, 4)}
    return out


def bench_shell(hyg, hip, st, steps, level, seed, solve_cycles):
    """Spherical shell ntan = 5, nrad = 2, r in [0.55, 1] (960 macro-cells), P2 level 5: the vector Laplacian
    -div(eta grad u_i), i = 1..3, with eta(x) = exp(2.99 (1 - |x|)) (1 + 0.1 xi_e), xi_e uniform in [-1, 1) per
    micro-element (P2ElementwiseOperator with p2_div_k_grad_affine_q4; the same scalar operator on each
    component). Apply throughput of the three components in FP64 and FP32 (hyg_apply_vector), with each point's
    stencil assembled from the coefficient values at the quadrature points in every apply (one assembly for the
    three components) and with stored per-point stencils (one pass over the stored weights for the three
    components' cell interiors); one component solved to 1e-10 with V(3,3) Chebyshev cycles in FP64
    and as FP32 cycles inside FP64 defect correction (stored stencils)."""
    mesh = hyg.Mesh("shell", 5, 2, 0.55, 1.0)
    L = level
    n = mesh.ndofs(L + 1)
    bnd = mesh.boundary(L + 1)
    inner = int((~bnd).sum())
    lib = hyg.load()
    op = hyg.P2DivKGradOperator(mesh, 0, L, kind=0, k0=1.0, alpha=2.99, beta=0.1, seed=seed)
    out = {"workload": f"P2 vector Laplacian -div(eta grad u_i), 3 components, spherical shell (5, 2, [0.55, 1]) P2 level {L}",
           "dofs_inner_per_component": inner, "components": 3,
           "algorithmic_bytes_per_dof_f64": 16,
           # a point's stencil: 110 FMA (row l of sum_q k_q M_q) per (micro-cell, node) pair, 7.5 pairs per point on
           # average (24 at vertex points, 36 over the seven edge-midpoint classes), then ~20 FMA per component
           "fma_per_point": "825 (assembly, shared by the components) + 20 per component"}
    rng = np.random.default_rng(seed)
    for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
        xs = [hyg.Function(mesh, L, L, dt, space="p2") for _ in range(3)]
        ys = [hyg.Function(mesh, L, L, dt, space="p2") for _ in range(3)]
        for x in xs:
            x.upload(L, rng.uniform(-1, 1, n))

        def ap():
            hyg.apply_vector(op, xs, ys, L, hyg.INNER, hyg.REPLACE, st)

        def ap1():
            hyg.check(lib.hyg_apply(op.h, xs[0].h, ys[0].h, L, hyg.INNER, hyg.REPLACE, st))

        ap()
        ap1()
        ms = timed(hip, st, ap, steps) / steps
        ms1 = timed(hip, st, ap1, steps) / steps
        peak = 78.6e12 if dt == hyg.F64 else 157.3e12
        micro = int(mesh.info()[3]) * (1 << L) ** 3
        out[f"vector_apply_{name}"] = {"ms": round(ms, 3), "GDoF/s": round(3 * inner / (ms / 1e3) / 1e9, 3),
                                       "valu_frac": round(inner * 2.0 * (825 + 3 * 20) / (ms / 1e3) / peak, 4),
                                       # the reference's integrateAll: 4118 adds + 5862 muls per P2 micro-element and component
                                       "reference_equivalent_tflops": round(3 * micro * 9982.0 / (ms / 1e3) / 1e12, 2)}
        out[f"component_apply_{name}"] = {"ms": round(ms1, 3), "GDoF/s": round(inner / (ms1 / 1e3) / 1e9, 3),
                                          "valu_frac": round(inner * 2.0 * (825 + 20) / (ms1 / 1e3) / peak, 4)}
        del xs, ys
    # the same operator with its per-point stencils stored (computeAndStoreLocalElementMatrices): the solves use it
    hip.sync(st)
    t0 = time.perf_counter()
    op.store_local_matrices()
    out["store_local_matrices_seconds"] = round(time.perf_counter() - t0, 3)
    for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
        xs = [hyg.Function(mesh, L, L, dt, space="p2") for _ in range(3)]
        ys = [hyg.Function(mesh, L, L, dt, space="p2") for _ in range(3)]
        for x in xs:
            x.upload(L, rng.uniform(-1, 1, n))

        def ap():
            hyg.apply_vector(op, xs, ys, L, hyg.INNER, hyg.REPLACE, st)

        ap()
        ms = timed(hip, st, ap, steps) / steps
        out[f"vector_apply_stored_{name}"] = {"ms": round(ms, 3), "GDoF/s": round(3 * inner / (ms / 1e3) / 1e9, 3)}
        del xs, ys
    b = hyg.Function(mesh, 0, L, space="p2")
    bv = rng.uniform(-1.0, 1.0, n)
    bv[bnd] = 0.0
    b.upload(L, bv)
    Lr = min(L, 3)
    pv, pt = hyg.Function(mesh, Lr, Lr, space="p2"), hyg.Function(mesh, Lr, Lr, space="p2")
    pv.upload(Lr, np.random.default_rng(5).uniform(-1, 1, mesh.ndofs(Lr + 1)))
    rho = op.power_iteration(pv, pt, 40, Lr)
    kw = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * rho, 1.2 * rho))
    g64 = hyg.GMG(op, 0, L, hyg.F64, **kw)
    g32 = hyg.GMG(op, 0, L, hyg.F32, **kw)
    x64, xmx = hyg.Function(mesh, 0, L, space="p2"), hyg.Function(mesh, 0, L, space="p2")
    hip.sync(st)
    t0 = time.perf_counter()
    r64 = g64.solve_tol(x64, b, L, solve_cycles, 1e-10, st)
    t64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    rmx = hyg.ir_solve(op, g32, xmx, b, L, inner=1, max_outer=solve_cycles, rel_tol=1e-10, stream=st)
    tmx = time.perf_counter() - t0
    u64, umx = x64.download(L), xmx.download(L)
    out["solve_one_component"] = {
        "cycle": f"V(3,3) Chebyshev order 2, rho(D^-1 A) = {rho:.5f}, coarse P2 level 0 CG",
        "fp64": {"seconds": round(t64, 3), "cycles": len(r64) - 1, "rel_residual": float(r64[-1] / r64[0]),
                 "rate_per_cycle": float((r64[-1] / r64[0]) ** (1.0 / max(1, len(r64) - 1)))},
        "mixed": {"seconds": round(tmx, 3), "outer_steps": len(rmx) - 1, "inner_cycles": 1, "rel_residual": float(rmx[-1] / rmx[0]),
                  "rel_error_vs_fp64": float(np.abs(umx - u64).max() / np.abs(u64).max())},
        "speedup_mixed_vs_fp64": round(t64 / tmx, 3)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mesh", default="cube_24el.msh")
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cheb-order", type=int, default=2)
    ap.add_argument("--no-tune", action="store_true", help="keep the default cell-kernel configuration")
    ap.add_argument("--no-divkgrad", action="store_true", help="skip the variable-coefficient workload")
    ap.add_argument("--divkgrad-level", type=int, default=7)
    ap.add_argument("--no-p2", action="store_true", help="skip the P2 workload")
    ap.add_argument("--no-shell", action="store_true", help="skip the spherical-shell P2 vector Laplacian workload")
    ap.add_argument("--shell-level", type=int, default=5)
    ap.add_argument("--shell-cycles", type=int, default=80, help="cycle cap of the shell solves to 1e-10")
    ap.add_argument("--p2-level", type=int, default=7)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched independent problems")
    ap.add_argument("--batch-problems", type=int, default=64)
    ap.add_argument("--ir-inner", type=int, default=3, help="FP32 V-cycles per FP64 residual in the mixed solves")
    args = ap.parse_args()
    global IR_INNER
    IR_INNER = max(1, args.ir_inner)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2

    hyg = load_pkg()
    hyg.init(0)
    hip = Hip()
    st = hip.stream()
    L = args.level
    mesh = hyg.Mesh(args.mesh)
    n = mesh.ndofs(L)
    bnd = mesh.boundary(L)
    inner = int((~bnd).sum())
    rng = np.random.default_rng(2024)
    op = hyg.P1ConstantOperator(mesh, 0, L)
    lib = hyg.load()
    # per-level choice of the cell-kernel configuration (rows x points per lane) for both value types
    tuned = {}
    if not args.no_tune:
        for lev in range(max(3, L - 4), L + 1):
            for dt, name in ((hyg.F64, "f64"), (hyg.F32, "f32")):
                tuned[f"L{lev}_{name}"] = op.tune(lev, dt, reps=3 if lev >= 7 else 10, stream=st)

    # ---------------- apply: the timed steps
    x = hyg.Function(mesh, L, L)
    y = hyg.Function(mesh, L, L)
    x.upload(L, rng.uniform(-1.0, 1.0, n))

    def apply_once():
        hyg.check(lib.hyg_apply(op.h, x.h, y.h, L, hyg.INNER, hyg.REPLACE, st))

    for _ in range(args.warmup):
        apply_once()
    ms = timed(hip, st, apply_once, args.steps)
    ms_step = ms / args.steps
    gdofs = inner * args.steps / (ms / 1e3) / 1e9
    alg_bytes = 16.0 * inner
    achieved = alg_bytes / (ms_step / 1e3) / 1e9

    # copy kernel of known size on the same stream (bandwidth reference)
    nval = int(mesh.info()[3]) * ((((1 << L) + 1) * ((1 << L) + 2) * ((1 << L) + 3) // 6 + 63) // 64 * 64)

    def copy_once():
        hyg.check(lib.hyg_convert(y.h, x.h, L, st))

    for _ in range(3):
        copy_once()
    ms_copy = timed(hip, st, copy_once, 20) / 20
    copy_gbs = 16.0 * nval / (ms_copy / 1e3) / 1e9

    # one multi-colour SOR sweep (relax 1, in place) for comparison with the fused Jacobi sweep
    ncol = mesh.sor_colours()[0]

    def sor_once():
        hyg.check(lib.hyg_smooth_sor(op.h, x.h, y.h, 1.0, L, hyg.INNER, 0, st))

    def jac_once():
        hyg.check(lib.hyg_smooth_jac(op.h, y.h, x.h, x.h, 2.0 / 3.0, L, hyg.INNER, st))

    def sor_ordered_once():
        hyg.check(lib.hyg_smooth_sor_ordered(op.h, x.h, y.h, 1.0, L, hyg.INNER, 0, st))

    sor_once()
    ms_sor = timed(hip, st, sor_once, 5) / 5
    sor_ordered_once()
    ms_sor_ord = timed(hip, st, sor_ordered_once, 3) / 3
    jac_once()
    ms_jac = timed(hip, st, jac_once, 10) / 10

    out = {
        "metric": "GDoF/s matrix-free apply + time-to-1e-10 V-cycle solve, mixed vs FP64, 1 GPU",
        "value": round(gdofs, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform[-1,1] source vector)",
        "config": {"workload": f"P1 constant-coefficient Laplace apply, {args.mesh} level {L}, FP64, Inner DoFs",
                   "mesh": args.mesh, "level": L, "dofs_total": n, "dofs_inner": inner, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": profile_traffic(),
                     "kernel": "P1 apply: kSweep (cell interiors, one-sided face sums, edges, vertices) + kFaceCombine",
                     "algorithmic_bytes_per_launch": alg_bytes},
        "copy_reference": {"bytes_per_launch": 16.0 * nval, "ms": round(ms_copy, 5), "GB/s": round(copy_gbs, 1)},
        "cell_kernel_config": {"apply": "rows %d x points %d" % op.config(L, hyg.F64, 0)[1:],
                               "tuned_per_mode": tuned},
        "smoothers_f64": {"jacobi_sweep_ms": round(ms_jac, 4),
                          "sor_sweep_ms": round(ms_sor, 4), "sor_colours": ncol,
                          "sor_reference_order_sweep_ms": round(ms_sor_ord, 4),
                          "note": "multi-colour SOR: one sweep-kernel pass plus a face pass per colour; reference "
                                  "order: vertices, edges, faces (anti-diagonals), then one launch per cell hyperplane "
                                  "x + 2y + 3z"},
    }

    # ---------------- time to 1e-10: FP64 vs mixed precision
    if not args.no_solve:
        b = hyg.Function(mesh, 0, L)
        bv = rng.uniform(-1.0, 1.0, n)
        bv[bnd] = 0.0
        b.upload(L, bv)
        # spectral radius of D^-1 A (power iteration, fixed start vector and 40 iterations), Chebyshev bounds [0.3, 1.2] rho
        Lr = min(L, 6)
        pv, pt = hyg.Function(mesh, Lr, Lr), hyg.Function(mesh, Lr, Lr)
        pv.upload(Lr, np.random.default_rng(5).uniform(-1, 1, mesh.ndofs(Lr)))
        rho = op.power_iteration(pv, pt, 40, Lr)
        kw = dict(pre=3, post=3, smoother="cheb", cheb_order=args.cheb_order, cheb_bounds=(0.3 * rho, 1.2 * rho))
        g64 = hyg.GMG(op, 0, L, hyg.F64, **kw)
        g32 = hyg.GMG(op, 0, L, hyg.F32, **kw)
        x64 = hyg.Function(mesh, 0, L)
        xmx = hyg.Function(mesh, 0, L)
        # warm-up solves (first-touch of all levels), then timed solves from x = 0
        g64.solve_tol(x64, b, L, 1, 1e-10, st)
        hyg.ir_solve(op, g32, xmx, b, L, inner=IR_INNER, max_outer=1, rel_tol=1e-10, stream=st)
        for f in (x64, xmx):
            f.interpolate(0.0, L)
        b.upload(L, bv)
        hip.sync(st)
        t0 = time.perf_counter()
        res64 = g64.solve_tol(x64, b, L, 60, 1e-10, st)
        t64 = time.perf_counter() - t0
        b.upload(L, bv)
        hip.sync(st)
        t0 = time.perf_counter()
        resmx = hyg.ir_solve(op, g32, xmx, b, L, inner=IR_INNER, max_outer=60, rel_tol=1e-10, stream=st)
        tmx = time.perf_counter() - t0
        u64, umx = x64.download(L), xmx.download(L)
        out["solve"] = {
            "problem": f"-Laplace u = f, {args.mesh} L{L}, zero Dirichlet, f ~ U[-1,1] seeded, u0 = 0",
            "cycle": f"V(3,3), Chebyshev order {args.cheb_order}, bounds [0.3, 1.2] x rho(D^-1 A) = {rho:.6f}, "
                     f"coarse level 0 CG",
            "fp64": {"seconds": round(t64, 5), "cycles": len(res64) - 1,
                     "rel_residual": float(res64[-1] / res64[0])},
            "mixed": {"seconds": round(tmx, 5), "outer_steps": len(resmx) - 1, "inner_cycles": IR_INNER,
                      "rel_residual": float(resmx[-1] / resmx[0]),
                      "rel_error_vs_fp64": float(np.abs(umx - u64).max() / np.abs(u64).max())},
            "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
            "note": "eager launches; FP64 residual norm after every FP64 cycle, mixed after every outer step",
        }
        out["solve_graph"] = graph_solves(hyg, hip, st, op, mesh, L, bv, kw, IR_INNER)

    # ---------------- variable-coefficient workload
    if not args.no_divkgrad:
        out["divkgrad"] = bench_divkgrad(hyg, hip, st, max(3, args.steps // 10), args.divkgrad_level, 2024)

    # ---------------- batched independent problems
    if not args.no_batched:
        out["batched"] = bench_batched(hyg, hip, st, max(3, args.steps // 10), args.batch_problems, 2027)

    # ---------------- P2 workload
    if not args.no_p2:
        out["p2"] = bench_p2(hyg, hip, st, max(3, args.steps // 10), args.p2_level, 2025)

    # ---------------- spherical shell, P2 vector Laplacian with variable viscosity
    if not args.no_shell:
        out["shell_blending"] = bench_blending(hyg, hip, st, max(5, args.steps // 10), 6)
        out["shell_p2"] = bench_shell(hyg, hip, st, max(2, args.steps // 25), args.shell_level, 2026, args.shell_cycles)

    # ---------------- the reference's CPU-runnable case on the GPU
    out["config1"] = bench_config1(hyg, hip, st)

    # ---------------- CPU baseline (oracle), rank 0, N = 1
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args.mesh, L, args.cpu_seconds, inner)
        except Exception as e:  # the baseline must not hide the GPU result
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
