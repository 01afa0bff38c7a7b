#!/usr/bin/env python3
"""Benchmark of the FE assembly hot path on one MI355X.

Workload (default): plain FP64 linear-elasticity stiffness on an H8 block of
100 x 100 x 100 = 1,000,000 elements (1,030,301 nodes, 3 DOF/node, 24 x 24
element matrices, 2x2x2 Gauss rule, E = 200e3, nu = 0.3, interior nodes
perturbed by seeded U[-1,1) * 0.15 h).  Mesh, DOFs and the CSC pattern are
resident in HBM before the timed region; one step is one numeric assembly
(element integration + global summation into the CSC values).

Prints one JSON line (the driver's contract); progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "finetools.jl_amd"))

import numpy as np  # noqa: E402

METRIC = "assembled elements/sec + nnz/sec (H8 elasticity FP64); p50/p99 latency served"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def iso_D(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    for i in range(3):
        D[i, i] = lam + 2 * mu
        D[3 + i, 3 + i] = mu
    return D


def algorithmic_bytes_numeric(npe, ne, nn, nnz):
    """Compulsory HBM bytes of one numeric assembly: int32 connectivity and
    FP64 coordinates read once, FP64 values written once."""
    return 4 * npe * ne + 24 * nn + 8 * nnz


FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak


def algorithmic_flops_numeric(npe, nqp, ne, nnz, ndn=3):
    """FP64 flops of one numeric elasticity assembly in the node-pair block
    formulation, counted once per element (the fused kernel recomputes the
    per-point work of its halo elements on top of this):
    per element and point: J = X^T dN (9 npe FMA), det J by cofactors (14),
    J w (1), the adjugate inverse scaled by 1/det (36 + 1 division), sqrt|J w|
    (1), u_a = sqrt|J w| * (dN_a inv J) (6 per component, 3 npe components);
    per element: the node-pair products sum_q u_m u_c^T over the npe(npe-1)/2
    off-diagonal pairs (9 FMA per point) and the npe diagonal ones (6 FMA, the
    symmetric upper triangle); per assembled node block (nnz / ndn^2 of them)
    the cubic-symmetry material contraction (32)."""
    per_point = 2 * 9 * npe + 14 + 1 + 37 + 1 + 6 * 3 * npe
    pairs = npe * (npe - 1) // 2 * nqp * 9 * 2 + npe * nqp * 6 * 2
    return ne * (nqp * per_point + pairs) + (nnz // (ndn * ndn)) * 32


def host_info():
    """nproc and, where lscpu answers, the CPU model and socket count."""
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() == "Model name":
                info["model"] = val.strip()
            elif key.strip() == "Socket(s)":
                info["sockets"] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(n_cells, seed, D, runs=5, threads=None):
    """CPU restatement of the reference's element loop + sparse() timed on a
    bounded sample of the same workload: one warm-up, then the median of
    `runs` runs, split into the element loop (COO triples) and sparse().
    The headline figure is serial (the reference's loop is serial); beside
    it the same element loop on `threads` OpenMP threads (sparse() serial)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    conn, xyz = O.block_mesh(O.H8, (n_cells,) * 3, L=1.0, amp_frac=0.15, seed=seed)
    ne, nn = conn.shape[0], xyz.shape[0]
    dofnums = np.arange(1, 3 * nn + 1, dtype=np.int64).reshape(nn, 3)
    coeff = D.ravel()

    def timed(nthreads):
        loop, spar = [], []
        for r in range(runs + 1):
            t0 = time.perf_counter()
            I, J, V = O.bilform_coo(O.H8, O.LIN_ELASTIC, conn, xyz, 3, dofnums, 3 * nn, coeff, nthreads=nthreads)
            t1 = time.perf_counter()
            O.sparse(I, J, V, 3 * nn, 3 * nn)
            t2 = time.perf_counter()
            del I, J, V
            if r:  # the first run warms up
                loop.append(t1 - t0)
                spar.append(t2 - t1)
        tl, ts = statistics.median(loop), statistics.median(spar)
        return tl, ts, statistics.median([a + b for a, b in zip(loop, spar)])

    tl, ts, tt = timed(1)
    nt = threads or max(1, min(16, os.cpu_count() or 1))
    tl2, ts2, tt2 = timed(nt)
    return {"value": ne / tt, "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": "H8 %d^3 = %d elements linear elasticity FP64, element loop + sparse(), one warm-up, median of "
                      "%d runs (%.2f s each)" % (n_cells, ne, runs, tt),
            "element_loop_s": tl, "sparse_s": ts, "host": host_info(),
            "threaded": {"value": ne / tt2, "unit": "elements/s", "cores": nt, "element_loop_s": tl2,
                         "sparse_s": ts2, "note": "OpenMP element loop on %d threads, sparse() serial" % nt}}


PROFILE = os.path.join(ROOT, "profiles", "k_fused_h8_elasticity_100cube.md")


def profiled_traffic(path=PROFILE):
    """HBM bytes per launch of the fused kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a
    streaming copy); None when the summary is absent."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def rel_fro(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bench_vec(feasm, core, n=100, steps=10):
    """Load vector (linform_dot, m = 3, + makevector!) of a constant body force
    on the C2 mesh (H8 100^3, 3 DOF/node): device kernel time per assembly and
    its algorithmic bytes (int32 connectivity and FP64 coordinates read once,
    FP64 vector written once)."""
    ctx = feasm.Context(0)
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / n, 2)
    ne, nn = conn.shape[0], xyz.shape[0]
    mesh = core.DeviceMesh(ctx, feasm.H8, conn, xyz)
    mesh.set_dofs(3)
    f = [0.0, -9.81, 2.0]
    Fv = mesh.linform_dot(feasm.RULE_GAUSS, 2, f, nn * 3)
    ctx.set_timing(steps)
    for _ in range(steps):
        mesh.linform_dot(feasm.RULE_GAUSS, 2, f, nn * 3)
    ctx.synchronize()
    ms = float(np.mean(ctx.timings()[:, 0]))
    ctx.set_timing(0)
    bytes_alg = 4 * 8 * ne + 24 * nn + 8 * 3 * nn
    out = {"workload": "H8 %d^3 body load, 3 DOF/node, 2x2x2 Gauss" % n, "kernels_ms": ms,
           "elements_per_s": ne / (ms * 1e-3), "achieved_GBps": bytes_alg / (ms * 1e-3) / 1e9,
           "algorithmic_bytes": bytes_alg, "total_load": [float(v) for v in Fv.reshape(-1, 3).sum(axis=0)]}
    mesh.close()
    ctx.close()
    return out


def bench_lumped(feasm, core, n=100, steps=10):
    """HRZ-lumped mass (bilform_dot into SysmatAssemblerSparseHRZLumpingSymm +
    makematrix!) of rho I3 on the C2 mesh (H8 100^3, 3 DOF/node, 2x2x2
    Gauss): stream time per assembly (element kernel, node gather, scan,
    compaction and the nnz read-back), and the total mass, which must equal
    ndn * rho * volume."""
    ctx = feasm.Context(0)
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / n, 2)
    ne, nn = conn.shape[0], xyz.shape[0]
    mesh = core.DeviceMesh(ctx, feasm.H8, conn, xyz)
    del conn, xyz
    mesh.set_dofs(3)
    rho = 7.8e-9
    c = (rho * np.eye(3)).ravel()
    M = mesh.bilform_dot_lumped(feasm.RULE_GAUSS, 2, c)
    total = float(M.csc_arrays(1)[2].sum())
    M.close()
    ctx.set_timing(steps)
    for _ in range(steps):
        mesh.bilform_dot_lumped(feasm.RULE_GAUSS, 2, c).close()
    ctx.synchronize()
    ms = float(np.mean(ctx.timings()[:, 0]))
    ctx.set_timing(0)
    out = {"workload": "H8 %d^3 HRZ-lumped mass rho I3, 3 DOF/node, 2x2x2 Gauss" % n, "ms": ms,
           "elements_per_s": ne / (ms * 1e-3), "total_mass_rel_err": abs(total - 3 * rho) / (3 * rho),
           "nodes": nn}
    mesh.close()
    ctx.close()
    return out


def bench_c2_mixed(feasm, core, D, n=100, steps=10):
    """C2 in mixed precision (FP32 element arithmetic with FP64 Jacobians,
    FP64 sums): device time per re-assembly and the relative Frobenius error
    of the values against the FP64 assembly of the same mesh."""
    ctx = feasm.Context(0)
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / n, 2)
    ne = conn.shape[0]
    mesh = core.DeviceMesh(ctx, feasm.H8, conn, xyz)
    del conn, xyz
    mesh.set_dofs(3)
    args = (feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D)
    K64 = mesh.bilform(*args)
    Km = mesh.bilform(*args, precision=feasm.MIXED)
    for _ in range(2):
        mesh.bilform_into(Km, *args, precision=feasm.MIXED)
    ctx.set_timing(steps)
    for _ in range(steps):
        mesh.bilform_into(Km, *args, precision=feasm.MIXED)
    ctx.synchronize()
    ms = float(np.mean(ctx.timings()[:, 0]))
    ctx.set_timing(0)
    nz64 = K64.csc_arrays(0)[2]
    nzm = Km.csc_arrays(0)[2]
    err = float(np.linalg.norm(nzm - nz64) / np.linalg.norm(nz64))
    del nz64, nzm
    out = {"workload": "H8 %d^3 linear elasticity, mixed precision, resident pattern" % n, "ms": ms,
           "elements_per_s": ne / (ms * 1e-3), "rel_fro_vs_fp64": err}
    K64.close()
    Km.close()
    mesh.close()
    ctx.close()
    return out


def bench_t10(feasm, core, n=60, steps=5):
    """C3: T10 n^3 x 6 tets, E=70e3 nu=0.33, rho=2.7e-9: FP64 and mixed stiffness
    plus the consistent mass; error of the mixed results against FP64."""
    ctx = feasm.Context(0)
    h = 1.0 / n
    t0 = time.perf_counter()
    conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, n, n, n, 0.1 * h, 0.03 * h, 3)
    tgen = time.perf_counter() - t0
    ne, nn = conn.shape[0], xyz.shape[0]
    log("T10 mesh: %d elements, %d nodes (generated in %.1f s)" % (ne, nn, tgen))
    mesh = core.DeviceMesh(ctx, feasm.T10, conn, xyz)
    del conn, xyz
    mesh.set_dofs(3)
    D = iso_D(70e3, 0.33)
    rho = 2.7e-9
    out = {"workload": "T10 %dx%dx%d x6 = %d elements, %d nodes, TetRule(4), E=70e3 nu=0.33, rho=2.7e-9" % (n, n, n, ne, nn)}
    res = {}
    for name, form, coeff, prec in (("K_fp64", feasm.LIN_ELASTIC, D, feasm.FP64),
                                    ("K_mixed", feasm.LIN_ELASTIC, D, feasm.MIXED),
                                    ("M_fp64", feasm.DOT, rho * np.eye(3), feasm.FP64),
                                    ("M_mixed", feasm.DOT, rho * np.eye(3), feasm.MIXED)):
        M = mesh.bilform(form, feasm.RULE_TET, 4, coeff, precision=prec)
        for _ in range(2):
            mesh.bilform_into(M, form, feasm.RULE_TET, 4, coeff, precision=prec)
        ctx.set_timing(steps)
        for _ in range(steps):
            mesh.bilform_into(M, form, feasm.RULE_TET, 4, coeff, precision=prec)
        ctx.synchronize()
        ms = float(np.mean(ctx.timings()[:, 2]))
        ctx.set_timing(0)
        nnz = M.nnz
        res[name] = M.csc_arrays(0)[2]
        gbs = algorithmic_bytes_numeric(10, ne, nn, nnz) / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "elements_per_s": ne / (ms * 1e-3), "nnz_per_s": nnz / (ms * 1e-3),
                     "achieved_GBps": gbs, "roofline_frac": gbs / HBM_PEAK_GBS}
        out["nnz"] = nnz
        M.close()
    out["K_mixed"]["rel_fro_vs_fp64"] = rel_fro(res["K_mixed"], res["K_fp64"])
    out["M_mixed"]["rel_fro_vs_fp64"] = rel_fro(res["M_mixed"], res["M_fp64"])
    mesh.close()
    ctx.close()
    return out


def bench_batched(feasm, core, steps=5, tune_db=None):
    """C4: 256 H8 blocks, nx,ny,nz ~ U{4..64}; half 1-DOF diffusion k~U[0.5,2],
    half 3-DOF elasticity E~U[1e3,1e5] nu~U[0.2,0.45]; mixed precision; one
    persistent launch per form replayed from a hipGraph; tile size autotuned.
    Batched and plain (one bilform_into per mesh, FP64) are compared on the
    same two clocks: host wall time around `steps` whole passes (each ending
    in a stream synchronisation), and device time from HIP events."""
    from feasm import workloads as W

    ctx = feasm.Context(0)
    specs = W.batched_specs()
    meshes = []
    t0 = time.perf_counter()
    ne_tot = bytes_alg = 0
    for (n, form, ndn, coeff, seed) in specs:
        conn, xyz = W.h8_block(n, 0.1, seed)
        m = core.DeviceMesh(ctx, feasm.H8, conn, xyz)
        m.set_dofs(ndn)
        meshes.append(m)
        ne_tot += conn.shape[0]
        nnz = ndn * ndn * (3 * n[0] + 1) * (3 * n[1] + 1) * (3 * n[2] + 1)
        bytes_alg += algorithmic_bytes_numeric(8, conn.shape[0], xyz.shape[0], nnz)
    log("batch: %d meshes, %d elements (setup %.1f s)" % (len(specs), ne_tot, time.perf_counter() - t0))
    results = {"workload": "256 H8 blocks, n ~ U{4..64}^3 (seeded), half diffusion half elasticity",
               "elements": ne_tot, "algorithmic_bytes": bytes_alg}
    vals = {}
    for label, prec in (("mixed", feasm.MIXED), ("fp64", feasm.FP64)):
        batch = core.Batch(ctx)
        for m, (n, form, ndn, coeff, _) in zip(meshes, specs):
            batch.add(m, form, feasm.RULE_GAUSS, 2, coeff, precision=prec)
        t0 = time.perf_counter()
        tuned = batch.autotune(db_path=tune_db, key="h8-batch256-%s" % label, reps=3)
        batch.prepare(use_graph=True, max_nodes=tuned["max_nodes"], budget_kib=tuned["budget_kib"],
                      threads=tuned["threads"])
        tprep = time.perf_counter() - t0
        batch.run()
        ctx.synchronize()
        dev = []
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.run()
            dev.append(batch.last_ms())  # waits for the replay
        wall_ms = 1e3 * (time.perf_counter() - t0) / steps
        st = batch.stats()
        ms = float(np.median(dev))
        results[label] = {"wall_ms": wall_ms, "device_ms": ms, "elements_per_s": st["elements"] / (wall_ms * 1e-3),
                          "nnz_per_s": st["nnz"] / (wall_ms * 1e-3), "tuned": tuned, "prepare_s": tprep,
                          "tiles": st["tiles"],
                          "roofline_frac": bytes_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        vals[label] = [batch.matrix(i).csc_arrays(0)[2] for i in range(0, len(specs), 16)]
        results["nnz"] = st["nnz"]
        batch.close()
    results["mixed"]["rel_fro_vs_fp64_max"] = max(rel_fro(a, b) for a, b in zip(vals["mixed"], vals["fp64"]))
    del vals
    # the plain per-mesh FP64 path: one bilform_into per mesh, in order
    mats = [m.bilform(form, feasm.RULE_GAUSS, 2, coeff) for m, (_, form, _, coeff, _) in zip(meshes, specs)]
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for m, M, (_, form, _, coeff, _) in zip(meshes, mats, specs):
            m.bilform_into(M, form, feasm.RULE_GAUSS, 2, coeff)
        ctx.synchronize()
    wall_plain = 1e3 * (time.perf_counter() - t0) / steps
    # device time in separate passes (the events add host work to the wall time)
    ctx.set_timing(len(specs) * steps)
    for _ in range(steps):
        for m, M, (_, form, _, coeff, _) in zip(meshes, mats, specs):
            m.bilform_into(M, form, feasm.RULE_GAUSS, 2, coeff)
    ctx.synchronize()
    dev_plain = float(ctx.timings()[:, 2].sum()) / steps
    ctx.set_timing(0)
    results["plain_fp64_per_mesh"] = {"wall_ms": wall_plain, "device_ms": dev_plain,
                                      "elements_per_s": ne_tot / (wall_plain * 1e-3)}
    results["speedup_vs_plain"] = {"wall": wall_plain / results["mixed"]["wall_ms"],
                                   "device": dev_plain / results["mixed"]["device_ms"]}
    for M in mats:
        M.close()
    for m in meshes:
        m.close()
    ctx.close()
    return results


def bench_coo(feasm, core, D, n, reps=3):
    """The COO path on C2: the element kernel writes every element's 24 x 24
    triples in assemble!'s order, then the device sparse() (radix sort of the
    64-bit keys, segment sums in input order, column pointers) builds the CSC
    matrix -- the path general DOF numberings and the assembler protocol take.
    Wall time of whole feasm_bilform_coo calls (new matrix each), and the
    result against the resident-pattern (fused) assembly: same pattern, values
    within rounding."""
    ctx = feasm.Context(0)
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / n, 2)
    ne = conn.shape[0]
    mesh = core.DeviceMesh(ctx, feasm.H8, conn, xyz)
    del conn
    mesh.set_dofs(3)
    mesh.bilform(feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D, path="coo").close()  # warm-up
    ts = []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        Mc = mesh.bilform(feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D, path="coo")
        ctx.synchronize()
        ts.append(1e3 * (time.perf_counter() - t0))
        if len(ts) < reps:
            Mc.close()
    Mf = mesh.bilform(feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D)
    cp_c, rv_c, nz_c = Mc.csc_arrays(0)
    cp_f, rv_f, nz_f = Mf.csc_arrays(0)
    same_pattern = bool(np.array_equal(cp_c, cp_f) and np.array_equal(rv_c, rv_f))
    del cp_c, rv_c, cp_f, rv_f
    rel = float(np.linalg.norm(nz_c - nz_f) / np.linalg.norm(nz_f))
    triples = ne * 24 * 24
    ms = float(np.median(ts))
    out = {"workload": "H8 %d^3 linear elasticity FP64 through the COO path (element triples + device sparse())" % n,
           "ms": ms, "runs_ms": ts, "elements_per_s": ne / (ms * 1e-3), "triples": triples,
           "triples_per_s": triples / (ms * 1e-3), "same_pattern_as_resident": same_pattern,
           "rel_fro_vs_resident": rel}
    Mc.close()
    Mf.close()
    mesh.close()
    ctx.close()
    return out


def bench_served(feasm, core, rate=200.0, duration=3.0, workers=8):
    """C5: Poisson arrivals (rate/s) of H8/T10 (50/50) block meshes with log-normal
    element counts (median 5e4, sigma 1.0, clipped to [1e3, 2e6]), diffusion or
    elasticity; 30 % re-assemble a resident mesh with new material values.
    Served (continuous batching, mixed precision, `workers` streams) vs the
    plain per-request FP64 path (one request at a time, in arrival order),
    paced at the arrival times and, for throughput at saturation, all at once."""
    from feasm import workloads as W

    t0 = time.perf_counter()
    meshes, rng = W.served_pool()
    stream = W.served_stream(rng, len(meshes), rate, duration)
    log("served: pool of %d meshes generated in %.1f s" % (len(meshes), time.perf_counter() - t0))
    out = {"workload": "Poisson %.0f req/s for %.1f s, H8/T10 50/50, median 5e4 elements (sigma 1.0), 30%% "
                       "re-assembly of resident meshes" % (rate, duration)}
    # one unmeasured round of each path at saturation first: the memory pool
    # and the pinned staging buffers reach the sizes the stream needs, as in
    # a server that has been running
    sat = {"saturated_served_mixed": (feasm.MIXED, workers), "saturated_plain_fp64_sequential": (feasm.FP64, 1)}
    for key, (prec, w) in sat.items():
        W.run_served(stream, meshes, prec, w, paced=False)
    for key, prec, w, paced in (("served_mixed", feasm.MIXED, workers, True),
                                ("plain_fp64_sequential", feasm.FP64, 1, True)):
        rec, wall = W.run_served(stream, meshes, prec, w, paced=paced)
        out[key] = W.served_summary(rec, wall)
        log("%s: %s" % (key, out[key]))
    # at saturation (all requests at once): two interleaved rounds of each,
    # summarised by their mean throughput (each round's throughput kept) and
    # the last round's latencies
    rounds = {k: [] for k in sat}
    for r in range(2):
        for key, (prec, w) in sat.items():
            rec, wall = W.run_served(stream, meshes, prec, w, paced=False)
            rounds[key].append(W.served_summary(rec, wall))
            log("%s: %s" % (key, rounds[key][-1]))
    for key, rs in rounds.items():
        out[key] = dict(rs[-1])
        out[key]["elements_per_s"] = float(np.mean([r["elements_per_s"] for r in rs]))
        out[key]["rounds_elements_per_s"] = [r["elements_per_s"] for r in rs]
    out["throughput_ratio"] = out["served_mixed"]["elements_per_s"] / out["plain_fp64_sequential"]["elements_per_s"]
    out["saturated_throughput_ratio"] = (out["saturated_served_mixed"]["elements_per_s"] /
                                         out["saturated_plain_fp64_sequential"]["elements_per_s"])
    return out


def bench_gram(feasm, ne=262144, reps=50, rounds=2):
    """The per-element node-pair Gram (8 nodes, 8 points, 3 x 3 blocks for the
    36 pairs a <= b) from element records in LDS, on the FP64 vector ALUs and
    on the FP64 matrix cores (v_mfma_f64_16x16x4f64, the 24 x 8 gradient
    table padded to 32 rows: six MFMA per element), interleaved rounds."""
    import ctypes

    ctx = feasm.Context(0)
    lib = feasm.lib()
    out = {"workload": "%d synthetic H8 element records x %d passes, LDS resident" % (ne, reps)}
    useful = 36 * 72 * 2  # flops of one element's upper node-pair Gram
    issued = {"vector": useful, "mfma": 6 * 16 * 16 * 4 * 2}
    best = {}
    for _ in range(rounds):
        for name, v in (("vector", 0), ("mfma", 1)):
            ms, cs = ctypes.c_double(), ctypes.c_double()
            feasm._lib.check(lib.feasm_probe_gram(ctx.h, v, ne, reps, ctypes.byref(ms), ctypes.byref(cs), None, 0),
                             ctx.h)
            best[name] = min(best.get(name, 1e30), ms.value)
    for name, ms in best.items():
        rate = ne * reps / (ms * 1e-3)
        out[name] = {"ms": ms, "element_grams_per_s": rate, "useful_fp64_tflops": rate * useful / 1e12,
                     "issued_fp64_tflops": rate * issued[name] / 1e12}
    out["mfma_over_vector"] = best["vector"] / best["mfma"]
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=100, help="cells per direction of the H8 block")
    ap.add_argument("--cpu-cells", type=int, default=50, help="cells per direction of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extras", default="mixed,vec,lumped,t10,batch,served,gram,coo",
                    help="comma list of extra workloads: mixed,vec,lumped,t10,batch,served,gram,coo (or 'none')")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB streaming copy (3 launches) to scale traffic counters")
    ap.add_argument("--path", default="auto", choices=["auto", "staged"],
                    help="numeric path of the C2 workload: auto (the fused kernel) or staged (element + gather "
                         "kernels, the fallback), for profiling the latter")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "multi-GPU sharding is a design note, not a run (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import feasm
    from feasm import core

    n = a.cells
    D = iso_D(200e3, 0.3)
    ctx = feasm.Context(0)
    copy_gbs = None
    if a.calibrate:
        import ctypes

        ms = ctypes.c_double()
        feasm._lib.check(feasm.lib().feasm_copy_calibration(ctx.h, 1 << 30, 3, ctypes.byref(ms)), ctx.h)
        copy_gbs = 2 * (1 << 30) / (ms.value * 1e-3) / 1e9
        log("copy calibration: 1 GiB in %.3f ms = %.0f GB/s (read + write)" % (ms.value, copy_gbs))
    # the process's first assembly loads the library's device code (every
    # kernel of the path, once per process): timed on a 4^3 mesh, so that
    # the C2 first assembly below is that of a new mesh in a running process
    t0 = time.perf_counter()
    c4, x4 = core.gen_h8block(1.0, 1.0, 1.0, 4, 4, 4)
    m4 = core.DeviceMesh(ctx, feasm.H8, c4, x4)
    m4.set_dofs(3)
    m4.bilform(feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D).close()
    ctx.synchronize()
    t_load = time.perf_counter() - t0
    m4.close()
    log("process's first assembly (4^3 mesh, device code loaded): %.1f ms" % (1e3 * t_load))
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, n, n, n)
    xyz = core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], 0.15 / n, 2)
    ne, nn = conn.shape[0], xyz.shape[0]
    log("mesh: %d elements, %d nodes" % (ne, nn))
    mesh = core.DeviceMesh(ctx, feasm.H8, conn, xyz)
    del conn
    mesh.set_dofs(3)
    if a.path != "auto":
        mesh.set_numeric_path(a.path)
    ctx.synchronize()
    t0 = time.perf_counter()
    mesh.build_pattern()
    ctx.synchronize()
    t_sym = time.perf_counter() - t0
    t0 = time.perf_counter()
    M = mesh.bilform(feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D)
    ctx.synchronize()
    t_first_numeric = time.perf_counter() - t0
    nnz = M.nnz
    ntiles, nhalo = mesh.tile_info()
    log("pattern: nnz %d, symbolic %.1f ms, first numeric %.1f ms (tiles %d, halo elements %d)"
        % (nnz, 1e3 * t_sym, 1e3 * t_first_numeric, ntiles, nhalo))

    for _ in range(a.warmup):
        mesh.bilform_into(M, feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D)
    ctx.synchronize()
    ctx.set_timing(max(a.steps, 1))
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        mesh.bilform_into(M, feasm.LIN_ELASTIC, feasm.RULE_GAUSS, 2, D)
    ctx.synchronize()
    t1 = time.perf_counter()
    tk = ctx.timings()
    ctx.set_timing(0)
    ms_step = 1e3 * (t1 - t0) / a.steps
    el_ms, ga_ms = float(np.mean(tk[:, 0])), float(np.mean(tk[:, 1]))
    kern_ms = el_ms + ga_ms
    log("step %.3f ms; element kernel %.3f ms, gather kernel %.3f ms" % (ms_step, el_ms, ga_ms))

    bytes_alg = algorithmic_bytes_numeric(8, ne, nn, nnz)
    achieved = bytes_alg / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": ne / (ms_step * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed H8 block)",
        "config": {"workload": "H8 %dx%dx%d linear elasticity FP64, 3 DOF/node, 2x2x2 Gauss, E=200e3 nu=0.3, "
                               "+-0.15h perturbation; numeric re-assembly into the resident CSC pattern"
                               % (n, n, n),
                   "elements": ne, "nodes": nn, "nnz": nnz, "parallelism": "single GPU"},
        "nnz_per_s": nnz / (ms_step * 1e-3),
        "first_assembly_ms": 1e3 * (t_sym + t_first_numeric),
        "process_first_assembly_ms": 1e3 * t_load,
        "symbolic_ms": 1e3 * t_sym,
        "kernels_ms": {"fused": el_ms} if ntiles > 0 else {"element": el_ms, "gather": ga_ms},
        "tiles": ntiles,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic() if (ntiles > 0 and n == 100) else None,
                     "kernel": "k_fused (fused tile kernel, whole numeric phase)" if ntiles > 0
                     else "k_element + k_gather (staged numeric phase)",
                     "algorithmic_bytes": bytes_alg,
                     # diagnostic: what a COO path would move (every element's 24 x 24
                     # triples written and read back at 16 bytes each)
                     "coo_path_bytes": 2 * 16 * (8 * 3) ** 2 * ne},
        "fp64_flops": None,
        "served": None,
    }
    flops = algorithmic_flops_numeric(8, 8, ne, nnz)
    out["fp64_flops"] = {"per_element": flops / ne, "achieved_tflops": flops / (kern_ms * 1e-3) / 1e12,
                         "peak_tflops": FP64_PEAK_TFLOPS,
                         "frac": flops / (kern_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                         "count": "node-pair block formulation, once per element (bench.algorithmic_flops_numeric)"}
    if copy_gbs is not None:
        out["copy_gbs_measured"] = copy_gbs
    del M
    mesh.close()
    ctx.close()
    extras = [x for x in a.extras.split(",") if x and x != "none"]

    def extra(key, name, fn):
        # an extra workload that fails is reported as such; the headline
        # line above is printed either way
        log("%s ..." % name)
        try:
            out[key] = fn()
        except Exception as e:  # noqa: BLE001
            log("%s failed: %r" % (name, e))
            out[key] = {"error": repr(e)}

    if "mixed" in extras:
        extra("c2_mixed", "C2 mixed", lambda: bench_c2_mixed(feasm, core, D, n))
        if "ms" in out["c2_mixed"]:
            out["c2_mixed"]["kernel_ratio_vs_fp64"] = el_ms / out["c2_mixed"]["ms"]
    if "vec" in extras:
        extra("vector", "load vector", lambda: bench_vec(feasm, core))
    if "lumped" in extras:
        extra("lumped", "lumped mass", lambda: bench_lumped(feasm, core))
    if "t10" in extras:
        extra("t10", "T10", lambda: bench_t10(feasm, core))
    if "batch" in extras:
        extra("batched", "batched", lambda: bench_batched(
            feasm, core, tune_db=os.path.join(ROOT, "finetools.jl_amd", "feasm", "autotune.txt")))
    if "served" in extras:
        extra("served", "served", lambda: bench_served(feasm, core))
    if "gram" in extras:
        extra("gram_probe", "element Gram, vector vs matrix cores", lambda: bench_gram(feasm))
    if "coo" in extras:
        extra("coo_path", "COO path", lambda: bench_coo(feasm, core, D, n))
        if "ms" in out["coo_path"]:
            out["coo_path"]["ms_over_resident_fused"] = out["coo_path"]["ms"] / el_ms
    if not a.no_cpu:
        extra("cpu_baseline", "cpu baseline", lambda: cpu_baseline(a.cpu_cells, 2, D))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
