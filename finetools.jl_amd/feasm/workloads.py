"""Seeded synthetic workloads of the batched and served assembly modes.

batched_specs: 256 independent H8 blocks, nx, ny, nz ~ U{4..64}; half 1-DOF
diffusion with k ~ U[0.5, 2], half 3-DOF linear elasticity with
E ~ U[1e3, 1e5], nu ~ U[0.2, 0.45]; interior nodes perturbed by +-0.1 h.

served_stream: Poisson arrivals at `rate` requests/s; each request is an H8
or T10 block (50/50) whose element count is log-normal (median 5e4, sigma
1.0, clipped to [1e3, 2e6]), diffusion or elasticity (50/50); 30 % of the
requests re-assemble a resident mesh with new material values.  Meshes come
from a seeded pool (the first request for a (mesh, DOFs per node) pair keeps
it resident; later new-mesh requests for the pair upload it again).
"""
import numpy as np

from . import _lib as L
from . import core


def iso_D(E, nu):
    """Isotropic 6x6 material matrix, Voigt order xx yy zz xy xz yz."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    for i in range(3):
        D[i, i] = lam + 2 * mu
        D[3 + i, 3 + i] = mu
    return D


def batched_specs(nmesh=256, seed=4):
    """[(cells (nx, ny, nz), form, ndn, coeff, perturbation seed)] for the batched workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nmesh):
        n = tuple(int(v) for v in rng.integers(4, 65, 3))
        if i % 2 == 0:
            out.append((n, L.DIFFUSION_ISO, 1, np.array([float(rng.uniform(0.5, 2.0))]), 1000 + i))
        else:
            D = iso_D(float(rng.uniform(1e3, 1e5)), float(rng.uniform(0.2, 0.45)))
            out.append((n, L.LIN_ELASTIC, 3, D.ravel(), 1000 + i))
    return out


def h8_block(n, amp_frac, seed):
    conn, xyz = core.gen_h8block(1.0, 1.0, 1.0, *n)
    h = 1.0 / max(n)
    return conn, core.gen_perturb(xyz, [0, 0, 0], [1, 1, 1], amp_frac * h, seed)


def served_pool(pool=48, seed=5):
    """[(etype, conn, xyz, (rule, order))] meshes the served requests draw from."""
    rng = np.random.default_rng(seed)
    meshes = []
    for i in range(pool):
        ne_target = float(np.clip(np.exp(np.log(5e4) + 1.0 * rng.standard_normal()), 1e3, 2e6))
        if i % 2 == 0:
            n = max(1, int(round(ne_target ** (1 / 3))))
            conn, xyz = h8_block((n, n, n), 0.1, 2000 + i)
            meshes.append((L.H8, conn, xyz, (L.RULE_GAUSS, 2)))
        else:
            n = max(1, int(round((ne_target / 6) ** (1 / 3))))
            conn, xyz = core.gen_t10block(1.0, 1.0, 1.0, n, n, n, 0.1 / n, 0.03 / n, 2000 + i)
            meshes.append((L.T10, conn, xyz, (L.RULE_TET, 4)))
    return meshes, rng


def served_stream(rng, pool_size, rate=200.0, duration=3.0):
    """[(arrival s, pool index, form, coeff, wants re-assembly)]."""
    nreq = int(rate * duration)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    stream = []
    for k in range(nreq):
        diff = rng.random() < 0.5
        form = L.DIFFUSION_ISO if diff else L.LIN_ELASTIC
        coeff = np.array([float(rng.uniform(0.5, 2.0))]) if diff else \
            iso_D(float(rng.uniform(1e3, 1e5)), float(rng.uniform(0.2, 0.45))).ravel()
        stream.append((float(arrivals[k]), int(rng.integers(0, pool_size)), form, coeff, bool(rng.random() < 0.3)))
    return stream


def run_served(stream, meshes, precision, workers, paced=True, sequential=False, hbm_fraction=0.6,
               digests=False, keep_results=None):
    """Submit the stream to an assembly server and wait for every request.

    paced: submit each request at its arrival time (else all at once);
    sequential: wait for each request before submitting the next, so that
    every request runs alone on the device.  Returns per-request records
    (index, elements, status, latency ms, digest or None) and the wall time.
    keep_results: optional set of stream indices whose CSC arrays are copied
    out (base 1) and returned in the records."""
    import time

    srv = core.Server(0, workers=workers, hbm_fraction=hbm_fraction)
    resident = {}  # (pool index, ndn) -> mesh id
    pending = []
    records = []
    t_start = time.perf_counter()

    def collect(k, rid, ne):
        st, lat, _ = srv.wait(rid)
        rec = {"index": k, "elements": ne, "status": st, "latency_ms": lat, "digest": None, "csc": None,
               "error": None}
        if st == 2:
            if digests:
                rec["digest"] = srv.result(rid).checksum()
            if keep_results is not None and k in keep_results:
                rec["csc"] = srv.result(rid).csc_arrays(1)
        else:
            try:
                srv.result(rid)
            except L.FeasmError as e:  # the request's own error text
                rec["error"] = str(e)
        srv.release(rid)
        records.append(rec)

    try:
        for k, (t_arr, pi, form, coeff, re) in enumerate(stream):
            while paced and time.perf_counter() - t_start < t_arr:
                time.sleep(0.0002)
            et, conn, xyz, rule = meshes[pi]
            ndn = 1 if form == L.DIFFUSION_ISO else 3
            key = (pi, ndn)
            rid = None
            if re and key in resident:
                try:
                    rid = srv.submit_reassembly(resident[key], form, rule[0], rule[1], coeff, precision)
                except L.FeasmError:
                    # its mesh is not resident (its first request failed, or it
                    # was evicted): the client sends the mesh again
                    del resident[key]
            if rid is None:
                rid, mid = srv.submit_mesh(et, conn, xyz, ndn, form, rule[0], rule[1], coeff, precision,
                                           keep=key not in resident, borrow=True)
                if key not in resident:
                    resident[key] = mid
            if sequential:
                collect(k, rid, conn.shape[0])
            else:
                pending.append((k, rid, conn.shape[0]))
        for k, rid, ne in pending:
            collect(k, rid, ne)
        wall = time.perf_counter() - t_start
    finally:
        srv.close()
    records.sort(key=lambda r: r["index"])
    return records, wall


def served_summary(records, wall):
    lats = np.array([r["latency_ms"] for r in records if r["status"] == 2])
    ne = sum(r["elements"] for r in records if r["status"] == 2)
    out = {"requests": len(records), "completed": int(len(lats)), "wall_s": wall, "elements_per_s": ne / wall,
           "p50_ms": float(np.percentile(lats, 50)), "p99_ms": float(np.percentile(lats, 99)),
           "mean_ms": float(lats.mean())}
    errors = sorted({r["error"] or "status %d" % r["status"] for r in records if r["status"] != 2})
    if errors:
        out["errors"] = errors[:5]
    return out
