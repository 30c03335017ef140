"""Benchmark of the DDSCAT Green-tensor MATVEC (and Krylov solve) on one MI355X.

Workload (default): rectangular block of 128 x 128 x 128 dipoles (2,097,152
dipoles, padded FFT grid 256^3 x 3 components, complex64), m = 1.5+0.1i,
x = k a_eff = 10 (kd = 0.126), incidence +x, E || y, LDR polarizabilities.
Synthetic, seeded inputs; everything resident in HBM before timing starts.

A "step" is one MATVEC y = A x on the compact solver vectors (5 FFT passes).
Prints one JSON line: MATVECs/s, per-pass roofline of the dominant kernel,
and the CPU baseline (the oracle's complex64 FFT MATVEC on host cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "ddscat-cuda_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0
METRIC = "matvecs/s + time-to-Q_ext at 1e-5 resid: batched mixed vs plain FP32, 1 GPU"


def pass_bytes(nat: int, nat0: int, n: tuple[int, int, int]):
    """Algorithmic HBM bytes per MATVEC pass (complex64, 3 components, pruned
    5-pass minimum; green octant read once in pass 3)."""
    nx, ny, nz = n
    green = 48 * (nx + 1) * (ny + 1) * (nz + 1)
    return [9 * nat * 8, 18 * nat * 8, 24 * nat * 8 + green, 18 * nat * 8, 18 * nat * 8]


def setup_workload(lib, edge=128, m=1.5 + 0.1j, xsize=10.0, seed=20261202):
    import numpy as np
    import torch
    from ddscat_hip import targets as T
    tg = T.block(edge)
    kd = T.kd_for_size_parameter(tg, xsize)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = np.array([0, 1, 0], complex)
    adia, aoff = T.operators(tg, m, kd, khat, e0)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    occ = tg.occ
    comp = lambda v: np.ascontiguousarray(v.reshape(3, -1)[:, occ].reshape(-1))
    rng = np.random.default_rng(seed)
    n0 = 3 * tg.nat0
    x = (rng.normal(size=n0) + 1j * rng.normal(size=n0)).astype(np.complex64)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.float32)).cuda()
    bufs = dict(x=dev(x), y=torch.empty(2 * n0, dtype=torch.float32, device="cuda"),
                adia=dev(comp(adia)), aoff=dev(comp(aoff)))
    torch.cuda.synchronize()
    return tg, kd, bufs, (x, adia, aoff)


def time_to_qext(lib, tg, kd, m=1.5 + 0.1j, method="qmrccg", tol=1e-5, maxit=1000, graph=True, n_orient=1,
                 seed=20261202, oracle=None):
    """Time to Q_ext of B = 2 n_orient right-hand sides, everything resident
    in HBM.  n_orient = 1: incidence +x with E || y and E || z; otherwise
    n_orient seeded SO(3) orientations x their two polarizations (per-RHS
    LDR operators):
      plain_fp32   - one single-RHS solve per RHS, FP32 reductions and
                     updates (PREC_FP32), then EVALQ;
      mixed        - the same with FP64 reductions / scalars (PREC_MIXED);
      batched_mixed- all B in one batched solve (one HIP graph replay per
                     iteration), then the batched EVALQ.
    Wall time from the first solve call to all Q values on the host.  A
    speed-up is reported only when every RHS of both modes converged;
    ``oracle`` (Q_ext, Q_abs per polarization of the FP64 oracle, n_orient
    = 1) adds the batched result's relative error against it."""
    import numpy as np
    import torch
    from ddscat_hip import PREC_MIXED, PREC_FP32
    from ddscat_hip import targets as T
    if n_orient == 1:
        khat = np.array([1.0, 0.0, 0.0])
        inc = [(khat, np.array([0, 1, 0], complex)), (khat, np.array([0, 0, 1], complex))]
    else:
        rots = T.random_rotations(np.random.default_rng(seed), n_orient)
        inc = [(R[:, 0], R[:, c].astype(complex)) for R in rots for c in (1, 2)]
    occ = tg.occ
    comp = lambda v: np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)).astype(np.complex64)
    bs, ads, aos, aks = [], [], [], []
    for khat, e0 in inc:
        ad, ao = T.operators(tg, m, kd, khat, e0)
        bs.append(T_incident_compact(tg, khat * kd, e0))
        ads.append(comp(ad)); aos.append(comp(ao)); aks.append((khat * kd).astype(np.float32))
    # flat float32 views of [B][3 NAT0] complex64 arrays; RHS r is row(t, r)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1).view(np.float32)).cuda()
    B = len(inc)
    bd, add, aod = dev(np.stack(bs)), dev(np.stack(ads)), dev(np.stack(aos))
    del bs, ads, aos
    aks = np.asarray(aks, np.float32)
    xd = torch.zeros_like(bd)
    n2 = bd.numel() // B
    assert bd.dim() == 1 and n2 == 2 * 3 * tg.nat0
    row = lambda t, r: t[r * n2:(r + 1) * n2]
    geo = math.pi * (3.0 * tg.nat0 / (4.0 * math.pi)) ** (2.0 / 3.0)
    # untimed one-iteration solves: allocate the Krylov workspaces
    lib.solve_device(method, row(bd, 0).data_ptr(), row(xd, 0).data_ptr(), row(add, 0).data_ptr(),
                     row(aod, 0).data_ptr(), maxit=1, tol=tol, check=False)
    lib.solve_batch(method, B, bd.data_ptr(), xd.data_ptr(), add.data_ptr(), aod.data_ptr(), maxit=1, tol=tol,
                    graph=graph, check=False)
    out = {}
    for mode, prec in (("plain_fp32", PREC_FP32), ("mixed", PREC_MIXED)):
        lib.set_precision(prec)
        xd.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        per = []
        for r in range(B):
            it, tole, nc, rc = lib.solve_device(method, row(bd, r).data_ptr(), row(xd, r).data_ptr(),
                                                row(add, r).data_ptr(), row(aod, r).data_ptr(), maxit=maxit,
                                                tol=tol, check=False)
            cabs, cext, _, rq = lib.evalq(aks[r], check=False)
            per.append(dict(iterations=it, matvecs=nc, resid=float(tole), rc=rc or rq, qext=cext / geo,
                            qabs=cabs / geo))
        dt = time.perf_counter() - t0
        out[mode] = dict(seconds=round(dt, 4), rhs=per if B <= 2 else None,
                         mean_iterations=float(np.mean([p["iterations"] for p in per])),
                         converged=all(p["resid"] <= 1.01 * tol and p["rc"] == 0 for p in per),
                         unconverged_rhs=[dict(rhs=r, iterations=p["iterations"], resid=p["resid"], rc=p["rc"])
                                          for r, p in enumerate(per) if not (p["resid"] <= 1.01 * tol and p["rc"] == 0)])
        out[mode]["_per"] = per
    lib.set_precision(PREC_MIXED)
    xd.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it, te, nc, rc = lib.solve_batch(method, B, bd.data_ptr(), xd.data_ptr(), add.data_ptr(), aod.data_ptr(),
                                     maxit=maxit, tol=tol, graph=graph, check=False)
    q = lib.evalq_batch(B, bd.data_ptr(), xd.data_ptr(), add.data_ptr(), aod.data_ptr(), aks, check=False)
    dt = time.perf_counter() - t0
    per_b = [dict(iterations=int(it[r]), matvecs=int(nc[r]), resid=float(te[r]), qext=q[r, 1] / geo,
                  qabs=q[r, 0] / geo) for r in range(B)]
    out["batched_mixed"] = dict(seconds=round(dt, 4), graph=graph, rc=rc, rhs=per_b if B <= 2 else None,
                                mean_iterations=float(np.mean(it)),
                                converged=bool(rc == 0 and np.all(te <= 1.01 * tol)),
                                unconverged_rhs=[dict(rhs=r, iterations=int(it[r]), resid=float(te[r]))
                                                 for r in range(B) if not te[r] <= 1.01 * tol])
    f, mx, bm = out["plain_fp32"], out["mixed"], out["batched_mixed"]
    qdiff = max(abs(a["qext"] - b["qext"]) / abs(b["qext"]) for a, b in zip(f.pop("_per"), per_b))
    mx.pop("_per")
    ok = f["converged"] and bm["converged"]
    res = dict(method=method, tol=tol, nrhs=B,
               incidence="+x, E||y and E||z" if n_orient == 1 else f"{n_orient} seeded SO(3) orientations x 2",
               **out,
               speedup_batched_mixed_vs_plain_fp32=round(f["seconds"] / bm["seconds"], 3) if ok else None,
               speedup_mixed_vs_plain_fp32=round(f["seconds"] / mx["seconds"], 3) if ok and mx["converged"] else None,
               qext_max_rel_diff_fp32_vs_batched_mixed=qdiff)
    if oracle is not None and n_orient == 1:
        res["batched_vs_fp64_oracle"] = dict(
            qext_max_rel_err=max(abs(p["qext"] - oracle["qext"][k]) / abs(oracle["qext"][k]) for k, p in enumerate(per_b)),
            qabs_max_rel_err=max(abs(p["qabs"] - oracle["qabs"][k]) / abs(oracle["qabs"][k]) for k, p in enumerate(per_b)),
            oracle="complex128 FFT MATVEC + QMRCCG to 1e-8 (tests/golden/config_block128.npz)")
    del bd, add, aod, xd
    torch.cuda.empty_cache()
    return res


def batched_matvec_scaling(lib, tg, bufs, sizes=(1, 2, 4, 8, 16), reps=10):
    """Milliseconds per right-hand side of one batched MATVEC launch sequence
    (ddscat_hip_apply_batch, compact vectors, per-RHS operators) at B = 1, 2,
    4, ... right-hand sides of the prepared target.  The B copies of each
    pass-3 column unit run back to back on one XCD, so the Green octant is read
    from HBM about once per launch, not once per RHS."""
    import torch
    out = []
    st = torch.cuda.ExternalStream(lib.stream())
    n = bufs["x"].numel() // 2
    for B in sizes:
        x = bufs["x"].view(torch.complex64).repeat(B)
        ad = bufs["adia"].view(torch.complex64).repeat(B)
        ao = bufs["aoff"].view(torch.complex64).repeat(B)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(2):
            lib.apply_batch("N", B, x.data_ptr(), y.data_ptr(), ad.data_ptr(), ao.data_ptr())
        lib.sync()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(st):
            ev0.record(st)
            for _ in range(reps):
                lib.apply_batch("N", B, x.data_ptr(), y.data_ptr(), ad.data_ptr(), ao.data_ptr())
            ev1.record(st)
        ev1.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        out.append(dict(nrhs=B, ms_per_launch=round(ms, 3), ms_per_rhs=round(ms / B, 4)))
        del x, ad, ao, y
        torch.cuda.empty_cache()
    assert n == 3 * tg.nat0
    return out


def batched_small_target(lib, n_orient=16, method="pbcgs2", tol=1e-5, maxit=300, seed=20261202, reps=5):
    """Sphere of diameter 32 dipoles (17,256 dipoles), m = 1.33+0.01i, x = 2,
    n_orient incidence directions x 2 polarizations, PBCGS2 at tol 1e-5:
    the small-target workload where single-RHS solves are launch-bound.
    plain_fp32: one single-RHS device solve per RHS (PREC_FP32);
    batched_mixed: all RHS in one batched solve (PREC_MIXED, graph replay).
    Each is timed from the first solve call to all Q_ext on the host, reps
    times from a zero initial guess (a few ms each); the medians are
    reported, with every repetition."""
    import numpy as np
    import torch
    from ddscat_hip import PREC_MIXED, PREC_FP32
    from ddscat_hip import targets as T
    tg = T.sphere(32)
    kd = T.kd_for_size_parameter(tg, 2.0)
    rng = np.random.default_rng(seed)
    dirs = rng.normal(size=(n_orient, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    occ = tg.occ
    comp = lambda v: np.ascontiguousarray(np.asarray(v).reshape(3, -1)[:, occ].reshape(-1)).astype(np.complex64)
    bs, ads, aos, aks = [], [], [], []
    for khat in dirs:
        for e0 in T.polarization_pair(khat):
            ad, ao = T.operators(tg, 1.33 + 0.01j, kd, khat, e0)
            bs.append(comp(T.incident_field(tg, khat * kd, e0, tg.centroid_x0())))
            ads.append(comp(ad)); aos.append(comp(ao)); aks.append(khat * kd)
    B = len(bs)
    aks = np.asarray(aks, np.float32)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, aks[0])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1).view(np.float32)).cuda()
    bd, add, aod = dev(np.stack(bs)), dev(np.stack(ads)), dev(np.stack(aos))
    xd = torch.zeros_like(bd)
    n2 = bd.numel() // B
    assert bd.dim() == 1 and n2 == 2 * 3 * tg.nat0
    row = lambda t, r: t[r * n2:(r + 1) * n2]
    geo = math.pi * (3.0 * tg.nat0 / (4.0 * math.pi)) ** (2.0 / 3.0)
    # warm-up (workspace allocation, graph code paths)
    lib.solve_device(method, row(bd, 0).data_ptr(), row(xd, 0).data_ptr(), row(add, 0).data_ptr(),
                     row(aod, 0).data_ptr(), maxit=2, tol=tol, check=False)
    lib.solve_batch(method, B, bd.data_ptr(), xd.data_ptr(), add.data_ptr(), aod.data_ptr(), maxit=2, tol=tol,
                    check=False)
    ts_single, ts_batch = [], []
    for _ in range(reps):
        lib.set_precision(PREC_FP32)
        xd.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        q_single, it_single = [], []
        for r in range(B):
            it, tole, nc, rc = lib.solve_device(method, row(bd, r).data_ptr(), row(xd, r).data_ptr(),
                                                row(add, r).data_ptr(), row(aod, r).data_ptr(), maxit=maxit,
                                                tol=tol, check=False)
            cabs, cext, _, _ = lib.evalq(aks[r], check=False)
            q_single.append(cext / geo)
            it_single.append(it)
        ts_single.append(time.perf_counter() - t0)
        lib.set_precision(PREC_MIXED)
        xd.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, te, nc, rc = lib.solve_batch(method, B, bd.data_ptr(), xd.data_ptr(), add.data_ptr(), aod.data_ptr(),
                                         maxit=maxit, tol=tol, check=False)
        q = lib.evalq_batch(B, bd.data_ptr(), xd.data_ptr(), add.data_ptr(), aod.data_ptr(), aks, check=False)
        ts_batch.append(time.perf_counter() - t0)
    t_single, t_batch = float(np.median(ts_single)), float(np.median(ts_batch))
    qb = q[:, 1] / geo
    qs = np.asarray(q_single)
    return dict(workload=f"sphere D=32 ({tg.nat0} dipoles), m=1.33+0.01i, x=2, {n_orient} orientations x 2 "
                         f"polarizations = {B} RHS, {method.upper()} tol {tol}",
                plain_fp32=dict(seconds=round(t_single, 4), rhs_per_s=round(B / t_single, 2),
                                mean_iterations=float(np.mean(it_single)),
                                seconds_each=[round(v, 4) for v in ts_single]),
                batched_mixed=dict(seconds=round(t_batch, 4), rhs_per_s=round(B / t_batch, 2),
                                   mean_iterations=float(np.mean(it)), rc=int(rc),
                                   seconds_each=[round(v, 4) for v in ts_batch]),
                speedup_batched_mixed_vs_plain_fp32=round(t_single / t_batch, 3),
                qext_max_rel_diff=float(np.max(np.abs(qb - qs) / np.abs(qs))),
                qext_mean=float(np.mean(qb)))


def batched_targets(lib, n_targets=48, n_orient=8, kd=0.1, method="pbcgs2", tol=1e-5, maxit=500,
                    single_sample=3, seed=20261202, group=8):
    """Uneven batch: n_targets seeded unions of 1-4 ellipsoids (semi-axes
    U[6,40] dipoles, grids 24-96 per axis), m = U[1.2,2.0] + i U[0,0.3] per
    target, n_orient orientations uniform on SO(3) x 2 polarizations per
    target, PBCGS2 at tol 1e-5, mixed precision.  Targets go through in
    groups of `group`: each target of a group is prepared in its own target
    slot (device Green build) with its inputs built on the device
    (incident fields + LDR operators of all its RHS), then one
    ddscat_hip_solve_targets call solves the whole group (one stream and host
    thread per target, graph replay), then the batched EVALQ per slot.  The
    same targets one after another (solve_batch per target) and plain-FP32
    single-RHS solves are timed on the first `single_sample` targets."""
    import numpy as np
    import torch
    from ddscat_hip import PREC_MIXED, PREC_FP32
    from ddscat_hip import targets as T
    rng = np.random.default_rng(seed)
    tgts = []
    for _ in range(n_targets):
        tg = T.ellipsoid_union(rng)
        m = complex(rng.uniform(1.2, 2.0), rng.uniform(0.0, 0.3))
        tgts.append((tg, m, T.random_rotations(rng, n_orient)))

    def inputs(tg, m, rots):
        aks, e0s = [], []
        for R in rots:
            for e0 in (R[:, 1], R[:, 2]):
                aks.append(R[:, 0] * kd); e0s.append(e0.astype(np.complex64))
        aks = np.asarray(aks, np.float32)
        B, n = len(aks), 3 * tg.nat0
        eps = torch.full((tg.nat0,), m * m, dtype=torch.complex64, device="cuda")
        b = torch.empty((B, n), dtype=torch.complex64, device="cuda")
        ad, ao = torch.empty_like(b), torch.empty_like(b)
        lib.setup_rhs(aks, np.asarray(e0s), tg.centroid_x0(), eps.data_ptr(), b.data_ptr(), ad.data_ptr(),
                      ao.data_ptr())
        return dict(aks=aks, B=B, b=b, ad=ad, ao=ao, x=torch.zeros_like(b))

    lib.set_precision(PREC_MIXED)
    rhs_total, iters, sizes = 0, [], []
    stage = dict(prepare=0.0, setup=0.0, solve=0.0, evalq=0.0)
    for g0 in range(0, n_targets, group):
        grp = tgts[g0:g0 + group]
        slots = list(range(1, len(grp) + 1))
        bufs = []
        torch.cuda.synchronize()
        for s_, (tg, m, rots) in zip(slots, grp):
            lib.select_target(s_)
            t0 = time.perf_counter()
            lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, rots[0][:, 0] * kd)
            t1 = time.perf_counter()
            bufs.append(inputs(tg, m, rots))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            stage["prepare"] += t1 - t0
            stage["setup"] += t2 - t1
            sizes.append((tg.nx, tg.ny, tg.nz, tg.nat0))
        t0 = time.perf_counter()
        per, rc = lib.solve_targets(method, slots, [d["B"] for d in bufs], [d["b"].data_ptr() for d in bufs],
                                    [d["x"].data_ptr() for d in bufs], [d["ad"].data_ptr() for d in bufs],
                                    [d["ao"].data_ptr() for d in bufs], maxit=maxit, tol=tol, check=False)
        t1 = time.perf_counter()
        for s_, d in zip(slots, bufs):
            lib.select_target(s_)
            lib.evalq_batch(d["B"], d["b"].data_ptr(), d["x"].data_ptr(), d["ad"].data_ptr(), d["ao"].data_ptr(),
                            d["aks"], check=False)
        t2 = time.perf_counter()
        stage["solve"] += t1 - t0
        stage["evalq"] += t2 - t1
        for d, (it, te, nc) in zip(bufs, per):
            rhs_total += d["B"]
            iters.extend(int(v) for v in it)
        for s_ in slots:
            lib.select_target(s_)
            lib.release()
        lib.select_target(0)
        del bufs
        torch.cuda.empty_cache()
    t_batch = sum(stage.values())
    # the first targets, one after another: batched solve per target, and
    # plain-FP32 single-RHS solves
    t_seq = t_single = 0.0
    rhs_single = 0
    for tg, m, rots in tgts[:single_sample]:
        lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, rots[0][:, 0] * kd)
        d = inputs(tg, m, rots)
        lib.set_precision(PREC_MIXED)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lib.solve_batch(method, d["B"], d["b"].data_ptr(), d["x"].data_ptr(), d["ad"].data_ptr(), d["ao"].data_ptr(),
                        maxit=maxit, tol=tol, check=False)
        lib.evalq_batch(d["B"], d["b"].data_ptr(), d["x"].data_ptr(), d["ad"].data_ptr(), d["ao"].data_ptr(),
                        d["aks"], check=False)
        t_seq += time.perf_counter() - t0
        lib.set_precision(PREC_FP32)
        d["x"].zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(d["B"]):
            lib.solve_device(method, d["b"][r].data_ptr(), d["x"][r].data_ptr(), d["ad"][r].data_ptr(),
                             d["ao"][r].data_ptr(), maxit=maxit, tol=tol, check=False)
            lib.evalq(d["aks"][r], check=False)
        t_single += time.perf_counter() - t0
        rhs_single += d["B"]
        lib.set_precision(PREC_MIXED)
        del d
    lib.release()
    return dict(workload=f"{n_targets} seeded ellipsoid unions (grids {min(min(s[:3]) for s in sizes)}-"
                         f"{max(max(s[:3]) for s in sizes)} per axis, {sum(s[3] for s in sizes)} dipoles), "
                         f"{n_orient} SO(3) orientations x 2 polarizations = {rhs_total} RHS, kd={kd}, "
                         f"{method.upper()} tol {tol}, groups of {group} targets per ddscat_hip_solve_targets call",
                batched_mixed=dict(seconds=round(t_batch, 3), rhs_per_s=round(rhs_total / t_batch, 1),
                                   solve_rhs_per_s=round(rhs_total / stage["solve"], 1),
                                   stage_seconds={k: round(v, 3) for k, v in stage.items()},
                                   mean_iterations=float(np.mean(iters)), max_iterations=int(np.max(iters))),
                solve_comparison=dict(sample_targets=single_sample, rhs=rhs_single,
                                      plain_fp32_single_seconds=round(t_single, 3),
                                      batched_mixed_one_target_at_a_time_seconds=round(t_seq, 3),
                                      speedup_batched_mixed_vs_plain_fp32=round(t_single / t_seq, 3)
                                      if t_seq > 0 else None))


def large_solve(lib, edge=256, xsize=20.0):
    """Time to Q_ext on a 256^3 block (16.8 M dipoles, padded grid 512^3 x 3),
    m = 1.5+0.1i, x = 20 (kd = 0.126, the 128^3 workload's lattice
    resolution), QMRCCG 1e-5: plain FP32, mixed, batched mixed."""
    from ddscat_hip import targets as T
    tg = T.block(edge)
    kd = T.kd_for_size_parameter(tg, xsize)
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (kd, 0.0, 0.0))
    r = time_to_qext(lib, tg, kd)
    r["workload"] = (f"block {edge}^3 ({tg.nat0} dipoles), padded {2 * edge}^3 x 3, m=1.5+0.1i, x={xsize} "
                     f"(kd={kd:.4f}), both polarizations")
    r["fft_kernels"] = {f"pass{p}": lib.info().get(f"kernel_pass{p}") for p in range(1, 6)}
    lib.release()
    return r


def large_target(lib, edge=384, kd=0.25, m=1.7 + 0.05j, steps=10, warmup=2, seed=20261202):
    """The large single target: sphere of diameter ``edge`` dipoles (384:
    29.6 M dipoles, padded grid 768^3 x 3 complex64), m = 1.7+0.05i, kd = 0.25,
    isotropic LDR operators.  MATVEC/s on compact device vectors for the
    FFT3D and SLICES backends (768-point axes: 32 x 24 register shapes for the
    single-component passes, 3 x 256 decimation for the Green pass, the
    LDS-staged kernel for FFT3D's pass 5), with the algorithmic bytes
    696 NAT + 48 (N+1)^3 per MATVEC."""
    import numpy as np
    import torch
    from ddscat_hip import targets as T, BACKEND_SLICES, BACKEND_FFT3D
    tg = T.sphere(edge)
    n0 = 3 * tg.nat0
    alpha_inv = (1.0 / T.ldr_alpha(np.array([m * m]), kd, (1, 0, 0), (0, 1, 0)))[0]
    x = torch.complex(torch.randn(n0, generator=torch.Generator().manual_seed(seed)),
                      torch.randn(n0, generator=torch.Generator().manual_seed(seed + 1))).to(torch.complex64).cuda()
    y = torch.empty_like(x)
    adia = torch.full((n0,), complex(alpha_inv), dtype=torch.complex64, device="cuda")
    aoff = torch.zeros(n0, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    pb = sum(pass_bytes(tg.nat, tg.nat0, (tg.nx, tg.ny, tg.nz)))
    out = dict(workload=f"sphere D={edge} ({tg.nat0} dipoles, grid {tg.nx}^3, padded {2 * tg.nx}^3 x 3), "
                        f"m=1.7+0.05i, kd={kd}, compact vectors", matvec_bytes=pb)
    for name, backend in (("fft3d", BACKEND_FFT3D), ("slices", BACKEND_SLICES)):
        lib.set_backend(backend)
        t0 = time.perf_counter()
        lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, (kd, 0.0, 0.0))
        prep_s = time.perf_counter() - t0
        st = torch.cuda.ExternalStream(lib.stream())
        args = (x.data_ptr(), y.data_ptr(), adia.data_ptr(), aoff.data_ptr())
        for _ in range(warmup):
            lib.apply_device("N", *args, layout=1)
        lib.sync()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(st):
            ev0.record(st)
            for _ in range(steps):
                lib.apply_device("N", *args, layout=1)
            ev1.record(st)
        ev1.synchronize()
        ms = ev0.elapsed_time(ev1) / steps
        info = lib.info()
        leg = dict(ms_per_matvec=round(ms, 3), matvecs_per_s=round(1000.0 / ms, 2),
                   algorithmic_gbs=round(pb / (ms * 1e-3) / 1e9, 1),
                   frac_of_hbm_peak=round(pb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                   prepare_seconds=round(prep_s, 3), green_build_ms=round(info["green_build_us"] / 1000.0, 1))
        if name == "fft3d":
            passes = lib.time_matvec("N", *args, layout=1, reps=5)
            leg["pass_ms"] = [round(v, 3) for v in passes[:5]]
        out[name] = leg
        lib.release()
    lib.set_backend(BACKEND_FFT3D)
    del x, y, adia, aoff
    torch.cuda.empty_cache()
    return out


def large_sphere_solve(lib, edge=384, kd=0.25, m=1.7 + 0.05j, tol=1e-5, maxit=1500, qmr_budget=300,
                       backend="fft3d", far_field=True):
    """The large single-target solve: sphere D=edge (384: 29.6 M dipoles,
    padded 768^3 x 3), m = 1.7+0.05i, kd = 0.25, incidence +x with E || y and
    E || z, inputs built on the device (ddscat_hip_setup_rhs), on the FFT3D or
    SLICES backend (device_gb: device memory in use at the end of the solve,
    the library's and the leg's own arrays).
      pbcgs2: batched mixed PBCGS2 to tol, then the batched EVALQ: time to
              Q_ext; both polarizations' Q_abs must agree (the lattice's
              cubic symmetry);
      far_field: the S-matrix on 181 theta x 4 phi directions (amplitudes and
              Mueller matrices of the device solutions);
      qmrccg: the configured solver, batched mixed QMRCCG, for a budget of
              qmr_budget iterations: the residual it reaches."""
    import numpy as np
    import torch
    from ddscat_hip import targets as T, scattering as DS, PREC_MIXED, BACKEND_FFT3D, BACKEND_SLICES
    tg = T.sphere(edge)
    khat = np.array([1.0, 0.0, 0.0])
    e0s = np.array([[0, 1, 0], [0, 0, 1]], np.complex64)
    aks = np.tile((khat * kd).astype(np.float32), (2, 1))
    x0 = tg.centroid_x0()
    lib.set_precision(PREC_MIXED)
    lib.release()
    lib.set_backend(BACKEND_SLICES if backend == "slices" else BACKEND_FFT3D)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    t0 = time.perf_counter()
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    prep = time.perf_counter() - t0
    n = 3 * tg.nat0
    eps = torch.full((tg.nat0,), complex(m) ** 2, dtype=torch.complex64, device="cuda")
    b = torch.empty((2, n), dtype=torch.complex64, device="cuda")
    ad, ao = torch.empty_like(b), torch.empty_like(b)
    lib.setup_rhs(aks, e0s, x0, eps.data_ptr(), b.data_ptr(), ad.data_ptr(), ao.data_ptr())
    geo = math.pi * tg.a_eff() ** 2
    out = dict(workload=f"sphere D={edge} ({tg.nat0} dipoles, padded {2 * edge}^3 x 3), m=1.7+0.05i, kd={kd} "
                        f"(x={kd * tg.a_eff():.1f}), +x incidence, E||y and E||z, {backend.upper()}",
               prepare_seconds=round(prep, 3))
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it, te, nc, rc = lib.solve_batch("pbcgs2", 2, b.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                     maxit=maxit, tol=tol, check=False)
    q = lib.evalq_batch(2, b.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(), aks, check=False)
    dt = time.perf_counter() - t0
    qabs, qext = q[:, 0] / geo, q[:, 1] / geo
    torch.cuda.synchronize()
    out["device_gb"] = round((free0 - torch.cuda.mem_get_info()[0]) / 1e9, 2)
    out["pbcgs2"] = dict(seconds_to_qext=round(dt, 3), iterations=it.tolist(), resid=te.tolist(), rc=int(rc),
                         converged=bool(rc == 0 and np.all(te <= 1.01 * tol)), qext=qext.tolist(), qabs=qabs.tolist(),
                         qabs_rel_diff_between_polarizations=float(abs(qabs[0] - qabs[1]) / abs(qabs[0])))
    if not far_field:
        del b, ad, ao, x, eps
        torch.cuda.empty_cache()
        lib.release()
        lib.set_backend(BACKEND_FFT3D)
        return out
    theta = np.radians(np.arange(181.0))
    phis = np.radians([0.0, 90.0, 180.0, 270.0])
    th, ph = np.repeat(theta, 4), np.tile(phis, 181)
    ks, epar, eperp = DS.directions(khat, e0s[0], th, ph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f, M = lib.scat_matrix(x.data_ptr(), 1, aks[:1], e0s[None], kd * ks, epar, eperp, x0)
    out["far_field"] = dict(directions=int(len(th)), seconds=round(time.perf_counter() - t0, 3),
                            s11_forward=float(M[0, 0, 0, 0]), s11_backward=float(M[0, 4 * 180, 0, 0]))
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it, te, nc, rc = lib.solve_batch("qmrccg", 2, b.data_ptr(), x.data_ptr(), ad.data_ptr(), ao.data_ptr(),
                                     maxit=qmr_budget, tol=tol, check=False)
    out["qmrccg"] = dict(iteration_budget=qmr_budget, seconds=round(time.perf_counter() - t0, 3),
                         iterations=it.tolist(), resid=te.tolist(),
                         replacements=[int(np.sum(lib.solve_log(r)["checks"][:, 3] == 2)) for r in range(2)])
    del b, ad, ao, x, eps
    torch.cuda.empty_cache()
    lib.release()
    lib.set_backend(BACKEND_FFT3D)
    return out


def adjoint_design(lib, edge=64, xsize=6.0, tol=1e-5, h=1e-2, fd_tol=1e-6, seed=20261202):
    """The inverse-design workload: a 64^3 block (262,144 dipoles) with a
    smooth random permittivity (eps_j = U[1,4] smoothed by a Gaussian of
    sigma = 3 dipoles, + 0.05i), J = mean Q_abs of both polarizations at +z
    incidence, x = 6.  Times the forward batched solve and the autograd
    backward (adjoint solve) at tolerance tol and checks dJ/d eps along a
    smooth random direction against a central difference with step h.  The
    two difference solves run at fd_tol = 1e-6, near the floor an FP32
    MATVEC reaches on this target; with h = 1e-2 their noise in J over 2h
    stays near 1e-4 of the derivative (at tol = 1e-5 and h = 1e-3 it would be
    several per cent)."""
    import numpy as np
    import torch
    import scipy.ndimage as ndi
    from ddscat_hip import targets as T
    from ddscat_hip.adjoint import design_q_abs
    tg = T.block(edge)
    kd = T.kd_for_size_parameter(tg, xsize)
    khat = np.array([0.0, 0.0, 1.0])
    pols = list(T.polarization_pair(khat))
    rng = np.random.default_rng(seed)
    field = ndi.gaussian_filter(rng.uniform(1.0, 4.0, size=(edge, edge, edge)), sigma=3.0)
    eps0 = field.reshape(-1)[tg.occ] + 0.05j
    v = ndi.gaussian_filter(rng.normal(size=(edge, edge, edge)), sigma=3.0).reshape(-1)[tg.occ]
    v = (v + 1j * np.roll(v, 7)) / np.abs(v).max()
    lib.prepare_green(tg.nx, tg.ny, tg.nz, tg.iocc, khat * kd)
    # warm-up forward and backward (graph capture, workspaces): an inverse-design
    # loop takes many gradient steps, the timed one is a steady-state step
    ew = torch.tensor(eps0, dtype=torch.complex128, requires_grad=True)
    design_q_abs(lib, tg, ew, kd, khat, pols, tol=tol).backward()
    eps = torch.tensor(eps0, dtype=torch.complex128, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = design_q_abs(lib, tg, eps, kd, khat, pols, tol=tol)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ana = float(np.sum(np.real(np.conj(eps.grad.numpy()) * v)))
    with torch.no_grad():
        jp = design_q_abs(lib, tg, torch.tensor(eps0 + h * v), kd, khat, pols, tol=fd_tol, maxit=1500).item()
        jm = design_q_abs(lib, tg, torch.tensor(eps0 - h * v), kd, khat, pols, tol=fd_tol, maxit=1500).item()
    num = (jp - jm) / (2 * h)
    lib.release()
    return dict(workload=f"block {edge}^3 ({tg.nat0} dipoles), smooth random eps in [1,4] + 0.05i, "
                         f"Q_abs at +z, both polarizations, x={xsize} (kd={kd:.4f}), QMRCCG tol {tol}",
                q_abs=round(J.item(), 6), forward_seconds=round(t1 - t0, 4), backward_seconds=round(t2 - t1, 4),
                gradient_seconds=round(t2 - t0, 4),
                directional_derivative=dict(adjoint=ana, central_difference=num, step=h, fd_solver_tol=fd_tol,
                                            rel_diff=abs(num - ana) / abs(ana)))


def T_incident_compact(tg, kvec, e0):
    import numpy as np
    from ddscat_hip import targets as T
    e = T.incident_field(tg, kvec, e0, tg.centroid_x0())
    return np.ascontiguousarray(e.reshape(3, -1)[:, tg.occ].reshape(-1))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_small_target(cores, budget_s=5.0):
    """The small workload on the host: sphere D=32 (17,256 dipoles), m =
    1.33+0.01i, x = 2, incidence +x, E || y: the complex64 CPU MATVEC rate and
    the CPU PBCGS2 time to tol 1e-5 with that MATVEC (oracle restatements)."""
    import numpy as np
    from oracle import ddscat_oracle as O
    from oracle import solvers as S
    from ddscat_hip import targets as T
    tg = T.sphere(32)
    kd = T.kd_for_size_parameter(tg, 2.0)
    khat = np.array([1.0, 0.0, 0.0])
    e0 = np.array([0, 1, 0], complex)
    adia, aoff = T.operators(tg, 1.33 + 0.01j, kd, khat, e0)
    goct = O.green_octant_separable(tg.nx, tg.ny, tg.nz, float(np.float32(kd)))
    mv = O.CpuMatvecFP32(goct, tg.nx, tg.ny, tg.nz, workers=cores)
    occ3 = np.flatnonzero(np.tile(tg.iocc != 0, 3))
    full = np.zeros(3 * tg.nat, np.complex64)

    def mvc(cw, v):
        full[occ3] = v
        return mv.apply(cw, full, adia, aoff, tg.iocc)[occ3].astype(np.complex128)

    b = T.incident_field(tg, khat * kd, e0, tg.centroid_x0()).astype(np.complex128)[occ3]
    mvc("N", b)
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < budget_s and n < 200:
        mvc("N", b)
        n += 1
    rate = n / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    res = S.pbcgs2(mvc, b, maxit=300, tol=1e-5)
    solve_s = time.perf_counter() - t0
    return dict(matvecs_per_s=round(rate, 2), pbcgs2_seconds=round(solve_s, 3),
                pbcgs2_iterations=int(res[1]), sample=f"{n} MATVECs, one PBCGS2 solve (both on {cores} workers)")


def cpu_baseline(tg, kd, host, budget_s=20.0):
    """Oracle FP32 (complex64) FFT MATVEC on the box's host cores (scipy.fft
    workers), 128^3 block; plus the small sphere's MATVEC rate and PBCGS2
    time to solution."""
    import numpy as np
    from oracle import ddscat_oracle as O
    x, adia, aoff = host
    xf = np.zeros(3 * tg.nat, np.complex64)
    xf.reshape(3, -1)[:, tg.occ] = x.reshape(3, -1)
    goct = O.green_octant_separable(tg.nx, tg.ny, tg.nz, float(np.float32(kd)))
    # the GPU box's host share is 16 cores whatever os.cpu_count() reports
    cores = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))
    mv = O.CpuMatvecFP32(goct, tg.nx, tg.ny, tg.nz, workers=cores)
    mv.apply("N", xf, adia, aoff)
    t0 = time.perf_counter()
    n = 0
    while True:
        mv.apply("N", xf, adia, aoff)
        n += 1
        dt = time.perf_counter() - t0
        if dt > budget_s or n >= 50:
            break
    return dict(value=round(n / dt, 4), unit="matvecs/s", cores=cores, kind="port",
                sample=f"{n} complex64 FFT MATVECs of the {tg.nx}^3 block (oracle CpuMatvecFP32: scipy.fft "
                       f"workers={cores} on the full padded grid, Green expanded once) in {dt:.1f} s",
                cpu_model=_cpu_model(), nproc=os.cpu_count(), fft_backend="scipy.fft (pocketfft), complex64",
                sphere32=cpu_small_target(cores))


TRAFFIC_FILE = ROOT / "profiles" / "matvec_fft3d_block128_traffic.json"


def measured_traffic(args, dom):
    """HBM bytes per launch of the dominant pass (FETCH_SIZE + WRITE_SIZE,
    calibrated on a copy of known size) from the committed rocprofv3 counter
    summary of this workload; None for workloads it does not cover."""
    if args.backend != "fft3d" or args.edge != 128 or not TRAFFIC_FILE.exists():
        return None
    passes = list(json.loads(TRAFFIC_FILE.read_text())["passes"].values())
    p = passes[dom]
    return p["hbm_read_bytes"] + p["hbm_write_bytes"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--edge", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-Q_ext leg")
    ap.add_argument("--solver", default="qmrccg")
    ap.add_argument("--targets", type=int, default=48, help="targets in the uneven batched leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--backend", choices=["fft3d", "slices"], default="fft3d")
    ap.add_argument("--pmc-calib", action="store_true", help="run one torch copy of known size (rocprof scale)")
    ap.add_argument("--no-large", action="store_true", help="skip the 384^3 sphere MATVEC leg")
    ap.add_argument("--batch-study", action="store_true",
                    help="time to Q_ext at 2 and 16 right-hand sides on the 128^3 and 256^3 blocks, then exit")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the 384^3 sphere MATVEC leg (FFT3D and SLICES) and print its JSON")
    ap.add_argument("--large-solve-study", action="store_true",
                    help="the 384^3 sphere's batched PBCGS2 to 1e-5 on FFT3D and on SLICES (time, memory), then exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (single-GPU box); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from ddscat_hip import Library, BACKEND_SLICES, BACKEND_FFT3D
    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    lib = Library()
    if args.large_only:
        print(json.dumps(dict(large_target=large_target(lib, steps=args.steps, warmup=args.warmup))), flush=True)
        return
    if args.large_solve_study:
        res = {}
        for be in ("fft3d", "slices"):
            res[be] = large_sphere_solve(lib, qmr_budget=0, backend=be, far_field=False)
            print(json.dumps({be: res[be]}), flush=True)
        print(json.dumps(dict(large_solve_study=res)), flush=True)
        return
    if args.batch_study:
        from ddscat_hip import targets as T
        res = {}
        for edge, xs in ((128, 10.0), (256, 20.0)):
            tgb = T.block(edge)
            kdb = T.kd_for_size_parameter(tgb, xs)
            lib.prepare_green(tgb.nx, tgb.ny, tgb.nz, tgb.iocc, (kdb, 0.0, 0.0))
            for no in (1, 8):
                res[f"block{edge}_rhs{2 * no}"] = time_to_qext(lib, tgb, kdb, method=args.solver, n_orient=no)
                print(json.dumps({f"block{edge}_rhs{2 * no}": res[f"block{edge}_rhs{2 * no}"]}), flush=True)
            lib.release()
        print(json.dumps(dict(batch_study=res)), flush=True)
        return
    lib.set_backend(BACKEND_SLICES if args.backend == "slices" else BACKEND_FFT3D)
    tg, kd, bufs, host = setup_workload(lib, edge=args.edge)
    info = lib.info()
    st = torch.cuda.ExternalStream(lib.stream())
    xp, yp, ap_, op = (bufs[k].data_ptr() for k in ("x", "y", "adia", "aoff"))

    if args.pmc_calib:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()

    for _ in range(args.warmup):
        lib.apply_device("N", xp, yp, ap_, op, layout=1)
    lib.sync()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(st):
        ev0.record(st)
        for _ in range(args.steps):
            lib.apply_device("N", xp, yp, ap_, op, layout=1)
        ev1.record(st)
    ev1.synchronize()
    lib.sync()
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    value = 1000.0 / ms_step

    passes = lib.time_matvec("N", xp, yp, ap_, op, layout=1, reps=max(10, min(args.steps, 100)))
    pb = pass_bytes(tg.nat, tg.nat0, (tg.nx, tg.ny, tg.nz))
    per_pass = []
    for i in range(5):
        gbs = pb[i] / (passes[i] * 1e-3) / 1e9 if passes[i] > 0 else float("nan")
        per_pass.append(dict(kernel=["pass1_xfwd", "pass2_yfwd", "pass3_z_green", "pass4_yinv",
                                     "pass5_xinv_finish"][i], ms=round(passes[i], 5), bytes=pb[i],
                             gbs=round(gbs, 1)))
    total_bytes = sum(pb)  # compact layout: no IOCC read
    if args.backend == "fft3d":
        dom = max(range(5), key=lambda i: passes[i])
        dom_bytes, dom_ms, dom_name = pb[dom], passes[dom], per_pass[dom]["kernel"]
    else:
        # SLICES interleaves its kernels over two streams: whole-MATVEC figure
        # (no per-pass counter summary: traffic is null)
        per_pass = []
        dom = None
        dom_bytes, dom_ms, dom_name = total_bytes, passes[5], "slices_matvec (all kernels)"
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(args, dom), kernel=dom_name)

    # the reference ABI's own MATVEC entry takes host buffers: x, adia, aoff
    # in and y out over PCIe at every call (cu:1027-1029); reported beside
    # the HBM-resident value, never as it
    x_h, adia_h, aoff_h = host
    xf = np.zeros(3 * tg.nat, np.complex64)
    xf.reshape(3, -1)[:, tg.occ] = x_h.reshape(3, -1)
    lib.apply("N", xf, adia_h, aoff_h)
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        lib.apply("N", xf, adia_h, aoff_h)
    abi_ms = (time.perf_counter() - t0) / reps * 1e3
    pcie = dict(matvecs_per_s=round(1000.0 / abi_ms, 2), ms_per_matvec=round(abi_ms, 3),
                host_bytes_per_matvec=4 * 3 * tg.nat * 8,
                note="ddscat_cuda_apply_f32 on host full-layout vectors (3 H2D + 1 D2H per call)")

    solve = small = None
    if not args.no_solve:
        oracle = None
        fx = ROOT / "tests" / "golden" / "config_block128.npz"
        if args.edge == 128 and fx.exists():
            f = np.load(fx, allow_pickle=False)
            oracle = dict(qext=f["qext"], qabs=f["qabs"])
        solve = time_to_qext(lib, tg, kd, method=args.solver, oracle=oracle)
        solve16 = time_to_qext(lib, tg, kd, method=args.solver, n_orient=8)
        scaling = batched_matvec_scaling(lib, tg, bufs)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(tg, kd, host, budget_s=args.cpu_budget)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"block {tg.nx}x{tg.ny}x{tg.nz} dipoles ({tg.nat0} dipoles), padded grid "
                               f"{2*tg.nx}x{2*tg.ny}x{2*tg.nz}x3 complex64, m=1.5+0.1i, x=10 (kd={kd:.4f}), "
                               f"{args.backend.upper()} MATVEC on compact vectors",
                   "backend": args.backend, "nat0": tg.nat0},
        "roofline": roofline,
        "fft_kernels": {f"pass{p}": info.get(f"kernel_pass{p}") for p in range(1, 6)} | {"autotuned": info.get("autotuned")},
        "matvec_gbs": round(total_bytes / (ms_step * 1e-3) / 1e9, 1),
        "matvec_bytes": total_bytes,
        "passes": per_pass,
        "abi_apply_pcie_inclusive": pcie,
        "time_to_qext": solve,
        "cpu_baseline": cpu,
    }
    if not args.no_solve:
        out["time_to_qext_16rhs"] = solve16
        out["batched_matvec_per_rhs"] = scaling
    if not args.no_solve:
        # small-target batched leg last: it re-prepares the library for its own target
        out["batched_sphere32"] = batched_small_target(lib)
        out["batched_targets"] = batched_targets(lib, n_targets=args.targets)
    if not args.no_solve:
        out["adjoint_design"] = adjoint_design(lib)
    if not args.no_large and not args.no_solve:
        out["time_to_qext_256"] = large_solve(lib)
    if not args.no_large:
        out["large_target"] = large_target(lib)
    if not args.no_large and not args.no_solve:
        out["large_sphere_solve"] = large_sphere_solve(lib)
    print(json.dumps(out))
    lib.release()


if __name__ == "__main__":
    main()
