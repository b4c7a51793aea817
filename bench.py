#!/usr/bin/env python3
"""Benchmark of the contraction hot path on one MI355X.

Workload: the U(1) two-site effective-Hamiltonian matvec of DMRG at bond dimension chi=1024
(31 charge sectors q in [-15, 15], sizes proportional to exp(-q^2/50); spin-1/2 Heisenberg MPO,
D=5 grouped as charges {0 x3, +2, -2}; iid N(0,1) blocks, environments symmetrised so the
operator is Hermitian).  The two MPO tensors are contracted once into W12, and each matvec is
two launches (--chain mix, default): the mix launch L.psi.W12 -> Y, whose L.psi slices stay on
chip, then the grouped GEMM Y.R.  --chain w12 runs three launches (L.psi, .W12, .R), --chain ref
the reference's four.  The flop count is always the reference's four-step
L.psi -> W1 -> W2 -> R count.

A "step" is one ground-state solve of that operator by Lanczos from a fixed start vector to
CvgCrit 1e-12 in the FP32 + refinement scheme (FP32 contractions, FP64 vectors and tridiagonal,
then an FP64 Lanczos restarted from the FP32 Ritz vector).  ``value`` is the Heff matvec
throughput of that solve, end to end: (matvecs executed) x (the reference's four-step flop
count) / (solve time), Lanczos vector work, host checks and the refinement included, inputs
resident in HBM.  The same solve in plain FP64 is timed beside it (``fp64_solve``), with the
energy difference (bar 1e-10) and the time ratio (``speedup_vs_fp64``).  The isolated matvec
rates (FP64, FP32, FP32x3), the roofline of the dominant kernel and the CPU baseline (the
oracle's restated per-pair contraction with OpenBLAS dgemm) are reported in the same JSON
line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp32x3": 157.3}
PEAK_HBM_GBS = 8000.0
METRIC = "U(1) Heff matvec TFLOP/s & %peak, FP32+refine vs FP64 at ΔE≤10^-10"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads() -> int:
    """Threads the CPU legs use: the process's CPU share (OMP_NUM_THREADS when the host sets it
    for one GPU's share of the machine, else every core this process may run on)."""
    share = os.environ.get("OMP_NUM_THREADS", "")
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(int(share), avail) if share.isdigit() and int(share) > 0 else avail or 1)


def host_cpu() -> dict:
    """The host the CPU baseline runs on: lscpu model, sockets, cores, and the thread share."""
    import subprocess
    info = {"nproc": os.cpu_count(), "threads_available": host_threads()}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        text = ""
    for line in text.splitlines():
        k, _, v = line.partition(":")
        k = k.strip()
        if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
            info[k.lower().replace("(s)", "s").replace(" ", "_")] = v.strip()
    return info


def time_matvec(H, x, y, steps, warmup, torch, graph=True):
    """Average matvec time over `steps` (HIP events on the launch stream).  With graph=True each
    matvec is one replay of the captured chain (one host launch per matvec); otherwise each
    grouped GEMM is launched from the host."""
    if graph:
        H.chain.prepare(H.precision)
        g = H.chain.graph(x, y, 1)
        run = g.launch
    else:
        def run():
            H.apply_flat(x, y)
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        run()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(end) / steps, wall


def time_steps(H, x, y, reps, torch):
    """Average duration of each chain step's grouped GEMM launch (HIP events on the launch
    stream, `reps` back-to-back launches of the same kernel)."""
    from cytnx_amd import _lib
    from cytnx_amd.device import stream_handle
    L = _lib.lib()
    s = stream_handle(0)
    H.apply_flat(x, y)  # leave every intermediate in place
    out = []
    for k in range(len(H.chain.steps)):
        for _ in range(3):
            _lib.check(L.cxa_chain_step(H.chain.handle, k, x.data_ptr(), y.data_ptr(), s))
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            _lib.check(L.cxa_chain_step(H.chain.handle, k, x.data_ptr(), y.data_ptr(), s))
        b.record()
        torch.cuda.synchronize()
        out.append(a.elapsed_time(b) / reps)
    return out


def step_bytes(H, n_x, prec):
    """Algorithmic bytes of each chain step: its running operand read once, its fixed operand
    read once, its output written once.  In FP32 mode the fixed operands and intermediates are
    FP32 copies, the input vector and the final output FP64."""
    from cytnx_amd.unitensor import PREC_FP32, meta_elems
    mid = 4 if prec == PREC_FP32 else 8
    out = []
    n_in, b_in = n_x, 8
    nsteps = len(H.chain.steps)
    for k, (plan, fixed, left, meta) in enumerate(H.chain.steps):
        n_out = meta_elems(meta)
        b_out = 8 if (k == nsteps - 1 or prec != PREC_FP32) else mid
        out.append(n_in * b_in + fixed._storage.numel() * mid + n_out * b_out)
        n_in, b_in = n_out, b_out
    return out


def bench_permute(torch, nperm=32):
    """Rank-6 permute sweep: shapes (16,8,16,8,16,8) and (32,4,32,4,32,4), FP64 and complex128,
    `nperm` seeded permutations.  Each permutation is timed as one replay of a captured graph
    of back-to-back permutes over rotating buffers (>= 1 GiB in + out, so the Infinity Cache
    cannot hold them); GB/s counts one read and one write of the tensor."""
    import numpy as np
    from cytnx_amd import _lib
    from cytnx_amd.device import stream_handle
    L = _lib.lib()
    rng = np.random.default_rng(20240)
    perms = [rng.permutation(6) for _ in range(nperm)]
    res = {}
    for shape in [(16, 8, 16, 8, 16, 8), (32, 4, 32, 4, 32, 4)]:
        for name, dt_id, tdt, es in [("f64", 3, torch.float64, 8),
                                     ("c128", 1, torch.complex128, 16)]:
            n = int(np.prod(shape))
            nb = max(2, -(-(1 << 30) // (2 * n * es)))
            src = torch.rand(nb * n, dtype=tdt, device="cuda:0")
            dst = torch.empty_like(src)
            shp = _lib.u64_array(shape)
            gbs = []
            for p in perms:
                mp = _lib.u64_array(p)
                g = torch.cuda.CUDAGraph()
                cs = torch.cuda.Stream()
                with torch.cuda.graph(g, stream=cs):
                    hs = stream_handle(0)
                    for i in range(nb):
                        _lib.check(L.cxa_permute(dt_id, 6, shp, mp, src.data_ptr() + i * n * es,
                                                 dst.data_ptr() + i * n * es, hs))
                g.replay()
                torch.cuda.synchronize()
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(3):
                    g.replay()
                b.record()
                torch.cuda.synchronize()
                per = a.elapsed_time(b) * 1e-3 / (3 * nb)
                gbs.append(2.0 * n * es / per / 1e9)
                del g
            # the same bytes moved by a plain device copy, timed the same way: the practical
            # ceiling of one launch at this size
            g = torch.cuda.CUDAGraph()
            cs = torch.cuda.Stream()
            with torch.cuda.graph(g, stream=cs):
                for i in range(nb):
                    dst[i * n:(i + 1) * n].copy_(src[i * n:(i + 1) * n])
            g.replay()
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(3):
                g.replay()
            b.record()
            torch.cuda.synchronize()
            copy_gbs = 2.0 * n * es / (a.elapsed_time(b) * 1e-3 / (3 * nb)) / 1e9
            del g
            # this library's own streaming copy (the identity permute), timed the same way: the
            # HBM rate a plain read-once / write-once kernel reaches at this size
            ident = _lib.u64_array(list(range(6)))
            g = torch.cuda.CUDAGraph()
            cs = torch.cuda.Stream()
            with torch.cuda.graph(g, stream=cs):
                hs = stream_handle(0)
                for i in range(nb):
                    _lib.check(L.cxa_permute(dt_id, 6, shp, ident, src.data_ptr() + i * n * es,
                                             dst.data_ptr() + i * n * es, hs))
            g.replay()
            torch.cuda.synchronize()
            a.record()
            for _ in range(3):
                g.replay()
            b.record()
            torch.cuda.synchronize()
            stream_gbs = 2.0 * n * es / (a.elapsed_time(b) * 1e-3 / (3 * nb)) / 1e9
            del g
            key = "%s_%s" % ("x".join(map(str, shape)), name)
            res[key] = {"mean_GBs": float(np.mean(gbs)), "min_GBs": float(np.min(gbs)),
                        "max_GBs": float(np.max(gbs)),
                        "frac_of_8TBs": float(np.mean(gbs)) / PEAK_HBM_GBS,
                        "copy_GBs": copy_gbs, "frac_of_copy": float(np.mean(gbs)) / copy_gbs,
                        "stream_copy_GBs": stream_gbs,
                        "frac_of_stream_copy": float(np.mean(gbs)) / stream_gbs,
                        "buffers_rotated": nb}
            del src, dst
    allg = [v["mean_GBs"] for v in res.values()]
    res["mean_GBs"] = float(np.mean(allg))
    res["permutations"] = nperm
    return res


def bench_c1(torch, cpu=True, iters=100):
    """Dense one-site Heff at chi=128 (L, A, R iid N(0,1), exact XXZ MPO D=5, d=2), order
    ((L.A).W).R with autotuned tilings, `iters` FP64 Lanczos iterations (CvgCrit below any
    reachable |dE|)."""
    import numpy as np
    from cytnx_amd import heisenberg
    from cytnx_amd.heff import Heff1Dense
    from cytnx_amd.lanczos import lanczos_gnd_native
    chi, D, d = 128, 5, 2
    rng = np.random.default_rng(1)
    Lh = rng.standard_normal((chi, D, chi))
    Ah = rng.standard_normal((chi, d, chi))
    Rh = rng.standard_normal((chi, D, chi))
    W = heisenberg.xxz_mpo_dense()
    H = Heff1Dense(torch.from_numpy(Lh).cuda(), torch.from_numpy(W).cuda(),
                   torch.from_numpy(Rh).cuda(), Ah.shape)
    x0 = torch.from_numpy(Ah.ravel().copy()).cuda()
    cvg = 1e-300
    tiles = H.tune(reps=10)  # the library's autotuner, on this operator's own buffers
    lanczos_gnd_native(H, x0, cvg, iters, batch=iters)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, v, k, _, _ = lanczos_gnd_native(H, x0, cvg, iters, batch=iters)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    flops = 4 * D * d * chi ** 3 + 2 * D * D * d * d * chi * chi
    out = {"iterations": int(k), "seconds": t, "flops_per_matvec": flops,
           "tflops": k * flops / t / 1e12, "ms_per_iteration": t / k * 1e3, "E": E,
           "tuned_tiles": tiles}
    if cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import pyoracle as po
        blas = po.use_openblas(nthreads=host_threads())

        def mv(x):
            A = x.reshape(chi, d, chi)
            X = po.tensordot(Lh, A, [0], [0])         # (w, l', s, r)
            Y = po.tensordot(X, W, [2, 0], [2, 0])    # (l', r, w', s')
            return po.tensordot(Y, Rh, [1, 2], [0, 1]).ravel()  # (l', s', r')

        t0 = time.perf_counter()
        Eo, _, ko = po.lanczos_gnd(mv, Ah.ravel().copy(), cvg, k)
        tc = time.perf_counter() - t0
        po.use_reference_loops()
        out["cpu"] = {"seconds": tc, "iterations": int(ko), "tflops": ko * flops / tc / 1e12,
                      "cores": host_threads(), "kind": "port", "host": host_cpu(),
                      "blas": blas, "E": Eo}
        out["speedup_vs_cpu"] = tc / t * (k / max(ko, 1))
    return out


def bench_mps_gradient(torch, L=32, chi=512):
    """Differentiable MPS energy: finite U(1) MPS, L sites, bond dimension up to chi, complex128
    and complex64 site tensors (iid N(0,1), seeded), Heisenberg MPO; one E = <psi|H|psi> /
    <psi|psi> evaluation plus its gradient with respect to every site tensor (hand-written
    adjoint contractions), timed after one warm-up."""
    from cytnx_amd import autograd
    from cytnx_amd.mps import mps_energy, random_mps
    out = {"L": L, "chi": chi}
    for name, dt in (("c128", torch.complex128), ("c64", torch.complex64)):
        As = random_mps(L, chi, seed=1, dtype=dt)
        for A in As:
            A._storage.requires_grad_(True)
        E, _, _ = mps_energy(As)
        E.backward()
        for A in As:
            A._storage.grad = None
        autograd.FLOPS.update(forward=0.0, adjoint=0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E, _, _ = mps_energy(As)
        E.backward()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        f = autograd.FLOPS["forward"] + autograd.FLOPS["adjoint"]
        gn = sum(float(torch.linalg.vector_norm(A._storage.grad).item()) ** 2 for A in As) ** 0.5
        out[name] = {"seconds": t, "E": float(E.item()), "grad_norm": gn,
                     "forward_flops": autograd.FLOPS["forward"],
                     "adjoint_flops": autograd.FLOPS["adjoint"], "tflops": f / t / 1e12,
                     "params": int(sum(A._storage.numel() for A in As))}
    return out


def measured_traffic(chi, chain, prec, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary of this
    workload (profiles/heff2_chi<chi>_traffic.json), or None when no measurement matches."""
    path = os.path.join(HERE, "profiles", "heff2_chi%d_traffic.json" % chi)
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("chi") != chi or t.get("chain") != chain or t.get("prec") != prec:
        return None
    return t.get("hbm_bytes_per_launch", {}).get(kernel)


def bench_dmrg(torch, L=100, chi=2048, sweeps=4, maxit=30, cpu=True):
    """Two-site U(1) DMRG of the open Heisenberg chain: L sites, bond dimension chi, `sweeps`
    full sweeps (left to right and back) from a random right-canonical MPS, `maxit` Lanczos
    iterations per bond.  Run twice from the same seed: contractions in FP32 + FP64 refinement
    (the reported run) and in FP64 (the energy reference); the truncation and environments are
    FP64 in both.  The CPU side times the oracle's restated per-pair contraction chain for one
    matvec at the centre bond of the final state and scales it by the run's matvec flops
    (a sampled estimate of the Lanczos share only)."""
    from cytnx_amd.dmrg import DMRG
    out = {"L": L, "chi": chi, "sweeps": sweeps, "lanczos_iterations_per_bond": maxit}
    runs = {}
    for prec in ("mixed", "fp64"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dm = DMRG(L, chi, seed=1, maxit=maxit, prec=prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dm.run(sweeps)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        st = dm.stats
        runs[prec] = {"E": dm.energies[-1], "init_s": t1 - t0, "sweeps_s": t2 - t1,
                      "wall_s": t2 - t0, "updates": st["updates"], "matvecs": st["matvecs"],
                      "lanczos_s": st["lanczos"], "lanczos_tflops":
                          st["flops"] / st["lanczos"] / 1e12 if st["lanczos"] else None,
                      "heff_plan_s": st["heff"], "split_s": st["svd"], "env_s": st["env"],
                      "fp64_updates": st["fp64_updates"],
                      "max_bond": max(dm.bond_dims())}
        log("dmrg", prec, json.dumps(runs[prec]))
        last = dm
    out["mixed"], out["fp64"] = runs["mixed"], runs["fp64"]
    out["dE_mixed_vs_fp64"] = runs["mixed"]["E"] - runs["fp64"]["E"]
    out["speedup_mixed_vs_fp64_lanczos"] = runs["fp64"]["lanczos_s"] / runs["mixed"]["lanczos_s"]
    if cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import pyoracle as po
        from cytnx_amd.unitensor import Contract

        def ob(ut):
            blocks = [ut.get_block_(i).detach().cpu().numpy() for i in range(ut.Nblocks())]
            bonds = [po.OBond(b.type(), b.qnums(), b.getDegeneracies()) for b in ut.bonds()]
            return po.OBlockTensor(bonds, ut.labels(), ut.rowrank(), itoi=ut.itoi(),
                                   blocks=blocks)

        from cytnx_amd.dmrg import left_env_step
        from cytnx_amd.heff import Heff2
        # after the closing right-to-left sweep LR[k >= 1] are right environments: rebuild the
        # left one of the centre bond from the boundary
        p = L // 2 - 1
        Lp = last.LR[0]
        for k in range(p):
            Lp = left_env_step(Lp, last.As[k], last.M)
        psi = Contract(last.As[p], last.As[p + 1]).relabel(["-1", "-2", "-3", "-4"])
        Lr = Lp.relabel(["-5", "-1", "0"])
        Rr = last.LR[p + 2].relabel(["-7", "-4", "3"])
        M1r = last.M.relabel(["-5", "-6", "-2", "1"])
        M2r = last.M.relabel(["-6", "-7", "-3", "2"])
        f = Heff2(psi, Lr, M1r, M2r, Rr).ref_flops
        Lo, Ro, M1, M2, P = ob(Lr), ob(Rr), ob(M1r), ob(M2r), ob(psi)
        ncpu = host_threads()
        blas = po.use_openblas(nthreads=ncpu)
        t0 = time.perf_counter()
        X = po.block_contract(Lo, P)
        X = po.block_contract(X, M1)
        X = po.block_contract(X, M2)
        po.block_contract(X, Ro)
        tc = time.perf_counter() - t0
        po.use_reference_loops()
        est = tc / f * last.stats["flops"]
        out["cpu"] = {"matvec_s_at_centre": tc, "centre_flops": f, "cores": ncpu, "kind": "port",
                      "host": host_cpu(),
                      "blas": blas, "lanczos_s_estimate": est,
                      "sample": "one centre-bond matvec of the oracle's restated "
                                "BlockUniTensor::contract chain (FP64), scaled by the run's "
                                "matvec flops"}
        out["speedup_lanczos_vs_cpu_estimate"] = est / runs["mixed"]["lanczos_s"]
    return out


def cpu_baseline(problem, flops, seconds):
    import numpy as np
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pyoracle as po

    ncpu = host_threads()
    psi, L, M1, M2, R = problem

    def ob(ut):
        blocks = [ut.get_block_(i).detach().cpu().numpy() for i in range(ut.Nblocks())]
        bonds = [po.OBond(b.type(), b.qnums(), b.getDegeneracies()) for b in ut.bonds()]
        return po.OBlockTensor(bonds, ut.labels(), ut.rowrank(), itoi=ut.itoi(), blocks=blocks)

    P, Lo, M1o, M2o, Ro = ob(psi), ob(L), ob(M1), ob(M2), ob(R)

    def matvec():
        X = po.block_contract(Lo, P)
        Y = po.block_contract(X, M1o)
        Z = po.block_contract(Y, M2o)
        return po.block_contract(Z, Ro)

    # OpenBLAS threading costs more than it gains on these small block GEMMs: give the CPU its
    # better setting (one thread or all cores), chosen by a short pilot
    best = None
    for cores in sorted({1, ncpu}):
        blas = po.use_openblas(nthreads=cores)
        matvec()
        t0 = time.perf_counter()
        matvec()
        matvec()
        dt = (time.perf_counter() - t0) / 2
        if best is None or dt < best[1]:
            best = (cores, dt)
    cores = best[0]
    blas = po.use_openblas(nthreads=cores)
    matvec()  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        matvec()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 2:
            break
    per = el / n
    po.use_reference_loops()
    return {"value": flops / per / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "host": host_cpu(), "threads_tried": sorted({1, ncpu}),
            "sample": "%d matvecs (%.1f s) of the same chi=1024 two-site U(1) Heff: the oracle's "
                      "restated BlockUniTensor::contract chain (per block pair: permute copies, "
                      "dgemm via %s, add pass), FP64" % (n, el, blas or "reference loops"),
            "ms_per_matvec": per * 1e3}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed ground-state solves (each to CvgCrit, FP32 + refinement and FP64)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--matvec-steps", type=int, default=200,
                    help="timed matvecs per precision for the isolated matvec rates")
    ap.add_argument("--chi", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the tiling autotuner (per-block default tiles)")
    ap.add_argument("--lanczos-cvg", type=float, default=1e-12)
    ap.add_argument("--lanczos-maxiter", type=int, default=600)
    ap.add_argument("--profile", action="store_true",
                    help="matvec loops only (for rocprofv3 runs), plus one 256 MiB device "
                         "copy whose counters calibrate FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--modes", default="fp64,fp32,fp32x3",
                    help="comma-separated precisions to time (fp32 is the reported value)")
    ap.add_argument("--save-tiles", default=None, help="write the tuned tilings (JSON)")
    ap.add_argument("--load-tiles", default=None,
                    help="apply tilings from a --save-tiles file instead of tuning")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the permute sweep and the dense one-site Lanczos lines")
    ap.add_argument("--dmrg-sweeps", type=int, default=4,
                    help="full sweeps of the L=100, chi=2048 two-site DMRG line (0: skip)")
    ap.add_argument("--dmrg-chi", type=int, default=2048)
    ap.add_argument("--dmrg-L", type=int, default=100)
    ap.add_argument("--chain", default="mix", choices=["mix", "w12", "ref"],
                    help="matvec program: mix (two launches), w12 (three), ref (four, the "
                         "reference's contraction order)")
    ap.add_argument("--permute-only", action="store_true",
                    help="run only the rank-6 permute sweep and print its JSON (for rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the contraction path of this build runs on a single "
              "device (multi-GPU sharding is a design note, not built); use --gpus 1",
              file=sys.stderr)
        return 2

    import torch
    from cytnx_amd import heisenberg
    from cytnx_amd.heff import Heff2
    from cytnx_amd.lanczos import ground_state_mixed, lanczos_gnd_native
    from cytnx_amd.unitensor import PREC_FP32, PREC_FP32X3, PREC_FP64
    import numpy as np

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    torch.cuda.set_device(0)
    if args.permute_only:
        print(json.dumps({"permute_sweep": bench_permute(torch)}), flush=True)
        return 0
    log("building chi=%d problem" % args.chi)
    host = heisenberg.two_site_problem(args.chi, args.seed, hermitian=True)
    problem = [t.to(0) for t in host]
    psi, L, M1, M2, R = problem
    H = Heff2(psi, L, M1, M2, R, fuse_mpo=args.chain != "ref", mix=args.chain == "mix")
    if H.mode != args.chain:
        print("bench.py: chain %s unavailable for this operator" % args.chain, file=sys.stderr)
        return 2
    flops = H.flops  # the reference order's four-step count
    step_flops = [p.flops for p in H.chain._plans]  # per launched kernel
    x = psi._storage
    y = torch.empty_like(x)
    log("flops/matvec %.4e, block pairs per step %s, psi elements %d" % (flops, H.pairs, x.numel()))

    all_modes = {"fp64": PREC_FP64, "fp32": PREC_FP32, "fp32x3": PREC_FP32X3}
    wanted = [m.strip() for m in args.modes.split(",") if m.strip()]
    if "fp32" not in wanted or any(m not in all_modes for m in wanted):
        print("--modes must list fp32 and only fp64 / fp32 / fp32x3", file=sys.stderr)
        return 2
    modes = {m: all_modes[m] for m in wanted}
    res = {}
    tuned = {}
    loaded = None
    if args.load_tiles:
        with open(args.load_tiles) as f:
            loaded = json.load(f)
    for name, prec in modes.items():
        H.set_precision(prec)
        if loaded is not None and name in loaded:
            H.set_tiles(loaded[name])
            tuned[name] = loaded[name]
        elif not args.no_tune:
            tuned[name] = H.tune(reps=20)
            log(name, "tuned", json.dumps(tuned[name]))
        ms_graph, wall_g = time_matvec(H, x, y, args.matvec_steps, 20, torch, graph=True)
        ms_eager, wall_e = time_matvec(H, x, y, args.matvec_steps, 20, torch, graph=False)
        per = time_steps(H, x, y, max(20, args.matvec_steps // 4), torch)
        # the reported rate is the faster way of issuing the K matvecs: the chain's launches
        # enqueued on the stream (what the Lanczos driver does by default) or one captured-graph
        # replay per matvec
        launch = "stream" if ms_eager <= ms_graph else "graph"
        ms = min(ms_eager, ms_graph)
        res[name] = {"ms_per_step": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                     "pct_peak": 100.0 * flops / (ms * 1e-3) / 1e12 / PEAK_TFLOPS[name],
                     "launch": launch, "ms_per_step_stream": ms_eager,
                     "ms_per_step_graph": ms_graph, "step_ms": per,
                     "wall_s": wall_e if launch == "stream" else wall_g}
        log(name, json.dumps(res[name]))
    if args.save_tiles:
        with open(args.save_tiles, "w") as f:
            json.dump(tuned, f)
    if args.profile:
        # calibration for HBM counters: one copy of exactly 256 MiB read + 256 MiB written
        a = torch.ones(32 << 20, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        # a short FP64 Lanczos run (20 iterations) for the vector kernels' trace
        H.set_precision(PREC_FP64)
        lanczos_gnd_native(H, x.clone(), 1e-300, 20, batch=20)
        torch.cuda.synchronize()
    # normwise error of the reduced-precision matvecs against FP64 on the benchmark input
    errs = {}
    H.set_precision(PREC_FP64)
    y64 = torch.empty_like(x)
    H.apply_flat(x, y64)
    for name in [m for m in ("fp32", "fp32x3") if m in modes]:
        H.set_precision(modes[name])
        yy = torch.empty_like(x)
        H.apply_flat(x, yy)
        errs[name] = float((torch.linalg.vector_norm(yy - y64) /
                            torch.linalg.vector_norm(y64)).item())
    H.set_precision(PREC_FP64)

    lanczos = None
    solve = None
    if not args.no_lanczos and not args.profile:
        x0 = x.clone()
        cvg = args.lanczos_cvg
        lanczos = {"CvgCrit": cvg}

        def solve_fp64(graph=False):
            return lanczos_gnd_native(H, x0, cvg, args.lanczos_maxiter, batch=16, graph=graph)

        def solve_mixed(graph=False):
            return ground_state_mixed(H, x0, cvg, args.lanczos_maxiter, batch=16, graph=graph)

        def timed_solves(fn, n):
            ts, out = [], None
            for _ in range(n):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                out = fn()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            return ts, out

        for _ in range(max(1, args.warmup)):  # contexts, FP32 copies, first launches
            solve_fp64()
            solve_mixed()
        t64s, (E64, v64, n64, _, _) = timed_solves(solve_fp64, args.steps)
        tmxs, (Emx, vmx, n1, n2) = timed_solves(solve_mixed, args.steps)
        t64, tmx = float(np.median(t64s)), float(np.median(tmxs))
        dE = abs(Emx - E64)
        lanczos.update({"E_fp64": E64, "iters_fp64": n64, "time_fp64_s": t64,
                        "E_fp32_refine": Emx, "iters_fp32": n1, "iters_fp64_refine": n2,
                        "time_fp32_refine_s": tmx, "dE": dE, "dE_ok": bool(dE <= 1e-10),
                        "speedup_fp32_refine_vs_fp64": t64 / tmx, "reorth": False,
                        "solves_timed": args.steps, "statistic": "median"})
        # the same solves replaying captured iteration graphs (one host launch per batch)
        tg64, _ = timed_solves(lambda: solve_fp64(True), 3)
        tgmx, _ = timed_solves(lambda: solve_mixed(True), 3)
        lanczos["time_fp64_graph_s"] = float(np.median(tg64))
        lanczos["time_fp32_refine_graph_s"] = float(np.median(tgmx))
        solve = {"ms": tmx * 1e3, "tflops": (n1 + n2) * flops / tmx / 1e12,
                 "fp64_ms": t64 * 1e3, "fp64_tflops": n64 * flops / t64 / 1e12}
        # the other drivers on the same chain: the four lowest eigenpairs (which="SA", k=4) and
        # one imaginary-time step exp(-0.001 H) psi (Lanczos_Exp; tau |H| ~ 1, at most 200 vectors)
        from cytnx_amd.lanczos import eigsh_native, exp_native
        eigsh_native(H, x0, "SA", 4, 1e-10, False)  # warm-up (context allocation)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev, _, m_sa, nconv = eigsh_native(H, x0, "SA", 4, 1e-10, True)
        torch.cuda.synchronize()
        t_sa = time.perf_counter() - t0
        lanczos["which_SA_k4"] = {"tol": 1e-10, "eigenvalues": [float(e) for e in ev],
                                  "krylov_vectors": m_sa, "nconv": nconv, "time_s": t_sa,
                                  "E0_minus_gnd": float(ev[0] - E64)}
        exp_native(H, x0, -0.001, 1e-12, 200)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y_exp, m_exp = exp_native(H, x0, -0.001, 1e-12, 200)
        torch.cuda.synchronize()
        t_exp = time.perf_counter() - t0
        lanczos["exp_tau_m0.001"] = {
            "CvgCrit": 1e-12, "krylov_vectors": m_exp, "time_s": t_exp,
            "norm_ratio": float(torch.linalg.vector_norm(y_exp) / torch.linalg.vector_norm(x0))}
        log("lanczos", json.dumps(lanczos))

    main_mode = "fp32"
    k_dom = max(range(len(step_flops)), key=lambda k: res[main_mode]["step_ms"][k])
    dom_ms = res[main_mode]["step_ms"][k_dom]
    sbytes = step_bytes(H, x.numel(), modes[main_mode])
    # the kernel's own bound: the larger of its MFMA time and its HBM time at peak
    t_mfma = step_flops[k_dom] / (PEAK_TFLOPS[main_mode] * 1e12)
    t_hbm = sbytes[k_dom] / (PEAK_HBM_GBS * 1e9)
    if t_hbm > t_mfma:
        achieved = sbytes[k_dom] / (dom_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": achieved / PEAK_HBM_GBS}
    else:
        achieved = step_flops[k_dom] / (dom_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[main_mode],
                    "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[main_mode]}
    roofline.update({"traffic": measured_traffic(args.chi, H.mode, main_mode,
                                                 H.step_names[k_dom]),
                     "kernel": "chain step %d of %d (%s)" % (k_dom, len(step_flops),
                                                             H.step_names[k_dom]),
                     "kernel_flops": step_flops[k_dom], "kernel_bytes": sbytes[k_dom],
                     "kernel_ms": dom_ms})

    cpu = None
    if not args.no_cpu and not args.profile:
        log("cpu baseline (~%.0f s)" % args.cpu_seconds)
        cpu = cpu_baseline(host, flops, args.cpu_seconds)
        log("cpu", json.dumps(cpu))

    extra = {}
    if not args.no_extra and not args.profile:
        log("permute sweep")
        extra["permute_sweep"] = bench_permute(torch)
        log("permute", json.dumps(extra["permute_sweep"]))
        log("dense one-site Heff, 100 Lanczos iterations")
        extra["one_site_dense_lanczos"] = bench_c1(torch, cpu=not args.no_cpu)
        log("one-site", json.dumps(extra["one_site_dense_lanczos"]))
        log("MPS energy gradient")
        extra["mps_energy_gradient"] = bench_mps_gradient(torch)
        log("mps", json.dumps(extra["mps_energy_gradient"]))
        if args.dmrg_sweeps > 0:
            log("two-site DMRG L=%d chi=%d, %d sweep(s)" % (args.dmrg_L, args.dmrg_chi,
                                                            args.dmrg_sweeps))
            extra["dmrg"] = bench_dmrg(torch, args.dmrg_L, args.dmrg_chi, args.dmrg_sweeps,
                                       cpu=not args.no_cpu)
            log("dmrg", json.dumps(extra["dmrg"]))

    if solve is not None:
        value, ms_step = solve["tflops"], solve["ms"]
        definition = ("Heff matvec TFLOP/s sustained by the whole FP32 + refinement ground-state "
                      "solve (matvecs executed x reference four-step flops / solve time, vector "
                      "work and refinement included), median of the timed solves")
    else:  # --no-lanczos / --profile: the isolated FP32 matvec
        value, ms_step = res[main_mode]["tflops"], res[main_mode]["ms_per_step"]
        definition = "isolated FP32 matvec TFLOP/s (no Lanczos solve was timed)"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps if solve is not None else args.matvec_steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 contractions + f64 refinement",
        "data": "synthetic (seeded iid N(0,1) blocks, hermitised environments, exact XXZ MPO)",
        "config": {"workload": "U(1) two-site Heff ground state by Lanczos to CvgCrit %g, "
                               "chi=%d, Gaussian sector profile, Heisenberg MPO D=5, matvec "
                               "order L.psi->W1->W2->R" % (args.lanczos_cvg, args.chi),
                   "step": "one ground-state solve" if solve is not None else "one matvec",
                   "value": definition,
                   "chi": args.chi, "sectors": len(heisenberg.gaussian_sectors(args.chi)[0]),
                   "psi_elements": int(x.numel()), "block_pairs_per_step": H.pairs,
                   "flops_per_matvec": flops, "chain": H.mode, "kernels": H.step_names,
                   "launch": res[main_mode]["launch"],
                   "kernel_flops": step_flops},
        "pct_peak_fp32": 100.0 * value / PEAK_TFLOPS["fp32"],
        "fp64_solve": ({"ms": solve["fp64_ms"], "tflops": solve["fp64_tflops"],
                        "pct_peak_fp64": 100.0 * solve["fp64_tflops"] / PEAK_TFLOPS["fp64"]}
                       if solve is not None else None),
        "speedup_vs_fp64": (lanczos["speedup_fp32_refine_vs_fp64"]
                            if lanczos is not None else None),
        "dE_vs_fp64": lanczos["dE"] if lanczos is not None else None,
        "matvec": {"fp64": res.get("fp64"), "fp32": res["fp32"], "fp32x3": res.get("fp32x3"),
                   "speedup_fp32_vs_fp64": (res["fp64"]["ms_per_step"] /
                                            res["fp32"]["ms_per_step"] if "fp64" in res
                                            else None)},
        "tuned_tiles": tuned,
        "matvec_rel_error_vs_fp64": errs,
        "lanczos": lanczos,
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    if cpu:
        # the same matvec flops on the CPU baseline: its rate over the GPU's sustained solve rate
        out["speedup_vs_cpu"] = value / cpu["value"]
        out["speedup_fp32_matvec_vs_cpu"] = res[main_mode]["tflops"] / cpu["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
