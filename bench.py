#!/usr/bin/env python3
"""Benchmark: Chebyshev-filter HBM throughput on one MI355X.

Workload: the operator-level filter case -- 96^3 periodic grid, h = 0.3 Bohr,
order-12 stencil, 512 vectors uniform[-0.5, 0.5] from a seed, Veff = 16 seeded
Gaussians, 64 projector sites x 8 projectors (2 radial x l<=1, cutoff 6 Bohr),
degree 24, bounds from a 10-step Lanczos.  A step is one full degree-24 filter
of the 512-vector block (inputs resident in HBM); consecutive steps filter the
previous output (the scaled recurrence keeps magnitudes bounded).

value = algorithmic bytes per filter / measured filter time, in GB/s, for the
FP32 (mixed-precision) filter; the FP64 filter is timed beside it.
Algorithmic bytes per filter:
    s*Nd*Ns*(2 + 3(m-1)) + m*s*Nd + 2*m*s*sum(Chi)        (s = 4 or 8)

The SCF half of the metric is reported under "scf": full device SCF runs of
diamond-cubic Si (64 atoms) and Si/Al (216 atoms) with the ONCV psp8
pseudopotentials (NLCC), LDA-PZ, Pulay/Kerker mixing and SPARC's defaults, in
FP64 and in mixed precision, with the mean SCF-iteration time over iterations
2..N and the energy and eigenvalue differences of the mixed run against the
FP64 run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SCF-iteration time & filter HBM GB/s, mixed vs FP64, at ΔE≤1e-5 Ha/atom"
HBM_PEAK_GBS = 8000.0
# MI355X dense matrix-core peaks (TFLOP/s)
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}


def host_threads():
    """(threads used by the CPU legs, host CPUs visible, CPU model).  The CPU
    share of a one-GPU box is 16 threads, so at most 16 are used."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return min(16, avail), avail, model


def mfma_entry(flops, ms, peak_key, issued_flops=None):
    """Achieved TFLOP/s of a dense kernel against the matrix-core peak of its
    arithmetic; issued_flops counts padded / split-precision MFMA work."""
    if not ms or ms <= 0:
        return None
    tf = flops / (ms * 1e-3) / 1e12
    out = {"ms": round(ms, 4), "tflops": round(tf, 3), "peak_tflops": PEAK_TFLOPS[peak_key],
           "peak": peak_key, "frac": round(tf / PEAK_TFLOPS[peak_key], 4)}
    if issued_flops is not None:
        ti = issued_flops / (ms * 1e-3) / 1e12
        out.update(issued_tflops=round(ti, 3), issued_frac=round(ti / PEAK_TFLOPS[peak_key], 4))
    return out


def dense_entry(key, f32, nd, ns, ms):
    """Matrix-core figures of the projection (Mp = Y'Y and Hp = Y'HY) or the
    rotation (Y Q) at nd points and ns states, against the peak of the
    arithmetic the kernel runs: the FP32 paths issue six BF16 products per
    value pair (three pieces per value), the Gram from 256 states computes the
    tiles on and above the diagonal only (128 x 128 tiles), FP64 is rocBLAS."""
    nprod = 2 if key == "project_gemm" else 1
    flops = nprod * 2.0 * nd * ns * ns
    gram_env = os.environ.get("SPARC_AMD_GRAM", "")
    rot_env = os.environ.get("SPARC_AMD_ROT", "")
    if not f32:
        e, kernel = mfma_entry(flops, ms, "fp64"), "rocblas dgemm"
    elif key == "rotate_gemm" and rot_env != "rocblas" and ns <= 512:
        nk = -(-ns // 32) * 32
        e, kernel = mfma_entry(flops, ms, "bf16", 6 * 2.0 * nd * nk * ns), "rot_bf16x3"
    elif key == "project_gemm" and gram_env != "rocblas" and (ns < 256 or gram_env == "bf16x3"):
        t = -(-ns // 64)
        e, kernel = mfma_entry(flops, ms, "bf16", 6 * 2 * 2.0 * nd * (64 * t) ** 2), "gram_bf16x3"
    elif key == "project_gemm" and gram_env != "rocblas":
        t = -(-ns // 128)
        issued = 6 * 2 * 2.0 * nd * 128 ** 2 * (t * (t + 1) // 2)
        e, kernel = mfma_entry(flops, ms, "bf16", issued), "gram3_bf16"
    else:
        e, kernel = mfma_entry(flops, ms, "fp32"), "rocblas sgemm"
    e["kernel"] = kernel
    return e


def algorithmic_bytes(nd, ns, m, s, chi_elems):
    return s * nd * ns * (2 + 3 * (m - 1)) + m * s * nd + 2 * m * s * chi_elems


def stencil_bytes(nd, ns, m, s):
    """Bytes the fused stencil/recurrence launches must move over one filter:
    first launch reads X and writes Y, the m-1 others read Y_k, Y_{k-1} and
    write Y_{k+1}; every launch reads Veff once per column block."""
    return s * nd * ns * (2 + 3 * (m - 1)) + m * s * nd


def profiled_traffic(kernel):
    """HBM bytes per launch of the dominant kernel (by its full name, e.g.
    stencil_tiled<float,32,16,2>), from the committed FETCH_SIZE / WRITE_SIZE
    summary (profiles/), or None."""
    import re
    p = ROOT / "profiles" / "filter_operator_96x512_deg24_hbm_traffic.md"
    if not p.exists():
        return None
    m = re.search(rf"traffic_bytes_per_launch {re.escape(kernel)}:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def bounds_lanczos(ctx, wl):
    from sparc_amd.workloads import uniform_block
    x0 = uniform_block(wl.grid.nd, 1, wl.lanczos_seed, 0.0, 1.0)[:, 0]
    emin, emax, _ = ctx.Lanczos(x0, 0.0, 0.0, 10)
    emax *= 1.01
    emin -= 0.1
    return emin, emax, 0.5 * (emin + emax)


def time_filter(ctx, gpu, wl, dtype, steps, warmup, bounds):
    a0, b, a = bounds
    X = ctx.block(wl.ncol, dtype).fill_uniform(wl.psi_seed, -0.5, 0.5)
    Y = ctx.block(wl.ncol, dtype)
    for _ in range(warmup):
        ctx.chebyshev_filter_dev(X, Y, wl.degree, a, b, a0)
        X, Y = Y, X
    ctx.sync()
    ctx.timer_reset()
    ctx.timer_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.chebyshev_filter_dev(X, Y, wl.degree, a, b, a0)
        X, Y = Y, X
    ctx.sync()
    t1 = time.perf_counter()
    ctx.timer_enable(False)
    out = {"wall_ms": (t1 - t0) * 1e3 / max(steps, 1)}
    for name in ("filter", "stencil", "nonlocal", "nl_project", "nl_apply"):
        tot, cnt = ctx.timer(name)
        out[name] = (tot, cnt)
    finite = bool(np.isfinite(X.download(1)).all())
    X.free()
    Y.free()
    out["finite"] = finite
    return out


def nonlocal_dense(ctx, wl, rec, peak_key):
    """Matrix-core figures of the nonlocal projection and apply of one filter
    step: useful FLOPs 2 * sum(Chi) * ncol per kernel, issued FLOPs 2 *
    stored * ncol (the layout's zero padding included)."""
    mode, stored = ctx.nonlocal_layout()
    chi, ns = int(wl.nloc.chi.size), wl.ncol
    out = {"layout": "bricks" if mode == 0 else "spheres", "stored_values": stored,
           "chi_values": chi}
    names = {"nl_project": "nlb_project_rs" if mode == 0 else "nls_project",
             "nl_apply": "nlb_apply_rs" if mode == 0 else "nls_apply / nls_apply_proj"}
    for key in ("nl_project", "nl_apply"):
        tot, cnt = rec[key]
        if cnt:
            out[key] = mfma_entry(2.0 * chi * ns, tot / cnt, peak_key, 2.0 * stored * ns)
            out[key]["kernel"] = names[key]
    return out


def cpu_baseline(wl, ncol_sample=48, degree=None):
    """FP64 oracle filter on the host cores over a bounded sample of the same
    workload (same grid, Veff, projectors; ncol_sample vectors)."""
    from oracle import oracle as O
    from sparc_amd.workloads import uniform_block
    threads, avail, model = host_threads()
    O.set_threads(threads)
    o = O.Oracle(wl.grid, wl.veff, wl.nloc)
    m = degree or wl.degree
    x0 = uniform_block(wl.grid.nd, 1, wl.lanczos_seed, 0.0, 1.0)[:, 0]
    emin, emax, _ = o.lanczos(x0, 0.0, 0.0, 10)
    emax *= 1.01
    emin -= 0.1
    a = 0.5 * (emin + emax)
    X = uniform_block(wl.grid.nd, ncol_sample, wl.psi_seed, -0.5, 0.5)
    t0 = time.perf_counter()
    o.chebyshev_filter(X, m, a, emax, emin)
    dt = time.perf_counter() - t0
    chi = int(wl.nloc.chi.size)
    gbs = algorithmic_bytes(wl.grid.nd, ncol_sample, m, 8, chi) / dt / 1e9
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "host_cpus": avail, "cpu_model": model,
            "sample": f"FP64 oracle filter (OpenMP over columns, {threads} of {avail} host CPUs: "
                      f"the CPU share of one GPU), same grid/Veff/projectors, {ncol_sample} of "
                      f"{wl.ncol} vectors, degree {m}, {dt:.2f} s"}


def scf_bench(name, unet_weights=None, kerker_weights=None):
    """FP64 and mixed-precision device SCF of one psp8 system, plus the mixed
    run with the learned Poisson preconditioner when trained weights are
    given, and the same with a second network as the Kerker solves'
    preconditioner."""
    from sparc_amd import gpu
    from sparc_amd.systems import scf_system
    t0 = time.perf_counter()
    st, p = scf_system(name)
    g = st.grid
    print(f"[bench] scf {name}: {st.n_atom} atoms, {g.nx}^3, {p['nstates']} states, "
          f"setup {time.perf_counter() - t0:.1f}s", flush=True)
    n_al = st.meta.get("n_al", 0)
    out = {"workload": f"{name}: {st.n_atom}-atom "
                       + ("SiC (reference test input)" if name == "sic8" else
                          "diamond-cubic Si" + (f" with {n_al} Al" if n_al else ""))
                       + f", ONCV psp8 with NLCC, {g.nx}^3 grid h={g.hx:.4f}, {p['nstates']} "
                       f"states, degree {p['cheb_degree']}, "
                       + ("LDA_PW" if p["xc"] == 1 else "LDA_PZ") + ", "
                       + ("Fermi-Dirac" if p["smear_type"] == 0 else "Gaussian")
                       + f" smearing {1.0 / p['beta']:.4g} Ha, tol_scf {p['tol_scf']:.3g}"}
    learned = unet_weights is not None
    plan = [("f64", gpu.F64, False, False), ("mixed", gpu.F32, False, False)]
    if learned:
        plan.append(("mixed_learned", gpu.F32, True, False))
        if kerker_weights is not None:
            plan.append(("mixed_learned_kerker", gpu.F32, True, True))
    runs = {}
    with gpu.Context(g) as ctx:
        ctx.set_nonlocal(st.nloc)
        net = knet = None
        if learned:
            net = ctx.unet(seed=1)
            net.set_weights(unet_weights)
            if kerker_weights is not None:
                knet = ctx.unet(seed=1)
                knet.set_weights(kerker_weights)
                knet.set_shift(-p["kerker_ktf"] ** 2)
        for tag, dt, use_net, use_knet in plan:
            ctx.timer_reset()
            ctx.timer_enable(True)
            want_f = tag in ("f64", "mixed")
            r = ctx.scf_run(st.b, st.rho_at, dict(p, dtype=dt), rho_core=st.rho_core,
                            poisson_net=net if use_net else None,
                            kerker_net=knet if use_knet else None, keep_psi=want_f)
            ctx.timer_enable(False)
            frec = None
            if want_f:
                from sparc_amd import forces as F
                from sparc_amd.sparc_io import occupations
                occ = occupations(r["lambda_"], r["Efermi"], p["beta"], p["smear_type"])
                t1 = time.perf_counter()
                ints = ctx.nonlocal_integrals(r["psi"])
                t_dev = time.perf_counter() - t1
                fnl = F.nonlocal_forces_from_integrals(st, ints, occ, p["occfac"])
                t1 = time.perf_counter()
                floc, fxc = F.local_forces(ctx, st, r["phi"], r["vxc"])
                t_loc = time.perf_counter() - t1
                r["forces"] = F.symmetrize(floc + fnl + fxc)
                r.pop("psi").free()
                fn = np.linalg.norm(r["forces"], axis=1)
                frec = {"nonlocal_integrals_ms": round(t_dev * 1e3, 2),
                        "local_and_nlcc_ms": round(t_loc * 1e3, 2),
                        "max_force": float(fn.max()), "mean_force": float(fn.mean())}
            phases = {}
            for k in ("scf_lanczos", "scf_chefsi", "scf_poisson", "scf_kerker",
                      "scf_density_energy", "filter", "stencil", "nonlocal", "project_gemm",
                      "eigensolve", "rotate_gemm"):
                tot, cnt = ctx.timer(k)
                if cnt:
                    phases[k] = round(tot / r["scf_iters"], 3)
            dense = {}
            ns_, nd_ = p["nstates"], g.nd
            for key in ("project_gemm", "rotate_gemm"):
                tot, cnt = ctx.timer(key)
                if cnt:
                    dense[key] = dense_entry(key, dt == gpu.F32, nd_, ns_, tot / cnt)
            tot, cnt = ctx.timer("eigensolve")
            if cnt:
                dense["eigensolve"] = {"ms": round(tot / cnt, 3), "n": ns_}
            runs[tag] = r
            out[tag] = {"ms_per_scf_iter": round(r["mean_iter_ms"], 3),
                        "dense_per_call": dense,
                        "total_ms": round(r["total_ms"], 1), "scf_iters": r["scf_iters"],
                        "converged": bool(r["converged"]), "Etot": r["Etot"],
                        "Etot_per_atom": r["Etot"] / st.n_atom,
                        "poisson_iters": int(r["poisson_iters"]),
                        "poisson_fallbacks": int(r["poisson_fallbacks"]),
                        "kerker_iters": int(r["kerker_iters"]),
                        "kerker_fallbacks": int(r["kerker_fallbacks"]),
                        "phase_ms_per_iter": phases,
                        "stencil_kernel": ctx.stencil_kernel(dt, p["nstates"])}
            if frec:
                out[tag]["forces"] = frec
            print(f"[bench] scf {name} {tag}: {r['scf_iters']} iters, "
                  f"{r['mean_iter_ms']:.2f} ms/iter, E/atom {r['Etot'] / st.n_atom:.10f}, "
                  f"Poisson AAR iterations {r['poisson_iters']}", flush=True)
        for n_ in (net, knet):
            if n_ is not None:
                n_.close()
    f = runs["f64"]
    for tag in runs:
        if tag == "f64":
            continue
        m = runs[tag]
        sfx = "" if tag == "mixed" else "_" + tag.split("_", 1)[1]
        out[f"dE_per_atom_mixed{sfx}_vs_f64"] = abs(m["Etot"] - f["Etot"]) / st.n_atom
        out[f"max_dlambda_mixed{sfx}_vs_f64"] = float(np.max(np.abs(m["lambda_"] - f["lambda_"])))
        out[f"speedup_mixed{sfx}_vs_f64"] = round(f["mean_iter_ms"] / m["mean_iter_ms"], 3)
    if "forces" in runs.get("mixed", {}) and "forces" in f:
        out["max_dforce_mixed_vs_f64"] = float(np.abs(runs["mixed"]["forces"] - f["forces"]).max())
    return out


def train_unet(train_steps=400):
    """Train the learned Poisson preconditioner on the GPU, on (residual,
    error) pairs harvested from Jacobi-AAR solves of the 8-atom (26^3, run on
    28^3 by interpolation) and 64-atom (52^3) Si SCFs' Poisson problems, Adam
    in blocks alternating between the two grids, the earliest weights with the
    fewest AAR iterations summed over held-out 8- and 64-atom problems kept
    (sparc_amd.learned).
    Returns the FP32 weights and a record of the training."""
    from sparc_amd.learned import train_poisson_unet
    w, info = train_poisson_unet(("si8", "si64"), steps=train_steps)
    print(f"[bench] unet: pairs {info['pairs']}, harvest {info['harvest_s']:.1f}s, train "
          f"{info['train_s']:.1f}s, loss {info['loss_first']:.3e} -> {info['loss_last']:.3e}, "
          f"validation AAR iterations {info['validation_aar_iterations']}, kept after "
          f"{info['kept_after_steps']} steps", flush=True)
    return w, info


def train_kerker_unet(train_steps=400):
    """Train a second network for the Kerker solves -(Lap - kTF^2) Pf = Lf:
    pairs from Jacobi-AAR solves (shift -kTF^2) whose right-hand sides are
    (Lap - 0.1 kTF^2) applied to the density changes of the 64-atom Si SCF's
    first iterations (the form of the mixing residuals Kerker_precond sees).
    Returns the FP32 weights and a record of the training."""
    from sparc_amd import gpu
    from sparc_amd.systems import scf_system
    st, p = scf_system("si64")
    k2 = p["kerker_ktf"] ** 2
    with gpu.Context(st.grid) as ctx:
        ctx.set_nonlocal(st.nloc)
        rhos = [st.rho_at]
        for k in range(1, 6):
            r = ctx.scf_run(st.b, st.rho_at, dict(p, max_scf=k, dtype=gpu.F32), rho_core=st.rho_core)
            rhos.append(r["rho"])
        net = ctx.unet(seed=1)
        net.set_shift(-k2)
        th = time.perf_counter()
        for a, b in zip(rhos[:-1], rhos[1:]):
            rhs = ctx.Lap_vec_mult(b - a, -k2 * p["kerker_thresh"]).ravel()
            net.harvest(rhs, np.zeros(st.grid.nd), tol=1e-8, max_iter=3000, stride=2, max_pairs=60)
        th = time.perf_counter() - th
        tt = time.perf_counter()
        hist = net.train(train_steps, batch=4, lr=2e-3, seed=5)
        tt = time.perf_counter() - tt
        w = net.get_weights()
        info = {"pairs": net.npairs, "steps": train_steps, "batch": 4,
                "harvest_s": round(th, 2), "train_s": round(tt, 2),
                "loss_first": float(hist[:5].mean()), "loss_last": float(hist[-20:].mean()),
                "shift": -k2,
                "source": "Jacobi-AAR Kerker solves (tol 1e-8) with right-hand sides "
                          "(Lap - 0.1 kTF^2) d_k, d_k the density changes of the si64 SCF's "
                          "first five iterations, 52^3"}
        net.close()
    print(f"[bench] kerker unet: {info['pairs']} pairs, train {tt:.1f}s, "
          f"loss {info['loss_first']:.3e} -> {info['loss_last']:.3e}", flush=True)
    return w, info


def poisson_bench(w, training):
    """AAR Poisson solve on the 128^3 / 216-charge case with the Jacobi
    preconditioner and with the learned U-Net preconditioner (weights w)."""
    from sparc_amd import gpu
    from sparc_amd.workloads import poisson_workload
    out = {"training": training}
    g, rhs = poisson_workload()
    out["workload"] = ("AAR Poisson, 128^3 periodic h=0.35, 216 Gaussian charges q=4 widths "
                       "0.8-1.5 neutralised, rel tol 1e-8, AAR 0.6/0.6/7/6, x0 = 0")
    with gpu.Context(g) as ctx:
        net = ctx.unet(seed=1)
        net.set_weights(w)
        x0 = np.zeros(g.nd)
        ctx.AAR(x0, rhs, tol=1e-2, max_iter=50)          # warm-up (allocations, code load)
        ctx.AAR_unet(net, x0, rhs, tol=1e-2, max_iter=50)
        ctx.sync()
        t = time.perf_counter()
        xj, itj, rrj, okj = ctx.AAR(x0, rhs, tol=1e-8, max_iter=10000)
        tj = time.perf_counter() - t
        t = time.perf_counter()
        xu, itu, rru, stu = ctx.AAR_unet(net, x0, rhs, tol=1e-8, max_iter=10000)
        tu = time.perf_counter() - t
        net.close()
    xj -= xj.mean()
    xu -= xu.mean()
    out["jacobi"] = {"iterations": itj, "ms": round(tj * 1e3, 2), "relres": rrj, "converged": okj}
    out["unet"] = {"iterations": itu, "ms": round(tu * 1e3, 2), "relres": rru,
                   "converged": stu == 0}
    out["solution_rel_diff"] = float(np.linalg.norm(xu - xj) / np.linalg.norm(xj))
    out["iteration_ratio"] = round(itj / max(itu, 1), 3)
    out["time_ratio"] = round(tj / tu, 3)
    print(f"[bench] poisson 128^3: Jacobi {itj} it {tj * 1e3:.1f} ms, U-Net {itu} it "
          f"{tu * 1e3:.1f} ms", flush=True)
    return out


def cpu_baseline_scf(name="si8"):
    """The FP64 oracle SCF of the 8-atom Si cell on the host cores (whole run)."""
    from oracle import oracle as O
    from oracle.scf import OracleSCF
    from sparc_amd.systems import scf_system
    from sparc_amd.workloads import uniform_block
    threads, avail, _ = host_threads()
    O.set_threads(threads)
    st, p = scf_system(name)
    nd = st.grid.nd
    psi0 = uniform_block(nd, p["nstates"], p["psi_seed"], -0.5, 0.5)
    x0 = uniform_block(nd, 1, p["lanczos_seed"], 0.0, 1.0)
    t0 = time.perf_counter()
    r = OracleSCF(st.grid, st.nloc, st.b, st.rho_at, p, psi0, x0, rho_core=st.rho_core).run()
    dt = time.perf_counter() - t0
    return {"value": round(dt * 1e3 / r["scf_iters"], 2), "unit": "ms/SCF-iteration",
            "cores": threads, "kind": "port", "host_cpus": avail,
            "sample": f"FP64 oracle SCF, {name} (8-atom Si, psp8, 26^3, 24 states), whole run: "
                      f"{r['scf_iters']} iterations in {dt:.2f} s (mean includes the first "
                      f"iteration)"}


def cpu_baseline_scf_iteration(name="si64", n_iter=2):
    """FP64 oracle SCF of the 64-atom Si cell on the host cores, run for
    n_iter iterations; the time of the last one is the per-iteration figure
    (iterations after the first hold one CheFSI pass, like the device's
    iterations 2..N)."""
    from oracle import oracle as O
    from oracle.scf import OracleSCF
    from sparc_amd.systems import scf_system
    from sparc_amd.workloads import uniform_block
    threads, avail, _ = host_threads()
    O.set_threads(threads)
    st, p = scf_system(name)
    nd = st.grid.nd
    psi0 = uniform_block(nd, p["nstates"], p["psi_seed"], -0.5, 0.5)
    x0 = uniform_block(nd, 1, p["lanczos_seed"], 0.0, 1.0)
    s = OracleSCF(st.grid, st.nloc, st.b, st.rho_at, p, psi0, x0, rho_core=st.rho_core)
    times = []
    for k in range(n_iter):
        t0 = time.perf_counter()
        s.step(k)
        times.append(time.perf_counter() - t0)
    return {"value": round(times[-1] * 1e3, 1), "unit": "ms/SCF-iteration", "cores": threads,
            "kind": "port", "host_cpus": avail,
            "sample": f"FP64 oracle SCF, {name} ({st.n_atom}-atom Si, psp8, {st.grid.nx}^3, "
                      f"{p['nstates']} states, degree {p['cheb_degree']}): iteration {n_iter} "
                      f"(one CheFSI pass, Poisson and Kerker solves); iteration 1 (Lanczos + "
                      f"{p['rho_trigger']} passes) took {times[0]:.2f} s"}


def eigensolve_bench(n=480, reps=5):
    """The FP64 generalized eigensolve (DP_Solve_Generalized_EigenProblem) at n
    states on a matrix pair with a Chebyshev-filtered block's spread of scales
    (Mp diagonal over eight decades, nearly diagonal Hp), per call and per
    phase (HIP events): blocked Cholesky + L^-1 + products, tridiagonal
    reduction (cooperative launch), tridiagonal divide and conquer, dormtr,
    back-transformation; and the same with rocSOLVER's dstedc, with
    rocSOLVER's dsyevd for the eigenproblem and with rocSOLVER's dpotrf +
    triangular solves, for comparison."""
    from sparc_amd import gpu
    from sparc_amd.workloads import Grid
    rng = np.random.default_rng(3)
    d = 10.0 ** rng.uniform(-4, 4, n)
    lam0 = np.sort(rng.uniform(-0.3, 0.5, n))
    E = 1e-3 * rng.standard_normal((n, n))
    Mp = np.sqrt(np.outer(d, d)) * (np.eye(n) + (E + E.T) / 2)
    F = 1e-3 * rng.standard_normal((n, n))
    Hp = np.sqrt(np.outer(d, d)) * (np.diag(lam0) + (F + F.T) / 2)
    phases = ("eig_potrf", "eig_cholesky", "eig_tridiag", "eig_stedc", "eig_ormtr", "eig_back",
              "eigensolve")
    out = {"n": n, "reps": reps}
    settings = {"default": {}, "rocsolver_stedc": {"SPARC_AMD_STEDC": "rocsolver"},
                "dsyevd": {"SPARC_AMD_COOP_TRIDIAG": "0"},
                "rocsolver_cholesky": {"SPARC_AMD_CHOL": "rocsolver"}}
    saved = {k: os.environ.get(k)
             for k in ("SPARC_AMD_COOP_TRIDIAG", "SPARC_AMD_CHOL", "SPARC_AMD_STEDC")}
    try:
        for name, env in settings.items():
            for k in saved:
                os.environ.pop(k, None)
            os.environ.update(env)
            with gpu.Context(Grid(8, 8, 8, 0.5, 0.5, 0.5)) as ctx:
                ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
                ctx.timer_reset()
                ctx.timer_enable(True)
                for _ in range(reps):
                    ctx.DP_Solve_Generalized_EigenProblem(Hp, Mp)
                ctx.timer_enable(False)
                out[name] = {k: round(ctx.timer(k)[0] / reps, 3) for k in phases}
            print(f"[bench] eigensolve n={n} {name}: " +
                  ", ".join(f"{k} {v:.3f} ms" for k, v in out[name].items()), flush=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--ncol", type=int, default=512)
    ap.add_argument("--degree", type=int, default=24)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-poisson", action="store_true",
                    help="skip the learned Poisson preconditioner: its training, the 128^3 "
                         "Jacobi vs learned comparison and the learned SCF run")
    ap.add_argument("--scf", default="si8,si64,sial216",
                    help="comma-separated SCF workloads to time (sic8, si8, si64, sial216) or "
                         "'none'")
    ap.add_argument("--only", choices=["scf", "poisson", "eigensolve"], default=None,
                    help="profiling aid: run only the SCF, the Poisson comparison or the "
                         "480-state eigensolve phases and print its JSON object (not the "
                         "benchmark line)")
    ap.add_argument("--eig-n", type=int, default=480,
                    help="state count of --only eigensolve (up to 180 the reduction runs in one "
                         "workgroup, without the cooperative launch)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one device copy of a 512-vector FP32 block (known bytes) so "
                         "that a profile holds a plain copy kernel to scale HBM counters by")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the path "
                                   "would shard by bands across GPUs, which is not built)"}))
        return 2

    if args.only == "scf":
        w = None if args.no_poisson else train_unet()[0]
        wk = None if args.no_poisson else train_kerker_unet()[0]
        print(json.dumps({name: scf_bench(name, w, wk) for name in args.scf.split(",")}))
        return 0
    if args.only == "poisson":
        print(json.dumps(poisson_bench(*train_unet())))
        return 0
    if args.only == "eigensolve":
        print(json.dumps(eigensolve_bench(n=args.eig_n)))
        return 0

    from sparc_amd import gpu
    from sparc_amd.workloads import operator_filter_workload

    t_setup = time.perf_counter()
    wl = operator_filter_workload(n=args.n, ncol=args.ncol, degree=args.degree)
    nd, ns, m = wl.grid.nd, wl.ncol, wl.degree
    chi = int(wl.nloc.chi.size)
    print(f"[bench] workload {args.n}^3 x {ns}, degree {m}, {wl.nloc.n_influence} sites, "
          f"chi {chi} values, setup {time.perf_counter() - t_setup:.1f}s", flush=True)
    ctx = gpu.Context(wl.grid)
    ctx.set_veff(wl.veff)
    ctx.set_nonlocal(wl.nloc)
    bounds = bounds_lanczos(ctx, wl)
    print(f"[bench] bounds eigmin {bounds[0]:.6f} eigmax {bounds[1]:.6f} cutoff {bounds[2]:.6f}",
          flush=True)

    if args.calib_copy:
        src = ctx.block(ns, gpu.F32).fill_uniform(1, 0.0, 1.0)
        dst = ctx.block(ns, gpu.F32)
        dst.copy_from(src)
        ctx.sync()
        print(f"[bench] calibration copy: {ns} x {nd} FP32 = {4 * ns * nd} bytes read and written",
              flush=True)
        src.free()
        dst.free()

    r32 = time_filter(ctx, gpu, wl, gpu.F32, args.steps, args.warmup, bounds)
    ms32 = r32["filter"][0] / max(r32["filter"][1], 1)
    gbs32 = algorithmic_bytes(nd, ns, m, 4, chi) / (ms32 * 1e-3) / 1e9
    st_ms = r32["stencil"][0] / max(r32["stencil"][1], 1)
    st_gbs = stencil_bytes(nd, ns, m, 4) / m / (st_ms * 1e-3) / 1e9
    print(f"[bench] fp32 filter {ms32:.2f} ms ({gbs32:.0f} GB/s), stencil launch {st_ms:.3f} ms "
          f"({st_gbs:.0f} GB/s), nonlocal {r32['nonlocal'][0] / max(r32['nonlocal'][1], 1):.3f} ms",
          flush=True)
    kernel = ctx.stencil_kernel(gpu.F32, ns)
    res = {
        "metric": METRIC,
        "value": round(gbs32, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["wall_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Chebyshev filter operator case: {args.n}^3 periodic grid h=0.3, "
                               f"order 12, {ns} vectors, degree {m}, 16-Gaussian Veff, "
                               f"{wl.nloc.n_influence} sites x 8 projectors rc 6, bounds from "
                               f"10-step Lanczos",
                   "nd": nd, "ncol": ns, "degree": m},
        "filter_ms": {"f32": round(ms32, 3)},
        "filter_gbs": {"f32": round(gbs32, 2)},
        "roofline": {"bound": "hbm", "achieved": round(st_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(st_gbs / HBM_PEAK_GBS, 4),
                     "traffic": (profiled_traffic(kernel)
                                 if (args.n, ns, m) == (96, 512, 24) else None),
                     "kernel": kernel,
                     "per_launch_algorithmic_bytes": stencil_bytes(nd, ns, m, 4) / m},
        "finite": r32["finite"],
        # the whole FP32 filter (stencil + nonlocal launches) against HBM peak
        "filter_roofline": {"bound": "hbm", "achieved": round(gbs32, 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(gbs32 / HBM_PEAK_GBS, 4)},
        "dense": {"nonlocal_f32": nonlocal_dense(ctx, wl, r32, "fp32")},
    }
    if not args.no_fp64:
        r64 = time_filter(ctx, gpu, wl, gpu.F64, args.steps, args.warmup, bounds)
        ms64 = r64["filter"][0] / max(r64["filter"][1], 1)
        gbs64 = algorithmic_bytes(nd, ns, m, 8, chi) / (ms64 * 1e-3) / 1e9
        res["dense"]["nonlocal_f64"] = nonlocal_dense(ctx, wl, r64, "fp64")
        res["filter_ms"]["f64"] = round(ms64, 3)
        res["filter_gbs"]["f64"] = round(gbs64, 2)
        res["speedup_f32_vs_f64"] = round(ms64 / ms32, 3)
        print(f"[bench] fp64 filter {ms64:.2f} ms ({gbs64:.0f} GB/s)", flush=True)
    ctx.close()
    trained = None if args.no_poisson else train_unet()
    if args.scf and args.scf != "none":
        w = trained[0] if trained else None
        ktrained = None if args.no_poisson else train_kerker_unet()
        wk = ktrained[0] if ktrained else None
        if ktrained:
            res["kerker_learned_precond"] = {"training": ktrained[1]}
        res["scf"] = {name: scf_bench(name, w, wk) for name in args.scf.split(",")}
    if trained:
        res["poisson_learned_precond"] = poisson_bench(*trained)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", flush=True)
        res["cpu_baseline"] = cpu_baseline(wl)
        res["cpu_baseline_scf"] = cpu_baseline_scf()
        res["cpu_baseline_scf_si64"] = cpu_baseline_scf_iteration()
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
