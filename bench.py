"""Benchmark: probe·slice steps/s of the batched CN slice stepping, forward + adjoint.

Workload (one "step" of this benchmark = one pass of the hot path over the
whole batch): 2D CN stepping with per-slice tridiagonal LU on four scans --
Nx = 1024 with 512 probes plus Nx = 384, 640, 896 with 128 probes each,
Nz = 1024 (1023 slices), Dirichlet rows, complex64 data with FP64 pivots.  A
pass factors every slice of every scan for the forward and for the adjoint
operator and sweeps all probes forward and adjoint in one persistent launch,
writing the full complex64 fields (916,608 forward + 916,608 adjoint
probe·slice steps).  Inputs (object coefficients, probes) are resident in HBM
before the timed region.

Also reported beside the headline: the plain form (one launch per slice), the
multislice propagator on the main scan (1024-point FFT per slice) and its
gradient, reduced-domain windows, the parallel-in-time GMRES solver, the 3PIE
and NLCG reconstruction drivers, the two-transverse-axis banded LU / learned
Krylov paths and reconstruction loop, the roofline of the dominant kernel
(the persistent sweep) and the FP64 CPU oracle timed on a bounded sample on
this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the persistent sweep on this workload (nz = 1024,
# 4 or 8 probes per workgroup, the 384-wide scan at 8 rows per lane), from
# rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a 1 GiB copy in the same process
# (profiles/tri_sweep_batched_cfg1_traffic.md).
MEASURED_SWEEP_TRAFFIC = {(1024, 4, True): 12.78e9, (1024, 8, True): 12.74e9}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nz", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extras", action="store_true", help="skip plain-form and multislice lines")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D (config 3) banded-LU and learned lines")
    ap.add_argument("--no-recon", action="store_true", help="skip the 3D reconstruction loop (256x256x128)")
    ap.add_argument("--recon-iters", type=int, default=50)
    ap.add_argument("--no-fixture", action="store_true",
                    help="skip the 3D reconstruction's 5-iteration comparisons with the FP64 fixture")
    ap.add_argument("--cpu-probes", type=int, default=128)
    ap.add_argument("--cpu-probes-per-proc", type=int, default=32)
    ap.add_argument("--ppb", default="auto", help="probes per workgroup of the persistent sweep, or 'auto'")
    ap.add_argument("--no-split-chunks", dest="split_chunks", action="store_false",
                    help="step every scan with 16 rows per lane (the 384-wide scan too)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (a kernel of known traffic for PMC scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: the probe-sharded multi-GPU path (probes split across GPUs, one "
            "all-reduce of the gradient per iteration) is not built.",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    from thick_ptycho_amd import workloads
    from thick_ptycho_amd.forward_model import ms_engine, operators, tri_engine

    wl = workloads.config1(nz=args.nz)
    log(f"workload {wl.name}: {wl.steps} forward steps per direction")
    cpu_results = None
    if not args.no_cpu:  # host-only work, before this process touches the GPU (the pool forks)
        cpu_results = (cpu_baseline(wl.scans[0], args.cpu_probes),
                       cpu_baseline_parallel(wl.scans[0], args.cpu_probes_per_proc),
                       cpu_baseline_widened())
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB write")
        del a, b

    scans = []
    for sc in wl.scans:
        A, B = operators.step_operators_1d(sc.nx, sc.k, sc.dx, sc.dz, sc.bc)
        k, dz = sc.k, sc.dz
        coef = (k / 1j) * (sc.n - 1)
        C = (dz / 2) * (coef[:, :-1] + coef[:, 1:]) / 2
        scans.append(
            dict(
                sc=sc,
                A=A,
                B=B,
                C=torch.from_numpy(np.ascontiguousarray(C)).to(dev),
                u0=torch.from_numpy(sc.probes).to(dev, torch.complex64).contiguous(),
                out_f=torch.empty((sc.probes.shape[0], sc.nz, sc.nx), dtype=torch.complex64, device=dev),
                out_a=torch.empty((sc.probes.shape[0], sc.nz, sc.nx), dtype=torch.complex64, device=dev),
            )
        )
    chunk = 16
    steps_per_dir = sum(s["sc"].steps for s in scans)
    steps_per_pass = 2 * steps_per_dir

    def make_specs(split: bool = True):
        # the eight factorisations (scan x direction), one launch per chunk width
        # (rows per lane: 16 for the 640/896/1024-wide scans, 8 for the 384-wide
        # one, or 16 for all), then the sweeps
        jobs = [(s["C"], s["A"], s["B"], mode) for mode in ("forward", "adjoint") for s in scans]
        split = split and args.split_chunks
        widths = [operators.chunk_for(s["sc"].nx) if split else chunk for s in scans + scans]
        facs = [None] * len(jobs)
        for w in sorted(set(widths), reverse=True):
            idx = [i for i, c in enumerate(widths) if c == w]
            for i, fac in zip(idx, tri_engine.factor_many([jobs[i] for i in idx], w)):
                facs[i] = fac
        specs = []
        for (_, _, _, mode), fac, s in zip(jobs, facs, scans + scans):
            key = "out_f" if mode == "forward" else "out_a"
            specs.append(tri_engine.TriScanSpec(fac, s["u0"], out=s[key], flip=(mode == "adjoint")))
        return specs

    def one_pass(plain: bool = False):
        # forward and adjoint of every scan are independent: one persistent
        # launch walks all eight (scan, direction) batches; the plain form takes
        # one chunk width per launch, so it steps every scan at 16 rows per lane
        specs = make_specs(split=not plain)
        if plain:
            tri_engine.sweep_plain(specs)
        else:
            tri_engine.sweep(specs, probes_per_block=args.ppb)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    tune = None
    if args.ppb == "auto":
        from thick_ptycho_amd import autotune

        tune = autotune.tune_sweep(make_specs())
        args.ppb = tune["ppb"]
        log(f"autotuned probes per workgroup: {tune}")
    else:
        args.ppb = int(args.ppb)

    # the plain form captured once into a HIP graph (all 1023 x 8 step launches)
    graph_ms = None
    if not args.no_extras:
        gspecs = make_specs(split=False)
        g = tri_engine.PlainGraph(gspecs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps // 2)):
            g.replay()
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - t0) / max(1, args.steps // 2) * 1e3
        log(f"plain form replayed from a HIP graph: {graph_ms:.3f} ms (sweeps only)")
        del g, gspecs

    log(f"persistent batched form: warmup {args.warmup}, steps {args.steps}")
    t_pass = timed(one_pass, args.steps, args.warmup)
    value = steps_per_pass / t_pass
    log(f"persistent: {t_pass * 1e3:.3f} ms/pass, {value:.4g} steps/s")

    # ---- dominant kernel: the persistent sweep launch (forward + adjoint of all
    # scans), timed with HIP events on the stream it runs on
    specs = make_specs()
    tri_engine.sweep(specs, probes_per_block=args.ppb)
    torch.cuda.synchronize()
    reps = max(3, args.steps)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        tri_engine.sweep(specs, probes_per_block=args.ppb)
    e1.record()
    torch.cuda.synchronize()
    t_kernel = e0.elapsed_time(e1) / 1e3 / reps
    alg_bytes = 0
    for s in scans:
        P, N, L = s["sc"].probes.shape[0], s["sc"].nx, s["sc"].nz - 1
        alg_bytes += 2 * (8 * N * P * L + 24 * N * L + 8 * N * P)
    achieved = alg_bytes / t_kernel / 1e9
    log(f"sweep kernel: {t_kernel * 1e3:.3f} ms/launch, {achieved:.1f} GB/s algorithmic")

    # ---- factor kernel share (all eight factorisations of a pass), on factor
    # buffers the allocator already holds
    del specs
    make_specs()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        make_specs()
    e1.record()
    torch.cuda.synchronize()
    t_factor = e0.elapsed_time(e1) / 1e3 / reps

    extras = {}
    if not args.no_extras:
        log("plain form (one launch per slice)")
        t_plain = timed(lambda: one_pass(plain=True), max(1, args.steps // 2), 1)
        extras["plain_value"] = steps_per_pass / t_plain
        extras["plain_ms_per_step"] = t_plain * 1e3
        log(f"plain: {t_plain * 1e3:.3f} ms/pass, {extras['plain_value']:.4g} steps/s")

        main = scans[0]
        sc = main["sc"]
        n_dev = torch.from_numpy(np.ascontiguousarray(sc.n)).to(dev)
        H = ms_engine.propagator(sc.nx, sc.dx, sc.k, sc.dz, dev)
        ms_out = torch.empty((sc.probes.shape[0], sc.nz, sc.nx), dtype=torch.complex64, device=dev)

        def ms_pass(plain=False):
            T = ms_engine.transmission(n_dev, sc.k, sc.dz, 1.0)
            spec = [ms_engine.MsScanSpec(T, H, main["u0"], out=ms_out)]
            (ms_engine.sweep_plain if plain else ms_engine.sweep)(spec)

        t_ms = timed(ms_pass, args.steps, 1)
        extras["multislice_value"] = sc.steps / t_ms
        t_ms_plain = timed(lambda: ms_pass(True), max(1, args.steps // 2), 1)
        extras["multislice_plain_value"] = sc.steps / t_ms_plain
        log(f"multislice: {extras['multislice_value']:.4g} steps/s persistent, "
            f"{extras['multislice_plain_value']:.4g} plain")
        # multislice amplitude gradient on the same scan: forward + conj(H) adjoint
        # sweeps with every plane kept, loss seed and per-slice accumulation
        from thick_ptycho_amd import gradient as tg

        sqrt_i = torch.sqrt(ms_engine.farfield(ms_out[:, -1].contiguous())) * 1.05
        del ms_out
        t_g = timed(lambda: tg.ms_amplitude_gradient(n_dev, main["u0"], sqrt_i, sc.dx, sc.k, sc.dz), 2, 1)
        extras["multislice_gradient_value"] = 2 * sc.steps / t_g
        extras["multislice_gradient_ms"] = t_g * 1e3
        log(f"multislice gradient: {t_g * 1e3:.2f} ms, {extras['multislice_gradient_value']:.4g} steps/s (fwd+adj)")

    if not args.no_extras:  # the drop-in solve(): full complex128 fields handed over in host memory
        from thick_ptycho_amd import SimulationSpace
        from thick_ptycho_amd.forward_model import PWEIterativeLUSolver

        nx_h, nz_h, P_h = 1024, 512, 256
        ss_h = SimulationSpace(nx_h, nz_h, 2.0, 4.0, num_probes=P_h)
        n_h = workloads.blob_object_2d(nx_h, nz_h, 24, np.random.default_rng(70))
        pr_h = workloads.gaussian_probes_1d(nx_h, np.linspace(nx_h / 4, 3 * nx_h / 4, P_h), 24.0)
        sol_h = PWEIterativeLUSolver(ss_h, pr_h[None], bc_type="impedance")
        t_dev = timed(lambda: sol_h.solve_device(n=n_h), 3, 1)
        t_host = timed(lambda: sol_h.solve(n=n_h), 2, 1)
        f_bytes = P_h * nx_h * nz_h * 16
        extras["host_api"] = {
            "workload": f"drop-in solve(), {nx_h}x{nz_h} x {P_h} probes, forward incl. factorisation, "
                        "complex128 (S, N, nz) fields returned in host memory",
            "device_ms": t_dev * 1e3,
            "host_ms": t_host * 1e3,
            "value_host": P_h * (nz_h - 1) / t_host,
            "handover_gb_per_s": f_bytes / (t_host - t_dev) / 1e9,
        }
        log(f"drop-in solve() with host hand-over: {t_host * 1e3:.1f} ms ({t_dev * 1e3:.2f} ms on the device, "
            f"{f_bytes / 1e9:.2f} GB of complex128 fields at {extras['host_api']['handover_gb_per_s']:.1f} GB/s)")
        del sol_h

    if not args.no_extras:  # reduced domain: every scan position on its own window of the object
        from thick_ptycho_amd import SimulationSpace
        from thick_ptycho_amd.forward_model import PWEIterativeLUSolver

        S_r, step_r, pad_r, nz_r = 64, 24, 232, 512
        nx_r = S_r * step_r + pad_r
        ss_r = SimulationSpace(nx_r, nz_r, 2.0, 4.0, num_probes=S_r,
                               reduced_windows=SimulationSpace.reference_scan_windows(nx_r, S_r, step_r, pad_r))
        rng_r = np.random.default_rng(50)
        n_r = workloads.blob_object_2d(nx_r, nz_r, 24, rng_r)
        pr_r = workloads.gaussian_probes_1d(ss_r.effective_nx, [ss_r.effective_nx / 2] * S_r, 24.0)
        sol_r = PWEIterativeLUSolver(ss_r, pr_r[None], bc_type="impedance", verbose=False)

        def reduced_pass():
            for mode in ("forward", "adjoint"):
                sol_r.solve_device(n=n_r, mode=mode)

        t_r = timed(reduced_pass, 2, 1)
        extras["reduced_domain"] = {
            "workload": f"reduced domain: {S_r} scans of a {nx_r}x{nz_r} object, window {ss_r.effective_nx}, "
                        "forward+adjoint incl. per-scan factorisation",
            "value": 2 * S_r * (nz_r - 1) / t_r,
            "ms_per_pass": t_r * 1e3,
        }
        log(f"reduced domain: {t_r * 1e3:.2f} ms per pass, {extras['reduced_domain']['value']:.4g} steps/s")
        from thick_ptycho_amd import gradient as tgr

        W_r = ss_r.effective_nx
        wins_r = list(ss_r._scan_windows)
        # exit-wave data (the window width, 255, is not a far-field transform length); synthetic: the probes
        t_g = timed(lambda: tgr.amplitude_gradient_reduced(n_r, pr_r, None, wins_r, ss_r.k, 2.0, 4.0, "impedance",
                                                           exit_data=pr_r), 2, 1)
        extras["reduced_domain"]["gradient_ms"] = t_g * 1e3
        extras["reduced_domain"]["gradient_value"] = 2 * S_r * (nz_r - 1) / t_g
        log(f"reduced-domain gradient (window width {W_r}): {t_g * 1e3:.2f} ms, "
            f"{extras['reduced_domain']['gradient_value']:.4g} steps/s (fwd+adj incl. factorisation, exit-wave loss)")
        del sol_r

        # parallel in time: whole-system GMRES with the alpha-circulant preconditioner
        from thick_ptycho_amd.forward_model.pint import PWEFullPinTSolver

        extras["pint"] = []
        for nx_p, nz_p, P_p in ((256, 256, 16), (1024, 512, 8)):
            ss_p = SimulationSpace(nx_p, nz_p, 2.0, 4.0, num_probes=P_p)
            n_p = workloads.blob_object_2d(nx_p, nz_p, 12, np.random.default_rng(60))
            pr_p = workloads.gaussian_probes_1d(nx_p, np.linspace(nx_p / 4, 3 * nx_p / 4, P_p), nx_p / 16)
            sol_p = PWEFullPinTSolver(ss_p, pr_p[None], bc_type="impedance", warm_start=False)
            sol_p.solve_device(n=n_p)  # operator and preconditioner factors (cached per mode)
            t_p = timed(lambda: sol_p.solve_device(n=n_p), 2, 1)
            res = sol_p.last_result
            extras["pint"].append({
                "workload": f"PiT GMRES, {nx_p}x{nz_p} x {P_p} probes, alpha 1e-2, rtol 1e-5, impedance, forward",
                "value": P_p * (nz_p - 1) / t_p,
                "ms_per_solve": t_p * 1e3,
                "iterations": res.iterations,
                "cycles": res.cycles,
                "converged": bool(res.converged.all()),
            })
            log(f"PiT {nx_p}x{nz_p}x{P_p}: {t_p * 1e3:.2f} ms per solve, {res.iterations} iterations, "
                f"{extras['pint'][-1]['value']:.4g} steps/s")
            del sol_p

        # reconstruction drivers: multislice 3PIE (scan after scan) and NLCG on the CN gradient
        from thick_ptycho_amd import reconstruct
        from thick_ptycho_amd.forward_model import ms_engine

        nx_e, nz_e, S_e = 256, 256, 16
        ss_e = SimulationSpace(nx_e, nz_e, 0.5, 1.0, num_probes=S_e)
        n_e = workloads.blob_object_2d(nx_e, nz_e, 12, np.random.default_rng(61), delta=(0, 2e-3), beta=(0, 2e-4))
        pr_e = workloads.gaussian_probes_1d(nx_e, np.linspace(24, nx_e - 24, S_e), 16.0)
        ex_e = torch.empty((S_e, nx_e), dtype=torch.complex64, device=dev)
        ms_engine.sweep([ms_engine.MsScanSpec(
            ms_engine.transmission(torch.from_numpy(n_e).to(dev), ss_e.k, 1.0, 1.0),
            ms_engine.propagator(nx_e, 0.5, ss_e.k, 1.0, dev), torch.from_numpy(pr_e).to(dev, torch.complex64),
            exit=ex_e)])
        I_e = np.abs(np.fft.fft(ex_e.cpu().numpy(), axis=-1)) ** 2
        rc_e = reconstruct.ReconstructorMS(ss_e, I_e, pr_e)
        rc_e.reconstruct(max_iters=1, alpha_obj=1e-3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, h_e = rc_e.reconstruct(max_iters=3, alpha_obj=1e-3)
        t_e = (time.perf_counter() - t0) / 3
        extras["pie_3pie"] = {
            "workload": f"multislice 3PIE, {nx_e}x{nz_e}, {S_e} scans, alpha 1e-3, phase retrieval",
            "ms_per_epoch": t_e * 1e3,
            "value": 2 * S_e * nz_e / t_e,
            "loss": h_e,
        }
        log(f"3PIE: {t_e * 1e3:.2f} ms per epoch ({extras['pie_3pie']['value']:.4g} slice steps/s)")

        sc2 = workloads.config2().scans[0]
        Ig = np.abs(np.fft.fft(sc2.probes, axis=-1)) ** 2  # synthetic data: the probes' free-space far field
        rc2 = reconstruct.Reconstruction2D(Ig, sc2.probes, sc2.k, sc2.dx, sc2.dz, sc2.bc)
        rc2.run(sc2.n, 1, method="nlcg", alpha=1e-9, wave_field=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc2.run(sc2.n, 3, method="nlcg", alpha=1e-9, wave_field=False)
        t_g2 = (time.perf_counter() - t0) / 4  # 3 iterations + the initial gradient
        extras["nlcg_2d"] = {
            "workload": f"NLCG on the CN amplitude gradient, {sc2.nx}x{sc2.nz}, {sc2.probes.shape[0]} probes",
            "ms_per_gradient": t_g2 * 1e3,
            "value": 2 * sc2.probes.shape[0] * (sc2.nz - 1) / t_g2,
        }
        log(f"NLCG 2D: {t_g2 * 1e3:.2f} ms per gradient ({extras['nlcg_2d']['value']:.4g} steps/s fwd+adj)")

    cnn = None
    if not args.no_3d:
        extras["config3"], cnn = bench_3d(dev, args.steps)
    if not args.no_recon:
        extras["config4"] = bench_recon(dev, args.recon_iters, cnn, fixture=not args.no_fixture)

    cpu, cpu_par = None, None
    if not args.no_cpu:
        cpu, cpu_par, cpu_wide = cpu_results
        if "pint" in extras:
            extras["pint"][0]["cpu_oracle"] = cpu_wide["pint_256"]
            log(f"cpu PiT oracle: {cpu_wide['pint_256']['value']:.4g} steps/s")
        if "pie_3pie" in extras:
            extras["pie_3pie"]["cpu_oracle"] = cpu_wide["pie_256"]
            log(f"cpu 3PIE oracle: {cpu_wide['pie_256']['ms_per_epoch']:.1f} ms per epoch")
        log(f"cpu baseline: {cpu['value']:.4g} steps/s ({cpu['sample']})")
        log(f"cpu baseline, probe-parallel: {cpu_par['value']:.4g} steps/s ({cpu_par['sample']})")

    line = {
        "metric": "probe·slice steps/s fwd+adj, batched/learned vs plain, ≤1e-4 rel-L2 vs FP64",
        "value": value,
        "unit": "probe·slice steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_pass * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 (FP64 pivots)",
        "data": "synthetic",
        "config": {
            "workload": "2D CN slice stepping, batched persistent form, forward+adjoint incl. per-slice factorisation",
            "scans": [f"{s['sc'].nx}x{s['sc'].nz} x {s['sc'].probes.shape[0]} probes" for s in scans],
            "steps_per_pass": steps_per_pass,
            "bc": "dirichlet",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"tri_sweep_kernel<16,{args.ppb},false,false,false> (forward+adjoint, 8 batches)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_SWEEP_TRAFFIC.get((args.nz, args.ppb, args.split_chunks)),
            "launch_ms": t_kernel * 1e3,
            "algorithmic_bytes": alg_bytes,
        },
        "factor_ms_per_pass": t_factor * 1e3,
        "autotune": tune,
        "plain_graph_sweep_ms": graph_ms,
        "sweep_only_value": steps_per_pass / t_kernel,
        **extras,
        "cpu_baseline": cpu,
        "cpu_baseline_parallel": cpu_par,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_3d(dev, steps: int) -> dict:
    """Config 3 (128 x 128 x 256, 64 probes): banded-LU pass (forward + adjoint
    factorisation and full-field sweeps), the CNN preconditioner's MFMA rate, and
    CN stepping with BiCGSTAB to 1e-6 (no preconditioner / trained CNN)."""
    import numpy as np
    import torch

    from thick_ptycho_amd import learned as tl
    from thick_ptycho_amd import workloads
    from thick_ptycho_amd.forward_model import band_engine as be

    sc = workloads.config3().scans[0]
    P = sc.probes.shape[0]
    nx, ny, nz = sc.nx, sc.ny, sc.nz
    N, L = nx * ny, nz - 1
    A, B = be.step_operators_3d(nx, ny, sc.k, sc.dx, sc.dy, sc.dz, sc.bc)
    C = tl._contribution(sc.n, sc.k, sc.dz, dev)
    u0 = torch.from_numpy(sc.probes.reshape(P, N).astype(np.complex64)).to(dev)
    out_f = torch.empty((P, nz, N), dtype=torch.complex64, device=dev)
    out_a = torch.empty((P, nz, N), dtype=torch.complex64, device=dev)
    w = be.factor(C, A, B, "forward").w
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def ev(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    facs = {}

    def factor_both():
        facs["f"] = be.factor(C, A, B, "forward", w=w)
        facs["a"] = be.factor(C, A, B, "adjoint", w=w)

    def sweep_both():
        be.sweep([be.BandScanSpec(facs["f"], u0, out=out_f)])
        be.sweep([be.BandScanSpec(facs["a"], u0, out=out_a, flip=True)])

    reps = max(1, min(steps, 3))
    t_factor = ev(factor_both, reps)
    t_sweep = ev(sweep_both, reps)
    units = 2 * P * L
    res = {
        "workload": f"3D CN stepping {nx}x{ny}x{nz}, {P} probes, {sc.bc}, forward+adjoint, full fields",
        "band_value": units / (t_factor + t_sweep),
        "band_sweep_value": units / t_sweep,
        "band_factor_ms": t_factor * 1e3,
        "band_sweep_ms": t_sweep * 1e3,
        "band_halfwidth": w,
    }
    log(f"3D banded LU: factor {t_factor * 1e3:.2f} ms, sweeps {t_sweep * 1e3:.2f} ms (fwd+adj), "
        f"{res['band_value']:.4g} steps/s")
    del out_f, out_a

    # CNN preconditioner: MFMA rate of one application to all probes
    cnn = tl.CNNPreconditioner(seed=0)
    t_cnn = ev(lambda: cnn.apply(u0, nx, ny), 5)
    flops = 39168.0 * N * P  # 2 * 9 * (2*32 + 32*32 + 32*32 + 32*2) per pixel
    res["cnn_apply_ms"] = t_cnn * 1e3
    res["cnn_tflops"] = flops / t_cnn / 1e12
    res["cnn_mfma_frac_bf16_dense"] = flops / t_cnn / 2.5e15
    log(f"CNN apply ({P} planes): {t_cnn * 1e3:.3f} ms, {res['cnn_tflops']:.1f} TFLOP/s")

    # learned stepping: train on 20480 banded-LU residual/correction pairs of two held-out
    # objects (seeds 31-32, 160 slices x 64 planes), step the config-3 object by BiCGSTAB to 1e-6
    objs = [workloads.ellipsoid_object_3d(nx, ny, 256, 32, np.random.default_rng(sd)) for sd in (31, 32)]
    t0 = time.perf_counter()
    pairs = tl.generate_pairs(objs, sc.k, sc.dx, sc.dy, sc.dz, sc.bc, per_slice=64, slices=160, seed=33)
    torch.cuda.synchronize()
    res["cnn_pairs"] = pairs.count
    res["cnn_pairs_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = cnn.train(pairs, steps=1500, batch=32, lr=2e-3)
    torch.cuda.synchronize()
    del pairs
    res["cnn_train_s"] = time.perf_counter() - t0
    res["cnn_train_loss"] = losses[-1]
    for pc in ("none", "jacobi", "cnn"):
        ks = tl.KrylovStepper(sc.n, sc.k, sc.dx, sc.dy, sc.dz, sc.bc, precond=pc, cnn=cnn, tol=1e-6)
        ks.sweep(u0)
        _, st = ks.sweep(u0)
        res[f"krylov_{pc}_value"] = P * L / st.seconds
        res[f"krylov_{pc}_ms"] = st.seconds * 1e3
        res[f"krylov_{pc}_iterations"] = float(np.mean(st.iterations))
        res[f"krylov_{pc}_fallbacks"] = st.fallback_probes
        log(f"3D BiCGSTAB ({pc}): {st.seconds * 1e3:.1f} ms forward sweep, mean iterations "
            f"{res[f'krylov_{pc}_iterations']:.2f}, fallbacks {st.fallback_probes}")
    res["harder_step"] = harder_step_ablation(dev, objs)
    return res, cnn


def harder_step_ablation(dev, objs, dx: float = 0.25, nz: int = 33) -> dict:
    """The config-3 step at dx = dy = 0.25 (dz = 4): the per-slice operator is worse
    conditioned (BiCGSTAB without a preconditioner needs ~20 iterations per slice) and
    the banded LU needs its widest band (half-width 32).  A CNN trained on 20480 pairs
    at this sampling; forward stepping of the 64 probes through 32 slices of the
    config-3 object with none / Jacobi / CNN (to 1e-6) and with the banded LU (factor +
    sweep)."""
    import numpy as np
    import torch

    from thick_ptycho_amd import learned as tl
    from thick_ptycho_amd import workloads
    from thick_ptycho_amd.forward_model import band_engine as be

    sc = workloads.config3(nz=nz).scans[0]
    P = sc.probes.shape[0]
    nx, ny = sc.nx, sc.ny
    L = nz - 1
    u0 = torch.from_numpy(sc.probes.reshape(P, -1).astype(np.complex64)).to(dev)
    pairs = tl.generate_pairs(objs, sc.k, dx, dx, sc.dz, sc.bc, per_slice=64, slices=160, seed=33)
    cnn = tl.CNNPreconditioner(seed=0)
    losses = cnn.train(pairs, steps=1500, batch=32, lr=2e-3)
    del pairs
    out = {"workload": f"CN stepping {nx}x{ny}x{nz} at dx = dy = {dx}, dz = {sc.dz}, {P} probes, forward, exit planes",
           "cnn_train_loss": losses[-1]}
    for pc in ("none", "jacobi", "cnn"):
        ks = tl.KrylovStepper(sc.n, sc.k, dx, dx, sc.dz, sc.bc, precond=pc, cnn=cnn, tol=1e-6, max_iter=400)
        ks.sweep(u0)
        _, st = ks.sweep(u0)
        out[f"{pc}_ms"] = st.seconds * 1e3
        out[f"{pc}_iterations"] = float(np.mean(st.iterations))
        out[f"{pc}_value"] = P * L / st.seconds
    A, B = be.step_operators_3d(nx, ny, sc.k, dx, dx, sc.dz, sc.bc)
    C = tl._contribution(sc.n, sc.k, sc.dz, dev)
    w = be.factor(C, A, B, "forward").w
    ex = torch.empty_like(u0)
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fac = be.factor(C, A, B, "forward", w=w)
        be.sweep([be.BandScanSpec(fac, u0, exit=ex)])
        torch.cuda.synchronize()
    t_band = time.perf_counter() - t0
    out.update({"band_halfwidth": w, "band_ms": t_band * 1e3, "band_value": P * L / t_band})
    log(f"3D step at dx = {dx}: BiCGSTAB none {out['none_iterations']:.2f} its {out['none_ms']:.1f} ms, jacobi "
        f"{out['jacobi_iterations']:.2f} its {out['jacobi_ms']:.1f} ms, CNN {out['cnn_iterations']:.2f} its "
        f"{out['cnn_ms']:.1f} ms; banded LU (w = {w}) {out['band_ms']:.1f} ms")
    return out


RECON_FIXTURE = os.path.join(ROOT, "tests", "golden", "recon3d_256x256x128_fp64.npz")


def bench_recon(dev, iters: int, cnn=None, fixture: bool = True) -> dict:
    """3D reconstruction loop on 256 x 256 x 128 with 400 probes: intensities of the
    seeded true object simulated on the GPU, start n0 = 1, ``iters`` steepest-descent
    iterations (forward + exact adjoint banded-LU stepping of every probe, 2D far-field
    loss seed, gradient, update; one HIP graph per iteration, step 0.5 / secant
    Lipschitz estimate).  Reports iterations/s, probe·slice steps/s, the object error
    against the true object and, against the FP64 CPU reconstruction fixture, the
    object after 5 iterations at the fixture's step.  The 2D-FFT multislice of the
    same object and probes is timed alongside."""
    import numpy as np
    import torch

    from thick_ptycho_amd import recon3d, workloads
    from thick_ptycho_amd.forward_model import ms_engine

    sc = workloads.config4().scans[0]
    nx, ny, nz = sc.nx, sc.ny, sc.nz
    P, L = sc.probes.shape[0], nz - 1
    probes = torch.from_numpy(sc.probes.reshape(P, -1).astype(np.complex64)).to(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    I = recon3d.simulate_intensity_3d(sc.n, probes, sc.k, sc.dx, sc.dy, sc.dz, sc.bc)
    e1.record()
    torch.cuda.synchronize()
    res = {"workload": f"3D reconstruction {nx}x{ny}x{nz}, {P} probes, {sc.bc}, n0 = 1, steepest descent",
           "simulate_ms": e0.elapsed_time(e1)}
    n0 = np.ones((nx, ny, nz), dtype=complex)
    n_true = torch.from_numpy(sc.n).to(dev)
    I = I.sqrt_()
    rc = recon3d.Reconstruction3D(I, probes, n0, sc.k, sc.dx, sc.dy, sc.dz, sc.bc)
    res["gradient_ms_by_sweep_shape"] = rc.tune()
    res["probes_per_block"], res["sweep_schedule"] = rc.ppb, rc.schedule
    log(f"3D reconstruction: gradient ms by sweep shape (probes per workgroup / schedule) "
        f"{res['gradient_ms_by_sweep_shape']}")
    alpha = 0.5 / rc.lipschitz()
    rc.capture()
    rc.run(1, alpha=alpha, check_band=False)  # warm-up replay
    rc.reset(n0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = rc.run(iters, alpha=alpha, graph=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    err = lambda: float(torch.linalg.vector_norm(rc.n - n_true) / torch.linalg.vector_norm(n_true - 1))  # noqa: E731
    res.update({
        "iterations": iters,
        "iteration_ms": dt / iters * 1e3,
        "iterations_per_s": iters / dt,
        "value": 2 * P * L * iters / dt,
        "unit": "probe·slice steps/s (forward + adjoint)",
        "alpha": alpha,
        "band_halfwidth": rc.w,
        "loss_first": out["loss"][0],
        "loss_last": out["loss"][-1],
        "object_error_vs_truth_start": 1.0,
        "object_error_vs_truth": err(),
    })
    log(f"3D reconstruction: {iters} iterations in {dt:.2f} s ({res['iteration_ms']:.1f} ms each), "
        f"loss {out['loss'][0]:.4e} -> {out['loss'][-1]:.4e}, object error vs truth {res['object_error_vs_truth']:.4f}")
    if fixture and os.path.exists(RECON_FIXTURE):
        fx = np.load(RECON_FIXTURE)
        rc.reset(n0)
        o5 = rc.run(5, alpha=float(fx["alpha"]), graph=True)
        n5 = rc.n.reshape(-1)[torch.from_numpy(fx["idx"]).to(dev)].cpu().numpy()
        ref = fx["n5"]
        res["fp64_5iter"] = {
            "alpha": float(fx["alpha"]),
            # ||n5 - n5_fp64|| / ||n5_fp64 - n0|| over the sample (n0 = 1): error relative to the update
            "update_rel_l2_vs_fp64": float(np.linalg.norm(n5 - ref) / np.linalg.norm(ref - 1)),
            "loss_rel_vs_fp64": float(np.max(np.abs(np.array(o5["loss"]) - fx["loss"]) / fx["loss"])),
            "object_error_vs_truth_gpu": err(),
            "object_error_vs_truth_fp64": float(fx["err5"]),
            "sample": f"{ref.size} seeded positions of the object",
        }
        log(f"vs FP64 CPU reconstruction (5 iterations): {res['fp64_5iter']}")
    hist_band = out["loss"]
    del rc
    if fixture and os.path.exists(RECON_FIXTURE):
        # refined exit plane (two FP64-residual correction sweeps) and FP64 loss seed,
        # with the FP64 intensities of refined exit planes as data (complex128 probes)
        fx = np.load(RECON_FIXTURE)
        p128 = torch.from_numpy(sc.probes.reshape(P, -1)).to(dev)
        parts = recon3d.simulate_exit_3d(sc.n, p128, sc.k, sc.dx, sc.dy, sc.dz, sc.bc, refine=2).cpu().numpy()
        amp64 = np.abs(np.fft.fft2(parts.astype(complex).sum(0).reshape(P, nx, ny), axes=(-2, -1)))
        del parts
        rr = recon3d.Reconstruction3D(torch.from_numpy(amp64).to(dev), p128, n0, sc.k, sc.dx, sc.dy, sc.dz, sc.bc,
                                      probes_per_block=res["probes_per_block"], refine=2)
        del amp64
        rr.schedule = res["sweep_schedule"]
        rr.alpha.fill_(float(fx["alpha"]))
        rr.capture()
        rr.run(1, alpha=float(fx["alpha"]), check_band=False)  # warm-up replay
        rr.reset(n0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        o5 = rr.run(5, alpha=float(fx["alpha"]), graph=True)
        torch.cuda.synchronize()
        dtr = (time.perf_counter() - t0) / 5
        n5 = rr.n.reshape(-1)[torch.from_numpy(fx["idx"]).to(dev)].cpu().numpy()
        ref = fx["n5"]
        res["refined"] = {
            "iteration_ms": dtr * 1e3,
            "value": 2 * P * L / dtr,
            "fp64_5iter": {
                "update_rel_l2_vs_fp64": float(np.linalg.norm(n5 - ref) / np.linalg.norm(ref - 1)),
                "loss_rel_vs_fp64": float(np.max(np.abs(np.array(o5["loss"]) - fx["loss"]) / fx["loss"])),
                "object_error_vs_truth_gpu": float(torch.linalg.vector_norm(rr.n - n_true)
                                                   / torch.linalg.vector_norm(n_true - 1)),
                "object_error_vs_truth_fp64": float(fx["err5"]),
            },
        }
        log(f"3D reconstruction with the refined exit plane: {dtr * 1e3:.1f} ms per iteration, {res['refined']}")
        del rr
    if cnn is not None:  # learned-preconditioned Krylov forward step (banded LU: adjoint and fallback)
        rk = recon3d.Reconstruction3D(I, probes, n0, sc.k, sc.dx, sc.dy, sc.dz, sc.bc, forward="krylov", cnn=cnn)
        rk.capture()  # one iteration in a HIP graph: fixed BiCGSTAB iterations per slice, flags on the device
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ok = rk.run(2, alpha=alpha, graph=True)
        torch.cuda.synchronize()
        dtk = (time.perf_counter() - t0) / 2
        flags = rk.krylov_done[rk.krylov_done > 0]  # BiCGSTAB iterations taken, last replay
        res["krylov_forward"] = {
            "iteration_ms": dtk * 1e3,
            "value": 2 * P * L / dtk,
            "mean_bicgstab_iterations": float(flags.float().mean()) if flags.numel() else float("nan"),
            "fallback_probes": int(sum(st.fallback_probes for st in rk.krylov_stats)),
            "graph_iterations_redone_eagerly": rk.krylov_redone,
            "bicgstab_iterations_per_slice_in_graph": rk.krylov_graph_iters,
            "loss_rel_vs_band": float(max(abs(a - b) / b for a, b in zip(ok["loss"], hist_band[:2]))),
        }
        log(f"3D reconstruction, CNN-preconditioned BiCGSTAB forward: {dtk * 1e3:.1f} ms per iteration, "
            f"{res['krylov_forward']}")
        del rk
    # 2D-FFT multislice of the true object, exit planes of all probes
    reps = 3
    ms_engine.multislice_2d(n_true, probes, sc.dx, sc.dy, sc.k, sc.dz)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        ms_engine.multislice_2d(n_true, probes, sc.dx, sc.dy, sc.k, sc.dz)
    e1.record()
    torch.cuda.synchronize()
    t_ms = e0.elapsed_time(e1) / 1e3 / reps
    res["multislice2d_ms"] = t_ms * 1e3
    res["multislice2d_value"] = P * L / t_ms
    log(f"2D-FFT multislice {nx}x{ny}, {P} probes, {L} slices: {t_ms * 1e3:.2f} ms, "
        f"{res['multislice2d_value']:.4g} steps/s")
    return res


def cpu_baseline(sc, probes: int) -> dict:
    """FP64 oracle on the host: per-slice SuperLU factorisation and, per probe
    and slice, a CSR mat-vec plus a SuperLU solve (the reference's loop),
    forward and adjoint, on the first ``probes`` probes of the main scan."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import pwe

    g = pwe.Grid(dimension=2, nx=sc.nx, nz=sc.nz, dx=sc.dx, dz=sc.dz, k=sc.k)
    with threadpool_limits(1):
        t0 = time.perf_counter()
        for mode in ("forward", "adjoint"):
            cache = pwe.build_lu_cache(g, sc.bc, sc.n, mode)
            for p in range(probes):
                pwe.solve_single_probe(g, cache, sc.probes[p], mode)
        dt = time.perf_counter() - t0
    steps = 2 * probes * (sc.nz - 1)
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": steps / dt,
        "unit": "probe·slice steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{probes} probes x {sc.nz - 1} slices of the {sc.nx}-wide scan, forward+adjoint incl. "
                  f"factorisation, complex128, 1 thread ({cpu_model}; host has {os.cpu_count()} cpus)",
        "seconds": dt,
    }


def cpu_baseline_widened() -> dict:
    """FP64 oracle, 1 thread, on the workloads of the widened lines: the
    parallel-in-time solve (scipy GMRES with the splu-factored alpha-circulant
    blocks, one probe at a time as the reference) on 4 of the 16 probes of the
    256 x 256 case, and one 3PIE epoch of the 256 x 256, 16-scan case."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import multislice as oms
    from oracle import pie, pint, pwe
    from thick_ptycho_amd import workloads

    out = {}
    with threadpool_limits(1):
        nx, nz, P = 256, 256, 16
        g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=2 * np.pi)
        n = workloads.blob_object_2d(nx, nz, 12, np.random.default_rng(60))
        pr = workloads.gaussian_probes_1d(nx, np.linspace(nx / 4, 3 * nx / 4, P), nx / 16)
        t0 = time.perf_counter()
        for p in range(4):
            pint.pit_solve(g, "impedance", n, pr[p])
        dt = (time.perf_counter() - t0) / 4
        out["pint_256"] = {"value": (nz - 1) / dt, "unit": "probe·slice steps/s", "cores": 1, "kind": "port",
                           "sample": "4 probes of the 256 x 256 PiT case, alpha 1e-2, rtol 1e-5, complex128"}
        nx_e, nz_e, S_e = 256, 256, 16
        n_e = workloads.blob_object_2d(nx_e, nz_e, 12, np.random.default_rng(61), delta=(0, 2e-3), beta=(0, 2e-4))
        pr_e = workloads.gaussian_probes_1d(nx_e, np.linspace(24, nx_e - 24, S_e), 16.0)
        ex = oms.solve_batched(n_e, pr_e, 0.5, 2 * np.pi, 1.0, keep="exit")
        I_e = np.abs(np.fft.fft(ex, axis=-1)) ** 2
        t0 = time.perf_counter()
        pie.reconstruct_3pie(pr_e, I_e, nx_e, nz_e, 0.5, 2 * np.pi, 1.0, max_iters=1, alpha_obj=1e-3)
        dt = time.perf_counter() - t0
        out["pie_256"] = {"value": 2 * S_e * nz_e / dt, "unit": "slice steps/s", "ms_per_epoch": dt * 1e3,
                          "cores": 1, "kind": "port", "sample": "one 3PIE epoch, 256 x 256, 16 scans, complex128"}
    return out


_CPU_SCAN = None


def _cpu_probe_share(probes) -> int:
    """One worker of the probe-parallel baseline: its own forward and adjoint
    factorisation, then its probes (the reference's loop, one thread)."""
    from threadpoolctl import threadpool_limits

    from oracle import pwe

    sc = _CPU_SCAN
    g = pwe.Grid(dimension=2, nx=sc.nx, nz=sc.nz, dx=sc.dx, dz=sc.dz, k=sc.k)
    with threadpool_limits(1):
        for mode in ("forward", "adjoint"):
            cache = pwe.build_lu_cache(g, sc.bc, sc.n, mode)
            for pr in probes:
                pwe.solve_single_probe(g, cache, pr, mode)
    return 2 * len(probes) * (sc.nz - 1)


def cpu_baseline_parallel(sc, per_proc: int) -> dict:
    """The same FP64 loop, probe-parallel over the host cores available to this
    job (at most 16): each process factors the slices itself and steps its own
    probes (factorisation replicated, as a multi-process CPU run would)."""
    import multiprocessing

    global _CPU_SCAN
    _CPU_SCAN = sc
    procs = max(1, min(16, os.cpu_count() or 1, sc.probes.shape[0] // max(1, per_proc)))
    shares = [sc.probes[i * per_proc:(i + 1) * per_proc] for i in range(procs)]
    ctx = multiprocessing.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        steps = sum(pool.map(_cpu_probe_share, shares))
    dt = time.perf_counter() - t0
    return {
        "value": steps / dt,
        "unit": "probe·slice steps/s",
        "cores": procs,
        "kind": "port",
        "sample": f"{procs} processes x {per_proc} probes x {sc.nz - 1} slices of the {sc.nx}-wide scan, "
                  f"forward+adjoint, factorisation replicated per process, complex128",
        "seconds": dt,
    }


if __name__ == "__main__":
    sys.exit(main())
