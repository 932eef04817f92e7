#!/usr/bin/env python3
"""Benchmark: batched Lindblad RK4 (QRT) propagation for linear absorption on one MI355X.

Workload (synthetic, seeded): 4096 disorder realisations of an 8x8 nearest-neighbour
exciton lattice (eps_i ~ N(0, 0.05), J = -0.08, pure dephasing gamma = 0.01 on every
site), Hilbert space = ground state + 64 sites (n = 65).  Per realisation three QRT
operators X_a(0) = mu_a |g><g| (a = x, y, z; mu_i unit vectors uniform on the sphere)
are propagated with RK4 (dt = 0.05) and C_a(t) = Tr[mu_a X_a(t)] is recorded after every
step in FP64; the isotropic C(t) is Fourier transformed to the absorption spectrum.

A bench "step" is one launch of the fused kernel advancing all 12288 operators by
``--rk-per-step`` RK4 steps (default 256); the default 16 steps cover the full 4096-step
trajectory.  Inputs are resident in HBM before timing starts.

Metric: rho-steps/s (one N x N operator advanced by one RK4 step = one rho-step).
Algorithmic FLOPs per rho-step: 64 * n^3 (dense Hilbert-space Lindblad with H_eff folding,
on-site dephasing elementwise).  Roofline peak: 157.3 TFLOP/s (MI355X dense FP32 MFMA).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_FP64_MFMA_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


RESIDENT_SOURCES = ("common.hpp", "lindblad_resident.hip", "lindblad_resident_decl.hpp")


def resident_source_hash():
    """First 16 hex digits of the SHA-256 of the resident kernel's source files."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lime_amd", "csrc")
    for name in RESIDENT_SOURCES:
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def profiled_traffic(args, variant):
    """HBM bytes per launch of the default workload from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by a copy-kernel calibration).  Returns
    (bytes or None, note): None when the run's configuration, launch variant or resident
    kernel source differ from the profiled ones (the figure would be stale)."""
    if (args.realisations, args.rk_per_step) != (4096, 256):
        return None, "not the profiled configuration (4096 realisations, 256 steps per launch)"
    if args.precision != "fp32":
        return None, f"profiled for the FP32 kernel; this run is {args.precision}"
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "lindblad_rk4_resident_lattice65_traffic.txt")
    fields = {}
    try:
        for line in open(path):
            if ":" in line:
                k, v = line.split(":", 1)
                fields[k.strip()] = v.strip()
    except OSError:
        return None, f"{path} missing"
    if fields.get("variant") != str(variant):
        return None, f"stale: profiled variant {fields.get('variant')}, this run {variant}"
    if fields.get("source_sha16") != resident_source_hash():
        return None, "stale: the resident kernel source changed since it was profiled"
    try:
        return float(fields["traffic_bytes_per_launch"]), "profiles/lindblad_rk4_resident_lattice65_traffic.txt"
    except (KeyError, ValueError):
        return None, "no traffic_bytes_per_launch line"


def flops_per_rho_step(n, k_dense=0):
    return 64.0 * n ** 3 * (1 + k_dense)


def cpu_baseline(cfg, budget_s=10.0, threads=None):
    """lime's NumPy operation sequence (2 + 5K zgemm per RHS), complex128, on this host."""
    from oracle import lindblad as ol
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
    except ImportError:                       # pragma: no cover
        threadpool_info, threadpool_limits = (lambda: []), None
    H = cfg["H"][0]
    c_ops = cfg["c_ops"]
    X = cfg["x0"][0, 0].copy()
    dt = cfg["dt"]
    info = threadpool_info()
    blas = info[0] if info else {}
    ctx = threadpool_limits(limits=threads) if (threads and threadpool_limits) else None
    try:
        if ctx is not None:
            ctx.__enter__()
        ol.rk4(X.copy(), ol.liouvillian_ref, dt, H, c_ops)          # warm-up
        t0 = time.perf_counter()
        steps = 0
        while time.perf_counter() - t0 < budget_s:
            X = ol.rk4(X, ol.liouvillian_ref, dt, H, c_ops)
            steps += 1
        el = time.perf_counter() - t0
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    used = threads or blas.get("num_threads", 1)
    return {"value": steps / el, "unit": "rho-steps/s", "cores": int(used), "kind": "port",
            "sample": f"realisation 0, component x: {steps} RK4 steps of lime's 2+5K "
                      f"operation sequence (n=65, K=64 on-site ops), NumPy complex128, "
                      f"{el:.1f} s; BLAS {blas.get('internal_api', '?')} "
                      f"{blas.get('version', '')}, host os.cpu_count()={os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rk-per-step", type=int, default=256)
    ap.add_argument("--realisations", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU comparator")
    ap.add_argument("--no-check", action="store_true", help="skip the accuracy checks")
    ap.add_argument("--no-bf3", action="store_true",
                    help="skip the 3xBF16 resident-kernel leg of the default workload")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense-operand leg of the default workload")
    ap.add_argument("--workload", default="lattice65",
                    choices=["lattice65", "ragged", "gradient", "dense512"],
                    help="lattice65 (default): 4096 x 3 QRT operators at n=65; ragged: 2048 "
                         "instances n in {16..256} in one graph; gradient: adjoint of the "
                         "128-site chain spectrum loss; dense512: n=512 with 16 dense L_k")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "3xbf16", "fp64", "fp32/fp64"],
                    help="GEMM arithmetic (ragged / dense512 / gradient; 'fp32/fp64' = FP32 forward "
                         "with FP64 adjoint applications, gradient only)")
    ap.add_argument("--batch", type=int, default=32, help="dense512: operators per launch")
    ap.add_argument("--instances", type=int, default=2048, help="ragged: number of instances")
    ap.add_argument("--ragged-mode", default="persistent", choices=["persistent", "classes"],
                    help="ragged: one persistent launch (default) or the per-class graph")
    ap.add_argument("--check-steps", type=int, default=100,
                    help="dense512: RK4 steps of the FP64 GPU reference comparison")
    ap.add_argument("--gradient-steps", type=int, default=0,
                    help="gradient: RK4 steps of the trajectory (default: the workload's 2000)")
    ap.add_argument("--variant", type=int, default=None,
                    help="launch-shape variant (default: autotuned on the workload)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a copy-like kernel (y = -x, float32) over MIB MiB three "
                         "times: the reference reading for the HBM byte counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "instances shard trivially across GPUs but no multi-GPU path is built.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    torch.cuda.set_device(torch.device("cuda", 0))
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.neg(x, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"kernel": "torch.neg float32 (read once, write once)",
                                               "read_bytes": nbytes, "write_bytes": nbytes,
                                               "launches": 3}}), flush=True)
        return
    if args.workload != "lattice65":
        rk_default = {"ragged": 10, "dense512": 10, "gradient": 0}[args.workload]
        if args.rk_per_step == ap.get_default("rk_per_step"):
            args.rk_per_step = rk_default
        result = {"ragged": bench_ragged, "gradient": bench_gradient,
                  "dense512": bench_dense512}[args.workload](args)
        print(json.dumps(result), flush=True)
        return
    from lime_amd import configs
    from lime_amd.propagator import LindbladPropagator

    dev = torch.device("cuda", 0)
    cfg = configs.lattice_spectrum(seed=args.seed, n_real=args.realisations)
    R, n = args.realisations, cfg["n"]
    B = 3 * R
    dt, chunk = cfg["dt"], args.rk_per_step
    prec = args.precision
    if prec not in ("fp32", "3xbf16"):
        print(f"bench.py: the lattice65 workload runs in fp32 or 3xbf16 (asked for {prec})", file=sys.stderr)
        sys.exit(2)
    log(f"[bench] building operators: {R} realisations x 3 components, n={n}, {prec}")
    prop = LindbladPropagator(cfg["H"], cfg["c_ops"], group=3, precision=prec)
    X = torch.as_tensor(cfg["x0"].reshape(B, n, n), device=dev).contiguous()
    E_t = torch.as_tensor(cfg["mu"].reshape(B, 1, n, n), device=dev).transpose(-1, -2).contiguous()
    total_steps = (args.warmup + args.steps) * chunk
    obs = torch.zeros((B, total_steps, 1), dtype=torch.complex128, device=dev)

    def bench_step(i):
        prop.step(X, dt, chunk, eops_t=E_t, eop_group=1, obs=obs, nout=1,
                  step_offset=i * chunk, rec_capacity=total_steps)

    if args.variant is None:
        from lime_amd import autotune
        autotune.tune(prop, X, dt, steps=16, eops_t=E_t, eop_group=1, verbose=True, force=True)
    else:
        prop.variant = args.variant
    log(f"[bench] launch variant {prop.variant}")

    for i in range(args.warmup):
        bench_step(i)
        torch.cuda.synchronize()
        log(f"[bench] warmup step {i} done")
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        bench_step(args.warmup + i)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    launch_ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    total_ms = ev[0].elapsed_time(ev[-1])
    ms_per_step = total_ms / args.steps
    rho_steps = B * chunk * args.steps
    value = rho_steps / (total_ms / 1e3)
    fl = flops_per_rho_step(n)
    achieved = fl * B * chunk / (np.mean(launch_ms) / 1e3) / 1e12
    log(f"[bench] timed {args.steps} launches: {total_ms:.1f} ms (wall {wall:.2f} s), "
        f"{value:.4g} rho-steps/s, {achieved:.2f} TFLOP/s")

    # spectrum of the isotropic ensemble average over the workload's nt steps (the tail of
    # the pipeline, lime's transform; not timed)
    from lime_amd import fft as gfft
    nspec = min(cfg["nt"], total_steps)
    C = obs[:, :nspec, 0].reshape(R, 3, nspec).sum(dim=1) / 3.0
    spec, _ = gfft.fft(C.mean(dim=0), dt * np.arange(1, nspec + 1))
    peak = float(torch.max(spec.real).item())

    extra = {}
    refs = qrt_references(cfg, min(cfg["nt"], total_steps), R) if not args.no_check else None
    if not args.no_check:
        extra.update(accuracy_checks(cfg, obs, total_steps, dt, R, refs, prec))
    if prec == "fp32" and not args.no_bf3:
        X0 = torch.as_tensor(cfg["x0"].reshape(B, n, n), device=dev).contiguous()
        b3 = bf3_leg(cfg, X0, E_t, args.steps, chunk, refs, check=not args.no_check)
        extra["bf16x3_resident"] = b3
    if not args.no_dense:
        d = dense_operand_leg(cfg, prop.variant, 3, prec)
        extra["dense_operand"] = d
        extra["dense_operand_rho_steps_per_s"] = d["value"]
    if not args.no_plain:
        extra.update(plain_fp64(cfg, B, dt, value))
        if "bf16x3_resident" in extra:
            b3 = extra["bf16x3_resident"]
            b3["speedup_vs_plain_fp64"] = b3["value"] / extra["plain_fp64_gpu"]["value"]

    traffic, traffic_note = profiled_traffic(args, prop.variant)
    if prec == "3xbf16":
        # BF16 MFMA work issued: 4 real products per complex product, 3 split products each,
        # over the nc = n - 1 core, two complex products per stage, plus the border row and
        # column (one 16-wide product set per tile row / column: 2 nc / 16 sets of 4 x 3
        # products of 16 x nc x 16 per complex product pair)
        nc = n - 1
        issued_per_rho_step = 4 * (2 * 4 * 3 * 2 * nc ** 3 + (2 * nc // 16) * 4 * 3 * 2 * 16 * nc * 16)
        issued = {"issued_bf16_tflops": achieved * issued_per_rho_step / fl,
                  "issued_bf16_frac": achieved * issued_per_rho_step / fl / PEAK_BF16_MFMA_TFLOPS}
        kernel = "lindblad_rk4_resident_bf3 (bordered n=65, 3xBF16 split products)"
    else:
        # FP32 MFMA work the kernel issues: the two complex products in the
        # three-multiplication form over the nc = n - 1 core, 6 real products
        # of nc^3 per stage = 48 nc^3 per rho-step (the border row and column
        # run on VALU), against the canonical 64 n^3
        issued = {"issued_mfma_tflops": achieved * 48.0 * (n - 1) ** 3 / (64.0 * n ** 3),
                  "issued_mfma_frac": achieved * 48.0 * (n - 1) ** 3 / (64.0 * n ** 3)
                  / PEAK_FP32_MFMA_TFLOPS}
        kernel = f"lindblad_rk4_resident (bordered n=65, variant {prop.variant})"
    result = {
        "metric": ("RK4 rho-steps/s (batched FP32-MFMA Lindblad QRT propagation)" if prec == "fp32" else
                   "RK4 rho-steps/s (batched 3xBF16-MFMA Lindblad QRT propagation)"),
        "value": value,
        "unit": "rho-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": _dtype(prec),
        "data": "synthetic (seeded 8x8 exciton-lattice disorder ensemble)",
        "config": {"workload": "lattice_spectrum: 4096 realisations x 3 dipole components, "
                               "n=65 (64 sites + ground), dt=0.05, C(t) every step",
                   "realisations": R, "operators": B, "n": n, "rk_steps_per_launch": chunk,
                   "rk_steps_total": total_steps, "seed": args.seed,
                   "flops_per_rho_step": fl},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_MFMA_TFLOPS,
                     "traffic": traffic,
                     "traffic_source": traffic_note + " (rocprofv3 FETCH_SIZE + WRITE_SIZE of "
                                       "the same kernel source and variant, copy-calibrated)",
                     "kernel": kernel,
                     "launch_ms_mean": float(np.mean(launch_ms)),
                     **issued},
        "spectrum_peak": peak,
    }
    result.update(extra)
    if not args.no_cpu:
        log("[bench] CPU baseline (all BLAS threads)")
        result["cpu_baseline"] = cpu_baseline(cfg, budget_s=10.0)
        log("[bench] CPU baseline (1 thread)")
        one = cpu_baseline(cfg, budget_s=6.0, threads=1)
        result["cpu_baseline"]["value_1core"] = one["value"]
    print(json.dumps(result), flush=True)


def qrt_references(cfg, nchk, R):
    """FP64 oracle C_a(t) of the first and the last realisation (three components each)
    over nchk steps: {r: (3, nchk) complex}."""
    from oracle import lindblad as ol
    refs = {}
    for r in (0, R - 1):
        refs[r] = np.stack([ol.propagate(cfg["H"][r], cfg["x0"][r, a], cfg["c_ops"], cfg["dt"], nchk,
                                         [cfg["mu"][r, a]])[1][:, 0] for a in range(3)])
    return refs


def qrt_errors(obs, refs, dt, nchk):
    """Max C(t) error of a run's records against the oracle (relative to max |C|), and the
    max error of each realisation's spectrum (lime's transform, fft.py:9-56: rocFFT on the
    GPU records against numpy FFT on the oracle's) relative to the spectrum's peak."""
    from oracle import spectra as osp
    from lime_amd import fft as gfft
    t = dt * np.arange(1, nchk + 1)
    err, scale, spec_err, spec_scale = 0.0, 0.0, 0.0, 0.0
    for r, ref in refs.items():
        got = obs[3 * r:3 * r + 3, :nchk, 0]
        err = max(err, float(np.max(np.abs(got.cpu().numpy() - ref))))
        scale = max(scale, float(np.max(np.abs(ref))))
        s_gpu, _ = gfft.fft(got.sum(dim=0) / 3.0, t)
        s_ref, _ = osp.fft(ref.sum(axis=0) / 3.0, t)
        spec_err = max(spec_err, float(np.max(np.abs(s_gpu.cpu().numpy() - s_ref))))
        spec_scale = max(spec_scale, float(np.max(np.abs(s_ref))))
    return err, err / scale, spec_err / spec_scale


def population_check(cfg, dt, prec):
    """Density-matrix populations of realisation 0 over the workload's nt steps against the
    oracle: rho0 = |mu_x><mu_x| / <mu_x|mu_x> in the site manifold (impulsive excitation)."""
    from oracle import lindblad as ol
    from lime_amd.propagator import LindbladPropagator
    n, nt = cfg["n"], cfg["nt"]
    v = cfg["mu"][0, 0][:, 0].copy()
    rho0 = np.outer(v, v.conj()) / np.vdot(v, v)
    P = [np.diag(np.eye(n)[i]).astype(np.complex128) for i in range(n)]
    p1 = LindbladPropagator(cfg["H"][0], cfg["c_ops"], precision=prec)
    rho, pobs, _ = p1.run(rho0, dt, nt, e_ops=P, nout=16)
    _, ref, _ = ol.propagate(cfg["H"][0], rho0, cfg["c_ops"], dt, nt, P, nout=16)
    r = rho[0].cpu().numpy()
    return {"pop_max_abs_err": float(np.max(np.abs(pobs[0].cpu().numpy() - ref))),
            "trace_residual": float(abs(np.trace(r) - 1)),
            "hermiticity_residual": float(np.max(np.abs(r - r.conj().T)))}


def accuracy_checks(cfg, obs, total_steps, dt, R, refs, prec="fp32"):
    """Max errors vs the FP64 oracle on the records of the timed run itself."""
    out = {}
    nchk = min(cfg["nt"], total_steps)
    # (1) C_a(t) of the first and the last realisation, all three components, over the whole
    #     nt-step trajectory of the workload, from the timed run's own records; (2) their spectra
    err, rel, spec = qrt_errors(obs, refs, dt, nchk)
    out["qrt_max_abs_err"] = err
    out["qrt_max_rel_err"] = rel
    out["qrt_checked"] = f"realisations 0 and {R - 1}, components x/y/z, records 1..{nchk}"
    out["spectrum_max_rel_err"] = spec
    log(f"[check] C(t) max error {err:.3e} (rel {rel:.3e}) over {nchk} steps, "
        f"realisations 0 and {R - 1}; spectrum rel error {spec:.3e}")
    # (3) density-matrix populations over the full trajectory for realisation 0
    out.update(population_check(cfg, dt, prec))
    log(f"[check] populations over {cfg['nt']} steps: max error {out['pop_max_abs_err']:.3e}; "
        f"trace residual {out['trace_residual']:.2e}; hermiticity {out['hermiticity_residual']:.2e}")
    return out


def bf3_leg(cfg, X0, E_t, launches, chunk, refs, check=True):
    """The same workload on the 3xBF16 resident kernel (split BF16 products on MFMA, FP32
    accumulation, FP64 state): one warm-up launch, then `launches` timed launches covering
    the trajectory; C(t), spectrum and population errors against the FP64 oracle."""
    import torch
    from lime_amd.propagator import LindbladPropagator
    R, n, dt = cfg["H"].shape[0], cfg["n"], cfg["dt"]
    B = 3 * R
    prop = LindbladPropagator(cfg["H"], cfg["c_ops"], group=3, precision="3xbf16")
    X = X0.clone()
    total = (launches + 1) * chunk
    obs = torch.zeros((B, total, 1), dtype=torch.complex128, device=X.device)
    step = lambda i: prop.step(X, dt, chunk, eops_t=E_t, eop_group=1, obs=obs, nout=1,
                               step_offset=i * chunk, rec_capacity=total)
    step(0)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(launches + 1)]
    ev[0].record()
    for i in range(launches):
        step(1 + i)
        ev[i + 1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[-1]) / launches
    value = B * chunk / (ms / 1e3)
    tflops = flops_per_rho_step(n) * B * chunk / (ms / 1e3) / 1e12
    out = {"value": value, "unit": "rho-steps/s", "tflops_canonical": tflops,
           "frac_of_fp32_peak": tflops / PEAK_FP32_MFMA_TFLOPS, "launch_ms_mean": ms,
           "kernel": "lindblad_rk4_resident_bf3 (n = 65, 8 waves x 2 tiles, 3xBF16 split products)"}
    log(f"[bf3] 3xBF16 resident: {value:.4g} rho-steps/s, {tflops:.2f} TFLOP/s canonical ({ms:.1f} ms per launch)")
    if check:
        nchk = min(cfg["nt"], total)
        err, rel, spec = qrt_errors(obs, refs, dt, nchk)
        out.update({"qrt_max_rel_err": rel, "spectrum_max_rel_err": spec})
        out.update(population_check(cfg, dt, "3xbf16"))
        log(f"[bf3] C(t) rel err {rel:.3e}, spectrum {spec:.3e}, populations {out['pop_max_abs_err']:.3e}")
    return out


def dense_operand_leg(cfg, variant, group, prec="fp32", launches=2, chunk=256, seed=7):
    """The headline kernel (same n, batch, launch shape and records) on dense data: GUE
    H_eff sets of the lattice's spectral radius with on-site dephasing, and dense random
    complex operators, so no operand is mostly zeros."""
    import torch
    from oracle import lindblad as ol
    from lime_amd import configs
    from lime_amd.propagator import LindbladPropagator
    rng = np.random.default_rng(seed)
    n = cfg["n"]
    R = cfg["H"].shape[0]
    B = group * R
    radius = float(np.max(np.abs(np.linalg.eigvalsh(cfg["H"][0]))))
    Hs = torch.as_tensor(configs.gue(rng, n, radius)).unsqueeze(0).repeat(R, 1, 1)
    # per-realisation diagonal shift keeps the sets distinct without R eigendecompositions
    Hs = Hs + torch.diag_embed(torch.as_tensor(rng.normal(0, 0.05, (R, n)).astype(np.complex128)))
    prop = LindbladPropagator(Hs, cfg["c_ops"], group=group, precision=prec)
    prop.variant = variant
    dev = prop.dev
    X0 = (rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))) / n
    X = torch.as_tensor(X0, device=dev).contiguous()
    Ew = (rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))) / n
    E_t = torch.as_tensor(Ew, device=dev).unsqueeze(1).transpose(-1, -2).contiguous()
    total = (launches + 1) * chunk
    obs = torch.zeros((B, total, 1), dtype=torch.complex128, device=dev)
    step = lambda i: prop.step(X, cfg["dt"], chunk, eops_t=E_t, eop_group=1, obs=obs, nout=1,
                               step_offset=i * chunk, rec_capacity=total)
    step(0)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(launches + 1)]
    ev[0].record()
    for i in range(launches):
        step(1 + i)
        ev[i + 1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[-1]) / launches
    value = B * chunk / (ms / 1e3)
    tflops = flops_per_rho_step(n) * B * chunk / (ms / 1e3) / 1e12
    # accuracy of one dense operator over the first launch
    _, ref, _ = ol.propagate(Hs[0].numpy(), X0[0], cfg["c_ops"], cfg["dt"], chunk, [Ew[0]])
    err = float(np.max(np.abs(obs[0, :chunk, 0].cpu().numpy() - ref[:, 0])) / np.max(np.abs(ref)))
    log(f"[dense] dense operands: {value:.4g} rho-steps/s, {tflops:.2f} TFLOP/s "
        f"({ms:.1f} ms per launch); C(t) rel err {err:.2e}")
    return {"value": value, "unit": "rho-steps/s", "tflops": tflops,
            "frac": tflops / PEAK_FP32_MFMA_TFLOPS, "launch_ms_mean": ms, "qrt_max_rel_err": err,
            "data": f"{R} GUE H_eff sets (spectral radius {radius:.3f} + N(0, 0.05) site shifts), "
                    f"on-site dephasing, {B} dense random complex operators and observables"}


def plain_fp64(cfg, B, dt, value, steps=48):
    """The plain complex128 GPU propagator (batched zgemm + elementwise) on the same batch."""
    import torch
    from lime_amd.plain import PlainFP64Propagator
    n = cfg["n"]
    p = PlainFP64Propagator(cfg["H"], cfg["c_ops"], group=3)
    X = torch.as_tensor(cfg["x0"].reshape(B, n, n), device="cuda").contiguous()
    X = p.step(X, dt)                       # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        X = p.step(X, dt)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    pv = B * steps / (ms / 1e3)
    log(f"[plain] FP64 GPU propagator: {pv:.4g} rho-steps/s ({ms / steps:.2f} ms per RK4 step)")
    return {"plain_fp64_gpu": {"value": pv, "unit": "rho-steps/s", "rk_steps_timed": steps},
            "speedup_vs_plain_fp64": value / pv}


# ----------------------------------------------------------------------------------------
# Further workloads (not the default line): uneven batch, adjoint gradient, n = 512 dense.

def _base(metric, value, args, ms_per_step, dtype, data, config, roofline):
    return {"metric": metric, "value": value, "unit": "rho-steps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": data, "config": config, "roofline": roofline}


def _roofline(kernel, tflops, launch_ms, precision):
    r = {"bound": "mfma", "achieved": tflops, "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
         "frac": tflops / PEAK_FP32_MFMA_TFLOPS, "traffic": None, "kernel": kernel,
         "launch_ms_mean": launch_ms}
    if precision == "3xbf16":
        # each complex multiply-add issues 12 BF16 slot products (24 flops) for the canonical 8
        r["issued_bf16_tflops"] = 3.0 * tflops
        r["issued_bf16_frac"] = 3.0 * tflops / PEAK_BF16_MFMA_TFLOPS
    if precision == "fp64":
        r["peak"] = PEAK_FP64_MFMA_TFLOPS
        r["frac"] = tflops / PEAK_FP64_MFMA_TFLOPS
    if precision in ("fp32", "fp64"):
        # the FP32 / FP64 tiled panels take 3 real products per complex product (the
        # three-multiplication form): 0.75x the canonical flops are issued on MFMA
        # (resident classes with dense operators, under 2 % of the ragged batch's time,
        # issue 1.0x)
        r["issued_mfma_tflops"] = 0.75 * tflops
        r["issued_mfma_frac"] = 0.75 * tflops / r["peak"]
    return r


def _dtype(precision):
    return {"fp32": "c64-fp32mfma+f64", "3xbf16": "c64-3xbf16mfma+f64", "fp64": "c128-fp64mfma",
            "fp32/fp64": "c64-fp32mfma+f64 forward, c128-fp64mfma adjoint"}[precision]


def _cpu_sample(H, c_ops, X, dt, budget_s, label):
    """lime's operation sequence on one operator (NumPy complex128, all BLAS threads)."""
    from oracle import lindblad as ol
    try:
        from threadpoolctl import threadpool_info
        blas = (threadpool_info() or [{}])[0]
    except ImportError:                       # pragma: no cover
        blas = {}
    ol.rk4(X.copy(), ol.liouvillian_ref, dt, H, c_ops)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        X = ol.rk4(X, ol.liouvillian_ref, dt, H, c_ops)
        steps += 1
    el = time.perf_counter() - t0
    return {"value": steps / el, "unit": "rho-steps/s", "cores": int(blas.get("num_threads", 1)),
            "kind": "port", "sample": f"{label}: {steps} RK4 steps of lime's 2+5K operation "
                                      f"sequence, NumPy complex128, {el:.1f} s; BLAS "
                                      f"{blas.get('internal_api', '?')}"}


def _timed(fn, steps, warmup):
    import torch
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        fn(warmup + i)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    return ev[0].elapsed_time(ev[-1]), per


def bench_dense512(args):
    """n = 512 GUE (spectral radius 1), 16 dense L_k (||L_k||_F = 0.05),
    rho0 = (1 - eta) I/n + eta psi psi^dag, dt = 0.01; batch 1 or 32."""
    import torch
    from lime_amd import configs
    from lime_amd.propagator import LindbladPropagator
    B, prec, chunk = args.batch, args.precision, args.rk_per_step
    distinct = min(B, 4)
    cfg = configs.dense_single(batch=distinct)
    reps = -(-B // distinct)
    H = np.concatenate([cfg["H"]] * reps)[:B]
    Ls = [np.concatenate([L] * reps)[:B] for L in cfg["c_ops"]]
    n, kd = 512, len(Ls)
    log(f"[dense512] batch {B} ({distinct} distinct systems), precision {prec}, {chunk} RK4 steps per launch")
    prop = LindbladPropagator(H, Ls, precision=prec)
    from lime_amd import autotune
    rho0 = torch.as_tensor(np.broadcast_to(cfg["rho0"], (B, n, n)).copy(), device="cuda")
    autotune.tune(prop, rho0, cfg["dt"], steps=2, verbose=True)
    X = rho0.clone()
    total_ms, per = _timed(lambda i: prop.step(X, cfg["dt"], chunk), args.steps, args.warmup)
    rho_steps = B * chunk * args.steps
    value = rho_steps / (total_ms / 1e3)
    fl = flops_per_rho_step(n, kd)
    tflops = fl * B * chunk / (np.mean(per) / 1e3) / 1e12
    log(f"[dense512] {value:.4g} rho-steps/s, {tflops:.1f} TFLOP/s canonical")
    out = _base(f"RK4 rho-steps/s (n=512, 16 dense L_k, {prec})", value, args, total_ms / args.steps,
                _dtype(prec), "synthetic (seeded GUE + dense jump operators)",
                {"workload": f"dense_single: n=512, 16 dense L_k, batch {B}, dt=0.01",
                 "batch": B, "n": n, "n_dense": kd, "eta": cfg["eta"], "seed": cfg["seed"],
                 "rk_steps_per_launch": chunk, "variant": prop.variant,
                 "flops_per_rho_step": fl},
                _roofline("lindblad_stage_tiled + lindblad_left_tiled", tflops, float(np.mean(per)) / (4 * chunk), prec))
    out["roofline"]["launch_ms_mean"] = None
    out["roofline"]["rk_step_ms_mean"] = float(np.mean(per)) / chunk
    if not args.no_check and prec != "fp64":
        m = min(B, 2)
        ns = args.check_steps
        ref = LindbladPropagator(H[:m], [L[:m] for L in Ls], precision="fp64")
        got = LindbladPropagator(H[:m], [L[:m] for L in Ls], precision=prec, variant=prop.variant)
        Xr, Xg = rho0[:m].clone(), rho0[:m].clone()
        ref.step(Xr, cfg["dt"], ns)
        got.step(Xg, cfg["dt"], ns)
        torch.cuda.synchronize()
        d = (Xg - Xr)
        pops = torch.diagonal(d, dim1=-2, dim2=-1).abs().max().item()
        out["check_vs_fp64_gpu"] = {
            "rk_steps": ns, "operators": m,
            "pop_max_abs_err": pops,
            "max_abs_err": d.abs().max().item(),
            "max_rel_err": d.abs().max().item() / Xr.abs().max().item(),
            "trace_residual": (torch.diagonal(Xg, dim1=-2, dim2=-1).sum(-1) - 1).abs().max().item(),
            "hermiticity_residual": (Xg - Xg.conj().transpose(-1, -2)).abs().max().item()}
        log(f"[dense512] vs FP64 GPU reference over {ns} steps: {out['check_vs_fp64_gpu']}")
    if not args.no_plain:
        from lime_amd.plain import PlainFP64Propagator
        p = PlainFP64Propagator(H[:min(B, 4)], [L[:min(B, 4)] for L in Ls])
        Y = rho0[:min(B, 4)].clone()
        Y = p.step(Y, cfg["dt"])
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(2):
            Y = p.step(Y, cfg["dt"])
        e1.record()
        torch.cuda.synchronize()
        pv = min(B, 4) * 2 / (e0.elapsed_time(e1) / 1e3)
        out["plain_fp64_gpu"] = {"value": pv, "unit": "rho-steps/s", "batch": min(B, 4)}
        out["speedup_vs_plain_fp64"] = value / pv
    if not args.no_cpu:
        out["cpu_baseline"] = _cpu_sample(cfg["H"][0], [L[0] for L in cfg["c_ops"]], cfg["rho0"].copy(),
                                          cfg["dt"], 10.0, "system 0 (n=512, K=16 dense)")
    return out


def bench_ragged(args):
    """2048 instances, n uniform in {16, 32, ..., 256}: GUE H (spectral radius 1), n on-site
    dephasing operators (gamma = 0.01) + 4 dense L_k (||L_k||_F = 0.1), rho0 = psi psi^dag,
    dt = 0.02.  Default: every instance in ONE persistent launch (``RaggedLindblad``), the
    launch captured in a hipGraph and replayed; ``--ragged-mode classes``: the size classes
    on the per-class kernels captured into one graph (``RaggedClasses``)."""
    import torch
    from lime_amd import configs
    from lime_amd.ragged import RaggedClasses, RaggedLindblad
    t0 = time.perf_counter()
    cfg = configs.ragged_gue(count=args.instances)
    log(f"[ragged] {args.instances} instances generated in {time.perf_counter() - t0:.1f} s")
    persistent = args.ragged_mode == "persistent"
    if persistent:
        if args.precision != "fp32":
            raise SystemExit("the persistent ragged kernel is built for FP32 MFMA (fp64 state); "
                             "use --ragged-mode classes for other precisions")
        rb = RaggedLindblad(cfg["instances"], tune=True, verbose=True)
        shape = f"persistent launch, variant {rb.variant}"
    else:
        rb = RaggedClasses(cfg["instances"], precision=args.precision, tune=True, verbose=True)
        shape = "hipGraph of per-class resident + tiled launches"
    chunk, dt = args.rk_per_step, cfg["dt"]
    rb.capture(dt, chunk)
    total_ms, per = _timed(lambda i: rb.graph.replay(), args.steps, args.warmup)
    count = len(cfg["instances"])
    value = count * chunk * args.steps / (total_ms / 1e3)
    tflops = rb.flops_per_step * chunk / (np.mean(per) / 1e3) / 1e12
    log(f"[ragged] {value:.4g} rho-steps/s, {tflops:.1f} TFLOP/s canonical ({shape})")
    sizes = np.asarray(cfg["sizes"])
    config = {"workload": f"ragged_gue: {count} instances, n uniform in 16..256 step 16, dt=0.02, "
                          f"one hipGraph replay = {chunk} RK4 steps of every instance",
              "instances": count, "mean_n3": float(np.mean(sizes.astype(np.float64) ** 3)),
              "rk_steps_per_replay": chunk, "seed": cfg["seed"],
              "flops_per_batch_step": rb.flops_per_step, "mode": args.ragged_mode}
    if persistent:
        config["variant"] = rb.variant
    else:
        config["size_classes"] = len(rb.classes)
        config["variants"] = {str(c["n"]): c["prop"].variant for c in rb.classes}
    out = _base(f"RK4 rho-steps/s (uneven batch n in 16..256, {args.precision})", value, args,
                total_ms / args.steps, _dtype(args.precision),
                "synthetic (seeded GUE instances, on-site dephasing + 4 dense L_k)", config,
                _roofline("lindblad_ragged_persistent" if persistent else shape, tflops,
                          float(np.mean(per)), args.precision))
    if not args.no_check:
        from oracle import lindblad as ol
        done = (args.warmup + args.steps) * chunk
        states = rb.states()
        errs = {}
        for target in (16, 128, 256):
            idx = [i for i, n in enumerate(sizes) if n == target]
            if not idx:
                continue
            i = idx[0]
            inst = cfg["instances"][i]
            n = target
            c_ops = list(inst["c_ops"]) + [np.sqrt(inst["dephasing"]) * np.diag(np.eye(n)[k]) for k in range(n)]
            ref, _, _ = ol.propagate(inst["H"], inst["rho0"], c_ops, dt, done)
            g = states[i]
            errs[str(n)] = {"max_abs_err": float(np.max(np.abs(g - ref))),
                            "pop_max_abs_err": float(np.max(np.abs(np.diag(g) - np.diag(ref)))),
                            "trace_residual": float(abs(np.trace(g) - 1))}
        out["check_vs_oracle"] = {"rk_steps": done, "by_n": errs}
        log(f"[ragged] vs oracle after {done} steps: {errs}")
    if not args.no_cpu:
        i = int(np.argmin(np.abs(sizes - 128)))
        inst = cfg["instances"][i]
        n = int(sizes[i])
        c_ops = list(inst["c_ops"]) + [np.sqrt(inst["dephasing"]) * np.diag(np.eye(n)[k]) for k in range(n)]
        out["cpu_baseline"] = _cpu_sample(inst["H"], c_ops, inst["rho0"].copy(), dt, 10.0,
                                          f"instance {i} (n={n}, K={len(c_ops)})")
    return out


def bench_gradient(args):
    """256 chains of 128 sites (+ ground state, n = 129), shared gamma = 0.02, dt = 0.05,
    2000 RK4 steps; spectrum L2 loss against a target from a second seeded parameter set;
    gradients w.r.t. eps, J and gamma by the adjoint kernels (checkpoints every 64 steps).
    One bench step = forward + backward of the whole batch."""
    import torch
    from lime_amd import configs
    from lime_amd.adjoint import ChainAbsorption, spectrum_loss
    cfg = configs.chain_gradient()
    tgt = configs.chain_gradient(seed=cfg["seed"] + 1000)
    dev = torch.device("cuda")
    B, ns, nt, K, dt = cfg["eps"].shape[0], cfg["eps"].shape[1], cfg["nt"], cfg["checkpoint"], cfg["dt"]
    if args.gradient_steps:
        nt = args.gradient_steps                 # a shorter trajectory (profiling runs)
    mu = torch.tensor(cfg["mu"][:, 0], device=dev)
    prec = args.precision
    with torch.no_grad():
        target = ChainAbsorption.apply(torch.tensor(tgt["eps"], device=dev), torch.tensor(tgt["J"], device=dev),
                                       torch.tensor(cfg["gamma"], dtype=torch.float64, device=dev), mu, dt, nt, K, prec)
    torch.cuda.synchronize()
    times = []
    for i in range(args.warmup + args.steps):
        eps = torch.tensor(cfg["eps"], device=dev, requires_grad=True)
        J = torch.tensor(cfg["J"], device=dev, requires_grad=True)
        gamma = torch.tensor(cfg["gamma"], dtype=torch.float64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        C = ChainAbsorption.apply(eps, J, gamma, mu, dt, nt, K, prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss = spectrum_loss(C, target, dt).sum()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        log(f"[gradient] pass {i}: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s, loss {loss.item():.6e}")
        if i >= args.warmup:
            times.append((t1 - t0, t2 - t1))
    fwd = float(np.mean([t[0] for t in times]))
    bwd = float(np.mean([t[1] for t in times]))
    value = B * nt / (fwd + bwd)
    n = ns + 1
    out = _base(f"RK4 rho-steps/s of forward + adjoint gradient (n=129 chain, {prec})", value, args,
                1e3 * (fwd + bwd), _dtype(prec), "synthetic (seeded 128-site chain ensemble)",
                {"workload": f"chain_gradient: {B} chains x 128 sites (+ground, n=129), dt=0.05, "
                             f"{nt} RK4 steps, checkpoints every {K} steps; one step = forward + backward",
                 "batch": B, "n": n, "rk_steps": nt, "checkpoint": K, "seed": cfg["seed"]},
                _roofline("forward: lindblad_stage_tiled; backward: lindblad_apply_tiled + adjoint_contract",
                          flops_per_rho_step(n) * B * nt / fwd / 1e12, None, prec))
    out["forward_s"] = fwd
    out["backward_s"] = bwd
    out["forward_rho_steps_per_s"] = B * nt / fwd
    out["grad_norms"] = {"eps": float(eps.grad.norm()), "J": float(J.grad.norm()), "gamma": float(gamma.grad)}
    if not args.no_cpu:
        from lime_amd.configs import onsite_dephasing
        Hc = np.zeros((n, n), dtype=np.complex128)
        Hc[1:, 1:] = np.diag(cfg["eps"][0]) + np.diag(cfg["J"][0], 1) + np.diag(cfg["J"][0], -1)
        c_ops = onsite_dephasing(n, cfg["gamma"], range(1, n))
        X = np.zeros((n, n), dtype=np.complex128)
        X[1:, 0] = cfg["mu"][:, 0]
        out["cpu_baseline"] = _cpu_sample(Hc, c_ops, X, dt, 10.0, "chain 0 forward (n=129, K=128 on-site)")
    return out


if __name__ == "__main__":
    main()
