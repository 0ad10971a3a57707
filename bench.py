"""Benchmark: full incompressible RHS evaluations per second at 512^3 on one
MI355X, FP32 storage with FP64-residual refinement versus FP64.

One step = one full RHS evaluation (pdo_rhs_eval): 9 cd06 velocity-gradient
passes, 9 cf90 filter passes, the cd06 divergence (3 passes), the periodic
FFT Poisson solve with cd06 modified wavenumbers (5 passes) and the pressure
gradient subtraction (3 passes), on a triply periodic 512^3 grid of side 2 pi.
Input: seeded divergence-free isotropic turbulence, E(k) ~ k^4 exp(-2 (k/8)^2),
rms 1, generated on the device; resident in HBM before timing starts.

Prints ONE JSON line.  value = FP32+refine evaluations/s (whole job, 1 GPU).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8000.0  # GB/s, MI355X HBM3E
# algorithmic words per point per RHS evaluation: 18 gradient passes, 18 filter
# passes, 8 divergence, 6 + 6 FFT passes, 9 pressure-gradient subtraction
WORDS_PER_POINT = 65


def turbulence(n, k0, seed, dtype, device):
    """Divergence-free random-phase field with E(k) ~ k^4 exp(-2 (k/k0)^2), rms 1."""
    from padeops_amd.synthetic import turbulence as gen
    return gen(n, k0, seed, dtype, device)


def rel_l2(a, b):
    import torch
    a = a.double()
    b = b.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-breakdown", action="store_true", help="skip the per-pass timing loop")
    ap.add_argument("--profile-pass", default=None,
                    help="only launch this pass type (cd06_x .. cf90_z, poisson_5pass) --launches times, then a "
                         "1 GiB device copy as many times (the byte-counter calibration of rocprofv3 --pmc runs)")
    ap.add_argument("--launches", type=int, default=5)
    ap.add_argument("--only", choices=["wall_les"], default=None,
                    help="run only this secondary workload and print its JSON (the wall-modelled LES step)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (sharding is a design only)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import padeops_amd as P

    dev = torch.device("cuda:0")
    n = args.n
    if args.profile_pass:
        profile_pass(P, n, dev, args.profile_pass, args.launches)
        return
    if args.only == "wall_les":
        print(json.dumps({"wall_les_512x512x256": wall_les_bench(P, dev, torch.cuda.current_stream().cuda_stream)}),
              flush=True)
        return
    N = n ** 3
    L = 2 * np.pi
    stream = torch.cuda.current_stream().cuda_stream

    # inputs (FP64 master field; the FP32 path gets it rounded)
    u64 = turbulence(n, 8.0, 1234, torch.float64, dev)
    u32 = [c.float() for c in u64]

    def buffers(dt):
        return (torch.empty(9 * N, dtype=dt, device=dev), [torch.empty(N, dtype=dt, device=dev) for _ in range(3)])

    def run(rhs, U, grad, out, steps):
        for _ in range(steps):
            rhs.eval(U[0], U[1], U[2], grad, *out, stream=stream)

    def timed(rhs, U, grad, out):
        run(rhs, U, grad, out, args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run(rhs, U, grad, out, args.steps)
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return e0.elapsed_time(e1) / args.steps, wall

    # FP64 path
    rhs64 = P.FullRHS(n, L, "f64")
    g64, o64 = buffers(torch.float64)
    ms64, _ = timed(rhs64, u64, g64, o64)
    div64 = rhs64.max_divergence(*o64)
    # FP32 + refinement path (the headline)
    rhs32 = P.FullRHS(n, L, "f32_refine")
    g32, o32 = buffers(torch.float32)
    ms32, wall32 = timed(rhs32, u32, g32, o32)

    # FP32 storage with FP64 arithmetic (no refinement needed), for comparison
    rhs32s = P.FullRHS(n, L, "f32_store")
    ms32s, _ = timed(rhs32s, u32, g32, o32)
    errs_s = {}
    for q in (0, 4, 8):
        errs_s["grad%d" % q] = rel_l2(g32[q * N:(q + 1) * N], g64[q * N:(q + 1) * N])
    for name, a, b in zip(("u_proj", "v_proj", "w_proj"), o32, o64):
        errs_s[name] = rel_l2(a, b)
    del rhs32s
    ms32, wall32 = timed(rhs32, u32, g32, o32)

    # accuracy of the FP32+refine evaluation against the FP64 one
    errs = {}
    names = ["dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz"]
    for q in range(9):
        errs[names[q]] = rel_l2(g32[q * N:(q + 1) * N], g64[q * N:(q + 1) * N])
    for name, a, b in zip(("u_proj", "v_proj", "w_proj"), o32, o64):
        errs[name] = rel_l2(a, b)

    evals64 = 1000.0 / ms64
    evals32 = 1000.0 / ms32
    bytes64 = WORDS_PER_POINT * N * 8
    bytes32 = WORDS_PER_POINT * N * 4

    # per-pass breakdown and the dominant kernel (HIP events on the launch stream)
    breakdown = {}
    roofline = None
    if not args.no_breakdown:
        breakdown, roofline = pass_breakdown(P, n, u32[0], stream, args.steps)

    probe = None if args.no_breakdown else hbm_probe(args.steps)
    if roofline is not None:
        roofline["traffic"] = committed_traffic(roofline["kernel"])

    # the mixed path's error of each filter direction (standalone cf90 passes)
    errs.update(filter_errors(P, n, u64[0], u32[0], stream))
    del g64, o64, g32, o32, rhs64, rhs32
    torch.cuda.empty_cache()

    # the wall-bounded 512 x 512 x 256 box: per-operator times and errors
    wall = None if args.no_breakdown else wall_box(P, dev, stream, args.steps)
    torch.cuda.empty_cache()
    # the compressible (cgrid) right-hand side at 256^3
    comp = None if args.no_breakdown else cgrid_bench(P, dev, stream, max(3, args.steps // 4))
    torch.cuda.empty_cache()
    # the wall-modelled LES step of the 512 x 512 x 256 box with the learned
    # closure in the loop
    les = None if args.no_breakdown else wall_les_bench(P, dev, stream)
    torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()

    line = {
        "metric": "RHS evals/s at 512^3: FP32+refine vs FP64 (<=1e-6 rel err); HBM GB/s vs peak",
        "value": round(evals32, 3),
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 storage+solves, fp64 residual refinement",
        "data": "synthetic: seeded divergence-free isotropic turbulence, E(k)~k^4 exp(-2(k/8)^2), rms 1",
        "config": {"workload": "full incompressible RHS (cd06 grad-u x9, cf90 x9, cd06 div, FFT Poisson "
                               "with cd06 modified wavenumbers, cd06 grad p), triply periodic %d^3, L=2pi" % n,
                   "grid": n, "passes_per_eval": 27, "words_per_point": WORDS_PER_POINT},
        "fp64_evals_per_s": round(evals64, 3),
        "fp64_ms_per_eval": round(ms64, 4),
        "speedup_fp32r_vs_fp64": round(evals32 / evals64, 3),
        "rhs_achieved_GBps_fp32r": round(bytes32 / (ms32 * 1e-3) / 1e9, 1),
        "rhs_achieved_GBps_fp64": round(bytes64 / (ms64 * 1e-3) / 1e9, 1),
        "rhs_frac_of_peak_fp32r": round(bytes32 / (ms32 * 1e-3) / 1e9 / PEAK_HBM, 4),
        "rhs_frac_of_peak_fp64": round(bytes64 / (ms64 * 1e-3) / 1e9 / PEAK_HBM, 4),
        "fp32r_rel_l2_vs_fp64": {k: (None if v is None else float("%.3e" % v)) for k, v in errs.items()},
        "fp64_max_div_after_projection": float("%.3e" % div64),
        "fp32_store_fp64_arith_evals_per_s": round(1000.0 / ms32s, 3),
        "fp32_store_rel_l2_vs_fp64": {k: float("%.3e" % v) for k, v in errs_s.items()},
        "roofline": roofline,
        "pass_ms": breakdown,
        "wall_box_512x512x256": wall,
        "cgrid_rhs_256": comp,
        "wall_les_512x512x256": les,
        "hbm_probe": probe,
        "cpu_baseline": cpu,
        "wall_s_timed_fp32r": round(wall32, 4),
    }
    print(json.dumps(line))


def filter_errors(P, n, f64, f32, stream):
    """Relative L2 error of the FP32+refine cf90 filter along x, y and z
    against the FP64 one, on the first velocity component."""
    import torch
    c = P.cf90()
    c.init(n, True)
    out = {}
    for d, name in ((1, "cf90_x"), (2, "cf90_y"), (3, "cf90_z")):
        a = torch.empty_like(f64)
        b = torch.empty_like(f32)
        getattr(c, "filter%d" % d)(f64, a, n, n, prec="f64", stream=stream)
        getattr(c, "filter%d" % d)(f32, b, n, n, prec="f32_refine", stream=stream)
        torch.cuda.synchronize()
        out[name] = rel_l2(b, a)
        del a, b
    return out


# algorithmic words per point of one cgrid RHS evaluation: 69 compact passes
# (9 velocity gradients; the artificial shear viscosity's 6 second-derivative
# and 6 Gaussian passes; the bulk viscosity's 3 + 6 + 6 and the conductivity's
# 3 + 6 + 6; 3 temperature gradients; 15 flux derivatives) of 2 words each,
# and the pointwise kernels' reads and writes (primitives 12, transport 4,
# strain magnitude 10, mustar 5, dilatation 4, bulkstar 15, kapstar 10,
# coefficient sums 9, tau 17, heat flux 9, fluxes 3 x 15, rhs updates
# 10 + 2 x 15)
CGRID_WORDS = 69 * 2 + 12 + 4 + 10 + 5 + 4 + 15 + 10 + 9 + 17 + 9 + 45 + 40


def cgrid_bench(P, dev, stream, steps):
    """The compressible RHS of cgrid (one species, viscous, LAD, triply
    periodic cd10) at 256^3 with a smooth seeded state: ms per evaluation in
    FP64 and FP32+refine (HIP events on the launch stream), the algorithmic
    GB/s of CGRID_WORDS words per point, and the FP32+refine error of each
    RHS component against FP64 on the same FP32-rounded state."""
    import math

    import torch
    n = 256
    L = 2 * math.pi
    x = torch.arange(n, device=dev, dtype=torch.float64) * (L / n)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    g = torch.Generator().manual_seed(99)

    def smooth():
        a = (torch.rand(6, generator=g, dtype=torch.float64) * 2 - 1).tolist()
        return (a[0] * torch.sin(X + a[1]) * torch.cos(2 * Y) + a[2] * torch.cos(3 * Z + Y)
                + a[3] * torch.sin(2 * X - Z) + a[4] * torch.cos(Y + a[5]))
    rho = 1.0 + 0.15 * smooth()
    u, v, w = 0.3 * smooth(), 0.3 * smooth(), 0.3 * smooth()
    p = 1.0 + 0.1 * smooth()
    e = p / (0.4 * rho)
    # Fortran order (x fastest): permute so that the flat tensor runs x first
    comps = [rho, rho * u, rho * v, rho * w, rho * (e + 0.5 * (u * u + v * v + w * w))]
    W = torch.stack([c.permute(2, 1, 0).contiguous() for c in comps]).reshape(-1)
    del X, Y, Z, rho, u, v, w, p, e, comps
    W = W.float().double()
    N = n ** 3
    out = {"grid": n, "words_per_point": CGRID_WORDS, "compact_passes": 69}
    res = {}
    for prec, dt in (("f64", torch.float64), ("f32_refine", torch.float32)):
        cg = P.cgrid()
        cg.init((n, n, n), (L / n,) * 3, inviscid=False, mu_ref=0.01, T_ref=1.0, visc_n=0.75, Pr=0.72, prec=prec)
        Wd = W.to(dt)
        r = torch.empty_like(Wd)
        cg.getRHS(Wd, r, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            cg.getRHS(Wd, r, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out["ms_" + prec] = round(ms, 4)
        out["GBps_" + prec] = round(CGRID_WORDS * N * Wd.element_size() / (ms * 1e-3) / 1e9, 1)
        res[prec] = r.double()
        del cg, Wd, r
        torch.cuda.empty_cache()
    out["speedup_fp32r_vs_fp64"] = round(out["ms_f64"] / out["ms_f32_refine"], 3)
    out["fp32r_rel_l2_vs_fp64"] = {c: float("%.3e" % rel_l2(res["f32_refine"][c * N:(c + 1) * N],
                                                            res["f64"][c * N:(c + 1) * N]))
                                   for c in range(5)}
    return out


def wall_les_bench(P, dev, stream, steps=3):
    """One SSP-RK45 step of the wall-modelled LES on the 512 x 512 x 256 box
    (periodic x, y of side 2 pi; walls in z, dz = 2 pi / 512: the Moeng wall
    model below, a slip wall above, dpFdx = 1, the Stokes pressure; log-law mean profile with
    z0 = 1e-4 plus the seeded turbulence of spectrum k^4 exp(-2 (k/4)^2);
    FP32+refine), with the learned
    closure at every stage and the dynamic Smagorinsky model where its
    dissipation is negative, using the network trained on the solver's own
    256^3 DNS (results/closure_c4_dns256_params.npy).  Reports ms per step
    (five stages, each a projection, the right-hand side, the dynamic
    procedure and the closure on 67 M cell and edge points), the closure's
    fallback share at the last stage and whether the run stayed finite."""
    import math
    import os

    import numpy as np
    import torch
    from padeops_amd.synthetic import wall_bounded
    nx, nz = 512, 256
    dz = 2 * math.pi / nx
    Lz = nz * dz
    g = P.igrid()
    g.init(nx, nx, nz, 2 * math.pi, 2 * math.pi, Lz, Re=0.0, scheme=2, CFL=0.5,
           sgs=dict(model="smagorinsky", cglobal=0.01, wall_model="moeng", z0=1e-4),
           walls=dict(botWall=3, topWall=2, dpFdx=1.0), prec="f32_refine")
    uC, _, wE = wall_bounded(nx, nx, nz, 4.0, 1234, torch.float32, dev)
    wE.view(nz + 1, -1)[0].zero_()
    wE.view(nz + 1, -1)[nz].zero_()
    g.set_velocity(uC, torch.zeros_like(uC), wE)
    del uC, wE
    params = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "results",
                                  "closure_c4_dns256_params.npy")).astype(np.float32)
    net = P.SGSNet(64)
    net.set_params(params)
    npts = nx * nx * (nz + 1)
    cl = P.Closure(net, npts)
    delta = (2 * math.pi / nx * 2 * math.pi / nx * dz) ** (1.0 / 3.0)
    g.set_closure(2, cl, delta=delta, dratio=delta / (math.pi / 21))
    g.step(1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    g.step(steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    u = g.field("u")
    finite = bool(torch.isfinite(u).all().item())
    # the divergence the last projection leaves (DivergenceCheck on the
    # state's spectra), against the scale max |u| / dx
    sp = P.spectral()
    sp.init(nx, nx, nz, 2 * math.pi / nx, 2 * math.pi / nx, dz, init_periodicInZ=False)
    pp = P.padepoisson()
    pp.init(sp, dz, PeriodicInZ=False)
    div = torch.empty(nx * nx * nz, dtype=torch.float32, device=dev)
    pp.DivergenceCheck(g.field("uhat"), g.field("vhat"), g.field("what"), div, prec="f32_refine")
    umax = float(u.abs().max())
    div_rel = float(div.abs().max()) / (umax / dz)
    del sp, pp, div
    out = {"ms_per_step": round(ms, 2), "steps_per_s": round(1000.0 / ms, 3), "stages_per_step": 5,
           "prec": "f32_refine", "closure": "learned + dynamic fallback", "finite": finite,
           "max_div_over_umax_per_dx": float("%.3e" % div_rel),
           "fallback_share": round(cl.fallback_count() / npts, 4), "t": round(g.time()[0], 5)}
    del g, cl
    torch.cuda.empty_cache()
    # the network alone on the stage's 67 M points (k_mlp_infer): matrix-core
    # work of the padded layers 16-64-64-64-16, two flops per multiply-add
    X = torch.randn((npts, 10), dtype=torch.float32, device=dev)
    Y = torch.empty((npts, 6), dtype=torch.float32, device=dev)
    net.forward(X, Y)
    reps = 10
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        net.forward(X, Y)
    e1.record()
    torch.cuda.synchronize()
    mlp_ms = e0.elapsed_time(e1) / reps
    flop = 2.0 * npts * (16 * 64 + 2 * 64 * 64 + 64 * 16)
    out["mlp_infer"] = {"points": npts, "ms": round(mlp_ms, 3), "tflops": round(flop / mlp_ms / 1e9, 1),
                        "frac_bf16_peak": round(flop / mlp_ms / 1e9 / 2500.0, 4),
                        "gelu_evals_T_per_s": round(192.0 * npts / mlp_ms / 1e9, 3)}
    del net, Y
    # one training step (forward keeping the activations, backward, ordered
    # gradient sum, Adam) at the closure's training batch of 65,536 points
    B = 65536
    tnet = P.SGSNet(B)
    tnet.set_params(params)
    Xb = X[:B].contiguous()
    Tb = torch.randn((B, 6), dtype=torch.float32, device=dev)
    for _ in range(3):
        tnet.train_step(Xb, Tb, lr=1e-5)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        tnet.train_step(Xb, Tb, lr=1e-5)
    e1.record()
    torch.cuda.synchronize()
    tr_ms = e0.elapsed_time(e1) / reps
    out["mlp_train_step"] = {"batch": B, "ms": round(tr_ms, 4), "samples_per_s": round(B / tr_ms * 1e3, 0)}
    del tnet, X, Xb, Tb
    return out


def wall_box(P, dev, stream, steps):
    """The wall-bounded box of 512 x 512 x 256 cells (periodic x, y; walls in
    z; log-law mean profile plus the seeded turbulence of spectrum
    k^4 exp(-2 (k/4)^2)): time per launch of
    each operator in FP32+refine and FP64 (HIP events on the launch stream),
    the achieved algorithmic GB/s (read the input field, write the output
    field), and the FP32+refine error against FP64."""
    import math

    import torch
    from padeops_amd.synthetic import wall_bounded
    nx, nz = 512, 256
    dz = 2 * math.pi / nx
    F64 = dict(zip(("uC", "uE", "wE"), wall_bounded(nx, nx, nz, 4.0, 1234, torch.float64, dev)))
    F32 = {k: v.float() for k, v in F64.items()}
    c10z = P.cd10()
    c10z.init(nz, dz, False, 0, 0)
    c10x = P.cd10()
    c10x.init(nx, dz, True, 0, 0)
    sES = P.cd06stagg()
    sES.init(nz, dz, True, True, False, True)
    sOO = P.cd06stagg()
    sOO.init(nz, dz, False, False, False, False)
    ops = {
        "cd10_dd3_walls": ("uC", nz, lambda f, g, p: c10z.dd3(f, g, nx, nx, 0, 0, prec=p, stream=stream)),
        "cd10_d2d3_walls": ("uC", nz, lambda f, g, p: c10z.d2d3(f, g, nx, nx, 0, 0, prec=p, stream=stream)),
        "stagg_ddz_C2E": ("uC", nz + 1, lambda f, g, p: sES.ddz_C2E(f, g, nx, nx, prec=p, stream=stream)),
        "stagg_InterpZ_C2E": ("uC", nz + 1, lambda f, g, p: sES.InterpZ_C2E(f, g, nx, nx, prec=p, stream=stream)),
        "stagg_ddz_E2C": ("uE", nz, lambda f, g, p: sES.ddz_E2C(f, g, nx, nx, prec=p, stream=stream)),
        "stagg_d2dz2_E2E": ("uE", nz + 1, lambda f, g, p: sES.d2dz2_E2E(f, g, nx, nx, prec=p, stream=stream)),
        "stagg_InterpZ_E2C_w": ("wE", nz, lambda f, g, p: sOO.InterpZ_E2C(f, g, nx, nx, prec=p, stream=stream)),
        "cd10_dd1_periodic": ("uC", nz, lambda f, g, p: c10x.dd1(f, g, nx, nz, 0, 0, prec=p, stream=stream)),
        "cd10_dd2_periodic": ("uC", nz, lambda f, g, p: c10x.dd2(f, g, nx, nz, 0, 0, prec=p, stream=stream)),
    }
    res = {}
    for name, (key, nout, fn) in ops.items():
        row = {}
        outs = {}
        for prec, F, dt in (("f32_refine", F32, torch.float32), ("f64", F64, torch.float64)):
            f = F[key]
            g = torch.empty(nx * nx * nout, dtype=dt, device=dev)
            for _ in range(2):
                fn(f, g, prec)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(steps):
                fn(f, g, prec)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            nbytes = (f.numel() + g.numel()) * f.element_size()
            row["ms_" + prec] = round(ms, 5)
            row["GBps_" + prec] = round(nbytes / (ms * 1e-3) / 1e9, 1)
            outs[prec] = g
        row["fp32r_rel_l2_vs_fp64"] = float("%.3e" % rel_l2(outs["f32_refine"], outs["f64"]))
        res[name] = row
        del outs
    return res


def pass_calls(P, n, f32, stream):
    """One launch of each pass type of the FP32+refine evaluation, and how
    often each runs per evaluation."""
    import numpy as np
    import torch
    out = torch.empty_like(f32)
    L = 2 * np.pi
    d = P.cd06()
    d.init(n, L / n, True)
    c = P.cf90()
    c.init(n, True)
    ps = P.PoissonPeriodic()
    ps.init(L / n, L / n, L / n, n, n, n, ("cd06",) * 3, 1)
    calls = {
        "cd06_x": lambda: d.dd1(f32, out, n, n, prec="f32_refine", stream=stream),
        "cd06_y": lambda: d.dd2(f32, out, n, n, prec="f32_refine", stream=stream),
        "cd06_z": lambda: d.dd3(f32, out, n, n, prec="f32_refine", stream=stream),
        "cf90_x": lambda: c.filter1(f32, out, n, n, prec="f32_refine", stream=stream),
        "cf90_y": lambda: c.filter2(f32, out, n, n, prec="f32_refine", stream=stream),
        "cf90_z": lambda: c.filter3(f32, out, n, n, prec="f32_refine", stream=stream),
        "poisson_5pass": lambda: ps.poisson_solve(f32, out, prec="f32_refine", stream=stream),
    }
    # uses per evaluation of each pass type (the x derivatives of the divergence
    # and of the projection are taken inside the Poisson solve's x transforms)
    uses = {"cd06_x": 3, "cd06_y": 5, "cd06_z": 5, "cf90_x": 3, "cf90_y": 3, "cf90_z": 3, "poisson_5pass": 1}
    return calls, uses, (d, c, ps, out)


def pass_breakdown(P, n, f32, stream, steps):
    """Average duration of each pass type of the FP32+refine evaluation, timed
    with HIP events on the launch stream; returns (ms per pass, roofline of
    the dominant kernel)."""
    import torch
    N = n ** 3
    calls, uses, _keep = pass_calls(P, n, f32, stream)
    res = {}
    for name, fn in calls.items():
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / steps
    # time per evaluation of each pass type, the Poisson solve included; the
    # roofline entry describes the largest share
    share = {k: res[k] * uses[k] for k in res}
    dom = max(share, key=share.get)
    # algorithmic words per point of one launch: a compact pass reads and
    # writes one field; the Poisson solve's five passes move 10 (r2c x, y
    # forward, fused z, y inverse, c2r x: 2 each, the half spectrum counted
    # as one word per point)
    words = 10 if dom == "poisson_5pass" else 2
    algo = words * N * 4
    ach = algo / (res[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": PEAK_HBM, "unit": "GB/s",
            "frac": round(ach / PEAK_HBM, 4), "traffic": None,
            "bytes_per_launch": algo, "avg_launch_ms": round(res[dom], 5),
            "ms_per_eval_by_pass": {k: round(v, 4) for k, v in share.items()}}
    return {k: round(v, 5) for k, v in res.items()}, roof


def profile_pass(P, n, dev, name, launches):
    """Launches of one pass type and of a plain 1 GiB copy, nothing else timed:
    the command rocprofv3 counter runs profile."""
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    f32 = turbulence(n, 8.0, 1234, torch.float32, dev)[0]
    calls, _, _keep = pass_calls(P, n, f32, stream)
    fn = calls[name]
    for _ in range(launches):
        fn()
    a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(launches):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"profile_pass": name, "launches": launches, "n": n, "copy_bytes_each_way": 4 << 28}))


def hbm_probe(steps):
    """Plain device copy (1 GiB -> 1 GiB) and STREAM triad a = b + 2 c over a
    4.3 GB working set (above the 256 MB last-level cache), timed with events."""
    import torch
    dev = torch.device("cuda:0")
    out = {}
    n = 1 << 28  # 1 GiB of FP32
    a = torch.ones(n, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    c = torch.ones_like(a)
    for name, fn, nbytes in (("copy", lambda: b.copy_(a), 2 * 4 * n),
                             ("triad", lambda: torch.add(b, c, alpha=2.0, out=a), 3 * 4 * n)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out[name + "_GBps"] = round(nbytes / (ms * 1e-3) / 1e9, 1)
    del a, b, c
    return out


def committed_traffic(kernel):
    """HBM bytes per launch of the dominant pass, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The oracle's FP64 full RHS (OpenMP C, precomputed FFT twiddles) on the
    host: one warm-up evaluation at 256^3, then the median of 10 timed ones,
    scaled by the point count to a 512^3 evaluation rate."""
    import numpy as np

    import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    n = 256
    rng = np.random.default_rng(3)
    u, v, w = (np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(3))
    O.full_rhs(u, v, w)
    times = []
    for _ in range(10):
        t0 = time.perf_counter()
        O.full_rhs(u, v, w)
        times.append(time.perf_counter() - t0)
    per = float(np.median(times))
    return {"value": round(1.0 / (per * 8.0), 5), "unit": "RHS evals/s (512^3-equivalent)",
            "cores": O.num_threads(), "kind": "port",
            "sample": "median of 10 FP64 full-RHS evaluations of the C oracle at 256^3 after one warm-up "
                      "(%.2f s each, %.1f s in all), rate scaled by 1/8 to 512^3 points; host CPU: %s"
                      % (per, sum(times), cpu_model())}


if __name__ == "__main__":
    main()
