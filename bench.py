#!/usr/bin/env python3
"""Benchmark of the MI355X path for CompLB3D's per-timestep hot path.

Headline ("value"): coupled MLUPS of the kinetics + biomass case -- 256 x 128 x 128 sphere
pack (radius 3-5, seed 0, porosity 0.45), 4 D3Q7 solutes (tau 0.8) with dual-Monod
kinetics (Ks 1e-2, k_max dt 1e-3, yield 0.4) and biomass seeded in 1 % of pore voxels
(B ~ U[10, 100]), FP32 population store, advected by the D3Q19 flow solved on the same
pack. One step = collide + kinetics + biomass threshold rule + stream for every solute.

Beside it, on the same GPU: D3Q19 flow MLUPS, the FP64 store, and the speciation legs --
plain FP64 Newton solves/s on the 128^3 Gaussian-smoothed totals field of the 8 / 24
tableau (seed 3), and MLP-warm vs plain Newton solves/s at the same 1e-10 residual on
fresh samples of the 12 -> 128 -> 128 -> 8 warm-start setting (log10 totals + 4
jittered logK). The CPU checker is timed on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "complb3d-lattice-boltzmann-fba-surrogate-cobrapy-glpk_amd")
sys.path.insert(0, PKG)

METRIC = "coupled MLUPS + speciation solves/s, MLP-warm vs plain Newton @1e-10 residual"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kind_omega(tau=0.8):
    kind = np.full(256, 2, dtype=np.int32)
    kind[0], kind[1] = 0, 1
    return kind, np.full(256, 1.0 / tau)


def coupled_case(clb, nx, ny, nz, prec, flow_steps):
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.45, seed=0)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.01, bmin=10.0, bmax=100.0, seed_sites=1, seed_density=2)
    mask = pack.copy()
    mask[b >= 10.0] = 3
    kind, om = kind_omega(0.8)
    lat = clb.Lattice(nx, ny, nz, precision=prec)
    lat.set_dynamics(kind, om)
    lat.upload_mask(mask)
    lat.flow_init(1.0, 0.999)
    lat.set_timing(True)
    t0 = time.perf_counter()
    lat.flow_steps(flow_steps)
    lat.sync()
    flow_wall = time.perf_counter() - t0
    pf = lat.perf()
    lat.set_timing(False)
    omg = np.full((4, 256), 1.0 / 0.8)
    lat.set_solutes(omg, [0, 1, 0, 1, 1, 1, 1, 1], [1.0, 0, 0.5, 0, 0, 0, 0, 0])
    lat.couple_velocity()
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 10.0, 2, 3)
    lat.solutes_init([0.2, 0.1, 0.0, 0.0])
    return lat, mask, b, dict(flow_wall=flow_wall, flow_ms=pf.flow_ms, flow_launches=pf.flow_launches,
                              fluid=pf.fluid_nodes)


def timed_steps(lat, steps, warmup):
    lat.step(warmup)
    lat.sync()
    lat.reset_perf()
    lat.set_timing(True)
    lat.sync()
    t0 = time.perf_counter()
    lat.step(steps)
    lat.sync()
    wall = time.perf_counter() - t0
    p = lat.perf()
    lat.set_timing(False)
    return wall, p


def speciation_legs(clb, n_side, n_train, epochs):
    S, lk = clb.gen_tableau(8, 24, seed=3)
    out = {}
    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, list(range(8)), tol=1e-10, max_iter=200)
    # plain Newton on the smoothed field
    tot = clb.gen_totals_field(n_side, n_side, n_side, 8, sigma=4.0, seed=4)
    lat.batch_load(tot)
    lat.batch_solve(clb.WARM_REF)            # warm-up (also compiles / loads code objects)
    lat.sync()
    lat.reset_perf()
    lat.set_timing(True)
    t0 = time.perf_counter()
    lat.batch_solve(clb.WARM_REF)
    lat.sync()
    wall = time.perf_counter() - t0
    p = lat.perf()
    lat.set_timing(False)
    x, it, r, cv = lat.batch_fetch()
    n = tot.shape[0]
    out["field"] = dict(n=n, wall_s=wall, kernel_ms=p.spec_ms, solves_per_s=n / (p.spec_ms * 1e-3),
                        converged=float(cv.mean()), max_resid=float(r.max()), mean_iters=float(it.mean()),
                        max_iters=int(it.max()), sum_iters=int(it.sum()))
    # MLP trained on plain-Newton solves, then warm vs plain on fresh samples
    var = [8, 9, 10, 11]

    def samples(m, seed):
        g = np.random.default_rng(seed)
        return 10 ** g.uniform(-6, -2, (m, 8)), lk[var][None, :] + g.uniform(-1, 1, (m, 4))

    Ttr, Ltr = samples(n_train, 6)
    lat.batch_load(Ttr, var, Ltr)
    t0 = time.perf_counter()
    lat.batch_solve(clb.WARM_REF)
    lat.sync()
    gen_s = time.perf_counter() - t0
    lat.mlp_config(12, 128, 8)
    t0 = time.perf_counter()
    loss = lat.mlp_train_from_batch(epochs=epochs, batch=512, lr=2e-3, seed=7)
    train_s = time.perf_counter() - t0
    Tte, Lte = samples(n, 9)
    lat.batch_load(Tte, var, Lte)
    res = {}
    for name, mode in (("plain", clb.WARM_REF), ("mlp", clb.WARM_MLP)):
        lat.batch_solve(mode)
        lat.sync()
        lat.reset_perf()
        lat.set_timing(True)
        t0 = time.perf_counter()
        lat.batch_solve(mode)
        lat.sync()
        w = time.perf_counter() - t0
        pp = lat.perf()
        lat.set_timing(False)
        x, it, r, cv = lat.batch_fetch()
        ms = pp.spec_ms + pp.mlp_ms
        res[name] = dict(wall_s=w, kernel_ms=ms, newton_ms=pp.spec_ms, mlp_ms=pp.mlp_ms,
                         solves_per_s=n / (ms * 1e-3), converged=float(cv.mean()), max_resid=float(r.max()),
                         mean_iters=float(it.mean()))
    lat.close()
    out["warm"] = dict(n=n, train_solves=n_train, gen_s=gen_s, train_s=train_s, train_loss=loss, **res,
                       gain=res["mlp"]["solves_per_s"] / res["plain"]["solves_per_s"])
    return out


def cpu_baseline(oracle_mod, clb, nx, ny, nz, steps):
    """The FP64 checker on the same coupled workload (full domain, a few steps)."""
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.45, seed=0)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.01, bmin=10.0, bmax=100.0, seed_sites=1, seed_density=2)
    mask = pack.copy()
    mask[b >= 10.0] = 3
    kind, _ = kind_omega()
    u = np.zeros((pack.size, 3))
    u[pack == 2, 0] = 1e-3
    run = oracle_mod.AdeRun(nx, ny, nz, mask, kind, np.full((4, 256), 1.25), [0, 1, 0, 1, 1, 1, 1, 1],
                            [1.0, 0, 0.5, 0, 0, 0, 0, 0], u, [0.2, 0.1, 0.0, 0.0], kinetics=True, biomass=b,
                            mask_update=(10.0, 2, 3))
    run.run(1)
    t0 = time.perf_counter()
    run.run(steps)
    dt = time.perf_counter() - t0
    return dict(value=nx * ny * nz * steps / dt / 1e6, unit="MLUPS", cores=oracle_mod.num_threads(),
                kind="port", sample=f"FP64 checker, full {nx}x{ny}x{nz} coupled case, {steps} steps ({dt:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--flow-steps", type=int, default=400)
    ap.add_argument("--spec-side", type=int, default=128)
    ap.add_argument("--train-solves", type=int, default=2000000)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--skip-speciation", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the multi-GPU slab "
                                   "decomposition is a design, not built)"}))
        sys.exit(2)

    import clb

    log("device:", clb.device_name())
    nx, ny, nz = args.nx, args.ny, args.nz
    N = nx * ny * nz
    # ---- headline: FP32 coupled step
    lat, mask, b, fl = coupled_case(clb, nx, ny, nz, clb.PREC_FP32, args.flow_steps)
    log(f"flow: {args.flow_steps} steps, {fl['flow_ms'] / args.flow_steps:.3f} ms/step")
    wall, p = timed_steps(lat, args.steps, args.warmup)
    fluid = p.fluid_nodes
    S, nb, rs = 4, 1, 4
    ms_step = wall * 1e3 / args.steps
    mlups = N * args.steps / wall / 1e6
    ade_ms = p.ade_ms / max(p.ade_launches, 1)
    bytes_launch = fluid * (2 * 7 * S * rs + 3 * rs + nb * rs) + N * 1
    achieved = bytes_launch / (ade_ms * 1e-3) / 1e9
    lat.close()
    log(f"coupled fp32: {ms_step:.3f} ms/step wall, kernel {ade_ms:.3f} ms, {mlups:.1f} MLUPS, "
        f"{achieved:.0f} GB/s")
    flow_ms = fl["flow_ms"] / max(fl["flow_launches"], 1)
    flow_bytes = fluid * 2 * 19 * 4 + N
    extra = dict(flow_fp32=dict(ms_per_step=flow_ms, mlups=N / (flow_ms * 1e-3) / 1e6,
                                mflups=fluid / (flow_ms * 1e-3) / 1e6,
                                gbs=flow_bytes / (flow_ms * 1e-3) / 1e9),
                 fluid_nodes=int(fluid), nodes=N)
    # ---- FP64 store for comparison
    if not args.skip_fp64:
        lat64, *_ = coupled_case(clb, nx, ny, nz, clb.PREC_FP64, 2)
        w64, p64 = timed_steps(lat64, max(args.steps // 4, 10), 5)
        lat64.close()
        extra["coupled_fp64"] = dict(ms_per_step=w64 * 1e3 / max(args.steps // 4, 10),
                                     mlups=N * max(args.steps // 4, 10) / w64 / 1e6)
    # ---- speciation legs
    if not args.skip_speciation:
        sp = speciation_legs(clb, args.spec_side, args.train_solves, args.epochs)
        extra["speciation"] = sp
        f = sp["field"]
        # FLOP estimate per residual+Jacobian evaluation of the 8 / 24 kernel (exp10 counted as 20)
        flop_eval = 24 * (2 * 8 + 20 + 2 * 8) + 24 * (8 + 2 * 64) + 8 * 20 + (2 * 8 ** 3) // 3 + 2 * 64
        f["fp64_tflops"] = flop_eval * f["sum_iters"] / (f["kernel_ms"] * 1e-3) / 1e12
        f["fp64_frac"] = f["fp64_tflops"] / FP64_PEAK_TFLOPS
        log(f"speciation field: {f['solves_per_s'] / 1e6:.1f} M solves/s, iters {f['mean_iters']:.2f}; "
            f"warm gain {sp['warm']['gain']:.2f}x")
    # ---- CPU baseline
    cpu = None
    if not args.skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        cpu = cpu_baseline(orc, clb, nx, ny, nz, args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.2f} MLUPS on {cpu['cores']} threads")
    line = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"coupled kinetics + biomass, {nx}x{ny}x{nz} sphere pack phi 0.45, "
                               f"4 D3Q7 solutes dual-Monod, FP32 store", "nodes": N, "fluid_nodes": int(fluid),
                   "solutes": S, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "ade_step_kernel<float,4>", "kernel_ms": round(ade_ms, 4),
                     "bytes_per_launch": int(bytes_launch)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
