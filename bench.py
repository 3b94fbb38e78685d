#!/usr/bin/env python3
"""Benchmark of the MI355X path for CompLB3D's per-timestep hot path.

Headline ("value"): coupled MLUPS of the full coupled case -- 384 x 192 x 192 sphere pack
(radius 3-5, seed 0, porosity 0.35), D3Q19 flow solved on it with the FP32 deviation store,
6 D3Q7 solutes (tau 0.8) carrying dual-Monod kinetics on biomass seeded in 1 % of the pore
voxels and per-voxel FP64 equilibrium speciation on a 3-component / 6-species tableau
(seed 11) every step, precipitation / dissolution node conversion every 100 steps followed
by a flow re-solve, steps replayed from captured HIP graphs. One coupled step = collide +
kinetics + biomass threshold rule + speciation + stream for every solute. One bench step is
one segment: 100 coupled steps, the node conversion, and the flow re-solve of the
reference's loop (until the driver's ValueTracer converges, the tracer kept across
re-solves); every segment of the timed region is inside the timing.

Beside it, on the same GPU: the kinetics + biomass case at 256 x 128 x 128 (4 solutes, FP32
and FP64 stores), D3Q19 flow MLUPS, plain FP64 Newton solves/s on the 128^3 Gaussian-
smoothed totals field of the 8 / 24 tableau (seed 3), and MLP-warm vs plain Newton solves/s
at the same 1e-10 residual on fresh samples of the 12 -> 128 -> 128 -> 8 warm-start setting.
The CPU oracle is timed on a bounded slab of the headline workload on the host cores.
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
# algorithmic bytes per voxel-step (FP32 store): 7 populations read + written per solute
# (56 B), 3 velocity components (12 B), the mask byte
BYTES_D3Q19 = 2 * 19 * 4 + 1
BYTES_COUPLED6 = 6 * 56 + 12 + 1
BYTES_KIN4 = 4 * 56 + 12 + 1 + 8  # + biomass read and write
# CA sweep, FP32 biomass, one species, per voxel (the dense part: the pushes themselves touch
# few voxels): total/mask update 4 b + 1 mask + 16 total r/w; push decision 1 mask + 8 total
# + 4 age + 1 flags; gather 1 mask + 1 flags; the sweep's biomass copy 8
BYTES_CA_SWEEP = 21 + 14 + 2 + 8
# Surrogate processor kernel, FP32 store, 4 solutes, 2 biomass fields (one surrogate): every
# voxel's mask byte and surrogate field are scanned; a voxel holding the organism also reads its
# 28 incoming populations and writes its 4 + 2 increments.
BYTES_SRG_SCAN = 1 + 4
BYTES_SRG_PRESENT = 28 * 4 + 6 * 4
# FP64 flops of the 2-10-10-10-10-1 network per evaluation (mapminmax, 4 tansig layers
# with exp counted as 20, linear output)
FLOP_SRG_NET = 2 * 2 + 2 * (2 * 10 + 3 * 100) + 4 * 10 * (20 + 4) + 2 * 10 + 4
# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on the default
# workloads, scaled by a copy kernel of known size; mean of the odd and even steps. Not
# measured in this run (the counters need the profiler): the file they come from is named.
# Reported only when the workload is the default one.
MEASURED_TRAFFIC_SOURCE = "profiles/hbm_traffic_full_coupled_384x192x192_kinetics_256x128x128.md"
MEASURED_TRAFFIC = {
    "full_coupled_384x192x192": (3.301e9 + 2.773e9) / 2,
    "flow_384x192x192": (1.533e9 + 0.826e9) / 2,
    "kinetics_256x128x128": (0.699e9 + 0.548e9) / 2,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kind_omega(tau=0.8):
    kind = np.full(256, 2, dtype=np.int32)
    kind[0], kind[1] = 0, 1
    return kind, np.full(256, 1.0 / tau)


def equilibrium(clb, S, lk, comp, totals):
    """Species concentrations of the equilibrium with the given component totals."""
    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, comp, tol=1e-12, max_iter=200)
    lat.batch_load(np.atleast_2d(totals))
    lat.batch_solve(clb.WARM_REF)
    x, _, _, cv = lat.batch_fetch()
    lat.close()
    if not cv.all():
        raise RuntimeError("equilibrium of the boundary water did not converge")
    return 10.0 ** (lk + S @ x[0])


def full_case_inputs(clb, nx, ny, nz):
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.35, seed=0)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.01, bmin=10.0, bmax=100.0, seed_sites=1, seed_density=2)
    mask = pack.copy()
    mask[b >= 10.0] = 3
    S, lk = clb.gen_tableau(3, 6, seed=11, log_t_min=-4.0, log_t_max=-2.0)
    c0 = equilibrium(clb, S, lk, [0, 1, 2], [1e-3, 5e-4, 1e-4])
    cin = equilibrium(clb, S, lk, [0, 1, 2], [4e-3, 2e-3, 2e-4])
    return pack, mask, b, S, lk, c0, cin


def full_coupled(clb, args):
    """The headline workload. One bench step is one segment of `interval` coupled steps,
    then the precipitation / dissolution node conversion and, when the mask changed, the
    flow re-solve of the reference's loop (src/complab.cpp:1782-1845): collideAndStream
    until the driver's tracer (ValueTracer(1, 1000, ns_converge_iT2), kept across re-solves)
    has converged, at most ns_max_iT2 steps, then the percolation test when it did not."""
    nx, ny, nz = args.nx, args.ny, args.nz
    N = nx * ny * nz
    pack, mask, b, S, lk, c0, cin = full_case_inputs(clb, nx, ny, nz)
    kind, om = kind_omega(0.8)
    lat = clb.Lattice(nx, ny, nz, precision=clb.PREC_FP32)
    if args.brick:
        lat.set_compact_order(*[int(v) for v in args.brick.split(",")])
    lat.set_dynamics(kind, om)
    lat.upload_mask(mask)
    lat.flow_init(1.0, 0.999)
    # set-up: the initial flow solve of the driver (src/complab.cpp:593, 614-620), to its own
    # tracer ns_convg1 = ValueTracer(1, 1000, ns_converge_iT1), at most ns_max_iT1 steps
    lat.sync()
    t_flow0 = time.perf_counter()
    lat.flow_tracer(1.0, 1000.0, args.ns_converge1)  # on the device: no host round trip per step
    flow0_steps, flow0_conv = lat.flow_resolve_device(args.ns_max1)
    lat.sync()
    flow0_s = time.perf_counter() - t_flow0
    lat.set_solutes(np.full((6, 256), 1.0 / 0.8), [0, 1] * 6, sum([[cin[s], 0.0] for s in range(6)], []))
    lat.couple_velocity()
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 10.0, 2, 3)
    lat.set_tableau(S, lk, [0, 1, 2], tol=1e-10, max_iter=200)
    lat.solutes_init(c0)
    lat.set_phase(np.zeros(N, dtype=np.uint8))
    # mineral threshold: the 0.01 % upper tail of species 5 after the start-up transient
    lat.step(args.startup)
    c5 = lat.conc()[5][mask >= 2]
    thr = float(np.quantile(c5, 0.9999))
    # launch shapes and fused budget picked on this state (set-up: the state is restored), or
    # the given ones
    tune = lat.autotune(args.tune_steps) if args.tune_steps > 0 else None
    if args.fused_block:
        lat.set_block_shape(clb.KERNEL_FUSED, args.fused_block, 1)
    if args.fused_budget:
        lat.set_fused_budget(args.fused_budget)
    if args.straggler_budget >= 0:
        lat.set_straggler_budget(args.straggler_budget)
    lat.set_graph(True, args.interval)
    lat.flow_tracer(1.0, 1000.0, args.ns_converge)  # the driver's ns_convg2, on the device
    st = dict(conversions=0, resolves=0, resolve_steps=0, unconverged=0, percolation=False)

    def segments(n):
        for _ in range(n):
            lat.step(args.interval)
            ch = lat.node_conversion(5, 2, 1, 9, thr, 0.5 * thr)
            st["conversions"] += ch
            if ch and not st["percolation"]:
                k, conv = lat.flow_resolve_device(args.ns_max)
                st["resolves"] += 1
                st["resolve_steps"] += k
                if not conv:
                    st["unconverged"] += 1
                    st["percolation"] = not lat.percolates()
                lat.couple_velocity()

    # warm-up segments (the first re-solve fills the tracer's window of 500 steps here)
    segments(args.warmup)
    warm = dict(st)
    for key in st:
        st[key] = 0 if key != "percolation" else st[key]
    lat.sync()
    t0 = time.perf_counter()
    segments(args.steps)
    lat.sync()
    wall = time.perf_counter() - t0
    steps = args.steps * args.interval
    # per-kernel durations on the library's stream (HIP events), uncaptured steps
    lat.reset_perf()
    s0 = lat.speciation_stats()
    lat.set_timing(True)
    lat.step(20)
    lat.flow_steps(20)
    lat.sync()
    p = lat.perf()
    lat.set_timing(False)
    s1 = lat.speciation_stats() - s0
    # coupled-step voxel states for the warm-start comparison: the solutes of the reactive
    # voxels after the last step (pore or biofilm, 1 <= x <= nx-2), a bounded sample
    states = None
    if not args.skip_warm_coupled:
        cf = lat.conc()
        mf = lat.download_mask()
        xs = np.arange(N) % nx
        sel = np.flatnonzero(((mf == 2) | (mf == 3)) & (xs >= 1) & (xs <= nx - 2))
        sel = np.random.default_rng(5).choice(sel, min(sel.size, args.warm_sample), replace=False)
        states = np.ascontiguousarray(cf[:, sel].T)
        del cf
    lat.close()
    ade_ms = p.ade_ms / max(p.ade_launches, 1)
    spec_ms = p.spec_ms / max(p.spec_launches, 1)
    flow_ms = p.flow_ms / max(p.flow_launches, 1)
    solves_step = s1[0] / 20.0
    fluid = int(p.fluid_nodes)
    # the whole configured run (5000 coupled steps) with the warm-up's first re-solve in it
    run_steps = 5000
    run_ms = (wall * 1e3 / steps) * run_steps + warm["resolve_steps"] * flow_ms
    return dict(N=N, flow0_steps=int(flow0_steps), flow0_converged=bool(flow0_conv), flow0_s=flow0_s, threshold=thr,
                segments=args.steps, steps=steps, wall=wall, ms_per_segment=wall * 1e3 / args.steps,
                ms_per_step=wall * 1e3 / steps, mlups=N * steps / wall / 1e6,
                ade_ms=ade_ms, straggler_ms=spec_ms, flow_ms=flow_ms, fluid=fluid, reactive=int(p.reactive_nodes),
                conversions=int(st["conversions"]), resolves=int(st["resolves"]),
                resolve_steps=int(st["resolve_steps"]), unconverged_resolves=int(st["unconverged"]),
                percolation=bool(st["percolation"]), warmup_resolve_steps=int(warm["resolve_steps"]),
                warmup_resolves=int(warm["resolves"]),
                run_5000_ms_per_step=run_ms / run_steps, run_5000_mlups=N * run_steps / (run_ms * 1e-3) / 1e6,
                solves_per_step=solves_step, stragglers_per_step=p.stragglers / 20.0,
                converged=int(s1[1]), solves=int(s1[0]), mean_evals=float(s1[2] / max(s1[0], 1)),
                coupled_solves_per_s=solves_step / (wall / steps), autotune=tune, states=states, tableau=(S, lk))


def coupled_warm_start(clb, S, lk, C, train_solves, epochs):
    """Warm starts for the coupled step's speciation, compared on voxel states of the coupled
    run: C holds the solutes (= the 3 / 6 tableau's species) of a sample of reactive voxels
    after a step. Each start solves the same totals (calc_component_totals) to the 1e-10 judge
    with the batch kernel:
      * component_slot: log10 of each component's species (the reference's start in
        run_equilibrium_biotic, src/complab3d_processors_part4_eqsolver.hh:977-994, which the
        fused step uses);
      * plain: log10 of the component totals;
      * mlp: the warm-start network trained on plain solves over the sample's box of totals.
    Reported: mean evaluations, kernel time (the network's inference included), and the same
    on the voxels that need more than 4 evaluations from the component slots (those the fused
    step hands to the persistent kernel)."""
    nc = S.shape[1]
    comp = list(range(nc))
    Cc = np.clip(C, 1e-30, 10.0)
    T = np.maximum(Cc @ S, 1e-30)
    x_slot = np.log10(Cc[:, comp])
    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, comp, tol=1e-10, max_iter=200)
    rng = np.random.default_rng(17)
    lo, hi = np.log10(T.min(axis=0)), np.log10(T.max(axis=0))
    lat.batch_load(10.0 ** rng.uniform(lo, hi, (train_solves, nc)))
    lat.batch_solve(clb.WARM_REF)
    lat.mlp_config(nc, 128, nc)
    t0 = time.perf_counter()
    loss = lat.mlp_train_from_batch(epochs=epochs, batch=512, lr=2e-3, seed=7)
    train_s = time.perf_counter() - t0
    lat.batch_load(T)
    out = dict(n=int(T.shape[0]), train_solves=train_solves, epochs=epochs, train_s=train_s, train_loss=loss)
    its = {}
    for name, mode, guess in (("component_slot", clb.WARM_USER, x_slot), ("plain", clb.WARM_REF, None),
                              ("mlp", clb.WARM_MLP, None)):
        lat.batch_solve(mode, guess)           # warm-up
        lat.sync()
        lat.reset_perf()
        lat.set_timing(True)
        lat.batch_solve(mode, guess)
        lat.sync()
        p = lat.perf()
        lat.set_timing(False)
        x, it, r, cv = lat.batch_fetch()
        its[name] = it
        out[name] = dict(mean_evals=float(it.mean()), kernel_ms=float(p.spec_ms + p.mlp_ms), mlp_ms=float(p.mlp_ms),
                         converged=float(cv.mean()), max_resid=float(r.max()))
    lat.close()
    hard = its["component_slot"] > 4
    out["hard_fraction"] = float(hard.mean())
    for name in its:
        out[name]["mean_evals_hard"] = float(its[name][hard].mean()) if hard.any() else None
    return out


def kinetics_case(clb, nx, ny, nz, prec, flow_steps):
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.45, seed=0)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.01, bmin=10.0, bmax=100.0, seed_sites=1, seed_density=2)
    mask = pack.copy()
    mask[b >= 10.0] = 3
    kind, om = kind_omega(0.8)
    lat = clb.Lattice(nx, ny, nz, precision=prec)
    lat.set_dynamics(kind, om)
    lat.upload_mask(mask)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(flow_steps)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1, 0, 1, 1, 1, 1, 1], [1.0, 0, 0.5, 0, 0, 0, 0, 0])
    lat.couple_velocity()
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 10.0, 2, 3)
    lat.solutes_init([0.2, 0.1, 0.0, 0.0])
    return lat


def outside_main_pore_space(mask, nx, ny, nz, walls=(0, 1)):
    """Voxels outside the largest face-connected open region (no wrap in y or z, as the CA's
    neighbourhood). A pocket sealed by walls that biomass fills to capacity has no empty voxel
    to hold its ages back, and the reference's age rule then raises them on every pass (an
    isolated voxel too: its rule holds for an empty neighbour set), so the age relaxation
    never settles and the reference's run ends there."""
    from scipy import ndimage
    open_ = ~np.isin(mask.reshape(nz, ny, nx), walls)
    lab, n = ndimage.label(open_)
    if n == 0:
        return np.ones(mask.size, dtype=bool)
    sizes = np.bincount(lab.ravel())
    sizes[0] = 0
    return (lab != int(np.argmax(sizes))).ravel()


def ca_leg(lat, N, bmax=100.0, grow=2.5):
    """CA biomass redistribution on the kinetics case: the biomass (1 % seeds up to Bmax,
    after the timed steps) is scaled by `grow`, so seeds hold up to 2.5 Bmax, then swept
    until every voxel is back under Bmax. Seeds outside the main pore space are removed first
    (outside_main_pore_space)."""
    b = lat.biomass()
    b[:, outside_main_pore_space(lat.download_mask(), lat.nx, lat.ny, lat.nz)] = 0.0
    lat.ca_init()
    lat.ca_redistribute(1e30)  # set-up: allocates the sweep buffers; nothing is over that cap
    lat.set_biomass(b * grow)
    lat.sync()
    t0 = time.perf_counter()
    rep = lat.ca_redistribute(bmax, half=0, seed=1)
    lat.sync()
    wall = time.perf_counter() - t0
    sweeps = max(rep["sweeps"], 1)
    ms = wall * 1e3 / sweeps
    log(f"CA: {rep['sweeps']} sweeps + {rep['age_passes']} age passes in {wall * 1e3:.2f} ms ({ms:.3f} ms/sweep), "
        f"max {rep['bmax_before']:.1f} -> {rep['bmax_after']:.4f}, status {rep['status']}, "
        f"{rep['mask_changes']} mask flips")
    # the FD biofilm step on the same state (two mask / total passes around one diffusion step)
    lat.fd_biomass(0.1)
    lat.sync()
    t0 = time.perf_counter()
    for _ in range(10):
        lat.fd_biomass(0.1)
    lat.sync()
    fd_ms = (time.perf_counter() - t0) * 1e3 / 10
    log(f"FD biofilm step: {fd_ms:.3f} ms per call")
    return dict(workload="CA redistribution, kinetics case 256x128x128, biomass x2.5", sweeps=rep["sweeps"],
                age_passes=rep["age_passes"], status=rep["status"], wall_ms=wall * 1e3, ms_per_sweep=ms,
                voxel_sweeps_per_s=N * sweeps / wall, gbs=BYTES_CA_SWEEP * N / (ms * 1e-3) / 1e9,
                bmax_before=rep["bmax_before"], bmax_after=rep["bmax_after"], mask_changes=rep["mask_changes"],
                fd_ms_per_call=fd_ms)


def surrogate_leg(clb, steps, tune_steps):
    """The kinetics workload (256 x 128 x 128, 4 solutes, FP32 store) with a second biomass
    field grown by the surrogate processor: the shipped organism's network shape
    (2-10-10-10-10-1, tansig, random weights) on 1 % seeds, Vmax-limited uptake. Times the
    whole step and the surrogate kernel alone."""
    kn = (256, 128, 128)
    N = kn[0] * kn[1] * kn[2]
    lat = kinetics_case(clb, *kn, clb.PREC_FP32, 400)
    mask = lat.download_mask()
    b0 = lat.biomass()[0]
    b1 = clb.gen_biomass_seed(mask, pore=2, fraction=0.01, bmin=5.0, bmax=60.0, seed_sites=3, seed_density=2)
    lat.set_biomass(np.stack([b0, b1]))
    rng = np.random.default_rng(11)
    sizes = [2, 10, 10, 10, 10, 1]
    params = np.concatenate([np.concatenate([rng.normal(scale=0.6, size=(sizes[k + 1], sizes[k])).ravel(),
                                             rng.normal(scale=0.3, size=sizes[k + 1])]) for k in range(5)])
    net = dict(n_in=2, in_sub=[0, 1], widths=sizes[1:], params=params, x_offset=[8.9e-4, 3.5e-5],
               x_gain=[0.2, 4.0], x_ymin=-1.0, y_ymin=-1.0, y_gain=37.9, y_offset=0.0)
    lat.set_surrogate([1], [[0.05, 0.02, 0.0, 0.0]], [6.94e-8], vmax=[[8.0, 0.4, 0.0, 0.0]])
    lat.set_surrogate_net(1, net)
    if tune_steps > 0:
        lat.autotune(tune_steps)
    wall, p = timed_steps(lat, steps, 20)
    m = lat.download_mask()
    kind, _ = kind_omega(0.8)
    x = np.arange(N) % kn[0]
    reactive = (kind[m] == 2) & (m != 0) & (m != 1) & (x >= 1) & (x <= kn[0] - 2)
    present = reactive & (lat.biomass()[1] > 1e-12)
    lat.close()
    sms = p.srg_ms / max(p.srg_launches, 1)
    nbytes = BYTES_SRG_SCAN * N + BYTES_SRG_PRESENT * present.sum()
    out = dict(workload="kinetics + surrogate 256x128x128, 4 solutes, 2 biomass fields (1 surrogate), FP32 store",
               ms_per_step=wall * 1e3 / steps, mlups=N * steps / wall / 1e6, srg_kernel_ms=sms,
               transport_kernel_ms=p.ade_ms / max(p.ade_launches, 1), reactive=int(reactive.sum()),
               present=int(present.sum()), srg_gbs=nbytes / (sms * 1e-3) / 1e9,
               srg_frac=nbytes / (sms * 1e-3) / 1e9 / HBM_PEAK_GBS,
               net_tflops=FLOP_SRG_NET * present.sum() / (sms * 1e-3) / 1e12)
    log(f"surrogate: {out['mlups']:.1f} MLUPS, surrogate kernel {sms:.3f} ms ({out['present']} voxels with the "
        f"organism, {out['srg_gbs']:.0f} GB/s algorithmic, net {out['net_tflops']:.2f} TFLOP/s FP64)")
    return out


def timed_steps(lat, steps, warmup):
    lat.step(warmup)
    lat.sync()
    t0 = time.perf_counter()
    lat.step(steps)
    lat.sync()
    wall = time.perf_counter() - t0
    lat.reset_perf()
    lat.set_timing(True)
    lat.step(20)
    lat.sync()
    p = lat.perf()
    lat.set_timing(False)
    return wall, p


def speciation_legs(clb, n_side, n_train, epochs):
    S, lk = clb.gen_tableau(8, 24, seed=3)
    out = {}
    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, list(range(8)), tol=1e-10, max_iter=200)
    tot = clb.gen_totals_field(n_side, n_side, n_side, 8, sigma=4.0, seed=4)
    lat.batch_load(tot)
    lat.batch_solve(clb.WARM_REF)            # warm-up (also loads the code objects)
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
    """The FP64 oracle (the reference's algorithms, PCF + Anderson speciation) on a slab of
    the headline workload: same generator and porosity, same 6 solutes, kinetics, biomass and
    tableau; a uniform pore velocity in place of the flow solve."""
    pack, mask, b, S, lk, c0, cin = full_case_inputs(clb, nx, ny, nz)
    kind, _ = kind_omega()
    u = np.zeros((pack.size, 3))
    u[mask >= 2, 0] = 1e-3
    run = oracle_mod.AdeRun(nx, ny, nz, mask, kind, np.full((6, 256), 1.25), [0, 1] * 6,
                            sum([[cin[s], 0.0] for s in range(6)], []), u, c0, kinetics=True, biomass=b,
                            tableau=(S, lk, [0, 1, 2]), eq_tol=1e-10, eq_max_iter=200, mask_update=(10.0, 2, 3))
    run.run(1)
    t0 = time.perf_counter()
    run.run(steps)
    dt = time.perf_counter() - t0
    return dict(value=nx * ny * nz * steps / dt / 1e6, unit="MLUPS", cores=oracle_mod.num_threads(), kind="port",
                sample=f"FP64 oracle, the coupled step (6 solutes, kinetics, PCF+Anderson speciation) on a "
                       f"{nx}x{ny}x{nz} slab of the same workload, {steps} steps ({dt:.1f} s)")


def speciation_cpu_baseline(oracle_mod, clb, n_side, budget_s=8.0):
    """The FP64 oracle's speciation (the reference's PCF + Anderson solve) on a bounded sample
    of the same smoothed totals field as the GPU's field leg (8 / 24 tableau, 1e-10 judge): a
    2 000-voxel probe sizes the sample to about `budget_s` seconds."""
    S, lk = clb.gen_tableau(8, 24, seed=3)
    tot = clb.gen_totals_field(n_side, n_side, n_side, 8, sigma=4.0, seed=4)

    def run(m):
        conc = np.zeros((m, 24))
        conc[:, :8] = tot[:m]
        t0 = time.perf_counter()
        _, it, _, cv, _ = oracle_mod.speciate_pcf(S, lk, list(range(8)), conc, tol=1e-10, max_iter=200)
        return time.perf_counter() - t0, it, cv

    dt, _, _ = run(2000)
    m = int(min(tot.shape[0], max(2000, 2000 * budget_s / max(dt, 1e-6))))
    dt, it, cv = run(m)
    return dict(value=m / dt, unit="solves/s", cores=oracle_mod.num_threads(), kind="port",
                sample=f"FP64 oracle (PCF + Anderson) on the first {m} voxels of the {n_side}^3 totals field, "
                       f"8 / 24 tableau, 1e-10 judge ({dt:.1f} s, mean {it.mean():.1f} iterations, "
                       f"converged {cv.mean():.4f})")


def copy_calibration():
    """Device-to-device copies of known size (the HIP runtime's copy kernel), for scaling the
    HBM traffic counters: each copy reads and writes 1 GiB."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(1 << 30)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), n) or hip.hipMalloc(ctypes.byref(dst), n):
        raise RuntimeError("hipMalloc failed for the copy calibration")
    hip.hipMemset(src, 1, n)
    for _ in range(3):
        if hip.hipMemcpy(dst, src, n, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed for the copy calibration")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: 3 copies of {n.value} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed steps: segments of --interval coupled steps, each "
                                                         "followed by node conversion and the flow re-solve")
    ap.add_argument("--warmup", type=int, default=1, help="untimed segments before them")
    ap.add_argument("--nx", type=int, default=384)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=192)
    ap.add_argument("--interval", type=int, default=100, help="steps between node conversions")
    ap.add_argument("--startup", type=int, default=100, help="coupled steps before the mineral threshold is drawn")
    ap.add_argument("--ns-converge", type=float, default=1e-4, help="ns_converge_iT2 of the re-solve tracer")
    ap.add_argument("--ns-max", type=int, default=2000, help="ns_max_iT2: at most this many re-solve steps")
    ap.add_argument("--ns-converge1", type=float, default=1e-6, help="ns_converge_iT1 of the initial flow solve")
    ap.add_argument("--ns-max1", type=int, default=20000, help="ns_max_iT1: at most this many initial flow steps")
    ap.add_argument("--tune-steps", type=int, default=10, help="steps per autotuner trial (0: default shapes)")
    ap.add_argument("--fused-block", type=int, default=0, help="fused kernel threads per block (0: tuned/default)")
    ap.add_argument("--fused-budget", type=int, default=0, help="fused evaluation budget (0: tuned/default)")
    ap.add_argument("--straggler-budget", type=int, default=-1,
                    help="evaluations of the stragglers' first pass (-1: the library's default)")
    ap.add_argument("--brick", default="", help="bx,by,bz: order of the compact layout's voxels (default: the library's)")
    ap.add_argument("--kin-steps", type=int, default=200)
    ap.add_argument("--spec-side", type=int, default=128)
    ap.add_argument("--train-solves", type=int, default=2000000)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--cpu-nz", type=int, default=192)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--skip-speciation", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-kinetics", action="store_true")
    ap.add_argument("--skip-ca", action="store_true")
    ap.add_argument("--skip-surrogate", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--skip-warm-coupled", action="store_true")
    ap.add_argument("--warm-sample", type=int, default=2000000, help="voxel states for the coupled warm-start A/B")
    ap.add_argument("--warm-train", type=int, default=1000000, help="training solves of its network")
    ap.add_argument("--warm-epochs", type=int, default=30)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus must be 1 (the multi-GPU slab "
                                   "decomposition is a design, not built)"}))
        sys.exit(2)

    import clb

    log("device:", clb.device_name())
    if args.copy_calibration:
        copy_calibration()
    # ---- headline: the full coupled case
    fc = full_coupled(clb, args)
    N = fc["N"]
    # the fused kernel runs on the compact layout: its algorithmic bytes are those of the
    # fluid voxels it updates (349 B each: 6 solutes x 7 populations read + written, the
    # velocity, the mask byte)
    nf = fc["fluid"]
    achieved = BYTES_COUPLED6 * nf / (fc["ade_ms"] * 1e-3) / 1e9
    tune = fc["autotune"]
    fused_bt = 512 if tune is None or tune["fused_bx"] * tune["fused_by"] > 256 else 256
    if args.fused_block:
        fused_bt = 512 if args.fused_block > 256 else 256
    if tune:
        log(f"autotune ({tune['trials']} trials): flow {tune['flow_bx']}x{tune['flow_by']} "
            f"{tune['flow_ms']:.3f} ms, fused {tune['fused_bx']}x{tune['fused_by']} budget {tune['fused_budget']} "
            f"{tune['fused_ms']:.3f} ms")
    log(f"full coupled: {fc['ms_per_step']:.3f} ms/step, {fc['mlups']:.1f} MLUPS; transport+speciation kernel "
        f"{fc['ade_ms']:.3f} ms ({achieved:.0f} GB/s algorithmic over {nf} fluid voxels), stragglers "
        f"{fc['straggler_ms']:.3f} ms, {fc['solves_per_step']:.0f} solves/step, {fc['mean_evals']:.2f} "
        f"evaluations/solve; {fc['resolves']} re-solves of {fc['resolve_steps']} flow steps in the timed segments "
        f"(warm-up: {fc['warmup_resolve_steps']}); 5000-step run {fc['run_5000_ms_per_step']:.3f} ms/step; "
        f"initial flow {fc['flow0_steps']} steps in {fc['flow0_s']:.2f} s (converged {fc['flow0_converged']}); "
        f"{fc['conversions']} voxels "
        f"converted in the timed segments")
    states, (S6, lk6) = fc.pop("states"), fc.pop("tableau")
    extra = dict(full_coupled=fc)
    if states is not None:
        extra["coupled_warm_start"] = coupled_warm_start(clb, S6, lk6, states, args.warm_train, args.warm_epochs)
        cw = extra["coupled_warm_start"]
        log(f"coupled warm start ({cw['n']} voxel states): evaluations component slots "
            f"{cw['component_slot']['mean_evals']:.2f} ({cw['component_slot']['kernel_ms']:.2f} ms), plain "
            f"{cw['plain']['mean_evals']:.2f} ({cw['plain']['kernel_ms']:.2f} ms), MLP {cw['mlp']['mean_evals']:.2f} "
            f"({cw['mlp']['kernel_ms']:.2f} ms); voxels beyond 4 evaluations {cw['hard_fraction']:.4f}")
        del states
    fl_ms = fc["flow_ms"]
    # the compact flow kernel moves the fluid voxels only: its roofline counts their bytes
    extra["flow_fp32"] = dict(ms_per_step=fl_ms, mlups=N / (fl_ms * 1e-3) / 1e6, layout="compact",
                              kernel="flow_step_compact_kernel<float,*,*>",
                              gbs_fluid=BYTES_D3Q19 * nf / (fl_ms * 1e-3) / 1e9,
                              frac_fluid=BYTES_D3Q19 * nf / (fl_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              traffic=(MEASURED_TRAFFIC["flow_384x192x192"]
                                       if (args.nx, args.ny, args.nz) == (384, 192, 192) else None))
    # ---- kinetics + biomass, 256 x 128 x 128, 4 solutes
    if not args.skip_kinetics:
        kn = (256, 128, 128)
        Nk = kn[0] * kn[1] * kn[2]
        lat = kinetics_case(clb, *kn, clb.PREC_FP32, 400)
        ktune = lat.autotune(args.tune_steps) if args.tune_steps > 0 else None
        wall, p = timed_steps(lat, args.kin_steps, 20)
        if not args.skip_ca:
            extra["ca"] = ca_leg(lat, Nk)
        lat.close()
        kms = p.ade_ms / max(p.ade_launches, 1)
        kf = int(p.fluid_nodes)
        # compact layout: the algorithmic bytes are the fluid voxels'
        extra["kinetics_fp32"] = dict(workload="kinetics + biomass 256x128x128, 4 solutes, FP32 store",
                                      ms_per_step=wall * 1e3 / args.kin_steps, mlups=Nk * args.kin_steps / wall / 1e6,
                                      kernel_ms=kms, fluid=kf, gbs=BYTES_KIN4 * kf / (kms * 1e-3) / 1e9,
                                      frac=BYTES_KIN4 * kf / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                      gbs_dense=BYTES_KIN4 * Nk / (kms * 1e-3) / 1e9,
                                      traffic=MEASURED_TRAFFIC["kinetics_256x128x128"], autotune=ktune)
        log(f"kinetics fp32: {extra['kinetics_fp32']['mlups']:.1f} MLUPS, kernel {kms:.3f} ms, "
            f"{extra['kinetics_fp32']['frac']:.3f} of HBM peak (fluid bytes)")
        if not args.skip_surrogate:
            extra["surrogate_fp32"] = surrogate_leg(clb, args.kin_steps, args.tune_steps)
        if not args.skip_fp64:
            lat = kinetics_case(clb, *kn, clb.PREC_FP64, 2)
            w64, _ = timed_steps(lat, max(args.kin_steps // 4, 10), 5)
            lat.close()
            extra["kinetics_fp64"] = dict(ms_per_step=w64 * 1e3 / max(args.kin_steps // 4, 10),
                                          mlups=Nk * max(args.kin_steps // 4, 10) / w64 / 1e6)
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
        cpu = cpu_baseline(orc, clb, args.nx, args.ny, args.cpu_nz, args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.2f} MLUPS on {cpu['cores']} threads")
        if not args.skip_speciation:
            sc = speciation_cpu_baseline(orc, clb, args.spec_side)
            extra["speciation"]["cpu_baseline"] = sc
            log(f"cpu speciation baseline: {sc['value']:.0f} solves/s on {sc['cores']} threads ({sc['sample']})")
    line = {
        "metric": METRIC,
        "value": round(fc["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": fc["segments"],
        "warmup": args.warmup,
        "ms_per_step": round(fc["ms_per_segment"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"full coupled case {args.nx}x{args.ny}x{args.nz}, sphere pack phi 0.35, D3Q19 + "
                               f"6 D3Q7 solutes (FP32 deviation store), dual-Monod kinetics + biomass, FP64 "
                               f"speciation 3/6 tableau every step; one step = {args.interval} coupled steps "
                               f"(captured graphs) + node conversion + flow re-solve to the driver's tracer "
                               f"(ValueTracer(1, 1000, {args.ns_converge}), at most {args.ns_max} steps)",
                   "segment_steps": args.interval, "coupled_steps": fc["steps"],
                   "coupled_ms_per_step": round(fc["ms_per_step"], 4),
                   "nodes": N, "fluid_nodes": fc["fluid"], "solutes": 6, "speciation": "fp64",
                   "speciation_solves_per_s": round(fc["coupled_solves_per_s"], 1), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (MEASURED_TRAFFIC["full_coupled_384x192x192"]
                                 if (args.nx, args.ny, args.nz) == (384, 192, 192) else None),
                     "traffic_source": MEASURED_TRAFFIC_SOURCE,
                     "kernel": f"ade_spec_step_compact_kernel<float,6,*,3,{fused_bt}>",
                     "kernel_ms": round(fc["ade_ms"], 4), "units": "fluid voxels", "units_per_launch": nf,
                     "bytes_per_unit": BYTES_COUPLED6, "bytes_per_launch": int(BYTES_COUPLED6 * nf)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
