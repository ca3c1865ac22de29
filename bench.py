"""Benchmark of the constrained-HMC hot path on one MI355X.

Workload (the headline configuration): 32^4 periodic lattice, Iwasaki action
(c1 = -0.331, beta = 2.6) with 3 levels of isotropic stout smearing
(rho = 0.1) feeding the force, rectangle force on, constrained Metropolis with
a demon window of +-0.5% of the initial plaquette sum, links stored FP32 with
12-real reconstruction, force-gradient integrator with tau = 1 and 12 steps,
FP64 Hamiltonian / accept. One step = one full trajectory (momentum refresh,
H_old, 12 FG steps = 48 force evaluations and 49 link updates, H_new, accept).
Inputs: Haar-random links from a fixed seed (synthetic), resident in HBM.

Run order: thermalisation (untimed, its own fixed rule, independent of the
flags), then the measured FP32-MD chain (--warmup untimed trajectories, then
--steps timed ones), the FP64 chain from the same links and seed (same
momenta, so the two chains' trajectories pair one by one while their accept
decisions agree), an FP64 reversibility round trip, the serving stream, and
the FP64 CPU oracle at the same lattice (force and plaquette errors of the
FP32 and FP64 device paths on the thermalised links, and the CPU baseline).

Prints one JSON line: traj/s (value), link-updates/s, force GB/s, the FP64
path's traj/s on the same workload, both acceptance rates with the paired
Metropolis-probability difference, the roofline of the dominant kernel
measured with HIP events on the library's stream, the serving stream's
latency and throughput, and the CPU oracle timed on the host cores.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FP32_UNPACKED_TFLOPS = 78.6  # 256 CUs x 64 lanes x 2 flops (one FMA) x 2.4 GHz
M_FLOPS = 198  # one 3x3 complex matrix product

# Per-link figures of each kernel family, FP32 storage with 12-real links.
# Bytes: every operand the kernel touches read once and written once. Flops:
# the canonical per-link counts of the force in units of one 3x3 complex
# product (M = 198 flops): Iwasaki staples + force 85M, each stout level
# 15M forward and 60M backward (310M for 3 levels). The backward chain is
# pipelined across kernels: "stout_top" holds the Iwasaki force of the
# smeared links plus the Lambda step of the top level, each "stout_bwd" the
# six-term sum of level k plus the Lambda step of level k-1 (one full
# backward level), "stout_bwd_force" the six-term sum of level 0 and the
# momentum update. The Lambda step is 1/6 of a backward level (10M), the
# six-term sum 5/6 (50M), in proportion to the products each executes.
KERNEL_UNITS = {
    # family: (bytes per link, 3x3 products per link); links 12 reals (48 B)
    "stout_fwd": (48 + 48 + 32 + 72, 15),          # read U_k; write U_k+1, Q_k, C_k
    "stout_top": (48 + 48 + 32 + 72 + 32 + 72, 85 + 10),   # read U_n, U_n-1, Q, C; write Lambda, W
    "stout_bwd": (48 + 32 + 72 + 48 + 32 + 72 + 32 + 72, 60),  # U_k, Lambda_k, W_k, U_k-1, Q, C; Lambda, W
    "stout_bwd_force": (48 + 32 + 72 + 32 + 32, 50),  # read U_0, Lambda_0, W_0, P; write P
    "link_update": (48 + 32 + 48, 3),
    "force_iwasaki": (48 + 32 + 32, 85),
}
# Canonical bytes per link of one Iwasaki + 3-stout force evaluation with
# FP32 12-real thin links (the numerator of "force GB/s").
FORCE_BYTES_LINK_C3 = 1520
FORCE_PRODUCTS_LINK = {0: 85, 1: 160, 2: 235, 3: 310}
# action: (kind, default beta, c1); c0 = 1 - 8 c1
ACTIONS = {"iwasaki": (1, 2.6, -0.331), "dbw2": (1, 1.04, -1.4069), "wilson": (0, 6.0, 0.0)}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5,
                   help="untimed trajectories of the measured chains before the timed ones (thermalisation "
                        "is separate: see THERM_*)")
    p.add_argument("--L", type=int, default=32)
    p.add_argument("--n-md", type=int, default=12)
    p.add_argument("--n-stout", type=int, default=3)
    p.add_argument("--fp64-steps", type=int, default=-1,
                   help="trajectories of the FP64 comparison chain (default: as many as --steps; 0: skip)")
    p.add_argument("--fp64-12-steps", type=int, default=-1,
                   help="trajectories of the FP64 chain with 12-real links (default: as --fp64-steps; 0: skip)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--serve-chains", type=int, default=512,
                   help="chains of the serving stream (0: skip); the full stream has 512")
    p.add_argument("--serve-max-traj", type=int, default=50, help="trajectories per chain drawn from 5..this")
    p.add_argument("--therm-min", type=int, default=40, help="fewest thermalisation trajectories")
    p.add_argument("--seed", type=int, default=32)
    p.add_argument("--action", choices=sorted(ACTIONS), default="iwasaki",
                   help="gauge action (dbw2 with --L 64 --n-md 16 --wloop is the 64^4 DBW2 workload)")
    p.add_argument("--beta", type=float, default=None, help="default: 2.6 Iwasaki, 1.04 DBW2, 6.0 Wilson")
    p.add_argument("--wloop", action="store_true", help="measure a 4x4 Wilson loop after every trajectory")
    p.add_argument("--local", action="store_true",
                   help="time the local updates: checkerboard Metropolis and checkerboard demon (FP64)")
    p.add_argument("--multi", action="store_true",
                   help="time the coupled coarse/fine md:: HMC (FP64) instead of the headline workload")
    p.add_argument("--no-reversibility", dest="reversibility", action="store_false",
                   help="skip the FP64 reversibility round trip after the timed run")
    p.add_argument("--shard", action="store_true",
                   help="time one GPU's share of the 64^4 DBW2 + 3-stout workload split over 8 GPUs along x3 "
                        "(a 64^3 x 8 block, periodic wrap in place of halos) and model its halo exchange")
    p.add_argument("--fp64-kernel-times", action="store_true",
                   help="also time the kernels of the FP64 chains (HIP events around each launch)")
    return p.parse_args()


def oracle_leg(M, ctx, L, arg, act_kind, beta, c1, n_stout, n_md, therm):
    """The FP64 CPU oracle at the measured lattice, on the thermalised links.

    1. Errors of the device forces against the oracle's FP64 force on the same
       links: the FP32 12-real links widened to FP64 and reunitarized in FP64
       (the device kernels form staple paths on two rows and rebuild the third,
       exact for links in SU(3); the widened FP32 links are off the group by
       ~1e-7), then rounded back to FP32 for the FP32-MD force. Reported as
       max |F_dev - F_ref| / max |F_ref| and as the RMS relative error, for
       the FP32-MD and the FP64 force; and the FP64 plaquette of both device
       fields against the oracle's.
    2. The CPU baseline (BASELINE.md §2 schedule): the oracle's force (the
       force of step 1 is its warm-up; then the median of 3), link update
       (median of 3) and Hamiltonian (median of 3) at the same lattice, composed
       into one trajectory with the FG schedule (4n forces, 4n + 1 link
       updates, 2 Hamiltonians)."""
    import statistics
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    dims = (L,) * 4
    a = O.action(act_kind, beta, c1, n_stout=n_stout, rho=0.1)
    U64 = M.Field(ctx, dims, M.LINK18, M.F64).copy_from(therm.links())
    M.reunitarize(U64)
    U32 = M.Field(ctx, dims, M.LINK12, M.F32).copy_from(U64)
    F32 = M.Field(ctx, dims, M.ALG8, M.F32)
    F64 = M.Field(ctx, dims, M.ALG8, M.F64)
    M.get_force(F32, U32, arg)
    M.get_force(F64, U64, arg)
    ctx.synchronize()
    U = U64.download()
    f32, f64 = F32.download(), F64.download()
    p32, p64 = M.avg_plaquette(U32), M.avg_plaquette(U64)
    del F32, F64, U64, U32
    print(f"bench: oracle force at {L}^4 ({O.num_threads()} threads)", file=sys.stderr, flush=True)

    def timed(fn):
        t = time.perf_counter()
        out = fn()
        return time.perf_counter() - t, out

    t_warm, Fref = timed(lambda: O.get_force(dims, U, a))
    fmax = np.abs(Fref).max()
    frms = np.sqrt(np.mean(Fref ** 2))

    def err(f):
        return {"max_rel": float(np.abs(f - Fref).max() / fmax),
                "rms_rel": float(np.sqrt(np.mean((f - Fref) ** 2)) / frms)}

    e32, e64 = err(f32), err(f64)
    del f32, f64
    pref = O.avg_plaquette(dims, U)
    errors = {"force_fp32": e32, "force_fp64": e64,
              "plaq_oracle": pref, "plaq_rel_fp32_links": abs(p32 - pref) / abs(pref),
              "plaq_rel_fp64_links": abs(p64 - pref) / abs(pref),
              "links": f"{L}^4 after thermalisation, widened to FP64 and reunitarized (FP32 force: those "
                       f"links rounded to FP32 12-real)"}
    print(f"bench: force errors FP32 {e32['max_rel']:.2e} FP64 {e64['max_rel']:.2e}; timing the oracle",
          file=sys.stderr, flush=True)
    tf = statistics.median(timed(lambda: O.get_force(dims, U, a))[0] for _ in range(3))
    del Fref
    P = O.init_momentum(dims, 7, 0)
    tu = statistics.median(timed(lambda: O.evolve_gauge_field(dims, U, P, 0.01, 0))[0] for _ in range(3))
    th = statistics.median(timed(lambda: O.hamiltonian(dims, U, P, a))[0] for _ in range(3))
    t_traj = 4 * n_md * tf + (4 * n_md + 1) * tu + 2 * th
    base = {"value": round(1.0 / t_traj, 8), "unit": "traj/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"FP64 oracle on the thermalised {L}^4 links, same action: force {tf:.2f} s "
                      f"(warm-up {t_warm:.2f} s), link update {tu:.3f} s, Hamiltonian {th:.2f} s (medians of 3); "
                      f"one trajectory = {4 * n_md} forces + {4 * n_md + 1} updates + 2 H = {t_traj:.0f} s"}
    return errors, base


def serve_stream(M, ctx, nchains, max_traj):
    """The serving stream (Poisson arrivals at 200/s; volumes {4,6,8,12,16,24}^4; Wilson or Iwasaki;
    0-3 stout levels; own seeds; FP32 MD, n_md = 10) for its first `nchains` chains with trajectory
    counts drawn from 5..max_traj, through the continuous-batching server."""
    from mag_amd import serving as S
    stream = S.poisson_stream(nchains=nchains, rate=200.0, seed=2026, traj_range=(5, max_traj))
    srv = S.Server(ctx, hbm_budget=200 << 30, max_batch=64, graphs=True)
    out, wall = S.drive(srv, stream, log=lambda m: print("bench: serving " + m, file=sys.stderr, flush=True))
    st = S.latency_stats(out, srv.extent)
    srv.close()
    total = sum(r.ntraj for _, r in stream)
    st.update({"chains": nchains, "wall_s": round(wall, 3), "traj_per_s": round(st["trajectories"] / wall, 2),
               "sample": (f"the full {nchains}-chain stream, {total} trajectories" if nchains == 512 and max_traj == 50
                          else f"first {nchains} chains of the 512-chain stream, trajectories 5..{max_traj}")})
    st["traj_latency_ms_by_extent_note"] = "p50, p99 of each lattice extent's per-trajectory gaps"
    return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}


def traffic_bytes(kernel, L):
    """HBM bytes per launch of `kernel` at L^4 from the calibrated rocprofv3
    FETCH_SIZE / WRITE_SIZE summary kept in profiles/traffic.json (None if
    that kernel and lattice were not profiled)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
        e = t["kernels"][kernel]
        if e["lattice"] != L:
            return None
        return int(e["fetch_bytes"] + e["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


# Thermalisation of the headline chain, independent of --warmup: at least
# THERM_MIN forced-accept FP32 trajectories from the Haar start, then blocks
# of THERM_BLOCK until the mean plaquette of the last block differs from the
# block before by less than THERM_DRIFT (relative), at most THERM_MAX. The
# measured chains accept within +-0.5% of their starting action; a drift of
# 2e-4 per 10 trajectories keeps a 25-trajectory run well inside it.
THERM_BLOCK, THERM_DRIFT, THERM_MAX = 10, 2e-4, 120


def thermalize(M, ctx, L, arg, seed, tmin):
    """Untimed: Haar start, then FP32 trajectories of the same workload with
    forced accepts until the plaquette has stopped drifting. Returns the chain
    holding the links and a summary."""
    ch = M.Chain(ctx, (L,) * 4, dataclasses.replace(arg, num_forced_accept_step=THERM_MAX), seed=seed,
                 precision=M.F32)
    ch.haar(seed)
    plaq = [r.avg_plaq for r in ch.run(tmin)]
    drift = None  # below two blocks (profiling runs) the drift rule is not applied
    while len(plaq) >= 2 * THERM_BLOCK:
        last = sum(plaq[-THERM_BLOCK:]) / THERM_BLOCK
        prev = sum(plaq[-2 * THERM_BLOCK:-THERM_BLOCK]) / THERM_BLOCK
        drift = abs(last - prev) / abs(last)
        if drift < THERM_DRIFT or len(plaq) + THERM_BLOCK > THERM_MAX:
            break
        plaq += [r.avg_plaq for r in ch.run(THERM_BLOCK)]
    ctx.synchronize()
    print(f"bench: thermalised in {len(plaq)} trajectories, plaquette {plaq[-1]:.6f}, block drift {drift}",
          file=sys.stderr, flush=True)
    return ch, {"trajectories": len(plaq), "avg_plaq": round(plaq[-1], 8), "block_drift": drift}


def timed_chain(M, ctx, L, arg, seed, precision, start, warmup, steps, timing, wloop=False, recon12=None,
                dominant_only=None):
    """A fresh chain (trajectory 0, Metropolis + demon window of 0.5% of the
    initial action sums) from the links of `start`: `warmup` untimed
    trajectories, then `steps` trajectories timed together between two device
    synchronisations (with `wloop`, one r = (4,0,0), T = 4 Wilson loop after
    each trajectory, inside the timed region). With `dominant_only` (a dict)
    and at least two warm-up trajectories, the warm-up after its first
    trajectory is kernel-timed family by family (per trajectory into
    `dominant_only`) and the timed region times only the family that took
    longest. Returns (seconds, warm-up results, timed
    results, chain, Wilson loops)."""
    ch = M.Chain(ctx, (L,) * 4, arg, seed=seed, precision=precision, accept_mode=2, demon_emax=-0.005,
                 recon12=recon12)
    ch.links().copy_from(start.links())
    ch.reset_demon()
    ch.demon_state()  # window centre from the starting links, outside the timed region
    if timing and dominant_only is not None and warmup >= 2:
        # the per-kernel breakdown from the warm-up trajectories after the
        # first (which also evaluates the starting Hamiltonian); inside the
        # timed region only the dominant kernel carries events
        pre = ch.run(1)
        ctx.synchronize()
        ctx.reset_timing()
        ctx.set_timing(True)
        pre += ch.run(warmup - 1)
        ctx.synchronize()
        warm = {}
        for f in list(KERNEL_UNITS) + ["stout_fwd_h", "action", "kinetic", "reduce", "mom_refresh", "reunitarize"]:
            ms, nl = ctx.kernel_time(f)
            if nl:
                warm[f] = (ms / (warmup - 1), nl / (warmup - 1))
        ctx.set_timing(False)
        dominant_only.update(warm)
        dom = max((f for f in warm if f in KERNEL_UNITS), key=lambda f: warm[f][0])
        ctx.set_timing_family(dom)
    else:
        pre = ch.run(warmup) if warmup else []
    ctx.synchronize()
    if timing:
        ctx.reset_timing()
        ctx.set_timing(True)
    loops = []
    t0 = time.perf_counter()
    if wloop:
        res = []
        for _ in range(steps):
            res += ch.run(1)
            loops.append(M.avg_wilson_loop(ch.links(), (4, 0, 0), 4))
    else:
        res = ch.run(steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_timing(False)
    ctx.set_timing_family(None)
    return dt, pre, res, ch, loops


def kernel_ms(ctx, steps):
    """HIP-event kernel time per trajectory of every timed kernel family."""
    out = {}
    for f in list(KERNEL_UNITS) + ["stout_fwd_h", "action", "kinetic", "reduce", "mom_refresh", "reunitarize"]:
        ms, nl = ctx.kernel_time(f)
        if nl:
            out[f] = round(ms / steps, 3)
    return out


def paired_acceptance(r32, r64):
    """The FP32-MD and FP64 chains start from the same links and draw the same
    momenta and die rolls, so trajectory i of both integrates the same initial
    state while every earlier decision agreed (the first trajectory whose
    decisions differ still pairs). Over those trajectories: the Metropolis
    probabilities min(1, e^-dH) of both, their paired difference (mean and
    standard error), |dH32 - dH64| (absolute, relative to |H|) and the
    plaquette difference of the resulting states."""
    import math
    m = min(len(r32), len(r64))
    n = 0
    while n < m and r32[n].accept == r64[n].accept:
        n += 1
    n = min(n + 1, m)
    p32 = [min(1.0, math.exp(-r.dh)) for r in r32[:n]]
    p64 = [min(1.0, math.exp(-r.dh)) for r in r64[:n]]
    d = [x - y for x, y in zip(p32, p64)]
    mean = sum(d) / n
    se = (sum((x - mean) ** 2 for x in d) / (n - 1) / n) ** 0.5 if n > 1 else None
    ddh = [abs(a.dh - b.dh) for a, b in zip(r32[:n], r64[:n])]
    dpl = [abs(a.avg_plaq - b.avg_plaq) for a, b in zip(r32[:n], r64[:n]) if a.accept and b.accept]
    return {"trajectories": n, "same_decisions": sum(a.accept == b.accept for a, b in zip(r32[:n], r64[:n])),
            "metropolis_p_fp32_mean": round(sum(p32) / n, 5), "metropolis_p_fp64_mean": round(sum(p64) / n, 5),
            "metropolis_p_diff_mean": mean, "metropolis_p_diff_se": se,
            "dh_abs_diff_max": max(ddh), "dh_abs_diff_mean": sum(ddh) / n,
            "dh_diff_rel_to_h_max": max(ddh) / abs(r64[0].h_old),
            "plaq_abs_diff_max": max(dpl) if dpl else None,
            "dh_fp32": [round(r.dh, 5) for r in r32[:n]], "dh_fp64": [round(r.dh, 5) for r in r64[:n]]}


# Coupled coarse/fine md:: HMC: 3x3 products per coarse site of one fine
# stout-term evaluation (4 links of class 1 and 3 with 63 products, 12 of
# class 2 and 4 with 11) and of one coarse-force evaluation (4 links, 12
# loops of 6 links at 3 products, the U2 pair, ~20 for DexpiQ and the tail).
MULTI_FFORCE_PRODUCTS_CSITE = 8 * 63 + 24 * 11
MULTI_CFORCE_PRODUCTS_CSITE = 4 * (12 * 6 * 3 + 2 * 3 + 20)


def multi_main(a):
    """run_hmc_multi (field-multi-md.h:785-988) at fine L^4 / coarse (L/2)^4,
    FP64: forced-accept warmup trajectories, then `steps` timed trajectories
    with HIP-event kernel timing. Prints one JSON line."""
    import mag_amd as M
    from mag_amd import multi_md as MD
    ctx = M.Context(0)
    L, n = a.L, a.n_md
    kind, beta, c1 = ACTIONS[a.action]
    beta = a.beta if a.beta is not None else beta
    arg = M.Arg_chmc(trajectory_length=n, dt=1.0 / n, beta=beta, gauge=M.Gauge(kind, c1),
                     num_forced_accept_step=a.warmup)
    mp = MD.params(0.1, 1.0)
    dims = (L, L, L, L)
    ch = MD.MultiChain(ctx, dims, arg, mp, seed=a.seed)
    UF, UC = ch.fields()
    UF.haar(a.seed)
    MD.get_Gb(UC, UF, mp)
    print(f"bench --multi: fine {L}^4 / coarse {L // 2}^4, n = {n}, warmup {a.warmup}, steps {a.steps}",
          file=sys.stderr, flush=True)
    if a.warmup:
        ch.run(a.warmup)
    ctx.synchronize()
    ctx.set_timing(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    res = ch.run(a.steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    fams = {}
    for f in ("force_wilson", "force_iwasaki", "multi_prep", "multi_fforce", "multi_cforce", "multi_heatbath",
              "multi_xi", "link_update", "mom_update", "copy", "kinetic", "action", "reunitarize", "reduce",
              "mom_refresh", "fill"):
        ms, nl = ctx.kernel_time(f)
        if nl:
            fams[f] = (ms, nl)
    Vc = (L // 2) ** 4
    # the fine stout term: one launch per evaluation (all four classes)
    ms, nl = fams["multi_fforce"]
    per_eval_s = ms * 1e-3 / nl
    tflops = MULTI_FFORCE_PRODUCTS_CSITE * M_FLOPS * Vc / per_eval_s / 1e12
    cms, cnl = fams["multi_cforce"]
    c_tflops = MULTI_CFORCE_PRODUCTS_CSITE * M_FLOPS * Vc / (cms * 1e-3 / cnl) / 1e12  # one launch per evaluation
    act_name = {"iwasaki": "Iwasaki", "dbw2": "DBW2", "wilson": "Wilson"}[a.action]
    out = {"metric": "md:: coarse/fine HMC trajectories/s (run_hmc_multi, FP64)", "value": round(a.steps / dt, 4),
           "unit": "traj/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(1e3 * dt / a.steps, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (Haar-random fine links, seed %d)" % a.seed,
           "config": {"workload": f"fine {L}^4 / coarse {L // 2}^4 {act_name}(beta={beta}), rho=0.1, XI0=1, "
                                  f"md_alpha=0.5, zeta_H tau=1 n={n}, heat bath every trajectory",
                      "lattice": [L] * 4, "parallelism": "single GPU"},
           "accept_rate": round(sum(r.accept for r in res) / len(res), 3),
           "dh": [round(r.dh, 5) for r in res],
           "kernel_ms_per_traj": {f: round(v[0] / a.steps, 3) for f, v in fams.items()},
           "roofline": {"bound": "valu", "kernel": "multi_fforce (classes 1-4, one launch)", "achieved": round(tflops, 3),
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
                        "traffic": None, "avg_eval_ms": round(per_eval_s * 1e3, 4)},
           "coarse_force_tflops": round(c_tflops, 3)}
    print(json.dumps(out))


def local_main(a):
    """Local updates at L^4, FP64 Wilson beta = 5.7: algGCtrnHeatBath in the
    checkerboard order (nHits = 10, small = 0.5, mag = 0 and 2) and the
    checkerboard demon (2 hits per link, E_max = 2). Hot start; `warmup`
    untimed iterations / sweeps, then `steps` timed ones (wall clock, no
    per-launch timing), then as many again with HIP-event kernel timing for
    the kernel times. Prints one JSON line."""
    import numpy as np
    import mag_amd as M
    ctx = M.Context(0)
    L = a.L
    dims = (L, L, L, L)
    nlinks = 4 * L ** 4
    out = {"metric": "local-update link updates/s (checkerboard Metropolis and demon, FP64)", "unit": "link hits/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (Haar-random links, seed %d)" % a.seed,
           "config": {"workload": f"{L}^4 Wilson beta=5.7", "lattice": [L] * 4, "parallelism": "single GPU"}}
    res = {}
    for mag in (0, 2):
        f = M.Field(ctx, dims, M.LINK18, M.F64).haar(a.seed)
        if a.warmup:
            M.local_metropolis(f, 5.7, mag, a.warmup, 10, 0.5, a.seed)
        ctx.synchronize()
        # wall time without per-launch event timing; kernel times from a second pass with it
        t0 = time.perf_counter()
        acc, p, _ = M.local_metropolis(f, 5.7, mag, a.steps, 10, 0.5, a.seed + 1)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        ctx.set_timing(True)
        ctx.reset_timing()
        M.local_metropolis(f, 5.7, mag, a.steps, 10, 0.5, a.seed + 2)
        ctx.synchronize()
        kms, kn = ctx.kernel_time("local_metropolis")
        ctx.set_timing(False)
        hits = sum(acc)
        res[f"metropolis_mag{mag}"] = {"hits_per_s": round(hits / dt, 1), "ms_per_iteration": round(1e3 * dt / a.steps, 3),
                                       "kernel_ms_per_iteration": round(kms / a.steps, 3),
                                       "accept": round((acc[0] + acc[2]) / max(hits, 1), 4),
                                       "plaquette": round(float(p[-1]), 6)}
    f = M.Field(ctx, dims, M.LINK18, M.F64).haar(a.seed)
    d, _ = M.demon_checkerboard(f, max(a.warmup // 5, 1), 2.0, a.seed)
    ctx.synchronize()
    nsw = max(a.steps // 5, 1)
    t0 = time.perf_counter()
    d, (acc, rej) = M.demon_checkerboard(f, nsw, 2.0, a.seed + 1, d)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_timing(True)
    ctx.reset_timing()
    M.demon_checkerboard(f, nsw, 2.0, a.seed + 2, d)
    ctx.synchronize()
    kms, kn = ctx.kernel_time("demon_checkerboard")
    ctx.set_timing(False)
    res["demon_checkerboard"] = {"hits_per_s": round((acc + rej) / dt, 1), "ms_per_sweep": round(1e3 * dt / nsw, 3),
                                 "kernel_ms_per_sweep": round(kms / nsw, 3), "accept": round(acc / (acc + rej), 4),
                                 "sweeps": nsw}
    out["value"] = res["metropolis_mag0"]["hits_per_s"]
    out["ms_per_step"] = res["metropolis_mag0"]["ms_per_iteration"]
    # roofline of the Metropolis class kernel: per link 12 staple products +
    # 10 hits x 3 products (U eps, two traces counted as one product each)
    prods = 12 + 10 * 3
    tflops = prods * M_FLOPS * nlinks / (res["metropolis_mag0"]["kernel_ms_per_iteration"] * 1e-3) / 1e12
    out["roofline"] = {"bound": "valu", "kernel": "local_class (checkerboard Metropolis)", "achieved": round(tflops, 3),
                       "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
                       "traffic": None}
    out["results"] = res
    print(json.dumps(out))


def shard_main(a):
    """One GPU's share of the 64^4 DBW2 + 3-stout, n = 16 workload decomposed over 8 GPUs along x3:
    the unchanged FP32-MD trajectory on a 64^3 x 8 block whose periodic wrap stands in for the halo.
    Hot start with forced accepts (a throughput measurement, like the 64^4 leg). Prints one JSON line
    with the measured force time per evaluation and a model of the halo traffic of the split: per
    force, each forward level needs U_k one layer deep, the top level U_3 two layers deep (rectangles)
    and each backward level Lambda_k one layer deep, from each of the two neighbours along x3, over one
    xGMI link each (153 GB/s)."""
    import mag_amd as M
    ctx = M.Context(0)
    dims = (64, 64, 64, 8)
    n = 16
    kind, beta, c1 = ACTIONS["dbw2"]
    arg = M.Arg_chmc(trajectory_length=n, dt=1.0 / n, beta=beta, gauge=M.Gauge(kind, c1), n_stout=3, rho=0.1,
                     num_forced_accept_step=1 << 30)
    ch = M.Chain(ctx, dims, arg, seed=a.seed, precision=M.F32)
    ch.haar(a.seed)
    warm = max(a.warmup, 1)
    ch.run(warm)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    ch.run(a.steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_timing(False)
    kms = kernel_ms(ctx, a.steps)
    force_ms = sum(kms.get(f, 0.) for f in ("stout_fwd", "stout_top", "stout_bwd", "stout_bwd_force")) / (4 * n)
    face = dims[0] * dims[1] * dims[2]  # sites per x3 layer
    link12, alg8 = 4 * 12 * 4, 4 * 8 * 4  # bytes per site: 4 links of 12 FP32 reals; 4 algebra elements
    per_side = 3 * face * link12 + 2 * face * link12 + 3 * face * alg8  # U_0..U_2, U_3 x 2 layers, Lambda_2..0
    xgmi = 153e9
    out = {"metric": "traj/s of one GPU's share of 64^4 DBW2+3stout over 8 GPUs (64^3 x 8 block)", "value":
           round(a.steps / dt, 4), "unit": "traj/s", "n_gpus": 1, "steps": a.steps, "warmup": warm,
           "ms_per_step": round(1e3 * dt / a.steps, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (Haar-random links, seed %d, forced accepts)" % a.seed,
           "config": {"workload": "64^3 x 8 DBW2(c1=-1.4069, beta=1.04) + 3-stout(rho=0.1), FG tau=1 n=16, FP32 "
                                  "12-real links, periodic wrap in place of halos", "lattice": list(dims),
                      "parallelism": "single GPU (one share of an 8-way split along x3)"},
           "force_ms": round(force_ms, 3), "kernel_ms_per_traj": kms,
           "halo_model": {"bytes_per_force_per_neighbour": per_side,
                          "xgmi_gbs_per_link": xgmi / 1e9,
                          "exchange_ms_per_force": round(1e3 * per_side / xgmi, 3),
                          "exchange_over_force": round(1e3 * per_side / xgmi / force_ms, 3),
                          "note": "two neighbours along x3 over two links at once; faces of 64^3 sites; "
                                  "interior layers 2..5 of 8 do not need the halo and can overlap it"}}
    print(json.dumps(out))


def main():
    a = parse()
    if a.shard:
        if a.gpus != 1:
            print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (single-GPU build)",
                  file=sys.stderr)
            sys.exit(2)
        shard_main(a)
        return
    if a.local:
        if a.gpus != 1:
            print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (single-GPU build)",
                  file=sys.stderr)
            sys.exit(2)
        local_main(a)
        return
    if a.multi:
        if a.gpus != 1:
            print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (single-GPU build)",
                  file=sys.stderr)
            sys.exit(2)
        multi_main(a)
        return
    if a.fp64_steps < 0:
        a.fp64_steps = a.steps
    if a.fp64_12_steps < 0:
        a.fp64_12_steps = a.fp64_steps
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (single-GPU build)", file=sys.stderr)
        sys.exit(2)
    import mag_amd as M
    ctx = M.Context(0)
    L, n = a.L, a.n_md
    kind, beta, c1 = ACTIONS[a.action]
    beta = a.beta if a.beta is not None else beta
    arg = M.Arg_chmc(trajectory_length=n, dt=1.0 / n, beta=beta, gauge=M.Gauge(kind, c1), n_stout=a.n_stout,
                     rho=0.1, num_forced_accept_step=0)
    act_name = {"iwasaki": "Iwasaki", "dbw2": "DBW2", "wilson": "Wilson"}[a.action]
    V = L ** 4
    nlinks = 4 * V
    print(f"bench: {L}^4 {act_name}+{a.n_stout}stout FP32-MD, warmup {a.warmup}, steps {a.steps}", file=sys.stderr,
          flush=True)
    therm, therm_info = thermalize(M, ctx, L, arg, a.seed, a.therm_min)
    warm = {}
    dt, pre, res, ch32, loops = timed_chain(M, ctx, L, arg, a.seed, M.F32, therm, a.warmup, a.steps, True, a.wloop,
                                            dominant_only=warm)
    print(f"bench: FP32 {a.steps} trajectories in {dt:.3f} s", file=sys.stderr, flush=True)
    fams = {}
    for f in list(KERNEL_UNITS) + ["stout_fwd_h", "action", "kinetic", "reduce", "mom_refresh", "reunitarize"]:
        ms, nl = ctx.kernel_time(f)
        if nl:
            fams[f] = (ms, nl)
    timed_families = sorted(fams)
    # families not timed inside the timed region: per-trajectory times of the warm-up, scaled to the steps
    for f, (ms, nl) in warm.items():
        if f not in fams:
            fams[f] = (ms * a.steps, nl * a.steps)
    traj_s = a.steps / dt
    nforce = 4 * n * a.steps
    force_ms = sum(fams.get(f, (0, 0))[0] for f in ("stout_fwd", "stout_top", "stout_bwd", "stout_bwd_force",
                                                     "force_iwasaki"))
    force_s = force_ms * 1e-3
    if a.n_stout == 3:
        force_bytes_link = FORCE_BYTES_LINK_C3
    else:
        force_bytes_link = sum(KERNEL_UNITS[f][0] * (a.n_stout if f == "stout_fwd" else
                                                     (a.n_stout - 1 if f == "stout_bwd" else 1))
                               for f in ("stout_fwd", "stout_top", "stout_bwd", "stout_bwd_force"))
    force_gbs = force_bytes_link * nlinks * nforce / force_s / 1e9 if force_ms else None
    force_tflops = (FORCE_PRODUCTS_LINK.get(a.n_stout, 85 + 75 * a.n_stout) * M_FLOPS * nlinks * nforce / force_s
                    / 1e12 if force_ms else None)
    if warm and len(timed_families) == 1:
        dom = timed_families[0]  # the one family with events inside the timed region
    else:
        dom = max((f for f in fams if f in KERNEL_UNITS), key=lambda f: fams[f][0])
    ms, nl = fams[dom]
    avg_s = ms * 1e-3 / nl
    b_link, m_link = KERNEL_UNITS[dom]
    gbs = b_link * nlinks / avg_s / 1e9
    tflops = m_link * M_FLOPS * nlinks / avg_s / 1e12
    roof = {"bound": "valu", "kernel": dom, "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "traffic": traffic_bytes(dom, L),
            "hbm_achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "avg_launch_ms": round(avg_s * 1e3, 4),
            # the library is built without packed FP32 instructions, so one FMA per lane per
            # cycle (half the published FP32 rate) is the ceiling these kernels can reach
            "valu_unpacked_peak": FP32_UNPACKED_TFLOPS,
            "valu_unpacked_frac": round(tflops / FP32_UNPACKED_TFLOPS, 4)}
    out = {"metric": "traj/s + force GB/s at 32^4 Iwasaki+3stout, FP32-MD vs FP64; chain p99 latency",
           "value": round(traj_s, 4), "unit": "traj/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(1e3 * dt / a.steps, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (Haar-random links, seed %d, thermalised by %d forced-accept FP32 trajectories)"
                   % (a.seed, therm_info["trajectories"]),
           "config": {"workload": f"{L}^4 {act_name}(c1={c1}, beta={beta}) + {a.n_stout}-stout(rho=0.1), "
                                  f"FG tau=1 n={n}, FP32 12-real links, FP64 dH, demon window 0.5%"
                                  + (", 4x4 Wilson loop per trajectory" if a.wloop else ""),
                      "lattice": [L] * 4, "parallelism": "single GPU"},
           "link_updates_per_s": round(4 * V * n * traj_s, 1),
           "force_gbs": round(force_gbs, 1) if force_gbs else None,
           "force_tflops": round(force_tflops, 2) if force_tflops else None,
           "accept_rate": round(sum(r.accept for r in res) / len(res), 3),
           "thermalisation": therm_info,
           "dh": [round(r.dh, 5) for r in res],
           "kernel_ms_per_traj": {f: round(v[0] / a.steps, 3) for f, v in fams.items()},
           "kernel_ms_source": (f"{dom}: HIP events over the timed region; the other kernels: HIP events over "
                                f"warm-up trajectories 2-{a.warmup}" if warm and len(timed_families) == 1 else
                                "HIP events over the timed region"),
           "roofline": roof}
    if loops:
        out["wilson_loop_4x4"] = [round(w, 8) for w in loops]
    if out["accept_rate"] == 0:
        out["invalid"] = "the measured chain accepted no trajectory"
    if a.fp64_steps > 0:
        dt64, pre64, r64, _, _ = timed_chain(M, ctx, L, arg, a.seed, M.F64, therm, a.warmup, a.fp64_steps,
                                             a.fp64_kernel_times, a.wloop)
        if a.fp64_kernel_times:
            out["fp64_kernel_ms_per_traj"] = kernel_ms(ctx, a.fp64_steps)
        out["fp64_traj_per_s"] = round(a.fp64_steps / dt64, 4)
        out["fp32_over_fp64"] = round(traj_s / (a.fp64_steps / dt64), 3)
        out["fp64_accept_rate"] = round(sum(r.accept for r in r64) / len(r64), 3)
        # warm-up and timed trajectories of both chains, paired one by one
        out["fp32_vs_fp64"] = paired_acceptance(pre + res, pre64 + r64)
        print(f"bench: FP64 {a.fp64_steps} trajectories in {dt64:.3f} s", file=sys.stderr, flush=True)
    if a.fp64_12_steps > 0:
        # the FP64 path with links stored as 12 reals (row 2 rebuilt on load)
        dt12, _, r12, _, _ = timed_chain(M, ctx, L, arg, a.seed, M.F64, therm, a.warmup, a.fp64_12_steps,
                                         a.fp64_kernel_times, a.wloop, recon12=True)
        if a.fp64_kernel_times:
            out["fp64_12real_kernel_ms_per_traj"] = kernel_ms(ctx, a.fp64_12_steps)
        out["fp64_12real_traj_per_s"] = round(a.fp64_12_steps / dt12, 4)
        out["fp32_over_fp64_12real"] = round(traj_s / (a.fp64_12_steps / dt12), 3)
        out["fp64_12real_accept_rate"] = round(sum(r.accept for r in r12) / len(r12), 3)
        print(f"bench: FP64 12-real {a.fp64_12_steps} trajectories in {dt12:.3f} s", file=sys.stderr, flush=True)
    if a.reversibility:
        dev, dh_rt = ch32.reversibility()
        out["reversibility_fp64"] = {"max_link_dev": dev, "dh_roundtrip": dh_rt}
    del ch32
    if a.serve_chains > 0:
        out["serving"] = serve_stream(M, ctx, a.serve_chains, a.serve_max_traj)
    if not a.no_cpu_baseline:
        errors, out["cpu_baseline"] = oracle_leg(M, ctx, L, arg, kind, beta, c1, a.n_stout, n, therm)
        out["errors_vs_oracle"] = errors
    print(json.dumps(out))


if __name__ == "__main__":
    main()
