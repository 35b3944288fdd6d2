"""Throughput of the operator-split reaction-diffusion period on one MI355X.

A step is one operator-split period (SSA + diffusion leaving + entering) over
every tet of the workload, with inputs already resident in HBM. The default
workload is the 4-species / 6-reaction model on a 40^3 x 6 Kuhn cube of side
40 um (384,000 tets) at the diffusion stability bound (2.96 ms per period);
338 steps are 1 s of simulated time. The JSON line also carries:
  - the batched ensemble (1,024 realizations of A + B <-> C on cubes of
    1,296..48,000 tets, 0.1 s each) and the same model run as one realization,
    so batched and plain throughput are measured on the same GPU in one run;
  - the 6,000,000-tet calcium membrane cube (pumps, influx on one face) and its
    EField on 1,030,301 vertices (mixed precision vs FP64 CG);
  - the 3,072-tet A + B <-> C protocol (run() every 1 ms to 0.1 s, totals
    read at every save) over 100 seeds;
  - the CPU oracle on bounded samples of these workloads on the host cores.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-batched]
                       [--no-membrane]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from steps_amd import workloads as W  # noqa: E402
from steps_amd.autotune import autotune  # noqa: E402
from steps_amd.engine import Engine  # noqa: E402

METRIC = "tet-steps/sec (batched vs plain, same GPU) + achieved HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0


def interior_faces_per_tet(mt) -> float:
    return float((mt.nbr >= 0).sum()) / len(mt.vol)


def bytes_per_tet_step(S, S_d, S_o, Fbar, T_over_N=0.0, S_s=0) -> float:
    """Compulsory HBM bytes per tet-step with state round-tripping HBM every
    period: int32 counts r/w, int32 neighbour + 8-bit face slot, FP64 face
    factor, int32 outgoing slot written and read, FP64 occupancy r/w for
    species that react and diffuse, membrane triangle terms."""
    return (8 * S + 5 * Fbar + 8 * Fbar + 8 * S_d * Fbar + 16 * S_o
            + T_over_N * (8 * S_s + 12))


def model_bytes(w) -> float:
    fm = w.fm
    S = len(fm.specs)
    diff = fm.spec_dcst > 0
    reacts = (fm.reac_upd != 0).any(axis=0)
    if len(fm.sreacs):
        reacts |= (fm.sreac_upd_i != 0).any(axis=0)
    S_d = int(diff.sum())
    S_o = int((diff & reacts).sum())
    N = w.n_tets
    T = len(w.mt.tri_tet)
    return bytes_per_tet_step(S, S_d, S_o, interior_faces_per_tet(w.mt), T / N, len(fm.sspecs))


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_period_four_species_cube_384k_hbm_traffic.md")
ENSEMBLE_TRAFFIC = os.path.join(ROOT, "profiles", "k_period_ab_binding_ensemble_1024_hbm_traffic.md")
MEMBRANE_TRAFFIC = os.path.join(ROOT, "profiles", "k_period_calcium_membrane_6M_hbm_traffic.md")


def measured_traffic(path=TRAFFIC_SUMMARY, key="traffic_bytes_per_launch"):
    """HBM bytes of the period kernel from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by the copy calibration): per launch of the
    plain workloads, per tet-step of the ensemble."""
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


# face slot width (bytes) of the fixed launch shapes (--block); the autotuner picks its own
FIXED_SLOT_WIDTH = 2


def load_engine(w, seed, block=256, graph=False, specialize=False, spec_waves=0):
    eng = Engine(w.fm, [w.mt], [seed])
    for s in range(len(w.fm.specs)):
        eng.set_tet_counts(s, w.counts[s])
    for s in range(len(w.fm.sspecs)):
        eng.set_tri_counts(s, w.tri_counts[s])
    eng.set_launch(block, graph)
    eng.set_slot_width(FIXED_SLOT_WIDTH)
    if specialize:
        eng.set_specialize(True, spec_waves)
    return eng


def time_plain(w, steps, warmup, seed=1, block=0, graph=False, specialize=False, spec_waves=0):
    """block = 0: launch shape and kernel chosen by the autotuner on this state."""
    eng = load_engine(w, seed, block or 256, graph, specialize and bool(block), spec_waves)
    tuned = autotune(eng, trial_periods=20) if not block else None
    dt = eng.update_period()
    if warmup:
        eng.run(warmup * dt)
    n0, t0 = eng.nsteps(), eng.time()
    tic = time.perf_counter()
    eng.run(t0 + steps * dt)  # returns after its stream has drained
    wall = time.perf_counter() - tic
    periods = eng.nsteps() - n0
    ms_launch, launches = eng.last_run_kernel_ms()
    return dict(eng=eng, wall=wall, periods=periods, tet_steps=periods * w.n_tets,
                ms_launch=ms_launch, launches=launches, dt=dt, tuned=tuned,
                block=tuned["block"] if tuned else block, graph=tuned["graph"] if tuned else graph,
                specialized=eng.specialized, lanes=eng.lanes, slot_width=eng.slot_width)


def time_ensemble(n_real=1024, end_time=0.1, seed=0, block=0, graph=True, specialize=False):
    fm, meshes, counts, reac_k, info = W.ab_binding_ensemble(n_real=n_real, seed=seed,
                                                             end_time=end_time)
    eng = Engine(fm, meshes, [seed + 1] * n_real, reac_k=reac_k)
    for r in range(n_real):
        for s in range(3):
            eng.set_tet_counts(s, counts[r][s], real=r)
    eng.set_launch(block or 256, graph)
    eng.set_slot_width(FIXED_SLOT_WIDTH)
    if block and specialize:
        eng.set_specialize(True)
    # launch shape from the autotuner (block = 0), timed on this ensemble's own state
    tuned = autotune(eng, trial_periods=3, blocks=(64, 128, 256)) if not block else None
    # warm-up pass (captures and instantiates the graph), then the timed pass
    eng.run(end_time)
    eng.reset()
    for r in range(n_real):
        for s in range(3):
            eng.set_tet_counts(s, counts[r][s], real=r)
    tic = time.perf_counter()
    eng.run(end_time)
    wall = time.perf_counter() - tic
    ms_launch, launches = eng.last_run_kernel_ms()
    tet_steps = sum(eng.nsteps(r) * len(meshes[r].vol) for r in range(n_real))
    tot_tets = sum(len(m.vol) for m in meshes)
    fb = np.mean([interior_faces_per_tet(m) for m in meshes])
    btet = bytes_per_tet_step(3, 3, 3, fb)
    # conservation check over the whole batch
    ok = True
    for r in range(n_real):
        tot = [eng.sum_tet_counts(s, r) for s in range(3)]
        ok &= (tot[0] + tot[2] == counts[r][0].sum() + counts[r][2].sum())
    return dict(wall=wall, tet_steps=tet_steps, n_tets=tot_tets, ms_launch=ms_launch,
                launches=launches, bytes_per_tet_step=btet, conserved=bool(ok), tuned=tuned,
                max_periods=max(eng.nsteps(r) for r in range(n_real)))


def time_membrane(w, steps=300, warmup=5, block=0, specialize=False):
    """Calcium buffering with membrane pumps and influx on the 100^3 x 6 cube."""
    p = time_plain(w, steps, warmup, block=block, specialize=specialize)
    btet = model_bytes(w)
    achieved = btet * w.n_tets / (p["ms_launch"] * 1e-3) / 1e9
    traffic = measured_traffic(MEMBRANE_TRAFFIC)
    return dict(
        workload="calcium_membrane_cube: 100^3 x 6 Kuhn cube, side 20 um, 6,000,000 tets, "
                 "120,000 membrane triangles, Ca + Buf <-> CaB, pumps, influx on one face",
        tet_steps_per_s=p["tet_steps"] / p["wall"], periods=p["periods"], period_s=p["dt"],
        ms_per_step=p["wall"] * 1e3 / max(p["periods"], 1),
        kernel_ms_per_launch=p["ms_launch"], bytes_per_tet_step=btet,
        achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS,
        traffic=traffic,
        traffic_frac=(traffic / (p["ms_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
        schedule={"block": p["block"], "graph": bool(p["graph"]), "specialized_kernel": bool(p["specialized"]),
                  "lanes_per_tet": p["lanes"], "slot_width": p["slot_width"]})


def time_efield(w, ef_steps=20, ef_dt=1e-5):
    """EField on the membrane cube's 1,030,301 vertices, coupled to the reaction-diffusion
    step: passive leak 1 S/m^2 at -65 mV, 1e-2 F/m^2, 1 ohm.m, 1 nA on the x = 0 face.
    Mixed-precision solve timed over coupled steps (reaction-diffusion periods included); then
    one bare step from the same state in mixed precision and in FP64 CG (tolerance 1e-11), both
    timed, their dV compared."""
    from steps_amd.efield import EField, FP64, JACOBI, MIXED, MULTIGRID

    mesh = w.mesh
    surf = mesh.getSurfTris()
    tic = time.perf_counter()
    ef = EField(mesh.verts, mesh.tets, mesh.tris[surf])
    setup_s = time.perf_counter() - tic
    ef.set_potential(-0.065)
    ef.set_capac(1e-2)
    ef.set_volres(1.0)
    ef.set_memb_res(1.0, -0.065)
    vx = mesh.verts[mesh.tris[surf]][:, :, 0]
    face = np.nonzero(np.all(np.abs(vx) < 1e-12, axis=1))[0]
    area = w.mt.tri_area[face]
    for t, a in zip(face, area):
        ef.set_tri_iclamp(int(t), 1e-9 * a / area.sum())
    eng = load_engine(w, 3)
    # the launch shape the solver interface would use for this mesh, chosen by
    # the autotuner on the reaction-diffusion state before the EField attaches
    tuned_rd = autotune(eng, trial_periods=5, blocks=(64, 128, 256), graphs=(False,))
    ef.run(eng, ef_dt, ef_dt)                      # warm-up step
    t0 = eng.time()
    iters, walls = [], []
    for k in range(1, ef_steps + 1):               # one coupled step per call, to read its iterations
        tic = time.perf_counter()
        ef.run(eng, t0 + k * ef_dt, ef_dt)
        walls.append(time.perf_counter() - tic)
        iters.append(ef.stats()["iterations"])
    wall = sum(walls)
    settled = walls[ef_steps // 2:]
    stats = ef.stats()
    # precision: one more step from the same state, mixed vs FP64 CG
    v0 = ef.verts_v()
    tic = time.perf_counter()
    ef.advance(ef_dt)
    mixed_s = time.perf_counter() - tic
    dv_mixed = ef.verts_v() - v0
    it_mixed = ef.stats()
    ef.set_verts_v(v0)
    ef.set_solver(FP64, 1e-11, 50000)
    tic = time.perf_counter()
    ef.advance(ef_dt)
    fp64_s = time.perf_counter() - tic
    dv_64 = ef.verts_v() - v0
    it_64 = ef.stats()
    ef.set_solver(MIXED)
    err = float(np.abs(dv_mixed - dv_64).max() / np.abs(dv_64).max())
    info = ef.precond_info()
    # a few more coupled steps with the Jacobi preconditioner, for comparison
    ef.set_precond(JACOBI)
    jsteps = 3
    t1 = eng.time()
    tic = time.perf_counter()
    ef.run(eng, t1 + jsteps * ef_dt, ef_dt)
    jacobi_s = (time.perf_counter() - tic) / jsteps
    it_j = ef.stats()
    ef.set_precond(MULTIGRID)
    nv, nt, nnz = ef.sizes()
    return dict(
        workload="EField on calcium_membrane_cube: 1,030,301 vertices, 120,000 membrane "
                 "triangles, 10 us steps coupled to the reaction-diffusion period",
        n_verts=nv, nnz=nnz, setup_s=setup_s, ms_per_ef_step=wall * 1e3 / ef_steps,
        ms_per_ef_step_second_half=sum(settled) * 1e3 / len(settled),
        ef_steps=ef_steps, cg_iterations_per_step=iters,
        rd_schedule={"block": tuned_rd["block"], "specialized_kernel": eng.specialized, "lanes_per_tet": eng.lanes,
                     "slot_width": eng.slot_width}, cg_iterations_last=stats["iterations"],
        refinements_last=stats["refinements"], rel_residual_last=stats["rel_residual"],
        mixed_vs_fp64_dv_rel_err=err, mixed_iterations=it_mixed["iterations"],
        mixed_step_ms=mixed_s * 1e3, fp64_iterations=it_64["iterations"], fp64_step_ms=fp64_s * 1e3,
        multigrid_levels=info["levels"], multigrid_coarsest=info["coarsest"],
        multigrid_setup_s=info["setup_s"], jacobi_step_ms=jacobi_s * 1e3,
        jacobi_iterations=it_j["iterations"])


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_periods(w, budget_s, seed=1):
    """Periods of w on the CPU oracle within about budget_s: (tet-steps/s, periods, wall)."""
    from oracle import bind

    o = bind.OracleSim(w.fm, w.mt, seed=seed)
    o.counts[:] = w.counts
    if len(w.fm.sspecs):
        o.tri_counts[:] = w.tri_counts
    dt = o.update_period()
    tic = time.perf_counter()
    o.run_periods([dt])
    one = time.perf_counter() - tic
    n = max(1, int(budget_s / max(one, 1e-6)))
    tic = time.perf_counter()
    o.run_periods([dt] * n)
    wall = time.perf_counter() - tic
    return n * w.n_tets / wall, n, wall


def config0_protocol_oracle(n_seeds, budget_s):
    """3,072-tet A + B <-> C (10,000 A and 10,000 B placed by seed), run() every
    1 ms to 0.1 s with the species totals read at every save, on the CPU
    oracle; seeds 0, 1, ... until n_seeds or the budget."""
    from oracle import bind

    done, steps, tic = 0, 0, time.perf_counter()
    for seed in range(n_seeds):
        w = W.ab_binding_cube(n=8, side=10e-6, n_a=10_000, n_b=10_000, seed=seed)
        o = bind.OracleSim(w.fm, w.mt, seed=seed)
        o.counts[:] = w.counts
        for k in range(1, 101):
            o.run(k * 1e-3, 0)
            o.counts.sum(axis=1)
        done += 1
        steps += o.nsteps * w.n_tets
        if time.perf_counter() - tic > budget_s:
            break
    wall = time.perf_counter() - tic
    return dict(tet_steps_per_s=steps / wall, seeds=done, wall_s=wall, periods_per_seed=100)


def config0_protocol_gpu(n_seeds=100):
    """The same protocol on the GPU: one engine per seed run one after another
    (plain), and all seeds as one batched engine (each seed a realization with its
    own Philox key); totals reduced on the device at every 1 ms save (one call for every
    species of every realization). Each engine's launch shape is autotuned on its initial
    state first (not timed)."""
    ws = [W.ab_binding_cube(n=8, side=10e-6, n_a=10_000, n_b=10_000, seed=sd) for sd in range(n_seeds)]
    eng = Engine(ws[0].fm, [ws[0].mt], [0])
    for s_ in range(3):
        eng.set_tet_counts(s_, ws[0].counts[s_])
    tuned_plain = autotune(eng, trial_periods=10, blocks=(64, 128, 256))
    tic = time.perf_counter()
    steps = 0
    for sd, w in enumerate(ws[:10]):
        eng.reset()
        eng.set_seed(sd)
        for s_ in range(3):
            eng.set_tet_counts(s_, w.counts[s_])
        for k in range(1, 101):
            eng.run(k * 1e-3)
            eng.sum_all_tet_counts()
        steps += eng.nsteps() * w.n_tets
    plain = steps / (time.perf_counter() - tic)
    beng = Engine(ws[0].fm, [w.mt for w in ws], list(range(n_seeds)), streams=np.zeros(n_seeds))
    for r, w in enumerate(ws):
        for s_ in range(3):
            beng.set_tet_counts(s_, w.counts[s_], real=r)
    tuned_batch = autotune(beng, trial_periods=10, blocks=(64, 128, 256))
    tic = time.perf_counter()
    for k in range(1, 101):
        beng.run(k * 1e-3)
        beng.sum_all_tet_counts()
    wall = time.perf_counter() - tic
    steps = sum(beng.nsteps(r) * w.n_tets for r, w in enumerate(ws))
    shape = lambda t: {"block": t["block"], "graph": t["graph"], "specialized_kernel": t["specialized"],
                       "lanes_per_tet": t["lanes"], "slot_width": t["slot_width"]}
    return dict(plain_one_seed_at_a_time_tet_steps_per_s=plain, plain_seeds=10,
                batched_tet_steps_per_s=steps / wall, batched_seeds=n_seeds, batched_wall_s=wall,
                plain_schedule=shape(tuned_plain), batched_schedule=shape(tuned_batch))


def cpu_baseline(w, budget_s=15.0, seed=1, membrane=None, n_seeds0=100):
    """The CPU oracle on this host: a bounded number of periods of the headline
    workload (the reported value), and bounded samples of the 3,072-tet
    protocol, the membrane cube and a loop over ensemble realizations."""
    from oracle import bind

    threads = bind.threads()
    value, n, wall = oracle_periods(w, budget_s, seed)
    out = dict(value=value, unit="tet-steps/s", cores=threads, kind="port", cpu_model=cpu_model(),
               sample=f"{n} periods x {w.n_tets} tets of the same workload on the CPU oracle "
                      f"(OpenMP, {threads} threads), {wall:.1f} s")
    out["config0_protocol"] = config0_protocol_oracle(n_seeds0, budget_s / 2)
    if membrane is not None:
        v, n, wall = oracle_periods(membrane, budget_s / 3)
        out["membrane_cube"] = dict(tet_steps_per_s=v, periods=n, wall_s=wall)
    fm, meshes, counts, reac_k, info = W.ab_binding_ensemble(n_real=16, seed=0)
    tic, steps = time.perf_counter(), 0
    for r in range(16):
        o = bind.OracleSim(fm, meshes[r], seed=1, stream=r, reac_k=reac_k[r])
        o.counts[:] = counts[r]
        o.run(0.1, 0)
        steps += o.nsteps * len(meshes[r].vol)
    out["ensemble_loop"] = dict(tet_steps_per_s=steps / (time.perf_counter() - tic), realizations=16)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=338)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--block", type=int, default=0,
                    help="threads per block of the period kernel; 0 = chosen by the autotuner")
    ap.add_argument("--graph", action="store_true", help="with --block: replay launches as a graph")
    ap.add_argument("--specialize", action="store_true",
                    help="with --block: use the model-specialised kernel (compiled at run time)")
    ap.add_argument("--slot-width", type=int, default=2, choices=(1, 2),
                    help="with --block: face slot width in bytes")
    ap.add_argument("--spec-waves", type=int, default=0,
                    help="with --specialize: wavefronts per SIMD the compiled kernel must allow")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-membrane", action="store_true",
                    help="skip the 6,000,000-tet membrane cube and its EField")
    ap.add_argument("--no-config0", action="store_true",
                    help="skip the 3,072-tet run-every-1-ms protocol")
    ap.add_argument("--part", choices=("all", "plain", "batched", "membrane", "efield"), default="all",
                    help="run one section only (counter profiling); with --block the section runs "
                         "that fixed schedule instead of the autotuned one")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    global FIXED_SLOT_WIDTH
    FIXED_SLOT_WIDTH = args.slot_width
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the whole mesh is resident on a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    if args.part != "all":
        out = {"part": args.part}
        if args.part == "plain":
            w = W.four_species_cube(n=40, side=40e-6, seed=0, device=True)
            p = time_plain(w, args.steps, args.warmup, block=args.block, graph=args.graph,
                           specialize=args.specialize, spec_waves=args.spec_waves)
            out.update(tet_steps_per_s=p["tet_steps"] / p["wall"], periods=p["periods"],
                       kernel_ms_per_launch=p["ms_launch"], launches=p["launches"])
        elif args.part == "batched":
            ens = time_ensemble(block=args.block, graph=args.graph if args.block else True,
                                specialize=args.specialize)
            out.update(tet_steps_per_s=ens["tet_steps"] / ens["wall"], tet_steps=ens["tet_steps"],
                       kernel_ms_per_launch=ens["ms_launch"], launches=ens["launches"],
                       max_periods=ens["max_periods"], conserved=ens["conserved"])
        elif args.part == "membrane":
            wm = W.calcium_membrane_cube(n=100, side=20e-6, seed=0, device=True)
            out.update(time_membrane(wm, args.steps, args.warmup, block=args.block,
                                     specialize=args.specialize))
        else:
            wm = W.calcium_membrane_cube(n=100, side=20e-6, seed=0, device=True)
            out.update(time_efield(wm))
        print(json.dumps(out))
        return

    # mesh topology built on the device (stg_mesh_build), as a large-mesh user would
    w = W.four_species_cube(n=40, side=40e-6, seed=0, device=True)
    btet = model_bytes(w)
    p = time_plain(w, args.steps, args.warmup, block=args.block, graph=args.graph, specialize=args.specialize,
                   spec_waves=args.spec_waves)
    value = p["tet_steps"] / p["wall"]
    # roofline of the period kernel: algorithmic bytes per launch / mean launch time
    achieved = btet * w.n_tets / (p["ms_launch"] * 1e-3) / 1e9
    traffic = measured_traffic()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tet-steps/s",
        "n_gpus": 1,
        "steps": p["periods"],
        "warmup": args.warmup,
        "ms_per_step": p["wall"] * 1e3 / max(p["periods"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 counts, f64 propensities",
        "data": "synthetic",
        "config": {
            "workload": "four_species_cube: 40^3 x 6 Kuhn cube, side 40 um, 384,000 tets, "
                        "4 species / 6 reactions, Poisson(50) initial, dt at the stability bound",
            "n_tets": w.n_tets,
            "period_s": p["dt"],
            "simulated_s": p["periods"] * p["dt"],
            "block": p["block"],
            "graph": bool(p["graph"]),
            "specialized_kernel": bool(p["specialized"]),
            "lanes_per_tet": p["lanes"],
            "slot_width": p["slot_width"],
            "spec_min_waves": p["tuned"].get("spec_min_waves") if p["tuned"] else (
                args.spec_waves if args.specialize else None),
            "schedule": "autotuned" if p["tuned"] else "fixed",
            "autotune_wall_ms_per_launch": p["tuned"]["timings"] if p["tuned"] else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
            "traffic_frac": (traffic / (p["ms_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "algorithmic_bytes_per_launch": btet * w.n_tets,
            "bytes_per_tet_step": btet,
            "kernel_ms_per_launch": p["ms_launch"],
            "kernel_timing": "HIP events around the full-period launches of the timed run (the "
                             "trailing entering-only pass is a separate kernel, not included)",
        },
    }
    if not args.no_batched:
        small = W.ab_binding_cube(n=13, side=10e-6, seed=0)
        ps = time_plain(small, 9, 1)
        ens = time_ensemble()
        sched = ens["tuned"] or {}
        ens_traffic = measured_traffic(ENSEMBLE_TRAFFIC, "traffic_bytes_per_tet_step")
        kern_s = ens["ms_launch"] * ens["launches"] * 1e-3
        out["batched"] = {
            "workload": "ab_binding_ensemble: 1,024 realizations, Kuhn cubes n=6..20 "
                        "(1,296..48,000 tets), kf/kb log-uniform +-1 decade, 0.1 s each",
            "tet_steps_per_s": ens["tet_steps"] / ens["wall"],
            "n_tets": ens["n_tets"],
            "kernel_ms_per_launch": ens["ms_launch"],
            "launches": ens["launches"],
            "achieved_gbs": ens["bytes_per_tet_step"] * ens["tet_steps"] / kern_s / 1e9,
            "frac": ens["bytes_per_tet_step"] * ens["tet_steps"] / kern_s / 1e9 / HBM_PEAK_GBS,
            "traffic_bytes_per_tet_step": ens_traffic,
            "traffic_frac": (ens_traffic * ens["tet_steps"] / kern_s / 1e9 / HBM_PEAK_GBS) if ens_traffic else None,
            "conserved": ens["conserved"],
            "plain_same_model_tet_steps_per_s": ps["tet_steps"] / ps["wall"],
            "plain_same_model": "ab_binding_cube n=13 (13,182 tets), one realization",
            "schedule": {"block": sched.get("block"), "graph": sched.get("graph"),
                         "specialized_kernel": sched.get("specialized"),
                         "lanes_per_tet": sched.get("lanes"),
                         "slot_width": sched.get("slot_width"),
                         "spec_min_waves": sched.get("spec_min_waves"),
                         "autotune_wall_ms_per_launch": sched.get("timings")},
            "plain_same_model_schedule": {"block": ps["block"], "graph": bool(ps["graph"]),
                                          "specialized_kernel": bool(ps["specialized"]),
                                          "lanes_per_tet": ps["lanes"],
                                          "slot_width": ps["slot_width"]},
        }
        out["batched"]["batched_over_plain"] = (out["batched"]["tet_steps_per_s"] /
                                                out["batched"]["plain_same_model_tet_steps_per_s"])
        out["batched"]["batched_over_plain_384k_cube"] = out["batched"]["tet_steps_per_s"] / value
        print(f"bench: batched {out['batched']['tet_steps_per_s']:.3e} tet-steps/s", file=sys.stderr, flush=True)
    wm = None
    if not args.no_membrane:
        wm = W.calcium_membrane_cube(n=100, side=20e-6, seed=0, device=True)
        out["membrane"] = time_membrane(wm)
        print(f"bench: membrane {out['membrane']['tet_steps_per_s']:.3e} tet-steps/s", file=sys.stderr, flush=True)
        out["efield"] = time_efield(wm)
        print(f"bench: efield {out['efield']['ms_per_ef_step']:.2f} ms per step", file=sys.stderr, flush=True)
    if not args.no_config0:
        out["config0_protocol"] = config0_protocol_gpu()
        out["config0_protocol"]["workload"] = ("ab_binding_cube: 8^3 x 6 Kuhn cube, side 10 um, 3,072 tets, "
                                               "10,000 A + 10,000 B by seed, run() every 1 ms to 0.1 s, "
                                               "totals read at every save")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_budget, membrane=wm)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
