"""Benchmark of the ground-state solve path on one MI355X.

Workload (default): 2D rotating vortex lattice -- s=2, Omega=0.9, beta2=200,
G_C=0, L=12, Ngrid=384 (3*2^7), complex128, res_tol=1e-6, nseeds=3 with the
reference's seed cycle (triangular, tf, gaussian; rng 0,1,2).  One step is one
multi-seed ground-state solve: the three solves from the pre-generated seeds
(inputs resident in HBM before timing) run as one batch through the batched
device loop, then the final residual/observables of each and the
lowest-energy selection.  res_tol=1e-6 is the tolerance at which the
reference's own lattice runs reach |dE/E| ~ 1e-9 (its pcg_variants results);
the line reports the selected energy's distance to the same state polished
to res_tol=1e-9.  linesearch_max=20 (the reference's default is 8): with 8
Armijo halvings the solver stops as "line_search_stalled" near res 2e-4 on
this workload, on the GPU and in the CPU oracle alike, when a stiff state
needs a step below alpha/256.

Prints one JSON line: value = selected ground states per second; also solves
and CG iterations per second, the same workload through the single-solve path
(one solve after the other, measured once after the timed region), the
roofline of the dominant kernel (achieved algorithmic GB/s from HIP events
around each of its launches on the launch stream, in a direct-launch replay of
the same batch after the timed region -- the timed region replays hipGraphs,
which cannot carry timing events -- with bytes counted per instance each
launch covered) and the CPU oracle timed on a bounded sample of the same
workload.
"""
import argparse
import ctypes
import json
import os
import platform
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "rotating_lattice_2d_n384": dict(
        cfg=dict(mode="single", dimension="2D", s=2, Omega=0.9, beta2=200.0, beta3=0.0, G_C=0.0,
                 kernel="none", L=12.0, Ngrid=384, nseeds=3, res_tol=1e-6, maxit=60000, linesearch_max=20),
        desc="2D rotating vortex lattice, Omega=0.9, beta2=200, L=12, N=384, nseeds=3, res_tol=1e-6, "
             "linesearch_max=20",
        check="grad_fd"),
    "harmonic_2d_n128": dict(
        cfg=dict(mode="single", dimension="2D", s=2, Omega=0.0, beta2=100.0, beta3=0.0, G_C=0.0,
                 kernel="none", L=16.0, Ngrid=128, nseeds=1, res_tol=1e-4, maxit=60000),
        desc="2D harmonic plain solve, beta2=100, L=16, N=128"),
    # 3D self-gravitating boson star: 1/r kernel, N=128 (convolution on 256^3), one solve
    "newton_3d_n128": dict(
        cfg=dict(mode="single", dimension="3D", s=2, trap_coeff=0.5, Omega=0.0, beta2=0.0, beta3=0.0, G_C=20.0,
                 kernel="newton", L=12.0, Ngrid=128, nseeds=1, seed="gaussian", seed_noise=1e-3,
                 res_tol=1e-6, maxit=60000, linesearch_max=20),
        desc="3D Newton boson star, G_C=20, beta2=0, L=12, N=128 (conv 256^3), Gaussian seed with 1e-3 "
             "noise, res_tol=1e-6, linesearch_max=20",
        check=("hf", "G_C")),
    # quasi-2D pancakes: K_eff(l_z=0.5); Omega x G_C = 8 points x 4 seeds = 32 solves on three grids,
    # instance i on N = (128, 192, 256)[i mod 3]
    "q2d_batch_uneven": dict(
        cfg=dict(mode="single", dimension="quasi2D", s=2, trap_coeff=0.5, Omega=0.0, beta2=50.0, beta3=0.0,
                 G_C=1.0, kernel="q2d", l_z=0.5, L=16.0, Ngrid=128, nseeds=4, res_tol=1e-6, maxit=60000,
                 linesearch_max=20),
        points=[dict(Omega=om, G_C=gc) for om in (0.0, 0.5) for gc in (1.0, 2.0, 5.0, 10.0)],
        instance_grid=lambda i: dict(Ngrid=(128, 192, 256)[i % 3]),
        desc="quasi-2D q2d kernel l_z=0.5, beta2=50, L=16, Omega in {0,0.5} x G_C in {1,2,5,10}, nseeds=4 "
             "(32 solves), N = 128/192/256 round-robin by instance id, one batch per grid (the three run "
             "concurrently), res_tol=1e-6, "
             "linesearch_max=20",
        replay_iters=250),
    # self-bound cubic-quintic droplets: no trap, 16 attractive beta2 x 4 seeds = 64 solves
    "droplet_sweep_2d_n256": dict(
        cfg=dict(mode="single", dimension="2D", s=2, trap_coeff=0.0, Omega=0.0, beta2=-120.0, beta3=1000.0,
                 G_C=0.0, kernel="none", L=20.0, Ngrid=256, nseeds=4, res_tol=1e-6, maxit=60000,
                 linesearch_max=20, warm_fp32=True, warm_res_tol=1e-4),
        points=[dict(beta2=float(b)) for b in
                [-200.0 + 160.0 * i / 15.0 for i in range(16)]],
        desc="2D cubic-quintic droplet sweep, no trap, beta2 = 16 values in [-200, -40], beta3=1e3, "
             "L=20, N=256, nseeds=4 (64 solves), FP32 warm phase to res 1e-4 then FP64 to res_tol=1e-6, "
             "linesearch_max=20",
        check=("hf", "beta2"),
        replay_iters=100),
}

# watched kernel: launch-site id -> (name, algorithmic bytes per instance of its
# launches in the solver loop).  The launches themselves report their exact
# algorithmic bytes (each element read or written once), which is what
# "achieved" divides by the time; these per-instance figures are the
# documented formulas for the loop's launches.
KERNELS = {
    114: ("k_kin<A2X,KH_LAST> (accepted state: psi update, x pass of H psi, local terms, mu psi, residual rows)",
          lambda n: 128 * n),   # reads psi, d, H share, old g; writes psi, s = change, g, y = change
    144: ("k_kin<A3X,KH_LAST> (accepted state: psi update, x pass of H psi, local terms, mu psi, residual rows)",
          lambda n: 144 * n),   # as 114, plus phi read and stored (long range)
    100: ("k_kin<A2Y,KE_FIRST> (trial energy: new direction, y pass, Parseval and local sums, H share)",
          lambda n: 80 * n),    # reads psi, d, z; writes d, H share (long range: +8n for |psi|^2)
    111: ("k_kin<A2X,KE_REST> (trial energy, x pass)", lambda n: 32 * n),
    2023: ("k_pass<strided,fwd*(k^2/2+sigma)^-1*inv> (Sobolev preconditioner, x pass)", lambda n: 32 * n),
    3000: ("k_pass<contig,fwd> (preconditioner, y pass forward)", lambda n: 32 * n),
    3015: ("k_pass<contig,inv,dots> (preconditioner y pass inverse + 16 CG inner products)",
           lambda n: 112 * n),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2,
                    help="timed steps (0: no timed region, only the event-timed replay -- for profiling)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rotating_lattice_2d_n384", choices=sorted(WORKLOADS))
    ap.add_argument("--watch", type=int, default=None,
                    help="launch-site id of the roofline kernel (default: the accepted-state H pass, "
                         "114 in 2D, 144 in 3D)")
    ap.add_argument("--measure-iters", type=int, default=None,
                    help="CG iterations of the event-timed replay of the roofline kernel (default: the "
                         "workload's replay_iters, else 300)")
    ap.add_argument("--cpu-sample-iters", type=int, default=60)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--maxit", type=int, default=None,
                    help="cap the CG iterations per solve (bounded profiling runs only)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-solve comparison run")
    ap.add_argument("--no-polish", action="store_true",
                    help="skip the accuracy check of the selected state (bounded profiling runs only)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner choose each plan's FFT stage orders and lines per workgroup before "
                         "timing (default: the fixed rules)")
    ap.add_argument("--no-autotune", action="store_true", help="the default: the fixed schedule rules")
    ap.add_argument("--fp64-only", action="store_true",
                    help="run a workload's FP32 warm phase in FP64 too (comparison)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device-to-device copy (counter calibration)")
    return ap.parse_args()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps(dict(error="this build runs on one GPU: the ground-state path has no multi-GPU "
                                    "code; run with --gpus 1")))
        sys.exit(2)
    import torch
    import becgpp_amd as B
    from becgpp_amd import _lib
    from becgpp_amd.seeds import make_seed

    wl = WORKLOADS[a.workload]
    cfg = B.default_cfg(**wl["cfg"])
    if a.fp64_only:
        cfg["warm_fp32"] = False
    if a.maxit is not None:
        cfg["maxit"] = int(a.maxit)
    if a.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    nseeds = int(cfg["nseeds"])
    if a.watch is None:
        a.watch = 144 if str(cfg["dimension"]).lower() == "3d" else 114
    points = wl.get("points", [{}])
    igrid = wl.get("instance_grid", lambda i: {})
    grids = {}                                  # grid key -> G
    runs = []                                   # (point index, cfg, psi0, grid key)
    for ip, over in enumerate(points):
        cp = dict(cfg, **over)
        rotating = abs(float(cp["Omega"])) >= 0.3
        cycle = ("triangular", "tf", "gaussian") if rotating else ("tf", "gaussian")
        for j in range(nseeds):
            c = dict(cp, **igrid(len(runs)))
            c.update(rng=int(cfg["rng"]) + j, seed=cycle[j % len(cycle)] if nseeds > 1 else cfg["seed"])
            gk = (int(c["Ngrid"]), float(c["L"]))
            if gk not in grids:
                grids[gk] = B.make_grid(c)
            runs.append((ip, c, make_seed(grids[gk], c).contiguous(), gk))
    torch.cuda.synchronize()
    cs = [c for _, c, _, _ in runs]
    psi0s = [p for _, _, p, _ in runs]
    nsol = len(runs)
    gk0 = runs[0][3]
    G = grids[gk0]
    n = int(G["N"]) ** int(G["ndim"])
    groups = {}
    for k, r in enumerate(runs):
        groups.setdefault(r[3], []).append(k)

    def select(res):
        """Per point: the lowest-energy seed (solvers.py:238-263)."""
        best = {}
        for (ip, _, _, _), (psi, _, obs) in zip(runs, res):
            if ip not in best or obs["E"] < best[ip][1]["E"]:
                best[ip] = (psi, obs)
        return best

    def solve_all():
        """Every instance through the batched solver: one batch set per grid,
        the sets of different grids running concurrently."""
        return B.ground_state_uneven(cs, psi0s=psi0s, grids=[grids[r[3]] for r in runs])

    def step():
        res = solve_all()
        best = select(res)
        Es = [obs["E"] for _, _, obs in res]
        stops = [f'{obs["stop_reason"]}@{obs["iters"]}' for _, _, obs in res]
        warm_iters[0] += sum(obs.get("warm_iters", 0) for _, _, obs in res)
        return best, sum(obs["iters"] for _, _, obs in res), Es, stops

    def step_single():
        res = [B.ground_state(c, G=grids[r[3]], psi0=p0, verbose=False) for c, p0, r in zip(cs, psi0s, runs)]
        return select(res), sum(obs["iters"] for _, _, obs in res), [obs["E"] for _, _, obs in res]

    lib = _lib.load()
    # schedule of each grid's plan: chosen by the autotuner before any timing
    schedules = {}
    if a.autotune and not a.no_autotune:
        from becgpp_amd.grid import plan_of
        from becgpp_amd.interactions import ensure_kernel
        for gk, ks in groups.items():
            Gg = grids[gk]
            lr = Gg["kernel"] != "none" and any(abs(float(cs[k]["G_C"])) > 1e-15 for k in ks)
            pl = ensure_kernel(Gg, "K") if lr else plan_of(Gg)
            schedules["N=%d" % gk[0]] = pl.autotune(min(len(ks), 64))
        torch.cuda.synchronize()
    warm_iters = [0]
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    warm_iters[0] = 0
    t0 = time.perf_counter()
    tot_iters = 0
    results = []
    for _ in range(a.steps):
        best, iters, Es, stops = step()
        tot_iters += iters
        results.append((best, Es, stops))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0

    # The timed steps replay hipGraphs, whose launches cannot carry timing
    # events; the roofline kernel is timed with HIP events on its stream in a
    # replay of the same batched workload with direct launches.
    # the replay runs the workload's solves from their seeds for fewer CG
    # iterations than they take to converge (droplets ~130, q2d ~280 per
    # solve), so the timed launches are those of solves in progress rather than
    # of instances retrying line searches at the FP64 floor past res_tol
    if a.measure_iters is None:
        a.measure_iters = int(wl.get("replay_iters", 300))
    # with several grid sizes, the replay takes the grid carrying the most
    # work (instances x points), where most of the step time goes
    gkr = max(groups, key=lambda g: len(groups[g]) * g[0] ** int(G["ndim"]))
    Gr = grids[gkr]
    nr = int(Gr["N"]) ** int(Gr["ndim"])
    lib.bec_prof_watch(int(a.watch))
    k0 = groups[gkr][:64]
    B.ground_state_batch([dict(cs[k], maxit=int(a.measure_iters), res_tol=0.0, warm_fp32=False) for k in k0], G=Gr,
                         psi0s=[psi0s[k] for k in k0])
    torch.cuda.synchronize()
    ms = ctypes.c_double()
    cnt = ctypes.c_longlong()
    nbytes = ctypes.c_double()
    lib.bec_prof_read(ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(nbytes))
    lib.bec_prof_watch(0)
    kname, kbytes = KERNELS.get(a.watch, (f"site {a.watch}", lambda n: None))
    avg_s = (ms.value / 1e3) / max(1, cnt.value)
    bpl = nbytes.value / max(1, cnt.value)             # algorithmic bytes per timed launch
    achieved = bpl / avg_s / 1e9 if cnt.value else None
    per_inst = kbytes(nr)
    roof = dict(bound="hbm", kernel=kname, site=a.watch, launches=cnt.value,
                timed_in=f"direct-launch replay of the batch on N = {gkr[0]}, {a.measure_iters} CG iterations",
                avg_launch_us=avg_s * 1e6, achieved=achieved,
                peak=PEAK_HBM_GBS, unit="GB/s", frac=(achieved / PEAK_HBM_GBS) if achieved else None,
                traffic=None, algorithmic_bytes_per_launch=bpl,
                algorithmic_bytes_per_instance=per_inst,
                instances_per_launch=(bpl / per_inst) if per_inst else None)
    # HBM bytes per launch of the roofline kernel, from the committed counter
    # summary of this workload's replay (rocprofv3 FETCH_SIZE / WRITE_SIZE,
    # calibrated on a device copy)
    tfile = os.path.join(HERE, "profiles", f"{a.workload}_hbm_traffic.md")
    if a.watch in (114, 144) and os.path.exists(tfile):
        m = re.search(r"traffic_bytes_per_dispatch: ([0-9.]+)", open(tfile).read())
        if m:
            roof["traffic"] = float(m.group(1))
            roof["traffic_source"] = os.path.relpath(tfile, HERE)
            roof["traffic_over_algorithmic"] = roof["traffic"] / bpl if bpl else None
    if a.steps == 0:
        # profiling runs: no timed region, so that a kernel trace of this
        # command holds exactly the launches the events timed
        print(json.dumps(dict(metric="roofline kernel only (--steps 0)", value=None, roofline=roof,
                              config=dict(workload=a.workload))), flush=True)
        return
    best, Es, stops = results[-1]
    npt = len(points)
    b0psi, b0 = best[0]
    out = dict(
        metric="ground states/sec at |dE/E|<=1e-8; CG iters/sec; HBM GB/s vs peak",
        value=a.steps * npt / wall, unit="ground states/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
        ms_per_step=1e3 * wall / a.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic",
        config=dict(workload=a.workload, desc=wl["desc"], Ngrid=sorted({gk[0] for gk in grids}) if len(grids) > 1
                    else int(G["N"]), ndim=int(G["ndim"]),
                    points=npt, nseeds=nseeds, solves_per_step=nsol, res_tol=float(cfg["res_tol"]),
                    path="batched"),
        solves_per_s=a.steps * nsol / wall, cg_iters_per_s=tot_iters / wall,
        iters_per_solve=tot_iters / (a.steps * nsol),
        E_selected=[best[i][1]["E"] for i in range(npt)] if npt > 1 else b0["E"],
        seed_energies=Es if npt == 1 else None, stop_reasons=stops if npt == 1 else None,
        converged=all(bool(best[i][1]["converged"]) for i in range(npt)),
        resid_rel=max(best[i][1]["resid_rel"] for i in range(npt)),
        roofline=roof,
    )
    if schedules:
        out["autotuned_schedule"] = schedules
    if cfg.get("warm_fp32", False):
        out["warm_phase"] = dict(dtype="f32", res_tol=float(cfg["warm_res_tol"]),
                                 iters_share=warm_iters[0] / max(1, tot_iters),
                                 note="complex64 fields and transforms, FP64 sums and Armijo test; the FP64 loop "
                                      "then runs from each warm state to res_tol")
    # |dE/E| of the first point's selected state: polish it to the FP64 residual floor
    if not a.no_polish:
        ib = min((k for k in range(nsol) if runs[k][0] == 0), key=lambda k: Es[k])
        pc = dict(cs[ib], res_tol=1e-9, maxit=20000, warm_fp32=False)
        _, _, pol = B.ground_state(pc, G=grids[runs[ib][3]], psi0=b0psi, verbose=False)
        out["E_polished"] = pol["E"]
        out["dE_rel_vs_polished"] = abs(b0["E"] - pol["E"]) / abs(pol["E"])
    # the configuration's own derivative check, at its full size (after timing)
    chk = wl.get("check")
    if chk == "grad_fd":
        # autograd backward (2 H psi dV) against a central difference of E on a
        # random normalised psi, along a random direction
        gen = torch.Generator(device="cpu").manual_seed(7)
        shp = tuple(b0psi.shape)
        rnd = lambda: torch.complex(torch.randn(shp, dtype=torch.float64, generator=gen),
                                    torch.randn(shp, dtype=torch.float64, generator=gen)).cuda()
        q = rnd()
        q = q / (q.abs().pow(2).sum() * float(G["dx"]) ** int(G["ndim"])).sqrt()
        v = rnd()
        par = [float(cfg[k]) for k in ("beta2", "beta3", "Omega", "G_C")]
        qq = q.clone().requires_grad_(True)
        B.energy_t(qq, G, *par).backward()
        an = float(torch.real(torch.vdot(qq.grad.reshape(-1), v.reshape(-1))))
        eps = 1e-5
        fd = (B.energy_components(q + eps * v, G, *par)["E"].item()
              - B.energy_components(q - eps * v, G, *par)["E"].item()) / (2 * eps)
        out["check"] = dict(kind="autograd backward vs central difference of E on a random normalised psi",
                            autograd=an, finite_difference=fd, rel_diff=abs(an - fd) / max(abs(fd), 1e-300))
    elif chk and chk[0] == "hf":
        # Hellmann-Feynman: the autograd derivative of E in the parameter at the
        # converged state against a central difference of converged energies
        pname = chk[1]
        ib = min((k for k in range(nsol) if runs[k][0] == 0), key=lambda k: Es[k])
        c0 = dict(cs[ib], res_tol=1e-9, maxit=20000, warm_fp32=False)
        Gb = grids[runs[ib][3]]
        par = {k: float(c0[k]) for k in ("beta2", "beta3", "Omega", "G_C")}
        t = torch.tensor(par[pname], dtype=torch.float64, requires_grad=True)
        args = [t if k == pname else par[k] for k in ("beta2", "beta3", "Omega", "G_C")]
        psi_c, _, _ = B.ground_state(c0, G=Gb, psi0=b0psi, verbose=False)
        B.energy_t(psi_c, Gb, *args).backward()
        h = 1e-3 * max(1.0, abs(par[pname]))
        Ep = B.ground_state(dict(c0, **{pname: par[pname] + h}), G=Gb, psi0=psi_c, verbose=False)[2]["E"]
        Em = B.ground_state(dict(c0, **{pname: par[pname] - h}), G=Gb, psi0=psi_c, verbose=False)[2]["E"]
        fd = (Ep - Em) / (2 * h)
        out["check"] = dict(kind=f"Hellmann-Feynman dE/d{pname}: autograd at the converged state vs central "
                                 f"difference of converged energies (h = {h:g}, res_tol 1e-9)",
                            autograd=t.grad.item(), finite_difference=fd,
                            rel_diff=abs(t.grad.item() - fd) / max(abs(fd), 1e-300))
    if not a.no_single:
        t0 = time.perf_counter()
        sbest, si, sE = step_single()
        torch.cuda.synchronize()
        ts = time.perf_counter() - t0
        out["single_solve_path"] = dict(ground_states_per_s=npt / ts, solves_per_s=nsol / ts,
                                        cg_iters_per_s=si / ts, iters_per_solve=si / nsol,
                                        E_selected=sbest[0][1]["E"], seed_energies=sE if npt == 1 else None)
        out["batched_speedup"] = out["value"] / out["single_solve_path"]["ground_states_per_s"]
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cs[0], psi0s[0].cpu().numpy(), a.cpu_sample_iters,
                                           out["iters_per_solve"], nsol / npt)
    print(json.dumps(out), flush=True)


def cpu_baseline(cfg, psi0, sample_iters, iters_per_solve, solves_per_state):
    """The oracle (numpy/scipy, all host threads) on the first seed of the same
    workload for a bounded number of CG iterations; scaled to ground states/s
    with the GPU's measured iterations per solve."""
    import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    c = dict(cfg, maxit=int(sample_iters), res_tol=0.0)
    Go = O.make_grid(c)
    O.ground_state(dict(c, maxit=3), Go, psi0)          # warm the FFT plans
    t0 = time.perf_counter()
    _, _, obs = O.ground_state(c, Go, psi0)
    dt = time.perf_counter() - t0
    its = obs["iters"] / dt
    return dict(value=its / (iters_per_solve * solves_per_state), unit="ground states/s", cores=threads,
                kind="port", cg_iters_per_s=its, cpu=cpu_model(),
                sample=f"{obs['iters']} CG iterations of the first solve of the same workload ({dt:.1f} s), "
                       f"scaled by the GPU's {iters_per_solve:.0f} iterations/solve x "
                       f"{solves_per_state:g} solves per ground state")


if __name__ == "__main__":
    main()
