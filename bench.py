"""Benchmark of the ground-state solve path on one MI355X.

Default workload (the line's top-level fields): the 3D Newton boson star --
s=2, Omega=0, beta2=0, G_C=20 with the 1/r kernel, L=12, Ngrid=128 (padded
convolution on 256^3), complex128, Gaussian seed with 1e-3 noise, one solve to
res_tol=1e-6.  The same JSON line carries, under "workloads", the other
batched configurations at their full sizes, each with the same fields:

* droplet_sweep_2d_n256 -- self-bound cubic-quintic droplets, no trap,
  beta2 = 16 values in [-200, -40], beta3 = 1e3, L = 20, N = 256, nseeds = 4
  (64 solves), FP32 warm phase to 1e-4 then FP64 to res_tol = 1e-8;
* q2d_batch_uneven -- quasi-2D q2d kernel l_z = 0.5, beta2 = 50, L = 16,
  Omega in {0, 0.5} x G_C in {1, 2, 5, 10}, nseeds = 4 (32 solves) on the
  grids N = 128, 192, 256 (instance i on (128, 192, 256)[i mod 3]);
* rotating_lattice_2d_n384 -- Omega = 0.9, beta2 = 200, L = 12, N = 384,
  nseeds = 3 with the reference's seed cycle.

A step is one pass of the workload: every solve from its pre-generated seed
(inputs resident in HBM before timing), batched on the device, then the
lowest-energy selection per point.  value = selected ground states per
second.  Each workload also reports the same solves through the single-solve
path (one after the other), its distance to a polished state (and to the
reference's published energy where one exists), its derivative check, the
roofline of its dominant kernel and the CPU oracle timed on a bounded sample.

Roofline: after the timed region the workload's largest grid group is replayed
with direct launches (the timed region replays hipGraphs, which cannot carry
timing events) and every launch of every kernel is timed with HIP events on
its stream, with the algorithmic bytes each launch moved (each element it must
read or write, counted once per instance it covered).  The kernel with the
largest share of the replay's time is the roofline kernel: achieved = its
bytes / its time.  step_roofline divides the algorithmic bytes of the whole
timed region (the per-iteration formula below, over every instance-iteration)
by its wall time.
"""
import argparse
import ctypes
import json
import math
import os
import platform
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
DEFAULT = "newton_3d_n128"
SECONDARY = ("droplet_sweep_2d_n256", "q2d_batch_uneven", "rotating_lattice_2d_n384")
SECONDARY_STEPS, SECONDARY_WARMUP = 2, 1

WORKLOADS = {
    "rotating_lattice_2d_n384": dict(
        cfg=dict(mode="single", dimension="2D", s=2, Omega=0.9, beta2=200.0, beta3=0.0, G_C=0.0,
                 kernel="none", L=12.0, Ngrid=384, nseeds=3, res_tol=1e-6, maxit=60000, linesearch_max=20),
        desc="2D rotating vortex lattice, Omega=0.9, beta2=200, L=12, N=384, nseeds=3 (triangular, tf, gaussian), "
             "res_tol=1e-6, linesearch_max=20",
        check="grad_fd",
        # the reference's own run of this grid (pr_precond, res 1e-6)
        published=(3.3642038171647553, "benchmarks/results/verification/pcg_variants.csv:6"),
        replay_iters=300),
    "harmonic_2d_n128": dict(
        cfg=dict(mode="single", dimension="2D", s=2, Omega=0.0, beta2=100.0, beta3=0.0, G_C=0.0,
                 kernel="none", L=16.0, Ngrid=128, nseeds=1, res_tol=1e-4, maxit=60000),
        desc="2D harmonic plain solve, beta2=100, L=16, N=128, res_tol=1e-4"),
    # 3D self-gravitating boson star: 1/r kernel, N=128 (convolution on 256^3), one solve
    "newton_3d_n128": dict(
        cfg=dict(mode="single", dimension="3D", s=2, trap_coeff=0.5, Omega=0.0, beta2=0.0, beta3=0.0, G_C=20.0,
                 kernel="newton", L=12.0, Ngrid=128, nseeds=1, seed="gaussian", seed_noise=1e-3,
                 res_tol=1e-6, maxit=60000),
        desc="3D Newton boson star, G_C=20, beta2=0, L=12, N=128 (conv 256^3), Gaussian seed with 1e-3 "
             "noise, res_tol=1e-6",
        check=("hf", "G_C"),
        mixed_variant=True,
        replay_iters=200),
    # quasi-2D pancakes: K_eff(l_z=0.5); Omega x G_C = 8 points x 4 seeds = 32 solves on three grids
    "q2d_batch_uneven": dict(
        cfg=dict(mode="single", dimension="quasi2D", s=2, trap_coeff=0.5, Omega=0.0, beta2=50.0, beta3=0.0,
                 G_C=1.0, kernel="q2d", l_z=0.5, L=16.0, Ngrid=128, nseeds=4, res_tol=1e-6, maxit=60000),
        points=[dict(Omega=om, G_C=gc) for om in (0.0, 0.5) for gc in (1.0, 2.0, 5.0, 10.0)],
        instance_grid=lambda i: dict(Ngrid=(128, 192, 256)[i % 3]),
        desc="quasi-2D q2d kernel l_z=0.5, beta2=50, L=16, Omega in {0,0.5} x G_C in {1,2,5,10}, nseeds=4 "
             "(32 solves), N = 128/192/256 round-robin by instance id; the instances of one grid form one "
             "batched loop and one hipGraph advances the three grids' loops together (a branch per grid); "
             "res_tol=1e-6",
        mixed_variant=True,
        replay_iters=200),
    # self-bound cubic-quintic droplets: no trap, 16 attractive beta2 x 4 seeds = 64 solves
    "droplet_sweep_2d_n256": dict(
        cfg=dict(mode="single", dimension="2D", s=2, trap_coeff=0.0, Omega=0.0, beta2=-120.0, beta3=1000.0,
                 G_C=0.0, kernel="none", L=20.0, Ngrid=256, nseeds=4, res_tol=1e-8, maxit=60000,
                 warm_fp32=True, warm_res_tol=1e-4),
        points=[dict(beta2=float(b)) for b in
                [-200.0 + 160.0 * i / 15.0 for i in range(16)]],
        desc="2D cubic-quintic droplet sweep, no trap, beta2 = 16 values in [-200, -40], beta3=1e3, "
             "L=20, N=256, nseeds=4 (64 solves), FP32 warm phase to res 1e-4 then FP64 to res_tol=1e-8",
        check=("hf", "beta2"),
        replay_iters=100),
}

# launch-site ids (include/becgpp_hip.h, bec_prof_watch) -> kernel names
_AXES = {0: "A2Y", 1: "A2X", 2: "A3Z", 3: "A3Y", 4: "A3X"}
_KINDS = {0: "KE_FIRST (trial energy, first axis)", 1: "KE_REST (energy / line forms, further axis)",
          2: "KH_FIRST (H of the accepted state, first axis)", 3: "KH_MID (H, middle axis)",
          4: "KH_LAST (H: last axis, local terms, gradient, residual rows)",
          5: "KL_FIRST (line coefficients of the polynomial search: new direction, local sums, q and r, "
             "first-axis forms of d)"}
_ELEM = {10: "k_dots", 11: "k_combine", 12: "k_norm2", 20: "k_sigma_local (Sobolev shift: bulk local potential)",
         22: "k_sigma_gather (Sobolev shift: candidates)", 23: "k_sigma_select (Sobolev shift: exact median)",
         30: "k_ctrl_dir", 31: "k_ctrl_armijo", 32: "k_ctrl_line (polynomial Armijo search)", 33: "k_post",
         34: "k_phi_comb"}


def site_name(sid):
    if sid in _ELEM:
        return _ELEM[sid]
    if 100 <= sid < 200:
        ax, kind = (sid - 100) // 10, (sid - 100) % 10
        fn = "k_lcoef" if kind == 5 else "k_kin"
        return f"{fn}<{_AXES.get(ax, ax)}> {_KINDS.get(kind, kind)}"
    if sid >= 2000:
        r = sid - 2000
        contig, r = r // 1000, r % 1000
        ld, r = r // 100, r % 100
        dr, mul = r // 10, r % 10
        cdot = mul >= 5 and mul - 5 == 0 and ld == 0 and dr == 1
        lds = {0: "complex", 1: "real, padded", 2: "half spectrum", 3: "real pair, padded",
               4: "complex, pair spectrum"}.get(ld, ld)
        drs = {0: "fwd", 1: "inv", 2: "fwd*mul*inv"}.get(dr, dr)
        muls = {0: "", 1: " x K^ (2D)", 2: " x K^ (3D)", 3: " / (k^2/2+sigma) (2D)", 4: " / (k^2/2+sigma) (3D)",
                5: " + 16 CG inner products" if cdot else " x Gaussian",
                7: " -> K*q, K*r + 3 sums"}.get(mul, f" mul {mul}")
        return f"k_pass<{'contig' if contig else 'strided'}, {lds}, {drs}{muls}>"
    return f"site {sid}"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed steps of the default workload")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps of the default workload")
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS),
                    help="the workload of the top-level fields")
    ap.add_argument("--only", action="store_true", help="no secondary workloads")
    ap.add_argument("--cpu-sample-s", type=float, default=10.0,
                    help="target CPU seconds of each workload's oracle sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single-solve comparison runs")
    ap.add_argument("--no-checks", action="store_true",
                    help="skip the polish and derivative checks (bounded profiling runs only)")
    ap.add_argument("--replay-only", action="store_true",
                    help="no timed region: only the event-timed replay of the default workload (profiling)")
    ap.add_argument("--line-search", default=None, choices=("poly", "trial"),
                    help="override the workloads' line search (default: poly)")
    ap.add_argument("--linesearch-max", type=int, default=None,
                    help="override the workloads' Armijo halvings per iteration (comparison runs)")
    ap.add_argument("--fp64-only", action="store_true", help="run FP32 warm phases in FP64 too (comparison)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner choose each plan's FFT stage orders and lines per workgroup first")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device-to-device copy (counter calibration)")
    return ap.parse_args()


def cpu_info():
    """Host CPU as lscpu reports it, and the threads this process may use (its
    affinity, capped by OMP_NUM_THREADS when set: the box's share of its host
    for one GPU)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    model = info.get("Model name") or platform.processor()
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else avail
    desc = dict(model=model, sockets=info.get("Socket(s)"), cores_per_socket=info.get("Core(s) per socket"),
                threads_per_core=info.get("Thread(s) per core"), host_cpus=os.cpu_count(), affinity=avail)
    return threads, desc


def b_iter(n, N, ndim, rot, lr):
    """Algorithmic bytes of one CG iteration of one instance (n = N^ndim points,
    c128 fields): SURVEY's per-iteration formula with one line pass (n_E = 1)
    and, with long range, the two convolutions of the polynomial search."""
    r, l, nE = (1 if rot else 0), (1 if lr else 0), 1
    conv = (80 * n + 8 * (N + 1) ** 2) if ndim == 2 else (144 * n + 8 * (N + 1) ** 3)
    return 32 * n * (3 + 2 * r + nE * (1 + 2 * r)) + 16 * n * (10 + 3 * nE) + 16 * n * l + l * 2 * conv


def traffic_of(workload, sid):
    """HBM bytes per dispatch of the roofline kernel from the committed counter
    summary of this workload (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated on
    a device copy), which lists the leading launch sites of the same replay."""
    f = os.path.join(HERE, "profiles", f"{workload}_hbm_traffic.md")
    if not os.path.exists(f):
        return None, None
    for m in re.finditer(r"site: ([0-9]+) traffic_bytes_per_dispatch: ([0-9.eE+]+)", open(f).read()):
        if int(m.group(1)) == sid:
            return float(m.group(2)), os.path.relpath(f, HERE)
    return None, None


def run_workload(name, steps, warmup, a, cpu_threads, cpu_desc):
    import torch
    import becgpp_amd as B
    from becgpp_amd import _lib
    from becgpp_amd.seeds import make_seed

    wl = WORKLOADS[name]
    cfg = B.default_cfg(**wl["cfg"])
    if a.fp64_only:
        cfg["warm_fp32"] = False
    if a.line_search:
        cfg["line_search"] = a.line_search
    if a.linesearch_max is not None:
        cfg["linesearch_max"] = int(a.linesearch_max)
    nseeds = int(cfg["nseeds"])
    points = wl.get("points", [{}])
    igrid = wl.get("instance_grid", lambda i: {})
    grids, runs = {}, []
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
    nsol, npt = len(runs), len(points)
    groups = {}
    for k, r in enumerate(runs):
        groups.setdefault(r[3], []).append(k)
    ndim = int(grids[runs[0][3]]["ndim"])

    def select(res):
        best = {}
        for (ip, _, _, _), (psi, _, obs) in zip(runs, res):
            if ip not in best or obs["E"] < best[ip][1]["E"]:
                best[ip] = (psi, obs)
        return best

    def solve_all():
        return B.ground_state_uneven(cs, psi0s=psi0s, grids=[grids[r[3]] for r in runs])

    lib = _lib.load()
    schedules = {}
    if a.autotune:
        from becgpp_amd.grid import plan_of
        from becgpp_amd.interactions import ensure_kernel
        for gk, ks in groups.items():
            Gg = grids[gk]
            lr = Gg["kernel"] != "none" and any(abs(float(cs[k]["G_C"])) > 1e-15 for k in ks)
            pl = ensure_kernel(Gg, "K") if lr else plan_of(Gg)
            schedules["N=%d" % gk[0]] = pl.autotune(min(len(ks), 64))
        torch.cuda.synchronize()
    out = dict(workload=name)
    res = None
    wall = 0.0
    if not a.replay_only:
        for _ in range(warmup):
            solve_all()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            res = solve_all()
            best = select(res)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        its = [obs["iters"] for _, _, obs in res]
        Es = [obs["E"] for _, _, obs in res]
        out.update(
            value=steps * npt / wall, unit="ground states/s", steps=steps, warmup=warmup,
            ms_per_step=1e3 * wall / steps,
            config=dict(workload=name, desc=wl["desc"],
                        Ngrid=sorted({gk[0] for gk in grids}) if len(grids) > 1 else runs[0][3][0], ndim=ndim,
                        points=npt, nseeds=nseeds, solves_per_step=nsol, res_tol=float(cfg["res_tol"]),
                        linesearch_max=int(cfg["linesearch_max"]),
                        line_search=str(cfg.get("line_search", "poly")),
                        warm_fp32=bool(cfg.get("warm_fp32", False)),
                        path="batched" if len(grids) == 1 else
                        "batched per grid, every grid's steps in one hipGraph (bec_ground_state_uneven)"),
            solves_per_s=steps * nsol / wall, cg_iters_per_s=steps * sum(its) / wall,
            iters_per_solve=sum(its) / nsol, iters_min=min(its), iters_max=max(its),
            E_selected=[best[i][1]["E"] for i in range(npt)] if npt > 1 else best[0][1]["E"],
            seed_energies=Es if npt == 1 else None,
            stop_reasons=sorted({obs["stop_reason"] for _, _, obs in res}),
            converged=all(bool(best[i][1]["converged"]) for i in range(npt)),
            resid_rel_max=max(best[i][1]["resid_rel"] for i in range(npt)),
        )
        if cfg.get("warm_fp32", False):
            wi = sum(obs.get("warm_iters", 0) for _, _, obs in res)
            out["warm_phase"] = dict(dtype="f32", res_tol=float(cfg["warm_res_tol"]), iters_share=wi / max(1, sum(its)),
                                     note="complex64 fields and transforms; FP64 sums, convolutions and Armijo "
                                          "test; the FP64 loop then runs from each warm state to res_tol")
        if schedules:
            out["autotuned_schedule"] = schedules
        # whole timed region against the per-iteration algorithmic bytes
        tot_b = 0.0
        for (ip, c, _, gk), (_, _, obs) in zip(runs, res):
            Gg = grids[gk]
            nn = int(Gg["N"]) ** ndim
            lr = Gg["kernel"] != "none" and abs(float(c["G_C"])) > 1e-15
            tot_b += obs["iters"] * b_iter(nn, int(Gg["N"]), ndim, abs(float(c["Omega"])) > 0, lr)
        ach = tot_b / (wall / steps) / 1e9
        out["step_roofline"] = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s",
                                    frac=ach / PEAK_HBM_GBS,
                                    basis="SURVEY 8(d) B_iter with n_E = 1 (one line pass) and two convolutions "
                                          "with long range, summed over every instance-iteration of a step, / "
                                          "ms_per_step")
        # accuracy of the first point's selected state
        if not a.no_checks:
            ib = min((k for k in range(nsol) if runs[k][0] == 0), key=lambda k: Es[k])
            b0psi, b0 = best[0]
            pc = dict(cs[ib], res_tol=1e-9, maxit=20000, warm_fp32=False)
            _, _, pol = B.ground_state(pc, G=grids[runs[ib][3]], psi0=b0psi, verbose=False)
            out["E_polished"] = pol["E"]
            out["dE_rel_vs_polished"] = abs(b0["E"] - pol["E"]) / abs(pol["E"])
            if "published" in wl:
                Epub, src = wl["published"]
                out["dE_rel_vs_reference"] = dict(value=abs(b0["E"] - Epub) / abs(Epub), E_reference=Epub,
                                                  source=src)
            chk = wl.get("check")
            if chk == "grad_fd":
                G = grids[runs[ib][3]]
                gen = torch.Generator(device="cpu").manual_seed(7)
                shp = tuple(b0psi.shape)
                rnd = lambda: torch.complex(torch.randn(shp, dtype=torch.float64, generator=gen),
                                            torch.randn(shp, dtype=torch.float64, generator=gen)).cuda()
                q = rnd()
                q = q / (q.abs().pow(2).sum() * float(G["dx"]) ** ndim).sqrt()
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
                pname = chk[1]
                c0 = dict(cs[ib], res_tol=1e-9, maxit=20000, warm_fp32=False)
                Gb = grids[runs[ib][3]]
                par = {k: float(c0[k]) for k in ("beta2", "beta3", "Omega", "G_C")}
                tg = torch.tensor(par[pname], dtype=torch.float64, requires_grad=True)
                args = [tg if k == pname else par[k] for k in ("beta2", "beta3", "Omega", "G_C")]
                psi_c, _, _ = B.ground_state(c0, G=Gb, psi0=b0psi, verbose=False)
                B.energy_t(psi_c, Gb, *args).backward()
                h = 1e-3 * max(1.0, abs(par[pname]))
                Ep = B.ground_state(dict(c0, **{pname: par[pname] + h}), G=Gb, psi0=psi_c, verbose=False)[2]["E"]
                Em = B.ground_state(dict(c0, **{pname: par[pname] - h}), G=Gb, psi0=psi_c, verbose=False)[2]["E"]
                fd = (Ep - Em) / (2 * h)
                out["check"] = dict(kind=f"Hellmann-Feynman dE/d{pname}: autograd at the converged state vs central "
                                         f"difference of converged energies (h = {h:g}, res_tol 1e-9)",
                                    autograd=tg.grad.item(), finite_difference=fd,
                                    rel_diff=abs(tg.grad.item() - fd) / max(abs(fd), 1e-300))
        if not a.no_single:
            t0 = time.perf_counter()
            sres = [B.ground_state(c, G=grids[r[3]], psi0=p0, verbose=False) for c, p0, r in zip(cs, psi0s, runs)]
            torch.cuda.synchronize()
            ts = time.perf_counter() - t0
            sb = select(sres)
            si = sum(obs["iters"] for _, _, obs in sres)
            out["single_solve_path"] = dict(ground_states_per_s=npt / ts, solves_per_s=nsol / ts,
                                            cg_iters_per_s=si / ts, iters_per_solve=si / nsol,
                                            E_selected=sb[0][1]["E"])
            out["batched_speedup"] = out["value"] / out["single_solve_path"]["ground_states_per_s"]
        if wl.get("mixed_variant") and not a.no_checks:
            # the same solves with the optional FP32 warm phase (to res 1e-4) before
            # the FP64 loop, one untimed and one timed pass: its measured effect
            cm = [dict(c, warm_fp32=True, warm_res_tol=1e-4) for c in cs]
            mixed = lambda: B.ground_state_uneven(cm, psi0s=psi0s, grids=[grids[r[3]] for r in runs])
            mixed()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mres = mixed()
            torch.cuda.synchronize()
            tm = time.perf_counter() - t0
            mb = select(mres)
            mi = sum(obs["iters"] for _, _, obs in mres)
            mw = sum(obs.get("warm_iters", 0) for _, _, obs in mres)
            out["mixed_precision"] = dict(
                warm_fp32=True, warm_res_tol=1e-4, ground_states_per_s=npt / tm, ms_per_step=1e3 * tm,
                iters_per_solve=mi / nsol, warm_iters_share=mw / max(1, mi),
                speedup_vs_fp64=(npt / tm) / out["value"],
                dE_rel_vs_fp64=max(abs(mb[i][1]["E"] - best[i][1]["E"]) / abs(best[i][1]["E"]) for i in range(npt)),
                note="complex64 fields and transforms in the warm phase; FP64 sums, convolutions and Armijo test; "
                     "then the FP64 loop to res_tol from each warm state")

    # roofline: every kernel of the largest grid group timed in a direct-launch replay
    replay_iters = int(wl.get("replay_iters", 300))
    gkr = max(groups, key=lambda g: len(groups[g]) * g[0] ** ndim)
    Gr = grids[gkr]
    k0 = groups[gkr][:64]
    lib.bec_prof_watch(-1)
    B.ground_state_batch([dict(cs[k], maxit=replay_iters, res_tol=0.0, warm_fp32=False) for k in k0], G=Gr,
                         psi0s=[psi0s[k] for k in k0])
    torch.cuda.synchronize()
    nsites = ctypes.c_int()
    lib.bec_prof_sites(None, 0, ctypes.byref(nsites))
    ids = (ctypes.c_int * max(1, nsites.value))()
    lib.bec_prof_sites(ids, nsites.value, ctypes.byref(nsites))
    per = []
    for sid in list(ids)[:nsites.value]:
        ms, cnt, nb = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        lib.bec_prof_read_site(sid, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(nb))
        per.append((sid, ms.value, cnt.value, nb.value))
    lib.bec_prof_watch(0)
    tot_ms = sum(p[1] for p in per)
    per.sort(key=lambda p: -p[1])
    sid, ms, cnt, nb = per[0]
    avg_s = ms / 1e3 / max(1, cnt)
    bpl = nb / max(1, cnt)
    achieved = bpl / avg_s / 1e9 if cnt and bpl else None
    roof = dict(bound="hbm", kernel=site_name(sid), site=sid, launches=cnt,
                share_of_replay_kernel_time=ms / tot_ms if tot_ms else None,
                timed_in=f"direct-launch replay of {len(k0)} instance(s) on N = {gkr[0]}, {replay_iters} CG "
                         f"iterations, every launch timed with HIP events on its stream",
                avg_launch_us=avg_s * 1e6, achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s",
                frac=(achieved / PEAK_HBM_GBS) if achieved else None, traffic=None,
                algorithmic_bytes_per_launch=bpl)
    tr, tsrc = traffic_of(name, sid)
    if tr:
        roof["traffic"] = tr
        roof["traffic_source"] = tsrc
        roof["traffic_over_algorithmic"] = tr / bpl if bpl else None
    roof["replay_kernel_us_per_iter"] = 1e3 * tot_ms / replay_iters
    roof["breakdown"] = [dict(site=p[0], kernel=site_name(p[0]), share=p[1] / tot_ms if tot_ms else None,
                              avg_us=1e3 * p[1] / max(1, p[2]),
                              GBs=(p[3] / max(1, p[2])) / (p[1] / 1e3 / max(1, p[2])) / 1e9 if p[1] and p[3] else None)
                         for p in per[:16] if tot_ms and p[1] / tot_ms >= 0.01]
    out["roofline"] = roof

    if not a.no_cpu_baseline and not a.replay_only:
        out["cpu_baseline"] = cpu_baseline(cs[0], psi0s[0].cpu().numpy(), a.cpu_sample_s, out["iters_per_solve"],
                                           nsol / npt, cpu_threads, cpu_desc)
    return out


def cpu_baseline(cfg, psi0, target_s, iters_per_solve, solves_per_state, threads, desc):
    """The oracle (numpy/scipy pocketfft) on the first seed of the same
    workload for a bounded number of CG iterations sized to about target_s
    seconds; its set-up (initial energy and residual, kernel spectra) is timed
    on its own and subtracted.  Scaled to ground states/s with the GPU's
    iterations per solve."""
    import oracle as O
    O.set_threads(threads)
    c = dict(cfg, res_tol=0.0)
    Go = O.make_grid(c)
    O.ground_state(dict(c, maxit=2), Go, psi0)          # warm the FFT plans and kernel spectra

    def timed(k):
        t0 = time.perf_counter()
        _, _, obs = O.ground_state(dict(c, maxit=k), Go, psi0)
        return time.perf_counter() - t0, obs["iters"]

    t1, i1 = timed(1)
    t3, i3 = timed(3)
    per_it = max((t3 - t1) / max(1, i3 - i1), 1e-6)
    k = int(max(4, min(4000, target_s / per_it)))
    dt, it = timed(k)
    its = max(1, it - i1) / max(dt - t1, 1e-9)
    return dict(value=its / (iters_per_solve * solves_per_state), unit="ground states/s", cores=threads,
                kind="port", cg_iters_per_s=its, cpu=desc,
                sample=f"{it} CG iterations of the first solve of the same workload ({dt:.1f} s on {threads} "
                       f"threads; the set-up, timed alone at {t1:.2f} s with one iteration, subtracted), scaled "
                       f"by the GPU's {iters_per_solve:.0f} iterations/solve x {solves_per_state:g} solves per "
                       f"ground state")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps(dict(error="this build runs on one GPU: the ground-state path has no multi-GPU "
                                    "code; run with --gpus 1")))
        sys.exit(2)
    import torch
    import becgpp_amd  # noqa: F401
    if a.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    threads, desc = cpu_info()
    print(f"[bench] {a.workload}: {a.steps} steps after {a.warmup} warm-up", file=sys.stderr, flush=True)
    top = run_workload(a.workload, a.steps, a.warmup, a, threads, desc)
    if a.replay_only:
        print(json.dumps(dict(metric="roofline replay only (--replay-only)", value=None,
                              roofline=top["roofline"], config=dict(workload=a.workload))), flush=True)
        return
    line = dict(metric="ground states/sec at |dE/E|<=1e-8; CG iters/sec; HBM GB/s vs peak",
                value=top.pop("value"), unit=top.pop("unit"), n_gpus=1, steps=a.steps, warmup=a.warmup,
                ms_per_step=top.pop("ms_per_step"), higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype="f64", data="synthetic", config=top.pop("config"))
    top.pop("steps", None)
    top.pop("warmup", None)
    line.update(top)
    if not a.only:
        line["workloads"] = {}
        for name in SECONDARY:
            if name == a.workload:
                continue
            print(f"[bench] {name}: {SECONDARY_STEPS} steps after {SECONDARY_WARMUP} warm-up", file=sys.stderr,
                  flush=True)
            line["workloads"][name] = run_workload(name, SECONDARY_STEPS, SECONDARY_WARMUP, a, threads, desc)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
