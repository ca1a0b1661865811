"""Benchmark of the selected-basis Davidson hot path on one MI355X.

Default workload: norb=36, 9 alpha + 9 beta electrons, 20,000 x 20,000 sampled strings
(400M determinants), seed 2 -- the configuration the mixed-vs-FP64 comparison is quoted
on; --config names another (the largest: norb=58, 12+12, 40,000 x 40,000, 1.6e9
determinants).  Synthetic integrals and strings (no datasets available offline).

A step is one sigma build (H c over all determinants, FP64) on vectors already resident
in HBM.  `value` is determinants x steps / seconds (det*H*v/s) over the K timed steps.
Also reported:
  * FP32 sigma throughput of the mixed path;
  * time-to-|dE| < 1e-6 Ha for the FP64 Davidson and the mixed Davidson at the workload's
    Davidson settings (SETTINGS), against the converged FP64 energy, and the mixed energy
    error: medians over three start vectors (the first determinant, two seeded random
    vectors; --tts-starts 1 keeps the first); the mixed runs refresh the exact sigma every
    refresh_k steps and, second, on the residual's schedule (refresh_k = 0; also every 8
    steps from the first determinant, except for the largest workload); for the largest
    workload the FP64 reference run takes the largest subspace that fits the free HBM, and
    the RDMs and their backward are timed after the mixed runs;
  * roofline of the dominant sigma kernel: its algorithmic HBM bytes per launch (model
    below) over its average duration from HIP events on the library stream, taken from
    sigma builds with the kernels in sequence (the timed builds run the alpha-beta kernel
    on a second stream beside the spin phases);
  * sigma_model: bytes per sigma under three models -- the compulsory bytes (c and the
    diagonal read once, sigma written once, tables once), the row-streamed model (one ket
    row per table entry) and, where a counter summary of this workload is committed, the
    HBM bytes rocprofv3 counted -- each with its rate at the measured sigma time;
  * cpu_baseline: the CPU restatement (oracle) timed on this host's cores on a bounded
    sample of rows of the same workload (OpenMP over the row-sample's elements), in
    det*H*v/s and Hamiltonian terms/s.

Run: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--no-tts] [--no-cpu]
                    [--sequential]   (all sigma kernels in one stream: per-kernel profiles)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sbd-ykt_amd"))

HBM_PEAK_GBS = 8000.0
# LDS: 256 CUs x 256 B/clk x 2.4 GHz.  gfx950's LDS serves 64 banks of 4 bytes per clock
# (the bank model the alpha-beta entry schedule is tuned to: 16 slots of 16 bytes); the
# larger of the two figures in use for CDNA parts, so the reported fraction is not flattered.
LDS_PEAK_TBS = 256 * 256 * 2.4e9 / 1e12


# Davidson settings of each workload: tolerance, subspace, Ritz vectors kept at a restart,
# FP64 refresh period of the mixed path
SETTINGS = {
    "c0_L12_4e4e_full": dict(max_nb=10, restart_keep=1, eps=1e-8, refresh_k=4),
    "c1_L26_5e5e_4000": dict(max_nb=24, restart_keep=4, eps=1e-8, refresh_k=4),
    "c2_L36_9e9e_20000": dict(max_nb=12, restart_keep=4, eps=1e-5, refresh_k=4),
    "c3_L16_5e5e_full": dict(max_nb=12, restart_keep=4, eps=1e-8, refresh_k=4),
    "c4_L58_12e12e_40000": dict(max_nb=8, restart_keep=3, eps=1e-5, refresh_k=4),
}


def describe(name):
    from sbdg import synthetic as syn

    c = syn.CONFIGS[name]
    space = (f"{c['na']} x {c['nb']} sampled strings" if c.get("na") else
             f"full space C({c['norb']},{c['nalpha']}) x C({c['norb']},{c['nbeta']})")
    return f"norb={c['norb']}, {c['nalpha']}+{c['nbeta']} electrons, {space}, seed {c['seed']}"


def traffic_summary(name):
    """The committed counter summary of this workload (HBM bytes per launch and per sigma from
    FETCH_SIZE scaled by the copy-kernel calibration, plus WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", f"sigma_{name.split('_')[0]}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t if t.get("config") == name else None


def measured_traffic(summary, kernel):
    if summary is None:
        return None
    v = summary.get("per_launch_bytes", {}).get(kernel)
    return None if v is None else int(v)


def kernel_bytes(na, nb, nnz, core, nel_b, s, direct=(0, 0)):
    """Algorithmic HBM bytes per launch of each sigma phase (s = bytes per element).

    transpose:     read c, write c^T                                   2 N s
    beta product:  on c^T: singles rows and their occupation-table
                   rows, diagonal-free write                           2 s na nnzSb + N s + 16 nnzSb
                   doubles by core (mb memberships of the larger
                   cores): member rows in, partial rows out, partial
                   rows gathered; the xb pairs of the smaller cores
                   read directly                                       + 3 s na mb + s na xb + 16 xb
                   (excitation-table form instead: s na nnzDb + 12 nnzDb)
    alpha product: same on c, with hii and c read for the diagonal     2 s nb nnzSa + 3 N s + ...
    transpose-add: read sigma^T and sigma, write sigma                 3 N s
    alpha-beta:    one ket row per alpha single (ia, ja), sigma read
                   and written, ELL entries                            s nb nnzSa + 2 N s + 4 nnzSb
    """
    N = na * nb
    sa, da, sb, db = nnz
    ma, mb = core
    xa, xb = direct

    def spin(ncols, singles, doubles, members, pairs):
        if members or pairs:
            return s * ncols * (2 * singles + 3 * members + pairs) + 16 * (singles + pairs)
        return s * ncols * (2 * singles + doubles) + 16 * (singles + doubles)

    return {
        "transpose": 2 * N * s,
        "beta": spin(na, sb, db, mb, xb) + N * s,
        "alpha": spin(nb, sa, da, ma, xa) + 3 * N * s,
        "transpose_add": 3 * N * s,
        "alpha_beta": s * nb * sa + 2 * N * s + 4 * sb,
    }


def ab_lds_bytes(na, nb, nnz, nel_b, s):
    """LDS bytes the alpha-beta kernel reads per launch: every alpha single (bra row, ket
    row) times every beta single of the bra column, one ket value and one integral each."""
    sa, _, sb, _ = nnz
    return sa * sb * 2 * s


def row_stream_alg_bytes(na, nb, nnz, s):
    """Per-sigma algorithmic bytes of the row-streamed model:
    N 3s + s Nb (nnzSa + nnzDa) + s Nb nnzSa + 4 nnz(tables)."""
    sa, da, sb, db = nnz
    return na * nb * 3 * s + s * nb * (sa + da) + s * nb * sa + 4 * (sa + da + sb + db)


def compulsory_bytes(na, nb, L, nnz, core, direct, s):
    """Bytes every sigma must move at least once: read c and the diagonal, write sigma (3 N s);
    read the row-product tables once (column 4 + value s + occupation index 4 per single or
    direct pair), the core memberships (row, pair, slot: 12 each), the alpha-beta entries
    (4 each) and the occupation tables of both spins ((L^2 + 1) (na + nb) s)."""
    sa, da, sb, db = nnz
    ma, mb = core
    xa, xb = direct
    rows = (sa + (xa if ma else da)) + (sb + (xb if mb else db))
    return 3 * na * nb * s + (8 + s) * rows + 12 * (ma + mb) + 4 * sb + (L * L + 1) * (na + nb) * s


def time_to_de(traj_t, traj_e, e_ref, tol=1e-6):
    for t, e in zip(traj_t, traj_e):
        if abs(e - e_ref) < tol:
            return float(t)
    return None


def cpu_baseline(wl, batches=4, budget_s=25.0):
    """Oracle sigma on a sample of alpha rows (all columns of each) on this host: batches of
    as many rows as threads (OpenMP over the batch's elements, so every thread works), until
    `batches` batches or the time budget; rows spread over the whole string set."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle

    p = oracle.Problem(wl.norb, wl.core, wl.h1, wl.astr, wl.bstr, eri=wl.eri)
    N, na, nb = p.ndet, p.na, p.nb
    threads = oracle.lib().orc_num_threads()
    rpSa, _ = p.table(0)
    rpDa, _ = p.table(1)
    nnz_sb = int(p.table(2)[0][-1])
    nnz_b = nnz_sb + int(p.table(3)[0][-1])
    x = np.random.default_rng(0).random(N, dtype=np.float32).astype(np.float64)
    done = terms = 0
    t_tot = 0.0
    step = max(1, na // (batches * threads))
    for b in range(batches):
        r0 = min(na - threads, b * threads * step)
        t0 = time.perf_counter()
        p.sigma_rows(x, None, r0, r0 + threads)
        t_tot += time.perf_counter() - t0
        for ia in range(r0, r0 + threads):
            sa, da = rpSa[ia + 1] - rpSa[ia], rpDa[ia + 1] - rpDa[ia]
            terms += nb * (1 + sa + da) + nnz_b + sa * nnz_sb
        done += threads
        print(f"[bench] cpu baseline: {done} rows in {t_tot:.1f} s", flush=True)
        if t_tot > budget_s:
            break
    return dict(value=done * nb / t_tot, unit="det\u00b7H\u00b7v/s", cores=threads, kind="port",
                terms_per_s=terms / t_tot,
                sample=f"{done} alpha rows x {nb} beta strings of the same workload ({terms:.3e} Hamiltonian "
                       f"terms), one FP64 sigma row block per batch of {threads} rows, {t_tot:.1f} s")


def fp64_subspace(s, st, na, nb, L, want, margin=4e9):
    """The largest FP64 subspace (at most `want`) whose Davidson vectors (the library's own plan
    of them) fit the device's free memory plus what the library holds beyond its tables (the
    Davidson releases the benchmark vectors and the other precision's vectors first): the FP64
    reference energy of the largest workload needs one."""
    import sbdg

    free, _ = sbdg.device_memory()
    free += s.stats()["mem_now"] - st["mem_now"]  # st: right after the build
    for m in range(want, 2, -1):
        need = sbdg.memory_plan(L, na, nb, st["nnz"], st["core_members"], st["core_direct"],
                                precision=0, max_nb=m)["davidson"]
        if need + margin < free:
            return m
    return 3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2_L36_9e9e_20000")
    ap.add_argument("--no-tts", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sequential", action="store_true", help="time sigma with its kernels in one stream")
    ap.add_argument("--tts-starts", type=int, default=3,
                    help="start vectors of the time-to-solution runs (1-3; speedups are medians over them)")
    ap.add_argument("--davidson-steps", action="store_true",
                    help="only run 20 FP64 and 24 mixed Davidson steps (for a kernel trace of the vector passes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy three times (scales the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(256 * 1024 * 1024, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("[bench] copy calibration: 3 x 1073741824 bytes read and written", flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import sbdg
    from sbdg import synthetic as syn

    cfg = SETTINGS[args.config]
    largest = args.config.startswith("c4")
    t0 = time.perf_counter()
    wl = syn.config_workload(args.config)
    print(f"[bench] workload {args.config}: {len(wl.astr)} x {len(wl.bstr)} = {wl.ndet} determinants "
          f"(generated in {time.perf_counter() - t0:.1f} s)", flush=True)
    s = sbdg.Solver(wl.norb, wl.core, wl.h1, wl.eri, wl.astr, wl.bstr, overlap=0 if args.sequential else 1)
    s.build()
    st = s.stats()
    na, nb, L = len(wl.astr), len(wl.bstr), wl.norb
    nnz = st["nnz"]
    print(f"[bench] tables+diagonal {st['t_tables']:.2f} s, nnz {nnz}, terms/sigma {st['terms']:.3e}, "
          f"device bytes {st['mem_now'] / 1e9:.1f} GB", flush=True)
    if args.davidson_steps:
        for prec in (0, 1):
            s.set_options(max_it=2, max_nb=12, restart_keep=4, eps=1e-12, precision=prec, refresh_k=0)
            s.run()
            sd = s.stats()
            print(f"[bench] davidson precision {prec}: {sd['iterations']} steps in {sd['t_davidson']:.3f} s, "
                  f"sigma {sd['t_sigma']:.3f} s", flush=True)
        s.close()
        return

    # warmup, then K timed FP64 sigma builds (library events per kernel; wall clock overall)
    s.bench_sigma(0, max(1, args.warmup))
    t0 = time.perf_counter()
    ms = s.bench_sigma(0, args.steps)
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    value = wl.ndet * args.steps / wall
    print(f"[bench] fp64 sigma {ms_step:.2f} ms/step (events {ms[0]:.2f} ms)", flush=True)
    s.bench_sigma(1, max(1, args.warmup))
    ms32 = s.bench_sigma(1, args.steps)
    print(f"[bench] fp32 sigma {ms32[0]:.2f} ms/step", flush=True)

    # per-kernel times for the roofline: the same sigma with every kernel in one stream, in
    # order (the timed builds above run the alpha-beta kernel beside the spin phases)
    s.set_options(overlap=0)
    ms = s.bench_sigma(0, 2)
    s.set_options(overlap=0 if args.sequential else 1)
    print(f"[bench] fp64 sigma, kernels in sequence: {ms[0]:.2f} ms", flush=True)
    names = ["transpose", "beta", "alpha", "transpose_add", "alpha_beta"]
    nel_b = bin(int(wl.bstr[0])).count("1")
    core, direct = st["core_members"], st["core_direct"]
    kb = kernel_bytes(na, nb, nnz, core, nel_b, 8, direct)
    per = {n: float(ms[i + 1]) for i, n in enumerate(names)}
    dom = max(per, key=per.get)
    achieved = kb[dom] / (per[dom] * 1e-3) / 1e9
    summary = traffic_summary(args.config)
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(summary, dom),
                    launch_ms=round(per[dom], 3), alg_bytes=int(kb[dom]))
    if dom == "alpha_beta":
        lb = ab_lds_bytes(na, nb, nnz, nel_b, 8)
        roofline["lds"] = dict(bytes=int(lb), achieved_TBs=round(lb / (per[dom] * 1e-3) / 1e12, 2),
                               peak_TBs=LDS_PEAK_TBS, frac=round(lb / (per[dom] * 1e-3) / 1e12 / LDS_PEAK_TBS, 4))

    def rate(b):
        gbs = b / (ms_step * 1e-3) / 1e9
        return dict(bytes=int(b), GBs=round(gbs, 1), frac_of_hbm_peak=round(gbs / HBM_PEAK_GBS, 4))

    sigma_model = dict(
        compulsory=rate(compulsory_bytes(na, nb, L, nnz, core, direct, 8)),
        row_streamed=dict(rate(row_stream_alg_bytes(na, nb, nnz, 8)),
                          note="bytes a sigma that streams one ket row per table entry would move; a "
                               "yardstick, not a bandwidth: this sigma moves fewer (doubles by core, rows "
                               "shared in cache), so its ratio to the peak can pass 1"),
        measured=(dict(rate(summary["per_sigma_bytes"]), source=f"profiles/{os.path.basename(summary['file'])}")
                  if summary and summary.get("per_sigma_bytes") else None),
        sequential_ms=round(float(ms[0]), 3), kernel_ms={k: round(v, 3) for k, v in per.items()})

    tts = None
    extra = {}
    if not args.no_tts:
        # Davidson time-to-|dE|<1e-6 from the first determinant (the reference's default)
        # and, with --tts-starts 3, two seeded random vectors.  Per start: the FP64 run, then
        # the mixed runs, each timed against that start's converged FP64 energy.  The early
        # steps of these synthetic Hamiltonians wander for a rounding-dependent number of
        # steps, so the speedups are medians over the starts.
        starts = [("first_determinant", dict(init=0)), ("random_seed1", dict(init=1, seed=1)),
                  ("random_seed2", dict(init=1, seed=2))][:max(1, args.tts_starts)]
        runs = []
        base = dict(max_it=200, max_nb=cfg["max_nb"], restart_keep=cfg["restart_keep"], eps=cfg["eps"],
                    max_time=900.0, verbose=1)
        for name, init in starts:
            if largest:
                # the FP64 reference energy with the largest subspace that fits beside the
                # tables; the mixed run at its own settings
                nb64 = fp64_subspace(s, st, na, nb, L, cfg["max_nb"])
                r = dict(start=name, fp64_max_nb=nb64)
                ks = (cfg["refresh_k"], 0)
            else:
                nb64 = cfg["max_nb"]
                r = dict(start=name)
                ks = (cfg["refresh_k"], 8, 0) if name == "first_determinant" else (cfg["refresh_k"], 0)
            s.set_options(precision=0, **dict(base, max_nb=nb64, restart_keep=min(cfg["restart_keep"], nb64 - 1)),
                          **init)
            e64 = s.run()
            t64, tr64 = s.trajectory()
            st64 = s.stats()
            a = time_to_de(t64, tr64, e64)
            r.update(e_fp64=e64, fp64_s=a, fp64_sigma=st64["sigma_fp64"], fp64_steps=st64["iterations"],
                     fp64_step_ms=1e3 * st64["t_davidson"] / max(1, st64["iterations"]),
                     fp64_mem_peak_GB=round(st64["mem_peak"] / 1e9, 2))
            for k in ks:  # 0: scheduled by the residual
                s.set_options(precision=1, refresh_k=k, **base, **init)
                emx = s.run()
                tmx, trmx = s.trajectory()
                stmx = s.stats()
                b = time_to_de(tmx, trmx, e64)
                key = "" if k == cfg["refresh_k"] else {8: "_k8", 0: "_adaptive"}[k]
                r.update({f"mixed_s{key}": b, f"speedup{key}": (a / b) if (a and b) else None,
                          f"mixed_error{key}": abs(emx - e64), f"mixed_steps{key}": stmx["iterations"],
                          f"mixed_sigma_fp32{key}": stmx["sigma_fp32"], f"mixed_sigma_fp64{key}": stmx["sigma_fp64"],
                          f"mixed_step_ms{key}": 1e3 * stmx["t_davidson"] / max(1, stmx["iterations"]),
                          f"mixed_mem_peak_GB{key}": round(stmx["mem_peak"] / 1e9, 2)})
                sched = f"refresh every {k}" if k > 0 else "adaptive refresh"
                print(f"[bench] {name}: time-to-|dE|<1e-6 fp64 {a} s, mixed ({sched}) {b} s; "
                      f"mixed error {abs(emx - e64):.2e}", flush=True)
            if largest and name == "first_determinant":
                # the outputs of the largest workload: RDMs of the converged mixed vector and
                # the Hellmann-Feynman backward
                t1 = time.perf_counter()
                s.rdm(compute=True)
                t_rdm = time.perf_counter() - t1
                t1 = time.perf_counter()
                dh, de = s.rdm_backward(1.0)
                t_bw = time.perf_counter() - t1
                extra = dict(rdm_s=round(t_rdm, 3), backward_s=round(t_bw, 3),
                             mem_peak_GB=round(s.stats()["mem_peak"] / 1e9, 2),
                             grad_norms=[float(np.linalg.norm(dh)), float(np.linalg.norm(de))])
                print(f"[bench] RDMs {t_rdm:.2f} s, backward {t_bw:.2f} s (with the L^4 copy to the host)", flush=True)
            runs.append(r)

        def median(key):
            v = sorted(r[key] for r in runs if r.get(key))
            return v[len(v) // 2] if len(v) % 2 else (0.5 * (v[len(v) // 2 - 1] + v[len(v) // 2]) if v else None)

        tts = dict(settings=cfg, speedup=median("speedup"), speedup_adaptive=median("speedup_adaptive"),
                   speedup_k8_first_determinant=runs[0].get("speedup_k8"),
                   fp64_step_ms=median("fp64_step_ms"), mixed_step_ms=median("mixed_step_ms"),
                   mixed_step_ms_adaptive=median("mixed_step_ms_adaptive"),
                   mixed_error_max=max(r[k] for r in runs for k in r if k.startswith("mixed_error")),
                   starts=runs, **extra)
    s.close()

    cpu = None if args.no_cpu else cpu_baseline(wl)

    line = {
        "metric": "Davidson time-to-|dE|<1e-6 Ha, mixed vs FP64 on 1 GPU; det\u00b7H\u00b7v/s",
        "value": value, "unit": "det\u00b7H\u00b7v/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": describe(args.config), "determinants": wl.ndet, "terms_per_sigma": st["terms"],
                   "name": args.config},
        "fp32_sigma": {"value": wl.ndet / (ms32[0] * 1e-3), "ms_per_step": float(ms32[0]), "dtype": "f32"},
        "time_to_solution": tts,
        "roofline": roofline,
        "sigma_model": sigma_model,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
