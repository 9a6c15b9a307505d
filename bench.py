"""Benchmark: beam.bin diffractions/s of the far-field diffraction path on one MI355X.

Workload (default): the seeded diffraction stress test -- 200,000 planar
triangular apertures (centroids in a 100 um cube, edges 0.3-5 um normal to a
random propagation vector, amplitudes (N+iN) 10^U(-6,0)), wavelength 0.532 um,
grid 1801 theta (0.1 deg) x 360 phi (1 deg) = 648,360 bins.

One step = one pass of the hot path over the whole beam set with the raw beam
arrays already resident in HBM: per-beam frame set-up (device), the coherent
far-field sum of every beam onto every bin (mixed-precision kernel), and the
Mueller epilogue (total / beam / ext Mueller matrices, phi integration, theta
integrals).  `value` = beams * bins / (seconds per step).

Also measured in the same run: the plain FP64 kernel on the same inputs (its
rate and the mixed/FP64 ratio), the accuracy of the mixed result against the
FP64 one per Mueller element, the dominant kernel's roofline fraction (HIP
events on the launch stream), and a CPU baseline (the FP64 C oracle, OpenMP,
on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_ALG = 250.0            # algorithmic FLOP per beam.bin (triangular aperture)
PEAK = {"mixed": 157.3e12, "fp64": 78.6e12}   # MI355X vector peaks, FLOP/s
METRIC = "beam·bin diffractions/s mixed-prec vs FP64 at ≤1e-5 Mueller rel err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--beams", type=int, default=200_000)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--fp64-steps", type=int, default=1, help="FP64 comparison steps (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=96, help="apertures in the CPU baseline sample (0: skip)")
    ap.add_argument("--orient", type=int, default=1000,
                    help="orientations of the batched orientation-average workload (0: skip)")
    ap.add_argument("--orient-fp64", type=int, default=1,
                    help="run the FP64 plain and FP64 batched comparisons of the orientation average (0: skip)")
    ap.add_argument("--adjoint-orient", type=int, default=64,
                    help="orientations of the adjoint workload (0: skip)")
    ap.add_argument("--large-orient", type=int, default=0,
                    help="orientations of the large absorbing column (20,736 facets; 0: skip)")
    ap.add_argument("--vis-beams", type=int, default=4096,
                    help="beams of the facet-visibility workload (0: skip)")
    ap.add_argument("--batch-tiles", default=None,
                    help="fixed tile shape of the orientation-batched kernel, e.g. 16x1 (default: autotune)")
    ap.add_argument("--json-out", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is a design (orientation/beam sharding + one RCCL reduce), not built.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from abt_amd import lib
    from abt_amd.workloads import GRIDS, grid, stress_beams

    lam = 0.532
    dev = torch.device("cuda:0")
    th, ph = grid(*GRIDS["stress"])
    t0 = time.time()
    beams = stress_beams(args.beams)
    log(f"workload: {beams.n} apertures, grid {len(th)} x {len(ph)} ({time.time() - t0:.1f} s to generate)")
    g = lib.DeviceGrid(th, ph, lam)
    g.waveno = 2 * lib.PI_R4 / lam
    nb, nbins = beams.n, g.nbins
    raw = {k: torch.from_numpy(getattr(beams, k)).to(dev) for k in
           ("ampl", "vk7", "prop", "fov", "vert_off", "verts")}
    R = lib.load().abt_record_doubles()
    recs = torch.empty((nb, R), dtype=torch.float64, device=dev)
    amp = torch.empty(8 * nbins, dtype=torch.float64, device=dev)
    amp_ext = torch.empty(8 * nbins, dtype=torch.float64, device=dev)
    mueller = torch.empty(48 * nbins, dtype=torch.float64, device=dev)
    m1d = torch.empty(48 * len(th), dtype=torch.float64, device=dev)
    params = torch.empty(16, dtype=torch.float64, device=dev)
    ws = lib.far_field_workspace(g, nb, args.precision)
    work = torch.empty(max(ws, 8) // 8 + 1, dtype=torch.float64, device=dev)
    fws = lib.load().abt_finalise_workspace(g.handle)
    fwork = torch.empty(fws // 8 + 1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    k_ev = []

    def step(prec, time_kernel=False):
        lib.prepare_beams_dev(raw, nb, lam, recs, fov_value=None, stream=stream)
        if time_kernel:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        lib.far_field_dev(g, recs, nb, amp, precision=prec, work=work, stream=stream)
        if time_kernel:
            e1.record(stream)
            k_ev.append((e0, e1))
        amp_ext.zero_()
        lib.finalise_dev(g, amp, amp_ext, mueller, m1d, params, work=fwork, stream=stream)

    def timed(prec, steps, warmup):
        k_ev.clear()
        for _ in range(warmup):
            step(prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            step(prec, time_kernel=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        kms = float(np.mean([a.elapsed_time(b) for a, b in k_ev]))
        return dt, kms

    dt, kms = timed(args.precision, args.steps, args.warmup)
    value = nb * nbins / dt
    log(f"{args.precision}: {dt * 1e3:.1f} ms/step, kernel {kms:.1f} ms, {value:.4e} beam.bin/s")
    res_main = amp.clone()
    m_main = mueller[:16 * nbins].clone()

    out = {
        "metric": METRIC, "value": value, "unit": "beam·bin/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": "diffraction stress test: 200,000 seeded triangular apertures, "
                               "1801 theta x 360 phi bins, lambda 0.532 um",
                   "beams": nb, "bins": nbins, "beam_bins": nb * nbins,
                   "precision": args.precision, "step": "device frame set-up + far-field sum + Mueller epilogue"},
    }
    ach = F_ALG * nb * nbins / (kms * 1e-3)
    kname = "k_far_field<%d,16>" % (1 if args.precision == "mixed" else 0)
    traffic = _traffic(kname) if nb == 200_000 else None
    out["roofline"] = {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK[args.precision] / 1e12,
                       "unit": "TFLOP/s", "frac": ach / PEAK[args.precision], "traffic": traffic,
                       "kernel": kname,
                       "kernel_ms": kms, "flop_per_beam_bin": F_ALG}

    if args.fp64_steps > 0 and args.precision == "mixed":
        dt64, kms64 = timed("fp64", args.fp64_steps, 0)
        v64 = nb * nbins / dt64
        log(f"fp64: {dt64 * 1e3:.1f} ms/step, kernel {kms64:.1f} ms, {v64:.4e} beam.bin/s")
        a64 = amp.cpu().numpy()
        m64 = mueller[:16 * nbins].cpu().numpy().reshape(16, nbins)
        am = res_main.cpu().numpy()
        mm = m_main.cpu().numpy().reshape(16, nbins)
        out["fp64"] = {"value": v64, "ms_per_step": dt64 * 1e3, "kernel_ms": kms64,
                       "roofline_frac": F_ALG * nb * nbins / (kms64 * 1e-3) / PEAK["fp64"]}
        out["speedup_vs_fp64"] = value / v64
        inc = torch.empty(nbins, dtype=torch.float64, device=dev)
        lib.far_field_incoherent_dev(g, recs, nb, inc, stream=stream)
        torch.cuda.synchronize()
        # the epilogue scaled the amplitudes by k in place; the incoherent sum is unscaled
        out["accuracy"] = accuracy(am, a64, mm, m64, nbins, inc.cpu().numpy() * g.waveno ** 2)

    if args.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(beams, th, ph, lam, args.cpu_sample)
    del recs, amp, amp_ext, mueller, work, raw
    torch.cuda.empty_cache()
    if args.orient > 0:
        out["orientation_average"] = orientation_average(args.orient, args.steps, lam, dev, bool(args.orient_fp64),
                                                         args.batch_tiles)
    if args.adjoint_orient > 0:
        out["adjoint"] = adjoint_bench(args.adjoint_orient, lam, dev)
    if args.large_orient > 0:
        out["large_column"] = orientation_average(args.large_orient, args.steps, lam, dev, True, None,
                                                  particle="large_column", grid_name="large", with_kappa=True)
    if args.vis_beams > 0:
        out["visibility"] = visibility_bench(args.vis_beams, max(args.steps, 20), max(args.warmup, 2), dev,
                                             args.cpu_sample > 0)
    line = json.dumps(out)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)


def _traffic(kname):
    """HBM bytes per launch of a kernel on its default workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by the copy-kernel calibration."""
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tpath):
        return None
    return json.load(open(tpath)).get(kname, {}).get("bytes_per_launch")


def accuracy(am, a64, mm, m64, nbins, incoherent):
    """Mixed vs FP64 on the full workload, per Mueller element.
    S11_rel: |dS11| / S11;  norm_elem: |dS_ij| / S11 (normalised element);
    rel_floored: |dS_ij| / max(|S_ij|, 1e-3 S11).
    kappa = sqrt(sum_b |A_b|^2) / |sum_b A_b| per bin (FP64 incoherent sum over the
    coherent one): a per-beam relative error eps becomes ~eps kappa in the bin, so
    bins of the coherent speckle minima (kappa >> 1) are reported separately and
    err / kappa gives the per-beam precision of the mixed kernel."""
    amp64 = np.sqrt((a64.reshape(8, nbins) ** 2).sum(0))
    kappa = np.sqrt(incoherent) / amp64
    d = np.abs(mm - m64)
    s11 = np.abs(m64[0])
    rel = d / np.maximum(np.abs(m64), 1e-3 * s11)
    norm = d / s11
    s11r = d[0] / s11
    q = lambda x: {"max": float(x.max()), "p999": float(np.quantile(x, 0.999))}
    out = {"reference": "FP64 GPU kernel on the same prepared beams",
           "S11_rel": q(s11r), "norm_elem": q(norm), "rel_floored": q(rel),
           "kappa": {"median": float(np.median(kappa)), "p999": float(np.quantile(kappa, 0.999)),
                     "max": float(kappa.max())},
           "per_beam_precision": q(np.abs(np.sqrt(np.abs(mm[0]) / s11) - 1) / kappa),
           "per_element_norm_max": [float(norm[e].max()) for e in range(16)]}
    for kmax in (3.0, 10.0):
        sel = kappa <= kmax
        out[f"kappa_le_{int(kmax)}"] = {"bins": int(sel.sum()), "S11_rel": q(s11r[sel]),
                                        "norm_elem": q(norm[:, sel])}
    return out


def _prepare_trees(w, lam, dev):
    import torch
    from abt_amd import lib
    R = lib.load().abt_record_doubles()
    nb, ne = w["beams"]["ampl"].shape[0], w["ext"]["ampl"].shape[0]
    rb = torch.empty((nb, R), dtype=torch.float64, device=dev)
    re = torch.empty((ne, R), dtype=torch.float64, device=dev)
    lib.prepare_beams_dev(w["beams"], nb, lam, rb, fov_value=np.pi / 2)
    lib.prepare_beams_dev(w["ext"], ne, lam, re, fov_value=lib.PI_R4 / 2)
    return rb, re, nb, ne


def orientation_average(norient, steps, lam, dev, with_fp64=True, fixed_tiles=None, particle="rough_column",
                        grid_name="orient", with_kappa=False):
    """Batched orientation averaging: `norient` orientations (uniform on SO(3), seeded) of
    a Gaussian-rough hexagonal column (radius 10 um, length 20 um, sigma 0.5 um, 2112
    facets), 10 internal reflections, 721 theta x 180 phi bins.  The whole job (every
    orientation's beam and ext trees, per-orientation Mueller matrices summed
    incoherently, phi integral, per-orientation integrals) is one graph-captured
    persistent launch with the autotuned tile shape, in mixed precision.  Beside it: the
    plain FP64 drop-in path run as the reference drivers do (per orientation diff_main,
    finalise, summation), and the FP64 batched kernel as the accuracy reference.  With
    particle="large_column", grid_name="large": the absorbing 20,736-facet column (radius
    50 um, length 100 um, m 1.31+0.01i, 12 reflections) on 451 theta (0.05 deg to 5 deg,
    then 0.5 deg) x 360 phi bins."""
    import torch
    from abt_amd import lib
    from abt_amd.workloads import GRIDS, grid, orientation_workload
    t0 = time.time()
    w = orientation_workload(norient, particle, device=dev)
    torch.cuda.synchronize()
    boff, eoff = w["beam_off"].cpu().numpy(), w["ext_off"].cpu().numpy()
    counts = np.diff(boff) + np.diff(eoff)
    th, ph = grid(*GRIDS[grid_name])
    g = lib.DeviceGrid(th, ph, lam)
    rb, re, nb, ne = _prepare_trees(w, lam, dev)
    del w
    torch.cuda.synchronize()
    nbins = g.nbins
    bb = (nb + ne) * nbins
    log(f"orientation average: {norient} orientations, {nb} beams + {ne} ext "
        f"({counts.min()}-{counts.max()} per orientation), {bb:.4e} beam.bins, "
        f"generated in {time.time() - t0:.1f} s")
    f64 = dict(dtype=torch.float64, device=dev)
    res = {}

    def buffers():
        return (torch.empty(16 * nbins, **f64), torch.empty(6 * norient, **f64),
                torch.empty(16 * len(th), **f64))

    # mixed precision: autotune (or a fixed tile shape), capture, replay
    plan = lib.BatchPlan(g, boff, eoff, "mixed")
    if fixed_tiles:
        tb, bpt = (int(v) for v in fixed_tiles.lower().split("x"))
        plan.set_tiles(tb, bpt)
        tune = {}
    else:
        tune = plan.autotune(rb, re, sample=32)
    tiles = plan.tiles()
    m, I, m1 = buffers()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.capture(rb, re, m, I, m1, stream=s)
    plan.launch(stream=s)
    s.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        plan.launch(stream=s)
    e1.record(s)
    s.synchronize()
    ms = e0.elapsed_time(e1) / steps
    v_mixed = bb / (ms * 1e-3)
    log(f"  mixed batched (graph): {ms:.1f} ms/step, {v_mixed:.4e} beam.bin/s, tiles {tiles}")
    mm, Im = m.cpu().numpy().reshape(16, -1), I.cpu().numpy().reshape(-1, 6)
    res = {
        "workload": (f"{norient} orientations (uniform SO(3), seed 1) of a rough hexagonal column "
                     "(r 10 um, l 20 um, sigma 0.5 um, 2112 facets), m 1.31, 10 reflections, "
                     "721 theta x 180 phi bins, lambda 0.532 um") if particle == "rough_column" else
                    (f"{norient} orientations (uniform SO(3), seed 1) of the absorbing hexagonal column "
                     "(r 50 um, l 100 um, 20736 facets), m 1.31+0.01i, 12 reflections, "
                     f"{len(th)} theta x {len(ph)} phi bins, lambda 0.532 um"),
        "beams": int(nb), "ext_beams": int(ne), "beams_per_orientation": [int(counts.min()), int(counts.max())],
        "beam_bins": bb, "metric": "beam·bin diffractions/s", "value": v_mixed, "ms_per_step": ms,
        "steps": steps, "launch": "one hipGraph replay per step (persistent kernel + reductions)",
        "tiles": tiles, "autotune_ms": {f"{k[0]}x{k[1]}": v for k, v in tune.items()},
        "roofline": {"bound": "valu", "achieved": F_ALG * bb / (ms * 1e-3) / 1e12, "peak": 157.3,
                     "unit": "TFLOP/s", "frac": F_ALG * bb / (ms * 1e-3) / PEAK["mixed"],
                     "kernel": "k_batch<1,%d,%d>" % (tiles["beams_per_tile"], tiles["bins_per_thread"]),
                     "traffic": _traffic("k_batch<1,%d,%d>" % (tiles["beams_per_tile"], tiles["bins_per_thread"]))
                     if norient == 1000 else None},
    }
    if not with_fp64:
        plan.close()
        return res
    # FP64 batched (accuracy reference)
    plan64 = lib.BatchPlan(g, boff, eoff, "fp64")
    m64, I64, m164 = buffers()
    torch.cuda.synchronize()
    t = time.perf_counter()
    plan64.run(rb, re, m64, I64, m164)
    torch.cuda.synchronize()
    ms64b = (time.perf_counter() - t) * 1e3
    log(f"  fp64 batched: {ms64b:.1f} ms, {bb / (ms64b * 1e-3):.4e} beam.bin/s")
    M64, I64h = m64.cpu().numpy().reshape(16, -1), I64.cpu().numpy().reshape(-1, 6)
    # plain FP64 drop-in path, orientation by orientation
    ab, ae = torch.empty(8 * nbins, **f64), torch.empty(8 * nbins, **f64)
    mu, m1d, par = torch.empty(48 * nbins, **f64), torch.empty(48 * len(th), **f64), torch.empty(16, **f64)
    tot = torch.zeros(16 * nbins, **f64)
    ws = max(lib.far_field_workspace(g, int(max(np.diff(boff).max(), np.diff(eoff).max())), "fp64"), 8)
    work = torch.empty(ws // 8 + 1, **f64)
    fwork = torch.empty(lib.load().abt_finalise_workspace(g.handle) // 8 + 1, **f64)

    def plain(o):
        b0, b1, q0, q1 = int(boff[o]), int(boff[o + 1]), int(eoff[o]), int(eoff[o + 1])
        lib.far_field_dev(g, rb[b0:b1], b1 - b0, ab, precision="fp64", work=work)
        lib.far_field_dev(g, re[q0:q1], q1 - q0, ae, precision="fp64", work=work)
        lib.finalise_dev(g, ab, ae, mu, m1d, par, work=fwork)
        tot.add_(mu[:16 * nbins])

    plain(0)
    tot.zero_()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for o in range(norient):
        plain(o)
    torch.cuda.synchronize()
    dt_plain = time.perf_counter() - t
    v_plain = bb / dt_plain
    log(f"  fp64 plain loop: {dt_plain * 1e3:.1f} ms, {v_plain:.4e} beam.bin/s")
    Mp = tot.cpu().numpy().reshape(16, -1)
    kappa = None
    if with_kappa:
        # conditioning of the coherent sums: kappa^2 = sum_o sum_b |A_b|^2 / sum_o |sum_b A_b|^2 (FP64),
        # with sum_o |X_o|_F^2 = 2 M11 for X = k A (the Mueller matrix's 1/2 convention)
        incs = torch.zeros(nbins, **f64)
        tmp = torch.empty(nbins, **f64)
        for o in range(norient):
            for rr, oo in ((rb, boff), (re, eoff)):
                a0, a1 = int(oo[o]), int(oo[o + 1])
                if a1 > a0:
                    lib.far_field_incoherent_dev(g, rr[a0:a1], a1 - a0, tmp)
                    incs.add_(tmp)
        torch.cuda.synchronize()
        kappa = np.sqrt(incs.cpu().numpy() * (2 * lib.PI_R4 / lam) ** 2 / (2 * np.abs(M64[0])))
    s11 = np.abs(M64[0])
    norm = np.abs(mm - M64) / s11
    s11r = np.abs(mm[0] - M64[0]) / s11
    # the acceptance metric |dS_ij| / max(1e-3 |S11|, |S_ij|) (floor 1e-3 S11), per element
    flo = np.abs(mm - M64) / np.maximum(1e-3 * s11, np.abs(M64))
    q = lambda x: {"max": float(x.max()), "p999": float(np.quantile(x, 0.999))}
    bad_e = np.nonzero((flo[1:] > 1e-5).any(0))[0]
    bad_s = np.nonzero(s11r > 1e-6)[0]
    asym_m, asym_64 = Im[:, 1] / Im[:, 0], I64h[:, 1] / I64h[:, 0]
    res.update({
        "fp64_plain": {"value": v_plain, "ms": dt_plain * 1e3,
                       "path": "per orientation: abt_far_field_dev x2 + abt_finalise_dev + sum"},
        "fp64_batched": {"value": bb / (ms64b * 1e-3), "ms": ms64b},
        "speedup_vs_fp64_plain": v_mixed / v_plain,
        "accuracy": {"reference": "FP64 batched kernel on the same prepared beams",
                     "S11_rel": q(s11r), "norm_elem": q(norm),
                     "per_element_norm_max": [float(norm[e].max()) for e in range(16)],
                     "rel_floored": {"floor": "1e-3 S11", "per_element_max": [float(flo[e].max()) for e in range(16)],
                                     "per_element_p999": [float(np.quantile(flo[e], 0.999)) for e in range(16)]},
                     "flagged_bins": {"elements_gt_1e-5": int(len(bad_e)), "S11_gt_1e-6": int(len(bad_s)),
                                      "first": [int(x) for x in np.union1d(bad_e, bad_s)[:20]]},
                     **({"kappa": _kappa_report(kappa, flo, s11r, s11, th, ph)} if kappa is not None else {}),
                     "plain_vs_batched_fp64_norm_max": float((np.abs(Mp - M64) / s11).max()),
                     "asymmetry_abs_max": float(np.abs(asym_m - asym_64).max()),
                     "scatt_rel_max": float(np.abs(Im[:, 0] / I64h[:, 0] - 1).max())},
    })
    for x in (plan, plan64):
        x.close()
    return res


def _kappa_report(kappa, flo, s11r, s11, th, ph):
    """Errors against the conditioning kappa of each bin's coherent sums (a per-beam relative
    error eps shows up as ~eps kappa): quantiles of kappa, the share of flagged bins that are
    ill-conditioned (kappa > 10), the errors divided by max(kappa, 1) (the per-beam precision;
    kappa < 1 where the beams add coherently), the flagged bins by the depth of S11 below its
    largest value, and where the largest errors sit."""
    q = lambda x: {"max": float(x.max()), "p999": float(np.quantile(x, 0.999))}
    smax = s11.max()
    where = lambda i: {"theta_deg": float(np.degrees(th[i // len(ph)])), "phi_deg": float(np.degrees(ph[i % len(ph)])),
                       "kappa": float(kappa[i]), "S11_rel": float(s11r[i]), "S11_over_max": float(s11[i] / smax)}
    worst = flo[1:].max(0)
    flag = (worst > 1e-5) | (s11r > 1e-6)
    well = kappa <= 10
    return {"median": float(np.median(kappa)), "p99": float(np.quantile(kappa, 0.99)), "max": float(kappa.max()),
            "flagged_with_kappa_gt_10": int((flag & ~well).sum()), "flagged_with_kappa_le_10": int((flag & well).sum()),
            "well_conditioned_bins": int(well.sum()),
            "S11_rel_kappa_le_10": q(s11r[well]) if well.any() else None,
            "floored_kappa_le_10": q(worst[well]) if well.any() else None,
            "S11_rel_over_kappa": q(s11r / np.maximum(kappa, 1.0)),
            "floored_over_kappa": q(worst / np.maximum(kappa, 1.0)),
            "flagged_by_S11_level": {"S11_ge_1e-4_max": int((flag & (s11 >= 1e-4 * smax)).sum()),
                                     "S11_1e-6_to_1e-4_max": int((flag & (s11 < 1e-4 * smax) & (s11 >= 1e-6 * smax)).sum()),
                                     "S11_lt_1e-6_max": int((flag & (s11 < 1e-6 * smax)).sum())},
            "bins_by_S11_level": {"S11_ge_1e-4_max": int((s11 >= 1e-4 * smax).sum()),
                                  "S11_1e-6_to_1e-4_max": int(((s11 < 1e-4 * smax) & (s11 >= 1e-6 * smax)).sum()),
                                  "S11_lt_1e-6_max": int((s11 < 1e-6 * smax).sum())},
            "worst_S11_bins": [where(i) for i in np.argsort(-s11r)[:5]],
            "worst_S11_over_kappa_bins": [where(i) for i in np.argsort(-(s11r / np.maximum(kappa, 1.0)))[:5]]}


def adjoint_bench(norient, lam, dev):
    """Adjoint workload: the rough column of the orientation average, `norient`
    orientations, 361 theta x 90 phi bins; loss = weighted L2 of the orientation-averaged
    S11 (relative) and -S12/S11 against the particle with vertices perturbed by
    N(0, 0.05 um); gradients with respect to every particle vertex and every beam's
    amplitude matrix.  Times the forward (mixed batched, per-orientation amplitudes) and
    the backward (mixed adjoint kernel) on HIP events."""
    import torch
    from abt_amd import autograd as AG
    from abt_amd import lib
    from abt_amd.workloads import GRIDS, grid, hex_column, random_rotations, trace_orientations
    verts, faces = hex_column(10.0, 20.0, 8, 14, 0.5, seed=0)
    rots = random_rotations(norient, 1)
    t = trace_orientations(verts, faces, rots, rec=10, device=dev)
    rng = np.random.default_rng(77)
    vt = verts + rng.normal(scale=0.05, size=verts.shape)
    tt = trace_orientations(vt, faces, rots, rec=10, device=dev)
    th, ph = grid(*GRIDS["adjoint"])
    g = lib.DeviceGrid(th, ph, lam)
    mk = lambda tr, prec: AG.FarFieldSystem(g, tr["beam_off"].cpu().numpy(), tr["ext_off"].cpu().numpy(),
                                            {k: tr["beams"][k] for k in ("prop", "vk7")},
                                            {k: tr["ext"][k] for k in ("prop", "vk7")}, precision=prec)
    st = mk(tt, "mixed")
    Xt = st.forward_amplitudes(tt["beams"]["verts"].reshape(-1, 3, 3), tt["beams"]["ampl"],
                               tt["ext"]["verts"].reshape(-1, 3, 3), tt["ext"]["ampl"])
    t11, t12 = AG.mueller_11_12(Xt)
    tp = -t12 / t11
    w = AG.solid_angle_weights(th, ph, dev)
    facesT = torch.as_tensor(faces, device=dev)
    R = torch.as_tensor(rots, dtype=torch.float64, device=dev)
    nb, ne = int(t["beam_off"][-1]), int(t["ext_off"][-1])
    out = {"workload": f"{norient} orientations of the rough column, 361 theta x 90 phi bins, "
                       "loss: weighted L2 of orientation-averaged S11 and -S12/S11",
           "beams": nb, "ext_beams": ne, "beam_bins": (nb + ne) * g.nbins}
    for prec in ("mixed", "fp64"):
        s = mk(t, prec)
        s.timing = True
        V = torch.tensor(verts, dtype=torch.float64, device=dev, requires_grad=True)
        ab = t["beams"]["ampl"].clone().requires_grad_(True)
        ae = t["ext"]["ampl"].clone().requires_grad_(True)

        gathers = {tree: AG.ApertureGather(facesT, t[tree]["fi"], V.shape[0]) for tree in ("beams", "ext")}

        def apert(tree):
            return gathers[tree](V, R, t[tree]["orient"])

        times = {"forward": [], "backward": [], "forward_kernel": [], "adjoint_kernel": []}
        for rep in range(3):
            V.grad = None
            ab.grad = ae.grad = None
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            vb, ve = apert("beams"), apert("ext")
            torch.cuda.synchronize()
            e[0].record()
            X = AG.far_field(s, vb, ab, ve, ae)
            e[1].record()
            L = AG.retrieval_loss(X, t11, tp, w)
            torch.cuda.synchronize()
            e[2].record()
            L.backward()
            e[3].record()
            torch.cuda.synchronize()
            if rep:
                times["forward"].append(e[0].elapsed_time(e[1]))
                times["backward"].append(e[2].elapsed_time(e[3]))
                times["forward_kernel"].append(s.kernel_ms("forward"))
                times["adjoint_kernel"].append(s.kernel_ms("adjoint"))
        grads = (V.grad, ab.grad, ae.grad)
        f, b = float(np.mean(times["forward"])), float(np.mean(times["backward"]))
        fk, ak = float(np.mean(times["forward_kernel"])), float(np.mean(times["adjoint_kernel"]))
        out[prec] = {"forward_ms": f, "backward_ms": b, "backward_over_forward": b / f,
                     "forward_kernel_ms": fk, "adjoint_kernel_ms": ak, "adjoint_over_forward_kernel": ak / fk,
                     "loss": float(L.detach()), "grad_vertex_norm": float(grads[0].norm())}
        log(f"  adjoint {prec}: forward {f:.1f} ms, backward {b:.1f} ms, ratio {b / f:.2f}")
        if prec == "mixed":
            gm = grads[0].detach().cpu().numpy()
        else:
            g64 = grads[0].detach().cpu().numpy()
    out["mixed_vs_fp64_grad_rel"] = float(np.linalg.norm(gm - g64) / np.linalg.norm(g64))
    return out


def visibility_bench(nbeam, steps, warmup, dev, with_cpu=True):
    """find_vis_int for `nbeam` internal beams of the rough column (2112 facets), one random
    orientation per beam, in one batched evaluation per step; inputs resident on the GPU.
    Metric: facet visibility results per second (beams x facets / step time).  Roofline of the
    blocker scan k_vis_main: the reference's FP64 work per evaluation, 1 comparison per
    candidate pair at the depth test plus 7 FLOP per blocker edge it evaluates (edge normal 2,
    offset 2, dot product 3), counted by the kernel itself in an untimed pass, over the scan's
    event-timed duration, against the FP64 vector peak."""
    import torch
    from abt_amd import visibility as V
    from abt_amd.workloads import vis_workload
    t0 = time.time()
    wl = vis_workload(nbeam, "int", seed=1)
    topo, v, mid, nz, apn, ino, fin, bmz = wl
    log(f"visibility workload: {nbeam} internal beams, {topo.nf} facets ({time.time() - t0:.1f} s to generate)")
    t = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
    B = V.VisBatch("int", topo, t(v), t(mid), t(nz), t(apn), t(ino, torch.int64), t(fin, torch.int64), t(bmz))
    B.run(counts=True)
    torch.cuda.synchronize()
    pairs, edges = (int(x) for x in B.counts.cpu().numpy())
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        B.run()
    ev = []
    torch.cuda.synchronize()
    tw = time.perf_counter()
    for _ in range(steps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        B.run(phases=1)
        e[1].record(stream)
        B.run(phases=2)
        e[2].record(stream)
        ev.append(e)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - tw) / steps
    setup_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    scan_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    value = nbeam * topo.nf / dt
    flop = pairs + 7 * edges
    ach = flop / (scan_ms * 1e-3)
    log(f"visibility: {dt * 1e3:.3f} ms/step (set-up {setup_ms:.3f} ms, scan {scan_ms:.3f} ms), "
        f"{value:.4e} facet results/s; {pairs} pairs, {edges} edges")
    out = {"metric": "facet visibility results per second (find_vis_int)", "value": value,
           "unit": "facet/s", "ms_per_step": dt * 1e3, "steps": steps, "warmup": warmup, "dtype": "f64+int",
           "config": {"workload": f"find_vis_int, {nbeam} internal beams of the Gaussian-rough hexagonal column "
                                  f"(r 10 um, l 20 um, sigma 0.5 um, {topo.nf} facets), one uniform random "
                                  "orientation per beam (seed 1), illuminating facets: one up-facing aperture",
                      "beams": nbeam, "facets": topo.nf, "candidate_pairs": pairs, "blocker_edges": edges},
           "setup_ms": setup_ms, "scan_ms": scan_ms,
           "roofline": {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                        "frac": ach / PEAK["fp64"], "traffic": _traffic("k_vis_main<0,false>"),
                        "kernel": "k_vis_main<0,false>", "kernel_ms": scan_ms, "flop_per_launch": flop}}
    if with_cpu:
        out["cpu_baseline"] = vis_cpu_baseline(wl, min(nbeam, 1024), value)
    return out


def vis_cpu_baseline(wl, nsample, gpu_value):
    """The visibility oracle (oracle/vis_oracle.c, a C restatement of find_vis_int) on the
    host, one thread, over the first `nsample` beams of the same workload."""
    import oracle as O
    topo, v, mid, nz, apn, ino, fin, bmz = wl
    n = nsample
    sub = (topo, v[:n], mid[:n], nz[:n], apn[:n], ino[:n + 1], fin[:ino[n]], bmz[:n])
    t = time.perf_counter()
    O.find_vis_batch("int", *sub, nthreads=1)
    dt = time.perf_counter() - t
    val = n * topo.nf / dt
    log(f"visibility cpu baseline: {n} beams in {dt:.2f} s, {val:.4e} facet/s (1 thread)")
    return {"value": val, "unit": "facet/s", "cores": 1, "kind": "port",
            "sample": f"the first {n} beams of the same workload, find_vis_int restated in C, 1 thread",
            "gpu_over_cpu": gpu_value / val}


def cpu_baseline(beams, th, ph, lam, nsample):
    """The FP64 C oracle (OpenMP over theta columns) on the first `nsample` apertures."""
    import oracle
    from abt_amd.beams import BeamSet
    cores = min(16, os.cpu_count() or 1)
    sub = beams.subset(np.arange(min(nsample, beams.n)))
    t = time.perf_counter()
    oracle.diff_main(th, ph, lam, sub, BeamSet.empty(), nthreads=cores, raw=True)
    dt = time.perf_counter() - t
    bb = sub.n * len(th) * len(ph)
    return {"value": bb / dt, "unit": "beam·bin/s", "cores": cores, "kind": "port",
            "sample": f"first {sub.n} apertures of the stress set on the full 1801x360 grid "
                      f"({bb:.3e} beam.bins, {dt:.1f} s, FP64 C restatement of diff_mod, OpenMP)"}


if __name__ == "__main__":
    main()
