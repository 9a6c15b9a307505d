"""Benchmark: beam.bin diffractions/s of the far-field diffraction path on one MI355X.

Headline workload: the batched orientation average -- 1000 orientations (uniform on
SO(3), seeded) of a Gaussian-rough hexagonal column (r 10 um, l 20 um, sigma 0.5 um,
2112 facets, m 1.31, 10 internal reflections; 4.3M beams + 1.06M external-diffraction
beams), 721 theta x 180 phi = 129,780 bins, lambda 0.532 um.  One step = one replay of
the plan's HIP graph: every orientation's beam tree and external-diffraction tree
diffracted onto every bin, each orientation's Mueller matrix summed into the
orientation sum, its phi integral and the per-orientation cross sections and asymmetry
parameters.  `value` = (beams + ext beams) * bins / (seconds per step).

The headline plan is the one that meets the accuracy gate -- every element of the
orientation-summed Mueller matrix within |dS_ij| <= 1e-5 max(|S_ij|, 1e-3 S11) and S11
within 1e-6 of the FP64 reference -- checked in the same run against the plain FP64
drop-in kernel (the reference's formulas) run as the reference drivers run it (per
orientation diff_main + finalise, summed).  Records are prepared on the host with the
reference's operation order and libm (bit-identical to the oracle's frames) and are
resident in HBM before the timed region.

Sections in the same JSON line: the mixed-precision kernel (FP32 body) and the gated
mixed plan on the same workload; the 200,000-aperture stress test (one orientation,
1801 x 360 bins); the large absorbing column (a stated number of orientations); the
adjoint; the reference-facing host entry points per call on the reference's CPU case;
facet visibility; CPU baselines (the C restatement of the reference, OpenMP, on bounded
samples).
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
PEAK_OF = {"mixed": "mixed", "mixed_gated": "mixed", "fp64": "fp64", "fp64_fast": "fp64"}
METRIC = "beam·bin diffractions/s mixed-prec vs FP64 at ≤1e-5 Mueller rel err"
LAM = 0.532
GATE = {"S11": 1e-6, "elem": 1e-5, "floor": 1e-3}
TILES = None             # (beams per tile, bins per lane) from --tiles: fixed shape, no autotuning


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _traffic(kname):
    """HBM bytes per launch of a kernel on its default workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by the copy-kernel calibration."""
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tpath):
        return None
    return json.load(open(tpath)).get(kname, {}).get("bytes_per_launch")


def gate_report(M, M64, th=None, ph=None):
    """M against the FP64 reference M64 (16 x nbins) under the accuracy gate: per element the
    floored relative error |dS_ij| / max(|S_ij|, 1e-3 S11) (max, p99.9), S11's relative error,
    the normalised error |dS_ij| / S11, and the bins that fail."""
    s11 = np.abs(M64[0])
    d = np.abs(M - M64)
    flo = d / np.maximum(np.abs(M64), GATE["floor"] * s11)
    s11r = d[0] / s11
    fail = (flo[1:] > GATE["elem"]).any(0) | (s11r > GATE["S11"])
    q = lambda x: {"max": float(x.max()), "p999": float(np.quantile(x, 0.999))}
    out = {"S11_rel": q(s11r), "floored_elem": q(flo[1:]), "norm_elem": q(d[1:] / s11),
           "floored_per_element_max": [float(flo[e].max()) for e in range(16)],
           "failing_bins": int(fail.sum()), "bins": int(len(s11)), "meets_gate": bool(not fail.any())}
    if th is not None and fail.any():
        idx = np.nonzero(fail)[0][:8]
        out["first_failing"] = [{"theta_deg": float(np.degrees(th[i // len(ph)])),
                                 "phi_deg": float(np.degrees(ph[i % len(ph)]))} for i in idx]
    return out


def _records_host(beams, lam, fov_value, dev):
    """Host preparation (reference operation order, glibc), then resident on the device."""
    import torch
    from abt_amd import lib
    recs, vrecs = lib.prepare_beams_host(beams, lam, fov_value=fov_value)
    poly = beams.n > 0 and bool((np.diff(beams.vert_off) > 3).any())   # vertex records only for polygons
    return torch.from_numpy(recs).to(dev), (torch.from_numpy(vrecs).to(dev) if poly else None)


class OrientationCase:
    """A multi-orientation workload resident on the device: records of both trees, grid,
    offsets, and the plain FP64 drop-in loop as the accuracy reference."""

    def __init__(self, norient, particle, grid_name, dev):
        import torch
        from abt_amd import lib
        from abt_amd.workloads import GRIDS, grid, orientation_workload, to_beamset
        t0 = time.time()
        w = orientation_workload(norient, particle, device=dev)
        self.boff, self.eoff = w["beam_off"].cpu().numpy(), w["ext_off"].cpu().numpy()
        B, E = to_beamset(w["beams"]), to_beamset(w["ext"])
        del w
        torch.cuda.empty_cache()
        self.th, self.ph = grid(*GRIDS[grid_name])
        self.g = lib.DeviceGrid(self.th, self.ph, LAM)
        self.rb, self.vb = _records_host(B, LAM, np.pi / 2, dev)
        self.re, self.ve = _records_host(E, LAM, lib.PI_R4 / 2, dev)
        self.B, self.E = B, E
        self.nb, self.ne = B.n, E.n
        self.norient = norient
        self.nbins = self.g.nbins
        self.bb = (self.nb + self.ne) * self.nbins
        self.dev = dev
        counts = np.diff(self.boff) + np.diff(self.eoff)
        self.counts = (int(counts.min()), int(counts.max()))
        log(f"{particle}: {norient} orientations, {self.nb} beams + {self.ne} ext ({counts.min()}-{counts.max()} "
            f"per orientation), {len(self.th)} x {len(self.ph)} bins, {self.bb:.4e} beam.bins "
            f"({time.time() - t0:.1f} s to trace and prepare)")

    def buffers(self):
        import torch
        f64 = dict(dtype=torch.float64, device=self.dev)
        return (torch.empty(16 * self.nbins, **f64), torch.empty(6 * self.norient, **f64),
                torch.empty(16 * len(self.th), **f64))

    def plan(self, precision, autotune=True):
        """A batched plan, autotuned on 32 orientations, or with the tile shape of --tiles
        (beams per tile x bins per lane) where that shape exists for the precision."""
        from abt_amd import lib
        p = lib.BatchPlan(self.g, self.boff, self.eoff, precision)
        tune = {}
        if autotune and TILES and precision != "fp64_accurate":
            tb, bpt = TILES
            p.set_tiles(tb, bpt if precision in ("mixed", "mixed_gated") else 1)
        elif autotune:
            tune = p.autotune(self.rb, self.re, self.vb, self.ve, sample=32)
        return p, tune

    def timed_graph(self, precision, steps, warmup, with_kernel_ms=True):
        """Replays of the plan's graph: (ms per step, kernel ms from HIP events on the plan's
        stream around an eager run, Mueller sum, integrals, plan info)."""
        import torch
        p, tune = self.plan(precision)
        m, I, m1 = self.buffers()
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            p.capture(self.rb, self.re, m, I, m1, self.vb, self.ve, stream=s)
        for _ in range(max(warmup, 1)):
            p.launch(stream=s)
        s.synchronize()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            p.launch(stream=s)
        s.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / steps
        kms = None
        if with_kernel_ms:
            # the persistent kernel and its short reductions, eager on the same stream
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            p.run(self.rb, self.re, m, I, m1, self.vb, self.ve, stream=s)
            e1.record(s)
            s.synchronize()
            kms = e0.elapsed_time(e1)
        info = {"tiles": p.tiles(), "autotune_ms": {f"{k[0]}x{k[1]}": v for k, v in tune.items() if v}}
        if precision == "mixed_gated":
            info["refined_bins"] = p.refined_bins(stream=s)
        M = m.cpu().numpy().reshape(16, -1)
        Ih = I.cpu().numpy().reshape(-1, 6)
        p.close()
        return ms, kms, M, Ih, info

    def plain_fp64(self):
        """The plain FP64 drop-in path orientation by orientation, as the reference drivers
        run it: per orientation abt_far_field_dev for both trees, abt_finalise_dev, sum."""
        import torch
        from abt_amd import lib
        f64 = dict(dtype=torch.float64, device=self.dev)
        nbins = self.nbins
        ab, ae = torch.empty(8 * nbins, **f64), torch.empty(8 * nbins, **f64)
        mu, m1d, par = torch.empty(48 * nbins, **f64), torch.empty(48 * len(self.th), **f64), torch.empty(16, **f64)
        ws = max(lib.far_field_workspace(self.g, int(max(np.diff(self.boff).max(), np.diff(self.eoff).max())),
                                         "fp64"), 8)
        work = torch.empty(ws // 8 + 1, **f64)
        fwork = torch.empty(lib.load().abt_finalise_workspace(self.g.handle) // 8 + 1, **f64)
        tot = torch.zeros(16 * nbins, **f64)
        params = []

        def one(o):
            b0, b1, q0, q1 = int(self.boff[o]), int(self.boff[o + 1]), int(self.eoff[o]), int(self.eoff[o + 1])
            lib.far_field_dev(self.g, self.rb[b0:b1], b1 - b0, ab, vrecs=self.vb, precision="fp64", work=work)
            lib.far_field_dev(self.g, self.re[q0:q1], q1 - q0, ae, vrecs=self.ve, precision="fp64", work=work)
            lib.finalise_dev(self.g, ab, ae, mu, m1d, par, work=fwork)
            tot.add_(mu[:16 * nbins])

        one(0)
        tot.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for o in range(self.norient):
            one(o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        return dt * 1e3, tot.cpu().numpy().reshape(16, -1)


def orientation_section(case, precision, steps, warmup, M64, plain_ms, gate=True):
    ms, kms, M, Ih, info = case.timed_graph(precision, steps, warmup)
    value = case.bb / (ms * 1e-3)
    peak = PEAK[PEAK_OF[precision]]
    kname = {"fp64_fast": "k_batch<4,%d,1>", "mixed": "k_batch<1,%d,%d>",
             "mixed_gated": "k_batch<1,%d,%d> (estimate) + k_batch<4,16,1> (listed bins)"}[precision]
    t = info["tiles"]
    kname = kname % ((t["beams_per_tile"],) if precision == "fp64_fast" else (t["beams_per_tile"], t["bins_per_thread"]))
    out = {"value": value, "unit": "beam·bin/s", "ms_per_step": ms, "steps": steps,
           "speedup_vs_fp64_plain": plain_ms / ms if np.isfinite(plain_ms) else None, **info,
           "roofline": {"bound": "valu", "achieved": F_ALG * case.bb / (kms * 1e-3) / 1e12, "peak": peak / 1e12,
                        "unit": "TFLOP/s", "frac": F_ALG * case.bb / (kms * 1e-3) / peak,
                        "traffic": _traffic(kname), "kernel": kname, "kernel_ms": kms,
                        "flop_per_beam_bin": F_ALG, "beam_bins_per_launch": case.bb}}
    if M64 is not None:
        out["accuracy"] = gate_report(M, M64, case.th, case.ph)
    log(f"  {precision}: {ms:.1f} ms/step, {value:.4e} beam.bin/s"
        + (f", {plain_ms / ms:.2f}x the FP64 plain loop" if np.isfinite(plain_ms) else "")
        + (f", gate {'met' if out['accuracy']['meets_gate'] else 'MISSED'} "
           f"(floored max {out['accuracy']['floored_elem']['max']:.2e}, S11 {out['accuracy']['S11_rel']['max']:.2e})"
           if M64 is not None else "") + (f", refined {info['refined_bins']} bins" if "refined_bins" in info else ""))
    return out, M, Ih


def cpu_baseline_orient(case, nsample_orient):
    """The C restatement of diff_main + finalise (oracle/, FP64, OpenMP over theta columns) on
    the first orientations of the workload, on the full grid."""
    import oracle
    cores = min(16, os.cpu_count() or 1)
    bb = 0
    t = time.perf_counter()
    for o in range(nsample_orient):
        b = case.B.subset(np.arange(case.boff[o], case.boff[o + 1]))
        e = case.E.subset(np.arange(case.eoff[o], case.eoff[o + 1]))
        rb, re = oracle.diff_main(case.th, case.ph, LAM, b, e, nthreads=cores, raw=True)
        oracle.finalise(case.th, case.ph, LAM, rb, re)
        bb += (b.n + e.n) * case.nbins
    dt = time.perf_counter() - t
    return {"value": bb / dt, "unit": "beam·bin/s", "cores": cores, "kind": "port",
            "sample": f"the first {nsample_orient} orientations of the same workload on the full grid "
                      f"({bb:.3e} beam.bins, {dt:.1f} s): diff_main + finalise of the C restatement of the "
                      "reference (oracle/diff_oracle.c; the reference itself needs a stand-in for GNU rand() to "
                      "build here, so it is not run)"}


def stress_section(nbeam, steps, lam, dev, cpu_sample):
    """The 200,000-aperture stress test: one orientation, 1801 x 360 bins.  Per precision
    one step = abt_far_field_dev over every aperture + abt_finalise_dev (records prepared
    on the host, resident in HBM); accuracy of the Mueller matrix against the FP64 kernel."""
    import torch
    from abt_amd import lib
    from abt_amd.beams import BeamSet
    from abt_amd.workloads import GRIDS, grid, stress_beams
    th, ph = grid(*GRIDS["stress"])
    t0 = time.time()
    beams = stress_beams(nbeam)
    g = lib.DeviceGrid(th, ph, lam)
    recs, vrecs = _records_host(beams, lam, None, dev)
    nbins = g.nbins
    log(f"stress: {nbeam} apertures, {len(th)} x {len(ph)} bins ({time.time() - t0:.1f} s to generate and prepare)")
    f64 = dict(dtype=torch.float64, device=dev)
    amp, amp_ext = torch.empty(8 * nbins, **f64), torch.zeros(8 * nbins, **f64)
    mu, m1d, par = torch.empty(48 * nbins, **f64), torch.empty(48 * len(th), **f64), torch.empty(16, **f64)
    ws = lib.far_field_workspace(g, nbeam, "fp64")
    work = torch.empty(max(ws, 8) // 8 + 1, **f64)
    fwork = torch.empty(lib.load().abt_finalise_workspace(g.handle) // 8 + 1, **f64)
    st = torch.cuda.current_stream()
    out = {"workload": f"diffraction stress test: {nbeam} seeded triangular apertures (edge directions uniform "
                       f"in the aperture plane), {len(th)} theta x {len(ph)} phi bins, lambda {lam} um, one orientation",
           "beams": nbeam, "bins": nbins, "beam_bins": nbeam * nbins}
    mus = {}
    for prec in ("mixed", "fp64_fast", "fp64"):
        kms = []
        for rep in range(steps + 1):
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record(st)
            lib.far_field_dev(g, recs, nbeam, amp, vrecs=vrecs, precision=prec, work=work, stream=st)
            e1.record(st)
            amp_ext.zero_()
            lib.finalise_dev(g, amp, amp_ext, mu, m1d, par, work=fwork, stream=st)
            e2.record(st)
            torch.cuda.synchronize()
            if rep:   # the first launch warms up
                kms.append((e0.elapsed_time(e1), e0.elapsed_time(e2)))
        k = float(np.mean([a for a, _ in kms]))
        s = float(np.mean([b for _, b in kms]))
        peak = PEAK[PEAK_OF[prec]]
        kname = {"mixed": "k_far_field<1,16>", "fp64_fast": "k_far_field<4,16>", "fp64": "k_far_field<0,16>"}[prec]
        out[prec] = {"value": nbeam * nbins / (s * 1e-3), "ms_per_step": s, "kernel_ms": k,
                     "roofline": {"bound": "valu", "achieved": F_ALG * nbeam * nbins / (k * 1e-3) / 1e12,
                                  "peak": peak / 1e12, "unit": "TFLOP/s",
                                  "frac": F_ALG * nbeam * nbins / (k * 1e-3) / peak, "kernel": kname,
                                  "traffic": _traffic(kname) if nbeam == 200_000 else None}}
        mus[prec] = mu[:16 * nbins].cpu().numpy().reshape(16, -1)
        log(f"  stress {prec}: {s:.1f} ms/step (kernel {k:.1f} ms), {nbeam * nbins / (s * 1e-3):.4e} beam.bin/s")
    for prec in ("mixed", "fp64_fast"):
        out[prec]["speedup_vs_fp64"] = out["fp64"]["ms_per_step"] / out[prec]["ms_per_step"]
        out[prec]["accuracy"] = gate_report(mus[prec], mus["fp64"])
    # the frame-free FP64 model (no kxx/kyy clamps, no split-sum cancellation; a one-orientation
    # batched plan) as a second reference for all three
    p = lib.BatchPlan(g, [0, nbeam], [0, 0], "fp64_accurate")
    mm, ii, m1m = torch.empty(16 * nbins, **f64), torch.empty(6, **f64), torch.empty(16 * len(th), **f64)
    t = time.perf_counter()
    p.run(recs, None, mm, ii, m1m, vrecs, None)
    torch.cuda.synchronize()
    Mm = mm.cpu().numpy().reshape(16, -1)
    out["vs_fp64_model"] = {"model": f"batched plan precision fp64_accurate (adjbody.h), one orientation, "
                                     f"{(time.perf_counter() - t) * 1e3:.0f} ms",
                            **{prec: gate_report(mus[prec], Mm, th, ph) for prec in ("fp64", "fp64_fast", "mixed")}}
    p.close()
    log("  stress against the FP64 model: " + ", ".join(
        f"{k} floored max {v['floored_elem']['max']:.2e} ({v['failing_bins']} failing bins)"
        for k, v in out["vs_fp64_model"].items() if k != "model"))
    if cpu_sample > 0:
        import oracle
        cores = min(16, os.cpu_count() or 1)
        sub = beams.subset(np.arange(min(cpu_sample, nbeam)))
        t = time.perf_counter()
        oracle.diff_main(th, ph, lam, sub, BeamSet.empty(), nthreads=cores, raw=True)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": sub.n * nbins / dt, "unit": "beam·bin/s", "cores": cores, "kind": "port",
                               "sample": f"first {sub.n} apertures on the full grid ({sub.n * nbins:.3e} beam.bins, "
                                         f"{dt:.1f} s): diff_main of the C restatement of the reference, OpenMP"}
    return out


def dropin_section(lam, dev):
    """The reference-facing host entry points on the reference's CPU-runnable case: the
    default hexagonal column (r 5 um, l 10 um, -tri) at Euler (0, 30, 0), 8 reflections,
    361 x 180 bins.  abt_diff_main per call (host beams in, host amplitudes out: the
    "diffraction took" span of diff_mod.f90:818-822) and abt_diffraction per beam (the
    -memory mode call, beam_loop_mod.f90:37-51), after the first call has cached the grid."""
    from abt_amd import lib
    from abt_amd.workloads import GRIDS, column_beams, euler_matrix, grid, hex_column
    v, f = hex_column(5.0, 10.0, 6, 12, 0.0)
    b, e, _ = column_beams(v, f, euler_matrix(0, 30, 0), rec=8)
    th, ph = grid(*GRIDS["small"])
    x, y, z = lib.make_far_field_bins(th, ph)
    nbins = len(th) * len(ph)
    out = {"workload": f"default hexagonal column (r 5 um, l 10 um) at Euler (0, 30, 0), 8 reflections: "
                       f"{b.n} beams + {e.n} ext beams, {len(th)} x {len(ph)} bins", "beams": b.n, "ext_beams": e.n}
    for prec in ("fp64", "fp64_fast", "mixed"):
        lib.diff_main(th, ph, lam, b, e, precision=prec, raw=True)
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            lib.diff_main(th, ph, lam, b, e, precision=prec, raw=True)
        dm = (time.perf_counter() - t) / reps
        lib.diffraction(th, ph, x, y, z, lam, b, 0, precision=prec)
        t = time.perf_counter()
        nd = 50
        for i in range(nd):
            lib.diffraction(th, ph, x, y, z, lam, b, i % b.n, precision=prec)
        dd = (time.perf_counter() - t) / nd
        out[prec] = {"diff_main_ms": dm * 1e3, "diff_main_beam_bins_per_s": (b.n + e.n) * nbins / dm,
                     "diffraction_ms_per_beam": dd * 1e3}
        log(f"  drop-in {prec}: diff_main {dm * 1e3:.1f} ms, diffraction {dd * 1e3:.2f} ms per beam")
    return out



def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed steps of the headline plan")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--orient", type=int, default=1000, help="orientations of the headline workload")
    ap.add_argument("--precision", default="fp64_fast", choices=["fp64_fast", "mixed_gated", "mixed"],
                    help="headline plan")
    ap.add_argument("--section-steps", type=int, default=2, help="timed steps of the other sections")
    ap.add_argument("--others", default="mixed,mixed_gated", help="other plans timed on the headline workload")
    ap.add_argument("--plain", type=int, default=1, help="run the plain FP64 loop (speedup and accuracy reference)")
    ap.add_argument("--beams", type=int, default=200_000, help="apertures of the stress test (0: skip)")
    ap.add_argument("--large-orient", type=int, default=24,
                    help="orientations of the large absorbing column (20,736 facets; 0: skip)")
    ap.add_argument("--adjoint-orient", type=int, default=64, help="orientations of the adjoint workload (0: skip)")
    ap.add_argument("--dropin", type=int, default=1, help="time the reference-facing host entry points (0: skip)")
    ap.add_argument("--vis-beams", type=int, default=4096, help="beams of the facet-visibility workload (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=2,
                    help="orientations of the CPU baseline sample (0: skip; the stress test samples 96 apertures)")
    ap.add_argument("--tiles", default=None,
                    help="fixed batched tile shape TBxBPT (e.g. 8x1) instead of autotuning (profiling runs)")
    ap.add_argument("--json-out", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the multi-GPU path "
              "(orientations sharded by beam count, one fixed-order reduction of the Mueller partials) "
              "is a design, not built.", file=sys.stderr)
        sys.exit(2)
    if args.tiles:
        global TILES
        TILES = tuple(int(v) for v in args.tiles.lower().split("x"))
    import torch
    dev = torch.device("cuda:0")
    out = {"metric": METRIC, "unit": "beam·bin/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    sections = {}
    if args.orient > 0:
        case = OrientationCase(args.orient, "rough_column", "orient", dev)
        plain_ms, M64 = case.plain_fp64() if args.plain else (None, None)
        if plain_ms:
            log(f"  fp64 plain loop: {plain_ms:.1f} ms, {case.bb / (plain_ms * 1e-3):.4e} beam.bin/s")
        head, Mh, Ih = orientation_section(case, args.precision, args.steps, args.warmup, M64, plain_ms or np.nan)
        out.update({"value": head.pop("value"), "ms_per_step": head.pop("ms_per_step"),
                    "dtype": {"fp64_fast": "f64", "mixed": "f32+f64", "mixed_gated": "f32+f64"}[args.precision]})
        head.pop("unit"); head.pop("steps")
        out["config"] = {"workload": f"batched orientation average: {args.orient} orientations (uniform SO(3), "
                                     "seed 1) of a Gaussian-rough hexagonal column (r 10 um, l 20 um, sigma 0.5 um, "
                                     "2112 facets), m 1.31, 10 reflections, 721 theta x 180 phi bins, lambda 0.532 um",
                         "beams": case.nb, "ext_beams": case.ne, "beams_per_orientation": list(case.counts),
                         "bins": case.nbins, "beam_bins": case.bb, "precision": args.precision,
                         "step": "one HIP-graph replay of the batched plan (persistent kernel + reductions)",
                         "records": "prepared on the host (reference operation order and libm), resident in HBM"}
        out["roofline"] = head.pop("roofline")
        out.update(head)
        if plain_ms:
            out["fp64_plain"] = {"value": case.bb / (plain_ms * 1e-3), "ms": plain_ms,
                                 "path": "per orientation: abt_far_field_dev x2 (ABT_FP64) + abt_finalise_dev + sum"}
        for prec in [p for p in args.others.split(",") if p and p != args.precision]:
            sec, _, _ = orientation_section(case, prec, args.section_steps, 1, M64, plain_ms or np.nan)
            sections[prec] = sec
        if args.cpu_sample > 0:
            out["cpu_baseline"] = cpu_baseline_orient(case, args.cpu_sample)
            log(f"  cpu baseline: {out['cpu_baseline']['value']:.4e} beam.bin/s on {out['cpu_baseline']['cores']} cores")
        del case
        torch.cuda.empty_cache()
    if args.beams > 0:
        sections["stress"] = stress_section(args.beams, args.section_steps, LAM, dev, 2000 if args.cpu_sample else 0)
        torch.cuda.empty_cache()
    if args.large_orient > 0:
        case = OrientationCase(args.large_orient, "large_column", "large", dev)
        plain_ms, M64 = case.plain_fp64()
        log(f"  large column fp64 plain loop: {plain_ms:.1f} ms")
        lc = {"workload": f"{args.large_orient} orientations (uniform SO(3), seed 1) of the absorbing hexagonal "
                          "column (r 50 um, l 100 um, 20736 facets), m 1.31+0.01i, 12 reflections, "
                          f"{len(case.th)} theta (0.05 deg to 5 deg, then 0.5 deg) x {len(case.ph)} phi bins",
              "beams": case.nb, "ext_beams": case.ne, "beam_bins": case.bb,
              "fp64_plain": {"value": case.bb / (plain_ms * 1e-3), "ms": plain_ms}}
        Ms = {}
        for prec in ("fp64_fast", "mixed"):
            lc[prec], Ms[prec], _ = orientation_section(case, prec, args.section_steps, 1, M64, plain_ms)
        # the frame-free FP64 model (no kxx/kyy clamps, no split-sum cancellation) as a second
        # reference: on this particle S11 spans 11 decades and the plain kernel's formula is
        # not exact in the deepest minima
        p, _ = case.plan("fp64_accurate", autotune=False)
        m, I, m1 = case.buffers()
        t = time.perf_counter()
        p.run(case.rb, case.re, m, I, m1, case.vb, case.ve)
        torch.cuda.synchronize()
        Mm = m.cpu().numpy().reshape(16, -1)
        lc["vs_fp64_model"] = {"model": "batched plan precision fp64_accurate (adjbody.h), "
                                        f"{(time.perf_counter() - t) * 1e3:.0f} ms",
                               "fp64_plain": gate_report(M64, Mm, case.th, case.ph),
                               "fp64_fast": gate_report(Ms["fp64_fast"], Mm, case.th, case.ph),
                               "mixed": gate_report(Ms["mixed"], Mm, case.th, case.ph)}
        log("  large column against the FP64 model: " + ", ".join(
            f"{k} floored max {v['floored_elem']['max']:.2e} ({v['failing_bins']} failing bins)"
            for k, v in lc["vs_fp64_model"].items() if k != "model"))
        p.close()
        sections["large_column"] = lc
        del case
        torch.cuda.empty_cache()
    if args.adjoint_orient > 0:
        sections["adjoint"] = adjoint_bench(args.adjoint_orient, LAM, dev)
    if args.dropin:
        sections["dropin"] = dropin_section(LAM, dev)
    if args.vis_beams > 0:
        sections["visibility"] = visibility_bench(args.vis_beams, 20, 2, dev, args.cpu_sample > 0)
    out.update(sections)
    line = json.dumps(out)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)


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
        out["cpu_baseline"] = vis_cpu_baseline(wl, min(nbeam, 4096), value)
    return out


def vis_cpu_baseline(wl, nsample, gpu_value):
    """The visibility oracle (oracle/vis_oracle.c, a C restatement of find_vis_int) on the
    host's cores (beams spread over OpenMP threads), over the first `nsample` beams of the
    same workload."""
    import oracle as O
    topo, v, mid, nz, apn, ino, fin, bmz = wl
    n = min(nsample, len(nz))
    cores = min(16, os.cpu_count() or 1)
    sub = (topo, v[:n], mid[:n], nz[:n], apn[:n], ino[:n + 1], fin[:ino[n]], bmz[:n])
    reps = 40   # ~10 s of host work
    t = time.perf_counter()
    for _ in range(reps):
        O.find_vis_batch("int", *sub, nthreads=cores)
    dt = time.perf_counter() - t
    val = reps * n * topo.nf / dt
    log(f"visibility cpu baseline: {reps} x {n} beams in {dt:.2f} s, {val:.4e} facet/s ({cores} threads)")
    return {"value": val, "unit": "facet/s", "cores": cores, "kind": "port",
            "sample": f"the first {n} beams of the same workload, {reps} evaluations, find_vis_int restated in C "
                      f"(oracle/vis_oracle.c; the reference's tracer is not built here), {cores} threads",
            "gpu_over_cpu": gpu_value / val}



if __name__ == "__main__":
    main()
