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
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and nb == 200_000:
        # HBM bytes per launch of this kernel on this workload, from the committed
        # rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by the copy-kernel calibration
        traffic = json.load(open(tpath)).get(kname, {}).get("bytes_per_launch")
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
    line = json.dumps(out)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)


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
