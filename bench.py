#!/usr/bin/env python3
"""Per-gulp search throughput on one MI355X.

One step = one hd_execute_device call (the whole per-gulp search: unpack,
dedispersion over the DM trials, baseline, RMS, normalisation, prefix sums,
boxcars 1..boxcar_max with block RMS, giant finding, clustering, merging and
the candidate read-back) over one synthetic 1024-channel gulp already
resident in HBM (262144 samples of 64 us, DM 0-2000).

Prints one JSON line:
  value          beam-seconds searched per second (graph replays, input in HBM)
  roofline       the whole gulp against HBM: algorithmic bytes B_in + 2P (the
                 input once, the DM plane written once and read once) over the
                 gulp time, with the HBM bytes rocprofv3 measured per gulp
  dominant_kernel the dedispersion kernel: integer adds against the int32 VALU
                 peak (it is compute-bound), its LDS read rate, its HBM view
  kernels_hbm    per-kernel HBM GB/s from the committed rocprofv3 counters
  host_path      hd_execute from page-locked host memory (the rate including
                 the PCIe copy; never the headline), and the gulp driver's
                 double buffering that hides the copy behind the search
  adaptive_dt    the same gulp with use_scrunching on (the reference's default
                 search: trials at their scrunch factor)
  batch          the ragged 352-beam batch (320 channels, 655.36 us) in one
                 hd_execute_batch per step, with its own roofline
  cpu_baseline   the CPU oracle on a bounded sample of the same gulp;
  cpu_baseline_small the CPU oracle on whole 512-channel gulps
  small_gpu      the same 512-channel gulps searched on the GPU
  proposal       proposal-gated vs full batched search on 4096-channel beams
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
VALU_INT_PEAK = 256 * 64 * 2.4e9  # int32 lane adds/s (256 CUs x 64 lanes x 2.4 GHz)
# LDS read rate of ds_read2st64_b32, measured by a standalone LDS probe
# (profiles/lds_read_bandwidth_probe.txt): 98.9 B/clk/CU at 2.4 GHz
LDS_READ2_PEAK = 98.9 * 256 * 2.4e9  # bytes/s


def host_cpu():
    """CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def cpu_threads():
    """Oracle threads: the cores available, at most 16 (one GPU's host share)."""
    return max(1, min(16, host_cpu()[1]))


def tile_json(t):
    """The dedispersion tile the autotuner chose (Pipeline.dedisp_tile())."""
    return {"trials_per_block": t[0], "xcd_remap": t[1], "byte_lanes": t[2],
            "channels_per_step": t[3], "block_threads": t[4],
            "chosen_by": "autotuner (first call of the shape)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="full_1024ch")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-stride", type=int, default=2,
                    help="CPU baseline: time every k-th DM trial of the same gulp")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-proposal", action="store_true",
                    help="skip the proposal-gated vs full search comparison")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the ragged-batch block of the default line")
    ap.add_argument("--no-adaptive", action="store_true",
                    help="skip the adaptive time resolution block")
    ap.add_argument("--no-host-path", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) timing")
    ap.add_argument("--all-ops", action="store_true",
                    help="gated_4096ch workload: every operating point, not only the default")
    ap.add_argument("--beams", type=int, default=0,
                    help="batch workloads: number of beams (default: the workload's)")
    ap.add_argument("--cpu-beams", type=int, default=16,
                    help="batch workloads: beams timed for the CPU baseline")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before the gulps, run one device-to-device copy of this many MiB "
                         "(the reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def load_kernel_traffic(name):
    """Per-kernel HBM bytes per call of the workload, from the committed
    rocprofv3 counter summary profiles/hbm_traffic_<workload>.txt (lines
    'kernel=<name> calls_per_step=<n> avg_us=<t> fetch_bytes=<f> write_bytes=<w>',
    bytes per launch, already scaled by the copy-kernel calibration)."""
    path = os.path.join(HERE, "profiles", "hbm_traffic_%s.txt" % name)
    if not os.path.exists(path):
        return None
    rows = []
    for line in open(path):
        if line.startswith("#") or "kernel=" not in line:
            continue
        d = {}
        for tok in line.split():
            if "=" in tok:
                k, v = tok.split("=", 1)
                d[k] = v
        try:
            rows.append({"kernel": d["kernel"], "calls_per_step": float(d["calls_per_step"]),
                         "avg_us": float(d["avg_us"]), "fetch_bytes": float(d["fetch_bytes"]),
                         "write_bytes": float(d["write_bytes"])})
        except (KeyError, ValueError):
            continue
    return rows or None


def kernels_hbm(rows):
    """Per-kernel HBM rate (measured bytes over measured time) and the bytes
    of one step."""
    if not rows:
        return None, None
    table = []
    per_step = 0.0
    for r in rows:
        b = r["fetch_bytes"] + r["write_bytes"]
        per_step += b * r["calls_per_step"]
        table.append({"kernel": r["kernel"], "avg_us": r["avg_us"],
                      "hbm_bytes": b, "hbm_gbs": round(b / (r["avg_us"] * 1e-6) / 1e9, 1),
                      "frac": round(b / (r["avg_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)})
    table.sort(key=lambda e: -e["avg_us"] * 1.0)
    return table, per_step


def gulp_roofline(b_alg, t_step, traffic):
    achieved = b_alg / t_step / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "scope": "whole step: algorithmic bytes B_in + 2P (input read once, DM plane "
                     "written once and read once) over the measured step time; traffic = "
                     "HBM bytes rocprofv3 measured per step (sum over every kernel)",
            "algorithmic_bytes": int(b_alg), "step_ms": round(t_step * 1e3, 4)}


def dedisp_block(name, t_dd, adds, b_in, plane, tile, t_step, rows):
    """The dedispersion kernel: integer-add bound (one exact add per sample,
    channel and trial), with its LDS and HBM views."""
    traffic = None
    if rows:
        # every launch of the stage (double- and single-buffered DM blocks)
        dd = [r for r in rows if r["kernel"].startswith("k_dedisperse")]
        if dd:
            traffic = sum((r["fetch_bytes"] + r["write_bytes"]) * r["calls_per_step"] for r in dd)
    per = 1 if tile[2] else 2     # LDS bytes read per sample-add (byte or 16-bit staging)
    return {"kernel": name, "ms": round(t_dd * 1e3, 4), "share_of_step": round(t_dd / t_step, 4),
            "bound": "valu", "unit": "int32 adds/s", "achieved": adds / t_dd,
            "peak": VALU_INT_PEAK, "frac": round(adds / t_dd / VALU_INT_PEAK, 4),
            "adds": adds,
            "lds_read_bytes": adds * per, "lds_read_frac": round(adds * per / t_dd / LDS_READ2_PEAK, 4),
            "hbm": {"algorithmic_bytes": int(b_in + plane),
                    "achieved_gbs": round((b_in + plane) / t_dd / 1e9, 2),
                    "frac": round((b_in + plane) / t_dd / 1e9 / HBM_PEAK_GBS, 5),
                    "traffic": traffic}}


def cpu_baseline(w, fb, stride, n_dm, beam_s):
    """The CPU oracle on the same gulp: every stride-th DM trial, scaled."""
    from oracle import oracle as O
    nthreads = cpu_threads()
    model, avail = host_cpu()
    p = O.make_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                      dm_max=w.dm_max, boxcar_max=w.boxcar_max, detect_thresh=w.detect_thresh,
                      nthreads=nthreads, dm_stride=stride)
    t0 = time.perf_counter()
    O.execute(p, fb, w.nsamps, w.nbits, 0, 1, want_text=False)
    dt = time.perf_counter() - t0
    n_sub = len(range(0, n_dm, stride))
    full_s = dt * n_dm / n_sub
    return {"value": beam_s / full_s, "unit": "beam-seconds searched/s", "cores": nthreads,
            "kind": "port", "cpu_model": model, "cores_available": avail,
            "sample": "CPU oracle (C, OpenMP over trials) on the same gulp, every %d-th of %d "
                      "DM trials (%d trials, %.1f s wall), scaled to all trials"
                      % (stride, n_dm, n_sub, dt)}


def cpu_baseline_small(seeds=(1, 2, 3), repeats=3):
    """The CPU oracle on whole 512-channel gulps (65536 samples of 256 us, DM
    0-500): the CPU-runnable configuration; median over repeats per seed."""
    from heimdall_multibeam_amd import synth
    from oracle import oracle as O
    w = synth.CONFIGS["small_512ch"]
    nthreads = cpu_threads()
    model, avail = host_cpu()
    p = O.make_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                      dm_max=w.dm_max, boxcar_max=w.boxcar_max, detect_thresh=w.detect_thresh,
                      nthreads=nthreads)
    rates = []
    total = 0.0
    for s in seeds:
        fb = synth.make_gulp(w, s)
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            r = O.execute(p, fb, w.nsamps, w.nbits, 0, 1, want_text=False)
            ts.append(time.perf_counter() - t0)
        total += sum(ts)
        beam_s = (w.nsamps - w.boxcar_max - r.max_delay) * w.dt
        rates.append(beam_s / float(np.median(ts)))
    return {"value": float(np.median(rates)), "unit": "beam-seconds searched/s",
            "cores": nthreads, "kind": "port", "cpu_model": model, "cores_available": avail,
            "workload": "small_512ch: 1 beam, 512 ch x 8-bit, fch1 1500 MHz, foff -0.5 MHz, "
                        "tsamp 256 us, gulp 65536, DM 0-500 (320 trials), boxcars 1..256",
            "sample": "whole gulps, seeds %s, median of %d calls each (%.1f s wall)"
                      % (list(seeds), repeats, total)}


def calibrate_copy(mib):
    """One device-to-device copy of `mib` MiB: the kernel of known size that
    scales rocprofv3's FETCH_SIZE and WRITE_SIZE to bytes."""
    if mib <= 0:
        return
    from heimdall_multibeam_amd import hip
    nb = mib << 20
    src, dst = hip.DeviceBuffer(nb), hip.DeviceBuffer(nb)
    hip.copy_device(dst.address, src.address, nb)
    hip.synchronize()
    src.free()
    dst.free()
    print("calibration copy: %d bytes" % nb, file=sys.stderr, flush=True)


def batch_measure(a, steps, warmup, cpu=True):
    """Ragged multibeam batch: one step = one hd_execute_batch over all beams,
    inputs resident in HBM."""
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import hip, synth
    w = synth.CONFIGS["batch_320ch"]
    nb = a.beams or synth.BATCH_BEAMS
    print("generating %d ragged beams of %s (seed %d)" % (nb, w.name, a.seed), file=sys.stderr,
          flush=True)
    beams = synth.make_batch(w, a.seed, nb)
    params = H.default_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                              dm_max=w.dm_max, boxcar_max=w.boxcar_max,
                              detect_thresh=w.detect_thresh, use_scrunching=False)
    calibrate_copy(a.calibrate_copy_mib)
    pl = H.Pipeline(params)
    pl.set_output(False)
    n_dm = len(pl.dm_list())
    max_delay = pl.max_delay()
    total = sum(fb.nbytes for fb, _ in beams)
    dbuf = hip.DeviceBuffer(total)
    dbuf.upload(np.concatenate([fb for fb, _ in beams]))
    gulps, off = [], 0
    for k, (fb, ns) in enumerate(beams):
        gulps.append((dbuf.address + off, ns, 1, 0, k))
        off += fb.nbytes
    beam_s = sum(max(ns - w.boxcar_max - max_delay, 0) for _, ns in beams) * w.dt
    for i in range(warmup):
        pl.execute_batch(gulps, w.nbits, device=True)
    hip.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        pl.execute_batch(gulps, w.nbits, device=True)
    hip.synchronize()
    elapsed = time.perf_counter() - t0
    graphs = pl.graph_stats()
    pl.set_graph(False)
    stage = {k: [] for k in H.Pipeline.STAGES}
    for i in range(steps):
        pl.execute_batch(gulps, w.nbits, device=True)
        for k, v in pl.timings().items():
            stage[k].append(v)
    n_cand = len(pl.candidates()["peak"])
    n_giants = len(pl.giants()["peak"])
    t_step = elapsed / steps
    t_dd = float(np.mean(stage["dedisperse"])) * 1e-3
    nout = sum(max(ns - max_delay, 0) for _, ns in beams)
    plane = 4.0 * n_dm * nout
    adds = float(n_dm) * nout * w.nchans
    rows = load_kernel_traffic(w.name)
    table, traffic = kernels_hbm(rows)
    out = {
        "value": round(beam_s / t_step, 3), "unit": "beam-seconds searched/s",
        "ms_per_step": round(t_step * 1e3, 3), "steps": steps, "warmup": warmup,
        "workload": "%s: %d beams, %d ch x %d-bit, fch1 %.4f MHz, foff %.6f MHz, tsamp %g us, "
                    "gulps %d-%d samples, live channels %d-%d (rest at the noise median), "
                    "DM %g-%g (%d trials, max delay %d), boxcars 1..%d"
                    % (w.name, nb, w.nchans, w.nbits, w.f0, w.df, w.dt * 1e6,
                       synth.BATCH_NSAMPS[0], synth.BATCH_NSAMPS[1], synth.BATCH_LIVE[0],
                       synth.BATCH_LIVE[1], w.dm_min, w.dm_max, n_dm, max_delay,
                       w.boxcar_max),
        "beam_seconds_per_step": round(beam_s, 3), "candidates": n_cand, "giants": n_giants,
        "stage_ms": {k: round(float(np.mean(v)), 4) for k, v in stage.items() if v},
        "hip_graph": {"captures": graphs[0], "replays": graphs[1]},
        "dedisp_tile": tile_json(pl.dedisp_tile()),
        "roofline": gulp_roofline(total + 2 * plane, t_step, traffic),
        "dominant_kernel": dedisp_block("k_dedisperse8", t_dd, adds, total, plane,
                                        pl.dedisp_tile(), t_step, rows),
        "kernels_hbm": table,
        "cpu_baseline": None,
    }
    if cpu and not a.no_cpu_baseline:
        from oracle import oracle as O
        nthreads = cpu_threads()
        model, avail = host_cpu()
        p = O.make_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                          dm_max=w.dm_max, boxcar_max=w.boxcar_max,
                          detect_thresh=w.detect_thresh, nthreads=nthreads)
        sample = beams[:a.cpu_beams]
        t0 = time.perf_counter()
        for fb, ns in sample:
            O.execute(p, fb, ns, w.nbits, 0, 1, want_text=False)
        dt = time.perf_counter() - t0
        bs = sum(max(ns - w.boxcar_max - max_delay, 0) for _, ns in sample) * w.dt
        out["cpu_baseline"] = {"value": bs / dt, "unit": "beam-seconds searched/s",
                               "cores": nthreads, "kind": "port", "cpu_model": model,
                               "cores_available": avail,
                               "sample": "CPU oracle on the first %d beams of the batch, one "
                                         "call per beam (%.1f s wall)" % (len(sample), dt)}
    pl.close()
    dbuf.free()
    return out


def batch_main(a):
    r = batch_measure(a, a.steps, a.warmup)
    out = {"metric": "beam-seconds searched/s/GPU at DM 0-2000 (ragged multibeam batch)",
           "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "u8 samples, exact u32 sums, f32 search",
           "data": "synthetic (seeded Gaussian noise, 0-3 dispersed pulses per beam, dead "
                   "channels held at the noise median)",
           "config": {"workload": r["workload"], "parallelism": "1 GPU"},
           **{k: r[k] for k in ("stage_ms", "hip_graph", "dedisp_tile", "roofline",
                                "dominant_kernel", "kernels_hbm", "cpu_baseline")}}
    print(json.dumps(out), flush=True)


def proposal_weights():
    """The trained proposal network shipped with the package: params and the
    operating points [(held-out recall target, logit threshold)], the
    default (highest recall) first."""
    path = os.path.join(HERE, "heimdall_multibeam_amd", "data", "proposal.npz")
    if not os.path.exists(path):
        return None, None
    z = np.load(path, allow_pickle=False)
    ops = [(None, float(z["threshold"]))]
    if "thresholds" in z.files:
        ops = sorted(zip([float(r) for r in z["recall_targets"]],
                         [float(t) for t in z["thresholds"]]), reverse=True)
    return z["params"].astype(np.float32), ops


def gated_eval(a, nbeams=64, per_call=16, steps=1, warmup=1, max_ops=None):
    """Batched full search vs proposal-gated search on the gated_4096ch beams
    (64 beams of 4096 channels x 2-bit, 131072 samples, 0-5 pulses each),
    rendered on the GPU, inputs resident in HBM.  One step searches every
    beam, per_call beams per hd_execute_batch call.  Recall: the full
    search's candidates with S/N >= 8 that the gated search reports with the
    same peak (beam, sample, trial, width, S/N bits)."""
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import hip, proposal as P, synth
    params, ops = proposal_weights()
    if ops is not None and max_ops is not None:
        ops = ops[:max_ops]
    if params is None:
        return {"error": "no trained network (heimdall_multibeam_amd/data/proposal.npz)"}
    w = synth.CONFIGS["gated_4096ch"]
    hp = H.default_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                          dm_max=w.dm_max, boxcar_max=w.boxcar_max,
                          detect_thresh=w.detect_thresh, use_scrunching=False,
                          min_tscrunch_width=w.min_tscrunch_width)
    pl = H.Pipeline(hp)
    pl.set_output(False)
    dms = pl.dm_list()
    max_delay = pl.max_delay()
    nbytes = w.nsamps * w.nchans * w.nbits // 8
    buf = hip.DeviceBuffer(nbytes * nbeams)
    rng = np.random.Generator(np.random.PCG64(a.seed + 55))
    for b in range(nbeams):
        synth.make_gulp_device(w, a.seed * 7777 + b, buf.address + b * nbytes, dm_trials=dms,
                               n_pulses=int(rng.integers(0, w.n_pulses + 1)))
    print("rendered %d gated_4096ch beams" % nbeams, file=sys.stderr, flush=True)
    calls = [[(buf.address + b * nbytes, w.nsamps, 1, 0, b)
              for b in range(c, min(c + per_call, nbeams))] for c in range(0, nbeams, per_call)]
    beam_s = nbeams * (w.nsamps - w.boxcar_max - max_delay) * w.dt

    def run(thr):
        if thr is not None:
            P.set_proposal(pl, params, thr)
        else:
            P.set_proposal(pl, None, 0.0)
        res = []
        for c in calls:
            pl.execute_batch(c, w.nbits, device=True)
            cd = pl.candidates()
            gi = pl.candidate_gulps()
            fl = P.last_proposal(pl) if thr is not None else None
            res.append((cd, gi, fl, pl.timings()))
        return res

    def timed(thr):
        for _ in range(warmup):
            run(thr)
        hip.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            res = run(thr)
        hip.synchronize()
        dt = (time.perf_counter() - t0) / steps
        print("%s search: %.1f ms per %d beams" % ("full" if thr is None else "gated (%.3f)" % thr,
                                                  dt * 1e3, nbeams), file=sys.stderr, flush=True)
        return dt, res

    full = timed(None)
    full_tile = pl.dedisp_tile()
    # stage times of the full search from calls launched kernel by kernel (a
    # graph replay is one span); gated calls always launch kernel by kernel
    pl.set_graph(False)
    full_eager = run(None)
    pl.set_graph(True)

    def cands(res, base_calls):
        keys = set()
        allc = []
        for ci, (cd, gi, _, _) in enumerate(res):
            for k in range(len(cd["peak"])):
                beam = base_calls[ci][int(gi[k])][4]
                key = (beam, int(cd["ind"][k]), int(cd["dm"][k]), int(cd["filter"][k]),
                       int(np.float32(cd["peak"][k]).view(np.uint32)))
                keys.add(key)
                allc.append((key, float(cd["peak"][k])))
        return keys, allc

    fk, fall = cands(full[1], calls)
    strong = [k for k, pk in fall if pk >= 8.0]
    t_full = full[0]
    st = {k: round(sum(r[3].get(k, 0.0) for r in full_eager), 3) for k in H.Pipeline.STAGES}
    res = {"workload": "gated_4096ch: %d beams, %d ch x %d-bit, tsamp %g us, gulp %d, DM %g-%g "
                       "(%d trials), boxcars 1..%d, min_tscrunch_width %d"
                       % (nbeams, w.nchans, w.nbits, w.dt * 1e6, w.nsamps, w.dm_min, w.dm_max,
                          len(dms), w.boxcar_max, w.min_tscrunch_width),
           "full_beam_s_per_s": round(beam_s / t_full, 3), "candidates_snr8": len(strong),
           "dedisp_tile": tile_json(full_tile),
           "full_stage_ms": st,
           "dedisp_share_of_full": round(st["dedisperse"] / max(st["total"], 1e-9), 4),
           "operating_points": []}
    for target, thr in ops:
        t_gated, gres = timed(thr)
        gk, _ = cands(gres, calls)
        rec = float(np.mean([k in gk for k in strong])) if strong else 1.0
        flags = np.concatenate([r[2] for r in gres])
        unit_hits = [bool(flags[k[0], k[2] // P.UNIT]) for k in strong]
        res["operating_points"].append({
            "heldout_recall_target": target, "threshold": thr,
            "gated_beam_s_per_s": round(beam_s / t_gated, 3),
            "speedup": round(t_full / t_gated, 3),
            "recall_snr8": round(rec, 4),
            "unit_recall_snr8": round(float(np.mean(unit_hits)) if strong else 1.0, 4),
            "units_searched_fraction": round(float(flags.mean()), 4),
            "gated_stage_ms": {k: round(sum(r[3].get(k, 0.0) for r in gres), 3)
                               for k in H.Pipeline.STAGES}})
    P.set_proposal(pl, None, 0.0)
    op0 = res["operating_points"][0]
    res.update({"gated_beam_s_per_s": op0["gated_beam_s_per_s"], "speedup": op0["speedup"],
                "recall_snr8": op0["recall_snr8"],
                "units_searched_fraction": op0["units_searched_fraction"]})
    pl.close()
    buf.free()
    return res


def gated_main(a):
    """Proposal-gated vs full batched search on the gated_4096ch beams as the
    main measurement (value: gated beam-seconds/s at the default threshold)."""
    nb = a.beams or 64
    r = gated_eval(a, nbeams=nb, max_ops=None if a.all_ops else 1)
    if "error" in r:
        print(json.dumps({"metric": "proposal", "error": r["error"]}), flush=True)
        sys.exit(1)
    op = r["operating_points"][0]
    out = {"metric": "beam-seconds searched/s/GPU at DM 0-3000 (proposal-gated batched search)",
           "value": op["gated_beam_s_per_s"], "unit": "beam-seconds searched/s", "n_gpus": 1,
           "steps": 1, "warmup": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "2-bit samples, exact u32 sums, f32 search",
           "data": "synthetic, rendered on the GPU (0-5 dispersed pulses per beam, 2% tone "
                   "channels)",
           "config": {"workload": r["workload"]}, "proposal": r, "roofline": None,
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def host_path(pl, w, fb, steps, beam_s):
    """hd_execute from page-locked host memory: the host-to-device copy of the
    gulp is part of every call (the rate a drop-in caller handing over host
    buffers sees)."""
    from heimdall_multibeam_amd import hip
    pin = hip.PinnedBuffer(fb.nbytes)
    pin.array()[:] = fb
    host = pin.array()
    pl.execute(host, w.nsamps, w.nbits, 0, 1)
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pl.execute(host, w.nsamps, w.nbits, 0, 1)
    hip.synchronize()
    t = (time.perf_counter() - t0) / steps
    # The gulp driver's way (bin/heimdall): the next gulp's bytes go to the
    # device on a second stream while the current gulp is searched from
    # device memory, so the copy hides behind the search.
    stream = hip.Stream()
    dev = [hip.DeviceBuffer(fb.nbytes), hip.DeviceBuffer(fb.nbytes)]
    stream.copy(dev[0].address, pin.address, fb.nbytes, hip.H2D)
    stream.synchronize()
    pl.execute(dev[0].address, w.nsamps, w.nbits, 0, 1, device=True)
    t0 = time.perf_counter()
    for k in range(steps):
        stream.copy(dev[(k + 1) % 2].address, pin.address, fb.nbytes, hip.H2D)
        pl.execute(dev[k % 2].address, w.nsamps, w.nbits, 0, 1, device=True)
        stream.synchronize()
    hip.synchronize()
    t_ov = (time.perf_counter() - t0) / steps
    for d in dev:
        d.free()
    stream.destroy()
    pin.free()
    return {"value": round(beam_s / t, 3), "unit": "beam-seconds searched/s",
            "ms_per_step": round(t * 1e3, 3), "bytes_copied_per_step": int(fb.nbytes),
            "note": "hd_execute on a page-locked host gulp: the host-to-device copy is inside "
                    "every call (reported beside value, which has the input in HBM)",
            "overlapped": {"value": round(beam_s / t_ov, 3), "ms_per_step": round(t_ov * 1e3, 3),
                           "note": "the gulp driver's double buffering: the next gulp's "
                                   "whole bytes copied from page-locked memory on a second "
                                   "stream while the current one is searched"}}


def small_path(steps, warmup, seed=1):
    """The CPU-runnable configuration (512 channels, 65536 samples of 256 us,
    DM 0-500) on the GPU: a whole gulp resident in HBM, graph replays."""
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import hip, synth
    w = synth.CONFIGS["small_512ch"]
    fb = synth.make_gulp(w, seed)
    params = H.default_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                              dm_max=w.dm_max, boxcar_max=w.boxcar_max,
                              detect_thresh=w.detect_thresh, use_scrunching=False,
                              min_tscrunch_width=w.min_tscrunch_width)
    pl = H.Pipeline(params)
    pl.set_output(False)
    beam_s = (w.nsamps - w.boxcar_max - pl.max_delay()) * w.dt
    dbuf = hip.DeviceBuffer(fb.nbytes)
    dbuf.upload(fb)
    for _ in range(max(warmup, 1)):
        pl.execute(dbuf.address, w.nsamps, w.nbits, 0, 1, device=True)
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pl.execute(dbuf.address, w.nsamps, w.nbits, 0, 1, device=True)
    hip.synchronize()
    t = (time.perf_counter() - t0) / steps
    n_dm = len(pl.dm_list())
    pl.close()
    dbuf.free()
    return {"value": round(beam_s / t, 3), "unit": "beam-seconds searched/s",
            "ms_per_step": round(t * 1e3, 3),
            "workload": "small_512ch: 1 beam, 512 ch x 8-bit, fch1 1500 MHz, foff -0.5 MHz, "
                        "tsamp 256 us, gulp 65536, DM 0-500 (%d trials), boxcars 1..256, "
                        "seed %d, input resident in HBM" % (n_dm, seed)}


def adaptive_path(w, dbuf, steps, warmup, beam_s):
    """The same HBM-resident gulp searched with adaptive time resolution
    (use_scrunching, the reference's default): each trial at its scrunch
    factor (1, 2, 4, ...)."""
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import hip
    from heimdall_multibeam_amd.dedisp_plan import DedispPlan
    params = H.default_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                              dm_max=w.dm_max, boxcar_max=w.boxcar_max,
                              detect_thresh=w.detect_thresh, use_scrunching=True,
                              min_tscrunch_width=w.min_tscrunch_width)
    pl = H.Pipeline(params)
    pl.set_output(False)
    for _ in range(warmup):
        pl.execute(dbuf.address, w.nsamps, w.nbits, 0, 1, device=True)
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pl.execute(dbuf.address, w.nsamps, w.nbits, 0, 1, device=True)
        n_cand = len(pl.candidates()["peak"])
    hip.synchronize()
    t = (time.perf_counter() - t0) / steps
    n_giants = len(pl.giants()["peak"])
    pl.close()
    plan = DedispPlan(w.nchans, w.dt, w.f0, w.df)
    plan.generate_dm_list(w.dm_min, w.dm_max, 40.0, 1.25)
    plan.enable_adaptive_dt(40.0, params.scrunch_tol)
    f, c = np.unique(plan.dt_factors(), return_counts=True)
    plan.close()
    return {"value": round(beam_s / t, 3), "unit": "beam-seconds searched/s",
            "ms_per_step": round(t * 1e3, 3), "candidates": n_cand, "giants": n_giants,
            "trials_per_factor": {str(int(k)): int(v) for k, v in zip(f, c)},
            "note": "use_scrunching on (scrunch_tol %.2f): the reference's default search, "
                    "trials at their scrunch factor; value keeps the full-resolution search"
                    % params.scrunch_tol}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: the search has no multi-GPU path "
              "(beams or DM ranges would shard one process per GPU); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if a.workload.startswith("batch"):
        return batch_main(a)
    if a.workload == "gated_4096ch":
        return gated_main(a)
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import hip, synth
    w = synth.CONFIGS[a.workload]
    print("generating %s gulp (seed %d)" % (w.name, a.seed), file=sys.stderr, flush=True)
    fb = synth.make_gulp(w, a.seed)
    params = H.default_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                              dm_max=w.dm_max, boxcar_max=w.boxcar_max,
                              detect_thresh=w.detect_thresh, use_scrunching=False,
                              min_tscrunch_width=w.min_tscrunch_width)
    pl = H.Pipeline(params)
    pl.set_output(False)
    n_dm = len(pl.dm_list())
    max_delay = pl.max_delay()
    nout = w.nsamps - max_delay
    beam_s = (w.nsamps - w.boxcar_max - max_delay) * w.dt  # pipeline.cu:370-372
    dbuf = hip.DeviceBuffer(fb.nbytes)
    dbuf.upload(fb)
    pl.preallocate(w.nsamps, 1)
    calibrate_copy(a.calibrate_copy_mib)
    print("n_dm=%d max_delay=%d beam_s/gulp=%.3f" % (n_dm, max_delay, beam_s), file=sys.stderr,
          flush=True)
    for i in range(a.warmup):
        pl.execute(dbuf.address, w.nsamps, w.nbits, 0, 1, device=True)
        print("warmup %d done" % i, file=sys.stderr, flush=True)
    hip.synchronize()
    n_cand = 0
    t0 = time.perf_counter()
    for i in range(a.steps):
        pl.execute(dbuf.address, w.nsamps, w.nbits, 0, 1, device=True)
        n_cand = len(pl.candidates()["peak"])
    hip.synchronize()
    elapsed = time.perf_counter() - t0
    n_giants = len(pl.giants()["peak"])
    graphs = pl.graph_stats()
    # per-stage times (and the dedispersion kernel's) from the same number of
    # calls with the kernels launched one by one: in the graph the chain is one
    # span
    pl.set_graph(False)
    stage = {k: [] for k in H.Pipeline.STAGES}
    for i in range(a.steps):
        pl.execute(dbuf.address, w.nsamps, w.nbits, 0, 1, device=True)
        for k, v in pl.timings().items():
            stage[k].append(v)
    pl.set_graph(True)
    t_step = elapsed / a.steps
    value = beam_s / t_step
    print("steps done: %.2f ms/step, %d candidates" % (t_step * 1e3, n_cand), file=sys.stderr,
          flush=True)
    t_dd = float(np.mean(stage["dedisperse"])) * 1e-3
    b_in = w.nsamps * w.nchans * w.nbits // 8
    plane = 4.0 * n_dm * nout
    adds = float(n_dm) * nout * w.nchans
    rows = load_kernel_traffic(w.name)
    table, traffic = kernels_hbm(rows)
    out = {
        "metric": "beam-seconds searched/s/GPU at DM 0-2000",
        "value": round(value, 3),
        "unit": "beam-seconds searched/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded Gaussian noise N(128,16^2) + %d dispersed pulses + "
                "%g%% tone channels + %d zero-DM bursts), input resident in HBM"
                % (w.n_pulses, 100 * w.tone_frac, w.n_bursts),
        "config": {"workload": "%s: 1 beam, %d ch x %d-bit, fch1 %.4f MHz, foff %.6f MHz, "
                               "tsamp %g us, gulp %d, DM %g-%g (%d trials, max delay %d), "
                               "boxcars 1..%d, threshold %g sigma"
                               % (w.name, w.nchans, w.nbits, w.f0, w.df, w.dt * 1e6, w.nsamps,
                                  w.dm_min, w.dm_max, n_dm, max_delay, w.boxcar_max,
                                  w.detect_thresh),
                   "arithmetic": "u8 samples, exact u32 dedispersion sums, f32 search",
                   "beam_seconds_per_gulp": round(beam_s, 4), "candidates": n_cand,
                   "giants": n_giants, "parallelism": "1 GPU"},
        "stage_ms": {k: round(float(np.mean(v)), 4) for k, v in stage.items() if v},
        "hip_graph": {"captures": graphs[0], "replays": graphs[1],
                      "note": "timed steps replay the captured stage chain; stage_ms comes "
                              "from as many calls launched kernel by kernel"},
        "dedisp_tile": tile_json(pl.dedisp_tile()),
        "roofline": gulp_roofline(b_in + 2 * plane, t_step, traffic),
        "dominant_kernel": dedisp_block("k_dedisperse8", t_dd, adds, b_in, plane,
                                        pl.dedisp_tile(), t_step, rows),
        "kernels_hbm": table,
        "host_path": None,
        "adaptive_dt": None,
        "cpu_baseline": None,
        "cpu_baseline_small": None,
        "small_gpu": None,
    }
    if not a.no_host_path:
        out["host_path"] = host_path(pl, w, fb, a.steps, beam_s)
    pl.close()
    if a.workload == "full_1024ch":
        try:
            out["small_gpu"] = small_path(a.steps, a.warmup)
        except Exception as e:  # reported, never hides the main measurement
            out["small_gpu"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if not a.no_adaptive:
        try:
            out["adaptive_dt"] = adaptive_path(w, dbuf, a.steps, a.warmup, beam_s)
        except Exception as e:  # reported, never hides the main measurement
            out["adaptive_dt"] = {"error": "%s: %s" % (type(e).__name__, e)}
    dbuf.free()
    if not a.no_cpu_baseline:
        print("timing CPU baselines", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, fb, a.cpu_stride, n_dm, beam_s)
        out["cpu_baseline_small"] = cpu_baseline_small()
        sg = out.get("small_gpu")
        if sg and "value" in sg and out["cpu_baseline_small"]["value"] > 0:
            sg["over_cpu_baseline_small"] = round(sg["value"] / out["cpu_baseline_small"]["value"], 1)
    if not a.no_batch:
        try:
            out["batch"] = batch_measure(a, a.steps, a.warmup)
        except Exception as e:  # reported, never hides the main measurement
            out["batch"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if not a.no_proposal:
        try:
            out["proposal"] = gated_eval(a)
        except Exception as e:
            out["proposal"] = {"error": "%s: %s" % (type(e).__name__, e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
