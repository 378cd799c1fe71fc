"""Trains the proposal network on data the pipeline makes itself.

Every gulp is rendered on the GPU by the project's generator (seeded pulses,
tones and bursts), searched in full by the library, and cut into coarse
DM-time tiles by the library's coarse pass.  A tile is labelled positive when
a full-search candidate at or above the detection threshold has its peak in
the tile's centre (its unit of 128 trials and its 512 samples).  The network
is trained by the library's own forward, backward and Adam kernels; PyTorch
only holds the tile set in device memory and gathers batches.

After training, held-out gulps (seeds never used for training) set the logit
threshold: the largest one at which at least `target_recall` of the full
search's candidates with S/N >= 8 lie in a proposed unit.  The same held-out
gulps measure the coarse-plane method without a network (a unit is proposed
when any of its centre rows holds a boxcar of 1..64 coarse samples above a
threshold, set by the same rule), so the two can be compared at equal recall.

Writes <out>.npz (float32 `params`, `threshold`) and <out>.json (metrics).

    python -m heimdall_multibeam_amd.train_proposal --gulps 400 --beams5 100 --out /tmp/p
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import Pipeline, default_params
from . import proposal as P
from . import synth
from .hip import DeviceBuffer

SNR_RECALL = 8.0


def log(*a):
    print("[train %7.1fs]" % (time.time() - T0), *a, flush=True)


T0 = time.time()


def params_for(w):
    return default_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                          dm_max=w.dm_max, boxcar_max=w.boxcar_max,
                          detect_thresh=w.detect_thresh, use_scrunching=False,
                          min_tscrunch_width=w.min_tscrunch_width)


class Setup:
    """One observing setup: its pipeline, device gulp buffer and trial list."""

    def __init__(self, name):
        self.w = synth.CONFIGS[name]
        self.pl = Pipeline(params_for(self.w))
        self.pl.set_output(False)
        self.dms = self.pl.dm_list()
        self.fb = DeviceBuffer(self.w.nsamps * self.w.nchans * self.w.nbits // 8)

    def gulp(self, seed, kind):
        """Render, search and coarse-grid one gulp; kind is 'pulses', 'noise' or 'rfi'."""
        w = self.w
        import dataclasses
        if kind == "noise":
            wk = dataclasses.replace(w, tone_frac=0.0, n_bursts=0)
            npul = 0
        elif kind == "rfi":
            wk = w if w.tone_frac > 0 or w.n_bursts > 0 else dataclasses.replace(w, tone_frac=0.02)
            npul = 0
        elif w.name == "learned_1024ch":
            # pulse gulps carry no RFI: the RFI-only gulps hold it (with the
            # tones of the full_1024ch setup the search finds almost none of
            # the injected pulses, and their tiles would be labelled empty)
            wk = dataclasses.replace(w, tone_frac=0.0, n_bursts=0)
            npul = None
        else:
            wk = w
            npul = int(np.random.default_rng(seed).integers(0, w.n_pulses + 1))
        truth = synth.make_gulp_device(wk, seed, self.fb.address, dm_trials=self.dms,
                                       n_pulses=npul)
        self.pl.execute(self.fb.address, w.nsamps, w.nbits, device=True)
        c = self.pl.candidates()
        nu, nv = P.coarse(self.pl, self.fb.address, w.nsamps, w.nbits)
        return c, nu, nv, truth


def cand_cells(c, nu, nv, snr_min):
    keep = c["peak"] >= snr_min
    u = (c["dm"][keep] // P.UNIT).astype(np.int64)
    v = (c["ind"][keep] // (P.TIME_DECIM * P.TSTEP)).astype(np.int64)
    ok = (u < nu) & (v < nv)
    return u[ok], v[ok]


def positive_cells(c, nu, nv, snr_min):
    """Tiles whose window holds a candidate peak: its unit, and every time
    position whose 64 coarse samples ([512v - 256, 512v + 768) in samples)
    contain the peak sample."""
    keep = c["peak"] >= snr_min
    pos = set()
    span = P.TIME_DECIM * P.TSTEP
    for dm, ind in zip(c["dm"][keep], c["ind"][keep]):
        u = int(dm) // P.UNIT
        if u >= nu:
            continue
        lo = max(0, -(-(int(ind) - 3 * span // 2 + 1) // span))
        hi = min(nv - 1, (int(ind) + span // 2) // span)
        for v in range(lo, hi + 1):
            pos.add((u, v))
    return pos


def choose_tiles(rng, c, nu, nv, snr_min):
    """Positives (tiles holding a candidate peak), the tiles of the same time
    positions in the neighbouring units, and random negatives.  Tiles next to
    a positive in time are not used as negatives: they hold the same pulse."""
    pos = positive_cells(c, nu, nv, snr_min)
    neg = set()
    for (a, b) in pos:
        for da in (-1, 1):
            q = (a + da, b)
            if 0 <= q[0] < nu and q not in pos:
                neg.add(q)
    n_rand = 2 * len(pos) + 8
    for _ in range(n_rand):
        q = (int(rng.integers(nu)), int(rng.integers(nv)))
        if q not in pos:
            neg.add(q)
    tiles = sorted(pos) + sorted(neg)
    labels = [1.0] * len(pos) + [0.0] * len(neg)
    return tiles, labels


def heuristic_unit_max(plane, nu):
    """Per unit: the largest normalised boxcar (1..64 coarse samples) of any of
    its 16 centre rows -- the coarse-plane method without a network."""
    from .proposal import UNIT, DM_DECIM
    x = (plane.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    S = np.concatenate([np.zeros((x.shape[0], 1)), np.cumsum(x, 1)], 1)
    best = np.full(x.shape[0], -np.inf)
    w = 1
    while w <= 64 and w < x.shape[1]:
        b = (S[:, w:] - S[:, :-w]) / np.sqrt(w)
        best = np.maximum(best, b.max(1))
        w *= 2
    rows = UNIT // DM_DECIM
    out = np.full(nu, -np.inf)
    for u in range(nu):
        seg = best[u * rows:(u + 1) * rows]
        if seg.size:
            out[u] = seg.max()
    return out


def evaluate(setup, seeds, kinds, params, need_heuristic=True):
    """Per held-out gulp: S/N>=8 candidate units, network unit maxima and
    heuristic unit maxima."""
    # load the network, then switch gating off: the searches here are the
    # full search (the arbiter of recall); scores use the loaded network
    P.set_proposal(setup.pl, params, 0.0)
    P.set_proposal(setup.pl, None, 0.0)
    out = []
    for s, k in zip(seeds, kinds):
        c, nu, nv, _ = setup.gulp(s, k)
        sc = P.scores(setup.pl, nu, nv)
        umax = sc.max(1)
        hmax = heuristic_unit_max(P.plane(setup.pl), nu) if need_heuristic else None
        u, _ = cand_cells(c, nu, nv, SNR_RECALL)
        out.append({"units": u, "net": umax, "heur": hmax, "nu": nu})
    return out


def ideal_fraction(evals):
    """Fraction of units that hold an S/N >= 8 candidate (a perfect proposal)."""
    tot = sum(e["nu"] for e in evals)
    return float(sum(len(set(e["units"].tolist())) for e in evals)) / max(tot, 1)


def threshold_for(evals, key, target):
    """Largest threshold keeping >= target of the S/N>=8 candidates in a
    proposed unit; returns (threshold, recall, fraction of units proposed)."""
    vals = np.concatenate([e[key][e["units"]] for e in evals if len(e["units"])])
    if vals.size == 0:
        return 0.0, 1.0, 0.0
    vals = np.sort(vals)
    k = int(np.floor((1.0 - target) * vals.size))   # candidates allowed below
    thr = np.nextafter(vals[k], -np.inf) if k < vals.size else vals[-1]
    rec = float(np.mean(vals > thr))
    frac = float(np.mean(np.concatenate([e[key] > thr for e in evals])))
    return float(thr), rec, frac


def collect(st, seeds, kinds, pick, dev, what):
    """Renders and searches the gulps; pick(c, nu, nv, scores_or_None) gives
    the (u, v) tiles and labels to keep.  Returns (tiles tensor, labels)."""
    import torch
    from .hip import copy_device
    buf = DeviceBuffer(4096 * 64 * 64 * 2 * 2)
    tiles, labels = [], []
    n_pos = n_tiles = n_cand = 0
    for i, (sd, kd) in enumerate(zip(seeds, kinds)):
        c, nu, nv, _ = st.gulp(sd, kd)
        pos, lab = pick(c, nu, nv)
        n_cand += int((c["peak"] >= st.w.detect_thresh).sum())
        for k0 in range(0, len(pos), 4096):
            chunk = pos[k0:k0 + 4096]
            P.extract(st.pl, chunk, buf.address)
            t = torch.empty((len(chunk), 64, 64, 2), dtype=torch.int16, device=dev)
            copy_device(t.data_ptr(), buf.address, len(chunk) * 64 * 64 * 2 * 2)
            tiles.append(t)
        labels.append(np.asarray(lab, np.float32))
        n_pos += int(sum(lab))
        n_tiles += len(lab)
        if (i + 1) % 100 == 0 or i + 1 == len(seeds):
            log("%s %s %d/%d  tiles %d (positive %d)  candidates %d" %
                (st.w.name, what, i + 1, len(seeds), n_tiles, n_pos, n_cand))
    return tiles, labels, {"gulps": len(seeds), "tiles": n_tiles, "positive_tiles": n_pos,
                           "candidates": n_cand}


def train(net, X, Y, epochs, lr0, batch, seed):
    import torch
    dev = X.device
    N = X.shape[0]
    pos_frac = float(Y.mean().item())
    W = torch.where(Y > 0.5, torch.tensor(0.5 / max(pos_frac, 1e-6), device=dev),
                    torch.tensor(0.5 / max(1 - pos_frac, 1e-6), device=dev)).float()
    log("training on %d tiles (%.1f%% positive), %d epochs" % (N, 100 * pos_frac, epochs))
    steps_per_epoch = N // batch
    total = steps_per_epoch * epochs
    step = 0
    g = torch.Generator(device="cpu").manual_seed(seed)
    lr = lr0
    for ep in range(epochs):
        perm = torch.randperm(N, generator=g).to(dev)
        losses = []
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            xb = X.index_select(0, idx).contiguous()
            yb = Y.index_select(0, idx).contiguous()
            wb = W.index_select(0, idx).contiguous()
            torch.cuda.current_stream().synchronize()
            net.forward_device(xb.data_ptr(), batch)
            losses.append(net.backward(yb.data_ptr(), wb.data_ptr()))
            lr = lr0 * 0.5 * (1 + np.cos(np.pi * step / max(total, 1)))
            net.adam(lr, weight_decay=1e-4)
            step += 1
        log("epoch %d  loss %.4f  lr %.2e" % (ep + 1, float(np.mean(losses)), lr))
    return pos_frac


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gulps", type=int, default=4000, help="learned_1024ch training gulps")
    ap.add_argument("--beams5", type=int, default=2000, help="gated_4096ch training beams")
    ap.add_argument("--val", type=float, default=0.1, help="held-out fraction")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--mine", type=int, default=600, help="gated beams re-scored for hard negatives")
    ap.add_argument("--mine-per", type=int, default=24, help="hard negatives kept per mined beam")
    ap.add_argument("--epochs2", type=int, default=6)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--target-recall", type=float, default=0.995)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--evaluate", default=None,
                    help="skip training: evaluate the network in this .npz on the held-out "
                         "gulps and write <out> with its thresholds")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    import torch
    rng = np.random.default_rng(a.seed)
    dev = torch.device("cuda:0")
    net = P.Cnn(max_batch=a.batch, seed=a.seed)
    plan = [("learned_1024ch", a.gulps), ("gated_4096ch", a.beams5)]
    setups, heldout, stats = {}, {}, {}
    tiles_all, labels_all = [], []
    for name, n in plan:
        if n <= 0:
            continue
        st = setups[name] = Setup(name)
        n_val = max(1, int(round(a.val * n)))
        n_tr = n - n_val
        kinds = ["noise" if (i * 7919) % 10 == 0 else ("rfi" if (i * 7919) % 10 == 1 else "pulses")
                 for i in range(n)]
        base = a.seed * 1000003 + (1 if name == "learned_1024ch" else 2) * 100000
        seeds = [base + i for i in range(n)]
        heldout[name] = (seeds[n_tr:], kinds[n_tr:])
        setups[name].train_seeds = (seeds[:n_tr], kinds[:n_tr])
        extra = 8 if name == "learned_1024ch" else 48

        def pick(c, nu, nv, st=st, extra=extra):
            pos, lab = choose_tiles(rng, c, nu, nv, st.w.detect_thresh)
            have = set(pos)
            for _ in range(extra):
                q = (int(rng.integers(nu)), int(rng.integers(nv)))
                if q not in have:
                    have.add(q)
                    pos.append(q)
                    lab.append(0.0)
            return pos, lab

        if a.evaluate:
            stats[name] = {"heldout_gulps": n_val, "n_dm": int(len(st.dms))}
            continue
        t, l, stt = collect(st, seeds[:n_tr], kinds[:n_tr], pick, dev, "train")
        tiles_all += t
        labels_all += l
        stats[name] = dict(stt, heldout_gulps=n_val, n_dm=int(len(st.dms)))
    if a.evaluate:
        z = np.load(a.evaluate, allow_pickle=False)
        return finish(a, z["params"], setups, heldout, stats, None)
    X = torch.cat(tiles_all)
    Y = torch.from_numpy(np.concatenate(labels_all)).to(dev)
    tiles_all = []
    pos_frac = train(net, X, Y, a.epochs, a.lr, a.batch, a.seed)
    # hard negatives: the highest-scoring negative tiles of re-rendered
    # training beams, then more epochs on the union
    if a.mine > 0 and "gated_4096ch" in setups:
        st = setups["gated_4096ch"]
        P.set_proposal(st.pl, net.params(), 0.0)
        P.set_proposal(st.pl, None, 0.0)
        seeds, kinds = st.train_seeds
        ms, mk = seeds[:a.mine], kinds[:a.mine]

        def pick_hard(c, nu, nv, st=st):
            sc = P.scores(st.pl, nu, nv)
            mask = np.zeros((nu, nv), bool)
            for (u, v) in positive_cells(c, nu, nv, st.w.detect_thresh):
                mask[u, max(v - 1, 0):v + 2] = True   # and the same pulse's neighbours
            sc = np.where(mask, -np.inf, sc)
            flat = np.argsort(sc, axis=None)[::-1][:a.mine_per]
            pos = [(int(f // nv), int(f % nv)) for f in flat if np.isfinite(sc.flat[f])]
            return pos, [0.0] * len(pos)

        t, l, stt = collect(st, ms, mk, pick_hard, dev, "mine")
        stats["gated_4096ch"]["mined_tiles"] = stt["tiles"]
        X = torch.cat([X] + t)
        Y = torch.cat([Y, torch.from_numpy(np.concatenate(l)).to(dev)])
        pos_frac = train(net, X, Y, a.epochs2, a.lr * 0.3, a.batch, a.seed + 1)
    params = net.params()
    del X, Y
    return finish(a, params, setups, heldout, stats, pos_frac)


RECALL_TARGETS = (0.95, 0.98, 0.99, 0.995)


def finish(a, params, setups, heldout, stats, pos_frac):
    """Held-out evaluation, thresholds per recall target, and the outputs."""
    metrics = {"positive_fraction": pos_frac, "epochs": [a.epochs, a.epochs2],
               "target_recall": a.target_recall, "setups": stats, "heldout": {}}
    thr_deploy = None
    targets = None
    for name, (seeds, kinds) in heldout.items():
        ev = evaluate(setups[name], seeds, kinds, params)
        thr_n, rec_n, frac_n = threshold_for(ev, "net", a.target_recall)
        thr_h, rec_h, frac_h = threshold_for(ev, "heur", a.target_recall)
        curve = {}
        thr_r = []
        for r in RECALL_TARGETS:
            tn = threshold_for(ev, "net", r)
            curve[str(r)] = {"network": tn[2], "network_threshold": tn[0],
                             "coarse_boxcar": threshold_for(ev, "heur", r)[2]}
            thr_r.append(tn[0])
        metrics["heldout"][name] = {
            "gulps": len(seeds), "candidates_snr8": int(sum(len(e["units"]) for e in ev)),
            "ideal_unit_fraction": ideal_fraction(ev),
            "network": {"threshold": thr_n, "recall": rec_n, "unit_fraction": frac_n},
            "coarse_boxcar": {"threshold": thr_h, "recall": rec_h, "unit_fraction": frac_h},
            "unit_fraction_at_recall": curve}
        log(name, json.dumps(metrics["heldout"][name]))
        if name == "gated_4096ch" or thr_deploy is None:
            thr_deploy = thr_n
            targets = np.array(thr_r, np.float32)
    metrics["threshold"] = thr_deploy
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.savez(a.out + ".npz", params=params, threshold=np.float32(thr_deploy),
             recall_targets=np.array(RECALL_TARGETS, np.float32), thresholds=targets)
    with open(a.out + ".json", "w") as f:
        json.dump(metrics, f, indent=1)
    log("wrote", a.out + ".npz", "threshold %.4f" % thr_deploy)
    return 0


if __name__ == "__main__":
    sys.exit(main())
