"""CPU restatement of Applications/candidate_profiler.C, the checker for
bin/candidate_profiler.

TEST INFRASTRUCTURE ONLY: nothing in the product imports it.

  normalise_rows   normalise_series_array (:83-132): the line through the
                   quarter-point order statistics of the two halves, then
                   division by 1.4826 x the median absolute value, with the
                   reference's float / double mix
  dm_time          gen_dm_time_plot (:232-409), dedispersion through
                   oracle/adaptive_dt.py (parity unpinned against dedisp)
  freq_time        gen_freq_time_plot (:136-230)
  outputs          main (:447-683): the texts of freq_time.dat, snr_dm.dat,
                   snr_time.dat and dm_time.pgm (ostream precision 6, then 10
                   after the first setprecision on each stream)
Where the reference reads outside its buffers the tool writes 0 or skips
the entry; so does this restatement.
"""
import numpy as np

from oracle import adaptive_dt as A
from oracle import oracle as O

f32 = np.float32
f64 = np.float64


def read_samples(data, nbytes_per, first, n):
    """Samples [first, first + n) of the file's data bytes, zero past the end."""
    out = np.zeros(n * nbytes_per, np.uint8)
    if first >= 2 ** 63 // max(nbytes_per, 1):
        return out
    a = first * nbytes_per
    seg = data[a:a + n * nbytes_per]
    out[:len(seg)] = seg
    return out


def normalise_rows(x):
    rows, n = x.shape
    for r in range(rows):
        v = x[r]
        m1 = f32(np.partition(v[:n // 2], n // 4)[n // 4])
        m2 = f32(np.partition(v[n // 2:], 3 * n // 4 - n // 2)[3 * n // 4 - n // 2])
        slope = f64(f32(f32(m2 - m1) / f32(n // 2)))
        m0 = f32(f64(m1) - slope * f64(n // 4))
        line = (f64(m0) + np.arange(n, dtype=f64) * slope).astype(f32)
        v = (v - line).astype(f32)
        mad = f32(np.partition(np.abs(v - f32(0)), n // 2)[n // 2])
        sigma = f32(f64(mad) * 1.4826)
        scale = f32(1.0 / f64(sigma))
        x[r] = (v * scale).astype(f32)
    return x


def _kc(h):
    return O.delay_table(h["nchans"], f32(h["tsamp"]), f32(h["fch1"]), f32(h["foff"]))


def _max_delay(dms, kc):
    return int((np.asarray(dms, f32)[:, None] * kc[None, :] + f32(0.5)).astype(np.int64).max())


def dm_time(h, data, dms, samp, filt, dm_idx, out_nsamps, out_dm_count, do_filter):
    dms = np.asarray(dms, f32)
    n_all = len(dms)
    lo = dm_idx - out_dm_count // 2 if dm_idx >= out_dm_count // 2 else 0
    bound = n_all - ((out_dm_count - 1) // 2 + 1)
    hi = dm_idx + (out_dm_count - 1) // 2 + 1 if (bound > 0 and dm_idx < bound) else n_all
    hi = min(hi, n_all)
    sub = dms[lo:hi]
    ndm = hi - lo
    kc = _kc(h)
    f = A.scrunch_list(sub, f32(h["tsamp"]), 40.0, f32(h["fch1"]), f32(h["foff"]), h["nchans"],
                       1.15)
    width = 1 << filt
    maxd = _max_delay(sub, kc)
    if not do_filter:
        nsamps = out_nsamps * width
        first = samp - out_nsamps // 2 * width if samp >= out_nsamps // 2 * width else 0
    else:
        nsamps = out_nsamps + width - 1
        first = samp - out_nsamps // 2 - width // 2 if samp > out_nsamps // 2 + width // 2 else 0
    in_nsamps = nsamps + maxd
    per = h["nchans"] * h["nbits"] // 8
    raw = read_samples(data, per, first, in_nsamps)
    dd = A.dedisperse_adaptive(raw, in_nsamps, h["nbits"], h["nchans"], sub, kc, maxd, f,
                               nthreads=8)
    out = np.zeros((ndm, out_nsamps), f32)
    if not do_filter:
        for d in range(ndm):
            k = width // int(f[d])
            for t in range(out_nsamps):
                acc = f32(0)
                seg = dd[d, t * k:t * k + k]
                for v in seg:
                    acc = f32(acc + v)
                out[d, t] = acc
    else:
        hoff = width // 2 + (width - 1) // 2 + 1
        for d in range(ndm):
            full = dd[d, np.arange(nsamps) // int(f[d])]
            scan = np.zeros(nsamps + 1, f32)
            scan[1:] = np.cumsum(full, dtype=f32)   # sequential float sums
            out[d] = (scan[hoff:hoff + out_nsamps] - scan[0:out_nsamps]).astype(f32)
    return normalise_rows(out), ndm


def freq_time(h, data, samp, filt, dm, out_nsamps, fscrunch):
    kc = _kc(h)
    maxd = _max_delay([dm], kc)
    delay = 2
    tscrunch = 1 << filt
    in_nsamps = out_nsamps * tscrunch
    if in_nsamps < maxd:
        while delay < maxd:
            delay *= 2
        in_nsamps = delay
        tscrunch = in_nsamps // out_nsamps
    first = (samp + maxd // 2 - in_nsamps // 2) % (1 << 64)
    nbits, nchans = h["nbits"], h["nchans"]
    per = nchans * nbits // 8
    raw = read_samples(data, per, first, in_nsamps)
    x = A.unpack(raw, in_nsamps, nbits, nchans)
    mask = f32((1 << nbits) - 1)
    out_nchans = nchans // fscrunch
    out = np.zeros((out_nchans, out_nsamps), f32)
    vals = (x.astype(f32) / mask * f32(255)).astype(f32)    # (float)x / mask * peak
    if in_nsamps % tscrunch == 0 and nchans % fscrunch == 0 and in_nsamps // tscrunch <= out_nsamps:
        # every bin inside the window: its sum in (sample, channel) order,
        # sequential in float32 (accumulate), for all bins at once
        nbt = in_nsamps // tscrunch
        blk = vals.reshape(nbt, tscrunch, out_nchans, fscrunch).transpose(0, 2, 1, 3)
        sums = np.add.accumulate(blk.reshape(nbt, out_nchans, -1), axis=2, dtype=f32)[:, :, -1]
        out[:, :nbt] = (sums / f32(tscrunch * fscrunch)).astype(f32).T
        return normalise_rows(out), in_nsamps
    for t in range(0, in_nsamps, tscrunch):
        if t // tscrunch >= out_nsamps:
            continue
        for c in range(0, nchans, fscrunch):
            s = f32(0)
            for ts in range(tscrunch):
                for k in range(fscrunch):
                    cc = c + k
                    if cc < nchans and t + ts < in_nsamps:
                        v = vals[t + ts, cc]
                    else:   # the next sample's first channels, as the word index runs on
                        tt, ch = divmod((t + ts) * nchans + cc, nchans)
                        v = vals[tt, ch] if tt < in_nsamps else f32(0)
                    s = f32(s + v)
            if c // fscrunch < out_nchans:
                out[c // fscrunch, t // tscrunch] = f32(s / f32(tscrunch * fscrunch))
    return normalise_rows(out), in_nsamps


def _g(v, prec):
    return ("%." + str(prec) + "g") % float(v)


def outputs(h, data, dms, samp, filt, dm_idx, out_nsamps, out_dm_count, fscrunch,
            do_filter=False):
    """{file name: text} of one run."""
    dms = np.asarray(dms, f32)
    img = np.zeros((out_dm_count, out_nsamps), f32)
    rows, ndm = dm_time(h, data, dms, samp, filt, dm_idx, out_nsamps, out_dm_count, do_filter)
    img[:ndm] = rows
    dm0, _ = dm_time(h, data, dms, samp, filt, 0, out_nsamps, 1, do_filter)
    ft, _ = freq_time(h, data, samp, filt - 2 if filt > 2 else 0, dms[dm_idx], out_nsamps,
                      fscrunch)
    res = {}
    lines = []
    for c in range(ft.shape[0]):
        row = ft[c].copy()
        row[np.isnan(row)] = 0
        if c * fscrunch < 150:
            row[:] = 0
        lines.append("".join(_g(v, 6) + "\t" for v in row) + "\n")
    res["freq_time.dat"] = "".join(lines)
    flat = img.reshape(-1)
    peak = int(np.argmax(flat))
    row0 = dm_idx - out_dm_count // 2 if dm_idx > out_dm_count // 2 else 0
    t = peak % out_nsamps
    s = "#dm_trial\tDM\tSNR\n"
    for d in range(min(out_dm_count, len(dms))):
        trial = row0 + d
        if trial >= len(dms):
            break
        s += "%d\t%s\t%s\n" % (trial, _g(dms[trial], 10), _g(img[d, t], 10))
    res["snr_dm.dat"] = s
    width = f32(1 << filt)
    d = peak // out_nsamps

    def at(r, tt):
        return img[r, tt] if 0 <= r < out_dm_count else f32(0)
    s = "#sample\ttime\tSNR\tSNR@DM-1\tSNR@DM+1\tSNR@DM0\n"
    for tt in range(out_nsamps):
        v = f32(f64(f32(f32(samp) + f32(f32(tt) - f32(out_nsamps // 2)) * width)) * h["tsamp"])
        p0 = 6 if tt == 0 else 10
        s += "%s\t%s\t%s\t%s\t%s\t%s\n" % (_g(v, p0), _g(v, 10), _g(at(d, tt), 10),
                                           _g(at(d - 1, tt), 10), _g(at(d + 1, tt), 10),
                                           _g(dm0[0, tt], 10))
    res["snr_time.dat"] = s
    if out_nsamps <= 16384:
        s = "P2 %d %d 255\n" % (out_nsamps, out_dm_count)
        for r in range(out_dm_count):
            vals = []
            for tt in range(out_nsamps):
                lv = f32(f32(f32(img[r, tt] - f32(0)) / f32(8)) * f32(255))
                lv = lv if lv > f32(0) else f32(0)
                lv = f32(255) if f32(255) < lv else lv
                vals.append("%d\t" % (0 if np.isnan(lv) else int(lv)))
            s += "".join(vals) + "\n"
        res["dm_time.pgm"] = s
    return res
