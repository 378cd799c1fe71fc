"""CPU restatement of the multibeam coincidencer (Applications/Candidates.C,
Applications/coincidencer.C).

TEST INFRASTRUCTURE ONLY: the checker for heimdall_multibeam_amd/app/
coincidencer.cpp; nothing in the product imports it.

  parse_line      Candidate::Candidate(const char*) (Candidates.C:29-50): the
                  ten fields, snr/sample_time/dm as float32, the rest integers
  is_coincident   Candidate::is_coincident (Candidates.C:63-77)
  compute         CandidateChunk::compute_coincidence (Candidates.C:257-301):
                  only the first coincident candidate of each other beam counts
  format_line     operator<< (Candidates.C:80-87), default stream precision
"""
import numpy as np

f32 = np.float32


def parse_line(line):
    t = line.split()
    if len(t) < 10:
        return None
    c = {"snr": f32(float(t[0])), "sample_idx": int(t[1]), "sample_time": f32(float(t[2])),
         "filter": int(t[3]), "dm_trial": int(t[4]), "dm": f32(float(t[5])),
         "members": int(t[6]), "begin": int(t[7]), "end": int(t[8]), "beam": int(t[9])}
    c["nbeams"] = 1
    c["primary_beam"] = c["beam"]
    c["max_snr"] = c["snr"]
    return c


def is_coincident(a, c, sep_time, sep_filter):
    """a.is_coincident(c): float tolerance sep_time * 2^max(filter), truncated
    to int64; dm_trial within 9999; filters within sep_filter; the relative
    S/N difference |c - a| / (c + a) at most 0.4 in float32."""
    tol = int(f32(sep_time) * f32(2.0 ** max(c["filter"], a["filter"])))
    rel = f32(abs(f32(c["snr"] - a["snr"]))) / f32(c["snr"] + a["snr"])
    return (abs(c["sample_idx"] - a["sample_idx"]) <= tol and
            abs(c["dm_trial"] - a["dm_trial"]) <= 9999 and
            abs(c["filter"] - a["filter"]) <= sep_filter and
            rel <= f32(0.4))


def compute(beams, sep_time=3, sep_filter=4):
    """beams: list (beam 1..n) of lists of parsed candidates, updated in place."""
    for i, bi in enumerate(beams):
        for cj in bi:
            max_snr = cj["snr"]
            for k, bk in enumerate(beams):
                if k == i:
                    continue
                for cl in bk:
                    if is_coincident(cj, cl, sep_time, sep_filter):
                        cj["nbeams"] += 1
                        if cl["snr"] > max_snr:
                            cj["primary_beam"] = cl["beam"]
                            cj["max_snr"] = cl["snr"]
                            max_snr = cl["snr"]
                        break
    return beams


def _g(v):
    return "%g" % float(v)


def format_line(c):
    return "\t".join([_g(c["snr"]), str(c["sample_idx"]), _g(c["sample_time"]),
                      str(c["filter"]), str(c["dm_trial"]), _g(c["dm"]), str(c["members"]),
                      str(c["begin"]), str(c["end"]), str(c["nbeams"]), "",
                      str(c["primary_beam"]), _g(c["max_snr"]), str(c["beam"])])


def coincide_lines(lines, nbeams, sep_time=3, sep_filter=4):
    """The _all.cand text of candidate lines (one gulp, beams 1..nbeams)."""
    beams = [[] for _ in range(nbeams)]
    for ln in lines:
        c = parse_line(ln)
        if c is not None and 1 <= c["beam"] <= nbeams:
            beams[c["beam"] - 1].append(c)
    compute(beams, sep_time, sep_filter)
    return "".join(format_line(c) + "\n" for b in beams for c in b)
