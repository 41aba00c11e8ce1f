"""Write tests/golden/ref_waveforms.json: values of the reference's own
ValueOrWaveform types (WaveformConfig.h / WaveformConfig.cpp, compiled unmodified
into oracle/_ref/ref_harness by `make -C oracle ref`) at times before, inside,
on the edges of and after each waveform's window.

Run from the repo root:  python tests/golden/make_ref_waveforms.py
"""
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.join(ROOT, "tests", "golden", "ref_waveforms.json")


def cases():
    rng = np.random.default_rng(17)
    out = []
    def add(type_, p, clamp=1, interp=0, points=()):
        p = list(p) + [0.0] * (8 - len(p))
        out.append(dict(type=type_, clamp=clamp, interp=interp, p=p,
                        points=[list(map(float, x)) for x in points]))
    add(0, [123.5])
    for clamp in (0, 1):
        add(1, [10.0, -250.0, 1e-6, 4e-6], clamp)
        add(3, [5.0, 20.0, -3e5, 2e-6, 9e-6], clamp)
        add(5, [-10.0, 40.0, 2.5e5, 0.3, 0.7, 1e-6, 30e-6], clamp)
    add(2, [1.0, 2e5, -3e10, 1e-6, 8e-6])
    add(4, [2.0, 150.0, 1.3e6, 0.25])
    add(6, [0.0, 75.0, 3e-6, 2.5e-6])
    ts = np.sort(rng.uniform(0, 20e-6, 9))
    vs = rng.uniform(-100, 100, 9)
    for interp in (0, 1, 2):
        add(7, [], 1, interp, zip(ts, vs))
    # evaluation times: a random spread plus the waveform's own edges
    for c in out:
        p = c["p"]
        edges = [p[2], p[3], p[4], p[5], p[6], p[2] + p[3]] + [x[0] for x in c["points"]]
        t = list(rng.uniform(-2e-6, 40e-6, 40)) + [e for e in edges if -1 < e < 1]
        c["t"] = [float(x) for x in t]
    return out


def main():
    cs = cases()
    text = []
    for c in cs:
        text.append("waveform %d %d %d %s %d %s %d %s" % (
            c["type"], c["clamp"], c["interp"], " ".join(repr(x) for x in c["p"]),
            len(c["points"]), " ".join("%r %r" % (a, b) for a, b in c["points"]),
            len(c["t"]), " ".join(repr(x) for x in c["t"])))
    res = subprocess.run([HARNESS], input="\n".join(text) + "\n", capture_output=True, text=True,
                         check=True)
    vals = [float(x) for x in res.stdout.split()]
    k = 0
    for c in cs:
        c["value"] = vals[k:k + len(c["t"])]
        k += len(c["t"])
    assert k == len(vals)
    with open(OUT, "w") as f:
        json.dump({"source": "WaveformConfig.h / WaveformConfig.cpp via oracle/_ref/ref_harness",
                   "cases": cs}, f)
    print(f"wrote {OUT}: {len(cs)} waveforms, {k} values")


if __name__ == "__main__":
    main()
