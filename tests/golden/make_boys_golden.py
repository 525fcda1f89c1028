"""Write boys_mpmath.json: F_n(t) = 1F1(n+1/2; n+3/2; -t)/(2n+1) at 40-digit
precision (mpmath), n = 0..12, t on a grid covering the series region, the
Taylor-grid region, the switch point and the asymptotic region.

Run: python tests/golden/make_boys_golden.py
"""

import json
import os

import mpmath as mp

mp.mp.dps = 40
ts = sorted(set([0.0, 1e-12, 1e-8, 1e-4, 1e-3, 5e-3, 0.0099, 0.01, 0.0101, 0.0125, 0.02, 0.05,
                 0.1, 0.2, 0.37, 0.5, 0.75, 1.0, 1.3, 2.0, 3.1415, 5.0, 7.77, 10.0, 12.5, 17.0,
                 20.0, 25.0, 29.97, 33.3, 39.99, 40.0, 40.01, 45.0, 60.0, 100.0, 250.0, 1000.0]))
out = {"t": ts, "nmax": 12, "F": []}
for t in ts:
    row = []
    for n in range(13):
        if t == 0.0:
            v = mp.mpf(1) / (2 * n + 1)
        else:
            v = mp.hyp1f1(n + mp.mpf(1) / 2, n + mp.mpf(3) / 2, -t) / (2 * n + 1)
        row.append(float(v))
    out["F"].append(row)
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys_mpmath.json")
with open(path, "w") as fh:
    json.dump(out, fh, indent=0)
print("wrote", path)
