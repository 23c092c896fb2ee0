"""Write h1_table.json: H1^(1)(x) = J1 + i Y1 at 30 significant digits
(mpmath) on arguments spanning the branches of the device evaluation and
the ranges the double-layer butterfly hits.  Run from this folder."""
import json

import mpmath as mp
import numpy as np

mp.mp.dps = 40
xs = np.concatenate([
    np.geomspace(1e-4, 4, 40), [3.999999, 4.0, 4.000001], np.linspace(4.05, 23.95, 40),
    [23.999999, 24.0, 24.000001], np.geomspace(24.5, 3.2e5, 60),
    [3.831705970207512, 7.015586669815619, 2.197141326031017, 5.429681040794135],  # near zeros of J1, Y1
])
rows = []
for x in xs:
    x = float(x)
    rows.append([repr(x), mp.nstr(mp.besselj(1, x), 30), mp.nstr(mp.bessely(1, x), 30)])
with open("h1_table.json", "w") as f:
    json.dump({"source": "mpmath %s besselj/bessely order 1 at 40 digits" % mp.__version__, "rows": rows}, f, indent=0)
print(len(rows), "rows")
