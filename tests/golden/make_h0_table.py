"""Write h0_table.json: H0^(1)(x) = J0 + i Y0 at 30 significant digits
(mpmath) on arguments spanning the branches of the device evaluation and
the ranges the butterfly hits (x = k r up to ~3e5).  Run from this folder."""
import json

import mpmath as mp
import numpy as np

mp.mp.dps = 40
xs = np.concatenate([
    np.geomspace(1e-4, 4, 40), [3.999999, 4.0, 4.000001], np.linspace(4.05, 23.95, 40),
    [23.999999, 24.0, 24.000001], np.geomspace(24.5, 3.2e5, 60),
    [2.404825557695773, 5.520078110286311, 0.8935769662791675, 3.957678419314858],  # near zeros of J0, Y0
])
rows = []
for x in xs:
    x = float(x)
    j = mp.besselj(0, x)
    y = mp.bessely(0, x)
    rows.append([repr(x), mp.nstr(j, 30), mp.nstr(y, 30)])
with open("h0_table.json", "w") as f:
    json.dump({"source": "mpmath %s besselj/bessely at 40 digits" % mp.__version__, "rows": rows}, f, indent=0)
print(len(rows), "rows")
