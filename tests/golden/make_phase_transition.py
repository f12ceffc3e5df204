"""Write bto16_phase_transition.json from the reference's BaTiO3 example output.

The reference ships the .avg files of its 16^3 BaTiO3 heating and cooling
scans (03example-BaTiO3-bulk-phase-transition/16x16x16-40000-10000/heating.avg
and cooling.avg, written by average_module.F:243-245 after 40000 thermalising
and 10000 averaging NP-MD steps per temperature, each temperature started from
the previous one's final state).  This script copies, per row, the temperature,
the averaged homogeneous strain (6) and <u> (3) into a small JSON fixture.
It reads data files only.

    python tests/golden/make_phase_transition.py [/root/reference]
"""
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
d = os.path.join(ref, "src", "03example-BaTiO3-bulk-phase-transition", "16x16x16-40000-10000")
out = {"_about": "BaTiO3 bulk 16^3, GPa -5, Q_Nose 0.1, dt 2 fs, NP-MD, 40000 + 10000 steps per temperature, "
                 "chained in temperature; init_dipo_avg 0.1 dev 0.02 at the first temperature. Columns: "
                 "kelvin, strain(xx yy zz yz zx xy), u(x y z) [A] from the reference's .avg files "
                 "(average_module.F:243-245).",
       "source": "src/03example-BaTiO3-bulk-phase-transition/16x16x16-40000-10000/{heating,cooling}.avg"}
for name in ("heating", "cooling"):
    rows = []
    with open(os.path.join(d, name + ".avg")) as f:
        for line in f:
            v = line.split()
            if len(v) < 13:
                continue
            rows.append([float(v[0])] + [float(x) for x in v[4:10]] + [float(x) for x in v[10:13]])
    out[name] = rows
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "bto16_phase_transition.json"), "w") as f:
    json.dump(out, f, indent=0)
print(f"wrote {len(out['heating'])} heating and {len(out['cooling'])} cooling rows")
