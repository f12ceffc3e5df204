"""Write bto32_hysteresis_360K.json from the reference's hysteresis-loop example.

22example-BaTiO3-new-param-hl/360.feram drives a 32^3 BaTiO3 supercell at
360 K ('hl': velocity scaling every step, hysteresis_loop.F) through a linear
field sweep from E_z = -0.003 to +0.003 V/A over n_average = 2000000 steps
after 20000 thermalising steps; 360.hl holds, every n_hl_freq = 5000 steps,
the step, temperature, E_field_add, strain(6) and <u>(3) over the sum_p cells
(hysteresis_loop.F:87-90).  This script parses 360.feram with
feram_amd.Param.from_feram (the reference's read_Param rules) and copies the
parameters and every .hl row into a JSON fixture.  It reads data files only.

    python tests/golden/make_hysteresis_loop.py [/root/reference]
"""
import dataclasses
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from feram_amd.params import Param  # noqa: E402

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
d = os.path.join(ref, "src", "22example-BaTiO3-new-param-hl")
with open(os.path.join(d, "360.feram")) as f:
    p = Param.from_feram(f.read())
rows = []
with open(os.path.join(d, "360.hl")) as f:
    for line in f:
        v = line.split()
        if len(v) == 14:
            rows.append([int(v[0])] + [float(x) for x in v[1:]])
out = {"_about": "22example-BaTiO3-new-param-hl/360.feram parsed with read_Param's rules, and every row of "
                 "360.hl: step, kelvin, E_field_add(3), strain(6), u(3) (hysteresis_loop.F:87-90).",
       "source": "src/22example-BaTiO3-new-param-hl/{360.feram,360.hl}",
       "param": dataclasses.asdict(p), "rows": rows}
with open(os.path.join(HERE, "bto32_hysteresis_360K.json"), "w") as f:
    json.dump(out, f, indent=0)
print(f"wrote {len(rows)} rows; method {p.method}, L {p.Lx}x{p.Ly}x{p.Lz}, kelvin {p.kelvin}, "
      f"n_thermalize {p.n_thermalize}, n_average {p.n_average}")
