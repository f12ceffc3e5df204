"""Run parameters of the per-step path, mirroring feram's ``Param_type``.

``Param`` holds the user-given members of ``Param_type``
(param_module.F:8-111) with the defaults of ``read_Param``
(param_module.F:129-194) and the derived members of ``make_Param``
(param_module.F:452-480).  ``Param.from_feram`` reads the ``tag = value``
lines of a ``.feram`` input with the same rules as ``read_Param``
(param_module.F:196-446): ``#`` lines and blank lines are skipped, species
tags take one or three values, an unknown tag raises ``ValueError`` (the
reference stops with "No such parameter").
"""
from __future__ import annotations

import dataclasses
import math
import re
from dataclasses import dataclass, field

AMU_IN_EV_A2_PS2 = 9648.53377163899   # param_module.F:26-33
KELVIN_EV = 11604.45                  # define.h:33
ELEMENTARY_CHARGE = 1.60217653        # define.h:34
RELAXOR_MAX = 3                       # define.h:17

METHODS = {"md": 0, "vs": 1, "lf": 2}
GEOMETRIES = {"bulk": 0, "film": 1, "epit": 2, "e110": 3}


def _three(v):
    return [float(v)] * RELAXOR_MAX


@dataclass
class Param:
    method: str = "md"
    GPa: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    kelvin: float = 0.0
    mass_dipo: list = field(default_factory=lambda: _three(100.0 / AMU_IN_EV_A2_PS2))
    mass_acou: list = field(default_factory=lambda: _three(-100.0))
    Q_Nose: float = 0.1
    bulk_or_film: str = "bulk"
    Lx: int = 32
    Ly: int = 32
    Lz: int = 32
    gap_id: int = 0
    padding_y: int = 1
    gap_dipole_u: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    a0: float = 4.0
    epi_strain: float = 0.0
    dt: float = 0.002
    n_thermalize: int = 10000
    n_average: int = 100000
    external_E_field: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    E_wave_type: str = ""
    n_E_wave_period: int = 0              # param_module.F:149
    n_hl_freq: int = 10000                # param_module.F:151
    n_hysteresis_loop_continue: int = 1   # param_module.F:152
    P_kappa2: list = field(default_factory=lambda: _three(0.0))
    P_alpha: list = field(default_factory=lambda: _three(0.0))
    P_gamma: list = field(default_factory=lambda: _three(0.0))
    P_k1: list = field(default_factory=lambda: _three(0.0))
    P_k2: list = field(default_factory=lambda: _three(0.0))
    P_k3: list = field(default_factory=lambda: _three(0.0))
    P_k4: list = field(default_factory=lambda: _three(0.0))
    j: list = field(default_factory=lambda: [0.0] * 7)
    B11: float = 0.0
    B12: float = 0.0
    B44: float = 0.0
    B1xx: float = 0.0
    B1yy: float = 0.0
    B4yz: float = 0.0
    modulation_constant: float = 0.0
    seed: list = field(default_factory=lambda: [123456789, 987654321])
    init_dipo_avg: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    init_dipo_dev: list = field(default_factory=lambda: [0.05, 0.05, 0.05])
    k_stripe: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    Z_star: list = field(default_factory=lambda: _three(0.0))
    epsilon_inf: float = 1.0
    kappa: float = 0.0
    delta_e: float = 0.0
    delta_e_acou: float = 0.0
    verbose: int = 1
    E_field_always_zero: bool = True

    # ---- derived members (make_Param, param_module.F:452-480) -------------
    @property
    def N(self) -> int:
        return self.Lx * self.Ly * self.Lz

    @property
    def Ninv(self) -> float:
        return 1.0 / self.N

    @property
    def dt_2(self) -> float:
        return self.dt / 2

    @property
    def n_freedom(self) -> int:
        return 3 * self.N

    @property
    def kappa_ewald(self) -> float:
        if self.kappa <= 0.0:
            return math.sqrt(math.pi) / self.a0 / float(self.N) ** (1.0 / 6)
        return self.kappa

    @property
    def unit_cell_volume(self) -> float:
        return self.a0 ** 3

    @property
    def pNa03(self) -> list:
        return [g * 0.01 / ELEMENTARY_CHARGE * self.N * self.unit_cell_volume for g in self.GPa]

    @property
    def temperature_eV(self) -> float:
        return self.kelvin / KELVIN_EV

    @property
    def target_kinetic_energy(self) -> float:
        return self.kelvin * 1.5 * self.N / KELVIN_EV

    @property
    def acoustic_md(self) -> bool:
        return self.mass_acou[0] > 0.0

    def copy(self, **changes) -> "Param":
        return dataclasses.replace(self, **{k: v for k, v in changes.items()})

    # ---- .feram reader --------------------------------------------------
    @classmethod
    def from_feram(cls, text: str) -> "Param":
        """Parse ``.feram`` text with the rules of read_Param (param_module.F:196-446)."""
        p = cls()
        species = {"P_kappa2", "P_alpha", "P_gamma", "P_k1", "P_k2", "P_k3", "P_k4", "Z_star"}
        for lineno, raw in enumerate(text.splitlines(), 1):
            s = raw.strip()
            if not s or s.startswith("#"):
                continue
            toks = re.findall(r"'[^']*'|\"[^\"]*\"|[^\s,]+", s)
            if len(toks) < 3 or toks[1] != "=":
                raise ValueError(f"line {lineno}: {raw!r}: Input error")
            name, vals = toks[0], toks[2:]

            def nums(n):
                out = []
                for v in vals[:n]:
                    out.append(float(v.replace("d", "e").replace("D", "e")))
                return out

            def first_float():
                return nums(1)[0]

            def species_vals():
                try:
                    v = nums(3)
                    if len(v) == 3:
                        return v
                except ValueError:
                    pass
                return _three(first_float())

            if name == "method":
                p.method = vals[0].strip("'\"").strip()
                if p.method == "hl":
                    p.E_field_always_zero = False
            elif name == "E_wave_type":
                p.E_wave_type = vals[0].strip("'\"").strip()
                p.E_field_always_zero = False
            elif name == "GPa":
                try:
                    v = nums(3)
                    if len(v) != 3:
                        raise ValueError
                except ValueError:
                    v = [first_float()] * 3
                p.GPa = v
            elif name == "kelvin":
                p.kelvin = first_float()
            elif name == "mass_amu":
                p.mass_dipo = [m / AMU_IN_EV_A2_PS2 for m in species_vals()]
            elif name == "acoustic_mass_amu":
                if p.method == "md":
                    raise ValueError("You cannot use method = 'md' togeter with acoustic_mass_amu.")
                p.mass_acou = [m / AMU_IN_EV_A2_PS2 for m in species_vals()]
            elif name in ("Q_Nose", "a0", "epi_strain", "dt", "B11", "B12", "B44", "B1xx",
                          "B1yy", "B4yz", "modulation_constant", "epsilon_inf", "kappa",
                          "delta_e", "delta_e_acou"):
                setattr(p, name, first_float())
            elif name == "bulk_or_film":
                p.bulk_or_film = vals[0].strip("'\"").strip()
                if p.bulk_or_film not in GEOMETRIES:
                    raise ValueError("Input error: bulk_or_film must be bulk, film or epit.")
            elif name == "L":
                p.Lx, p.Ly, p.Lz = (int(v) for v in vals[:3])
            elif name in ("padding_y", "gap_id", "n_thermalize", "n_average", "verbose"):
                setattr(p, name, int(vals[0]))
            elif name in ("gap_dipole_u", "init_dipo_avg", "init_dipo_dev", "k_stripe"):
                setattr(p, name, nums(3))
            elif name == "external_E_field":
                p.external_E_field = nums(3)
                if any(e != 0.0 for e in p.external_E_field):
                    p.E_field_always_zero = False
            elif name in species:
                setattr(p, name, species_vals())
            elif name == "j":
                p.j = nums(7)
            elif name == "seed":
                p.seed = [int(vals[0]), int(vals[1])]
            elif name in ("n_E_wave_period", "n_hl_freq", "n_hysteresis_loop_continue"):
                setattr(p, name, int(first_float()))
                if (name == "n_E_wave_period" and not p.E_wave_type.startswith("ramping")
                        and p.n_E_wave_period % 4 != 0):   # param_module.F:284-287
                    raise ValueError(f"line {lineno}: n_E_wave_period must be divided by 4.")
            elif name in ("n_coord_freq", "coord_directory", "slice_directory",
                          "distribution_directory", "continue", "plot_dispersion",
                          "output_polarization", "DDB_a", "DDB_b", "DDB_c", "DDB_d", "DDB_e",
                          "DDB_f", "DDB_g"):
                pass   # driver/output tags: owned by the caller of the step
            else:
                raise ValueError(f"line {lineno}: {raw!r}: Input error: No such parameter")
        return p


def batio3_zvr(L, *, method="md", kelvin=300.0, acoustic_mass_amu=None) -> Param:
    """The BaTiO3 effective-Hamiltonian set of the bulk phase-transition example
    (03example-BaTiO3-bulk-phase-transition/16x16x16-40000-10000/heating.sh:22-62)."""
    Lx, Ly, Lz = (L, L, L) if isinstance(L, int) else L
    p = Param(method=method, GPa=[-5.0] * 3, kelvin=kelvin, mass_dipo=_three(39.0 / AMU_IN_EV_A2_PS2),
              Q_Nose=0.1, Lx=Lx, Ly=Ly, Lz=Lz, a0=3.94, dt=0.002,
              P_kappa2=_three(5.502), P_alpha=_three(110.4), P_gamma=_three(-163.1),
              j=[-2.648, 3.894, 0.898, -0.789, 0.562, 0.358, 0.179],
              B11=126.0, B12=44.9, B44=50.3, B1xx=-211.0, B1yy=-19.3, B4yz=-7.75,
              Z_star=_three(9.956), epsilon_inf=5.24)
    if acoustic_mass_amu is not None:
        p.mass_acou = _three(acoustic_mass_amu / AMU_IN_EV_A2_PS2)
    return p
