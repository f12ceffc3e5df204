"""Temperature scans as the reference's example drivers run them, on replica batches.

The reference's heating and cooling scans (03example-BaTiO3-bulk-phase-transition/
16x16x16-40000-10000/heating.sh, cooling.sh) run one feram process per
temperature; each starts from the previous temperature's final state (its
.coord written with p / s_Nose, coord_module.F:474-479, read back as the
.restart by initialize-dipoR.F:22-37), thermalises n_thermalize steps and
averages n_average steps (molecular-dynamics.F:98-150).  ``temperature_chain``
runs that protocol for several independent seeds at once: the seeds are
replicas of one batch, so every temperature of every chain is one batched
call, and the seed-to-seed spread gives the error bars of <u>(T).
"""
from __future__ import annotations

import numpy as np

from .params import Param


# the nine potential-energy terms of Energy_type (energy_module.F:73-87)
POTENTIAL = ("short_range", "long_range", "dipole_E_field", "unharmonic", "homo_strain", "homo_coupling",
             "inho_strain", "inho_coupling", "inho_modulation")


def phase_of(u_sorted, threshold: float = 0.04) -> str:
    """Name of the ferroelectric phase from the sorted |<u>| (ascending):
    R (three components polarised), O (two), T (one) or C (none)."""
    n = int(np.sum(np.asarray(u_sorted) > threshold))
    return "CTOR"[n]


def temperature_chain(temps, seeds=(1, 2, 3, 4), base: Param | None = None, L: int = 16,
                      n_thermalize: int = 40000, n_average: int = 10000, precision: int = 0,
                      init_avg: float = 0.1, init_dev: float = 0.02, progress=None,
                      bases=None, variance: bool = False) -> dict:
    """Chained NP-MD scan over `temps` for every seed (or, with `bases`, one
    chain per given Param -- other fields, pressures -- seeded 1, 2, ...);
    returns per temperature and chain <u> (3), its sorted magnitudes and the
    averaged homogeneous strain (6), the chain means and standard errors of
    the sorted |<u>|, and with `variance` the per-component variance of u over
    the cells of the final state of each temperature; "pe" is the averaged
    potential energy per cell [eV]."""
    from .params import batio3_zvr
    from .system import Batch, System
    temps = [float(t) for t in temps]
    if bases is None:
        p0 = batio3_zvr(L, method="md", kelvin=temps[0]) if base is None else base.copy(kelvin=temps[0])
        bases = [p0] * len(seeds)
    else:
        seeds = tuple(range(1, len(bases) + 1))
    systems = []
    for sd, b in zip(seeds, bases):
        p = b.copy(kelvin=temps[0])
        g = System(p, precision=precision)
        rng = np.random.default_rng(sd)
        g.put("dipoR", init_avg + init_dev * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
        systems.append(g)
    u = np.zeros((len(temps), len(seeds), 3))
    var = np.zeros((len(temps), len(seeds), 3))
    pe = np.zeros((len(temps), len(seeds)))
    strain = np.zeros((len(temps), len(seeds), 6))
    with Batch(systems) as bt:
        for it, T in enumerate(temps):
            for g in systems:
                if it > 0:
                    # the next run reads p / s_Nose from the previous .coord
                    s = g.nose[0]
                    g.put("dipoP", g.get("dipoP") / s)
                g.set_kelvin(T)
                g.md_prologue()
            bt.step(n_thermalize)
            for g in systems:
                g.avg_reset()
                g.avg_auto(True)
            bt.step(n_average)
            for j, g in enumerate(systems):
                a = g.avg_get()
                g.avg_auto(False)
                u[it, j] = np.array(a["u"]) / (a["n_add"] * g.avg_u_norm())
                strain[it, j] = np.array(a["strain"]) / a["n_add"]
                pe[it, j] = sum(a["energy"][t] for t in POTENTIAL) / (a["n_add"] * g.N)
                if variance:                   # over the counted cells (films: the lower half)
                    var[it, j] = g.get("dipoR")[g.counted()].var(axis=0)
            if progress:
                progress(it, T, u[it])
    for g in systems:
        g.close()
    us = np.sort(np.abs(u), axis=2)
    err = us.std(axis=1, ddof=1) / np.sqrt(len(seeds)) if len(seeds) > 1 else np.full(us.shape[::2], np.nan)
    out = {"temps": temps, "seeds": list(seeds), "u": u, "u_sorted": us, "strain": strain, "pe": pe,
           "u_sorted_mean": us.mean(axis=1), "u_sorted_err": err,
           "phases": [[phase_of(us[i, j]) for j in range(len(seeds))] for i in range(len(temps))]}
    if variance:
        out["variance"] = var
    return out


def hl_field(amplitude, n_thermalize: int, n_average: int, i_step: int) -> np.ndarray:
    """E_field_add of step i_step of the 'hl' driver (hysteresis_loop.F:47-49;
    the amplitude before n_thermalize, initialize_E_field.F:58-66)."""
    A = np.asarray(amplitude, dtype=float)
    if i_step < n_thermalize:
        return A.copy()
    return A - 2 * A * float(i_step - n_thermalize) / float(n_average)


class _Each:
    """Steps each system through its own calls (the Batch interface)."""

    def __init__(self, systems):
        self.systems = systems

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False

    def step(self, nsteps: int = 1) -> list:
        return [g.step(nsteps) for g in self.systems]


def hysteresis_loop(systems, n_thermalize: int, n_average: int, n_hl_freq: int, amplitude,
                    i_start: int = 1, progress=None) -> list:
    """The 'hl' driver (hysteresis_loop.F:41-96) for one or more supercells:
    velocity-scaling steps i_start .. n_thermalize + n_average with the field
    E_add = amplitude before n_thermalize and swept linearly to -amplitude
    over n_average steps after it (evaluated on the device every step), and,
    every n_hl_freq steps of the sweep, a .hl row per system: step, kelvin,
    E_field_add, strain(6), <u> over the sum_p cells.  Several systems run as
    one replica batch.  Returns one list of rows per system."""
    from .system import Batch
    systems = list(systems)
    for g in systems:
        if g.p.method != "vs":
            raise ValueError("the hysteresis loop steps with velocity scaling (method 'vs')")
        g.set_E_schedule("hysteresis", amplitude, 0, n_thermalize, n_average, i_start)
    rows = [[] for _ in systems]
    n_end = n_thermalize + n_average
    outputs = [i for i in range(n_thermalize, n_end + 1, n_hl_freq) if i >= i_start]
    # one supercell steps through its own context, several as one replica batch
    with (_Each(systems) if len(systems) == 1 else Batch(systems)) as bt:
        i = i_start
        for out_step in outputs:
            bt.step(out_step - i + 1)
            i = out_step + 1
            E = hl_field(amplitude, n_thermalize, n_average, out_step)
            for j, g in enumerate(systems):
                u = g.mean_u()
                rows[j].append([out_step, g.p.kelvin] + E.tolist() + g.strain.tolist() + u.tolist())
            if progress:
                progress(out_step, [r[-1] for r in rows])
        if i <= n_end:
            bt.step(n_end - i + 1)
    return rows


def replica_ensemble(n: int = 48, seed: int = 20241203, edges=(12, 16, 24, 32)):
    """The batched-ensemble workload: n BaTiO3 bulk NP-MD replicas with edges
    drawn (seeded, uneven mix) from `edges`, T ~ U[150, 450] K and a static
    E_z ~ U[0, 100] kV/cm (0 to 1e-3 V/A, param_module.F:15 units;
    external_E_field = 0 0 E_z as initialize_E_field.F:58-66 applies it).
    Returns the Params (Q_Nose 0.1 at 16^3, the heating example's, scaled
    with N) and the initial u of each, uniform in [-0.05, 0.05] A."""
    from .params import batio3_zvr
    rng = np.random.default_rng(seed)
    Ls = rng.choice(list(edges), size=n)
    temps = rng.uniform(150.0, 450.0, size=n)
    fields = rng.uniform(0.0, 1e-3, size=n)
    params, u0 = [], []
    for L, T, Ez in zip(Ls, temps, fields):
        L = int(L)
        p = batio3_zvr(L, method="md", kelvin=float(T))
        p.Q_Nose = 0.1 * p.N / 16 ** 3
        p.external_E_field = [0.0, 0.0, float(Ez)]
        p.E_field_always_zero = False
        params.append(p)
        u0.append(rng.uniform(-0.05, 0.05, size=(L, L, L, 3)))
    return params, u0
