"""Learned warm start: equilibrium dipole statistics predicted from (T, E_z, p).

The reference starts every run from ``init_dipo_avg`` / ``init_dipo_dev``
(initialize-dipoR.F:40-58) or from the previous run's ``.restart`` file, as
its heating scans do (03example-BaTiO3-bulk-phase-transition/.../heating.sh),
and thermalises for tens of thousands of steps.  This module trains a small
network (``Mlp``: fhip_mlp, BF16 MFMA kernels) on equilibrated states that the
plain MD produces (``equilibrium_samples``, run as one replica batch) and
uses it to start new runs near equilibrium (``WarmStart.initialise``).  The
MD still decides when a run is equilibrated.

Inputs: temperature [K], field E_z [V/A] (along z) and the strain control:
hydrostatic pressure [GPa] for bulk supercells, the clamped in-plane strain
epi_strain for epitaxial films ('epit', optimize-homo-strain.F:34-41), one
network per geometry (``WarmStart(control=...)``).  Outputs (13): the mean
dipole <u> in a canonical frame (bulk: magnitudes sorted ascending, so the
largest lies along z, the field axis; films: the two in-plane components
sorted, z kept as the film normal), the per-component variance of u over the
cells in the same order, the homogeneous strain (xx yy zz yz zx xy) in that
frame, and the potential energy per cell.

A warm start draws the dipoles as the predicted mean plus Gaussian
fluctuations with the spectrum of the quadratic dipole energy
(``shaped_fluctuations``) and the predicted variances, and gives the momenta
the kinetic energy that makes the total energy the predicted equilibrium one,
so that neither the structure nor the thermostat has much to relax.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .params import Param
from .sweeps import POTENTIAL

NIN, NOUT = 3, 13
# inputs are mapped to about [-1, 1] over the trained ranges
T_MID, T_HALF = 300.0, 150.0
E_MID, E_HALF = 5e-4, 5e-4
P_MID, P_HALF = -5.0, 1.0                 # hydrostatic pressure [GPa] (bulk)
S_MID, S_HALF = -0.01, 0.01               # epitaxial strain (films)


class MlpError(RuntimeError):
    pass


def normalise_inputs(T, Ez, s, control: str = "pressure") -> np.ndarray:
    """The three network inputs; s is the pressure [GPa] (control "pressure")
    or the epitaxial strain (control "epi_strain")."""
    T, Ez, s = np.broadcast_arrays(np.asarray(T, float), np.asarray(Ez, float), np.asarray(s, float))
    mid, half = (P_MID, P_HALF) if control == "pressure" else (S_MID, S_HALF)
    return np.stack([(T - T_MID) / T_HALF, (Ez - E_MID) / E_HALF, (s - mid) / half], axis=-1)


class Mlp:
    """The device network (fhip_mlp): 3-128-128-13, GELU, BF16 MFMA, FP32 master weights."""

    def __init__(self, seed: int = 0):
        self._L = _lib.load()
        q = C.c_void_p()
        rc = self._L.fhip_mlp_create(C.c_uint64(seed), C.byref(q))
        self.q = q
        if rc != 0:
            msg = self._L.fhip_mlp_last_error(q).decode() if q.value else ""
            self.close()
            raise MlpError(f"fhip_mlp_create: {_lib.ERRORS.get(rc, rc)}: {msg}")
        self.nparams = int(self._L.fhip_mlp_nparams())

    def close(self):
        if getattr(self, "q", None) is not None and self.q.value:
            self._L.fhip_mlp_destroy(self.q)
        self.q = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            raise MlpError(f"{what}: {_lib.ERRORS.get(rc, rc)}: {self._L.fhip_mlp_last_error(self.q).decode()}")

    @staticmethod
    def _f32(a, shape=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if shape is not None and a.shape != shape:
            raise ValueError(f"expected shape {shape}, got {a.shape}")
        return a, a.ctypes.data_as(C.POINTER(C.c_float))

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        self._check(self._L.fhip_mlp_get_params(self.q, out.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return out

    @params.setter
    def params(self, p) -> None:
        a, ptr = self._f32(p, (self.nparams,))
        self._check(self._L.fhip_mlp_set_params(self.q, ptr), "set_params")

    def forward(self, x) -> np.ndarray:
        x, px = self._f32(x)
        n = x.shape[0]
        y = np.zeros((n, NOUT), np.float32)
        self._check(self._L.fhip_mlp_forward(self.q, n, px, y.ctypes.data_as(C.POINTER(C.c_float))), "forward")
        return y

    def gradient(self, x, y, w=None):
        x, px = self._f32(x)
        y, py = self._f32(y, (x.shape[0], NOUT))
        w, pw = self._f32(np.ones(NOUT) if w is None else w, (NOUT,))
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        self._check(self._L.fhip_mlp_gradient(self.q, x.shape[0], px, py, pw, C.byref(loss),
                                              g.ctypes.data_as(C.POINTER(C.c_float))), "gradient")
        return float(loss.value), g

    def train(self, x, y, w=None, iters: int = 1000, lr: float = 1e-3) -> np.ndarray:
        x, px = self._f32(x)
        y, py = self._f32(y, (x.shape[0], NOUT))
        w, pw = self._f32(np.ones(NOUT) if w is None else w, (NOUT,))
        hist = np.zeros(iters, np.float32)
        self._check(self._L.fhip_mlp_train(self.q, x.shape[0], px, py, pw, int(iters), float(lr),
                                           hist.ctypes.data_as(C.POINTER(C.c_float))), "train")
        return hist


# ---- fluctuations with the spectrum of the quadratic dipole energy ---------------
KELVIN_EV = 11604.45            # K per eV (define.h:33)


def short_range_kernel(p: Param) -> np.ndarray:
    """S(k) on numpy's rfftn grid [kz][ky][kx <= Lx/2][a][b]: the Fourier
    transform of the j1..j7 neighbour couplings plus 2 kappa2 on the diagonal
    (dipole-dipole-short-range.F:35-83), with c = cos, s = sin of 2 pi k / L."""
    def cs(n, L):
        a = 2 * np.pi * np.arange(n) / L
        return np.cos(a), np.sin(a)
    cx, sx = cs(p.Lx // 2 + 1, p.Lx)
    cy, sy = cs(p.Ly, p.Ly)
    cz, sz = cs(p.Lz, p.Lz)
    cx, sx = cx[None, None, :], sx[None, None, :]
    cy, sy = cy[None, :, None], sy[None, :, None]
    cz, sz = cz[:, None, None], sz[:, None, None]
    j1, j2, j3, j4, j5, j6, j7 = p.j
    k2t = 2 * p.P_kappa2[0]
    cxyz = cx * cy * cz
    S = np.zeros((p.Lz, p.Ly, p.Lx // 2 + 1, 3, 3))
    S[..., 0, 0] = k2t + 2 * j2 * cx + 2 * j1 * (cy + cz) + 4 * j3 * cx * (cy + cz) + 4 * j4 * cy * cz + 8 * j6 * cxyz
    S[..., 1, 1] = k2t + 2 * j2 * cy + 2 * j1 * (cz + cx) + 4 * j3 * cy * (cz + cx) + 4 * j4 * cz * cx + 8 * j6 * cxyz
    S[..., 2, 2] = k2t + 2 * j2 * cz + 2 * j1 * (cx + cy) + 4 * j3 * cz * (cx + cy) + 4 * j4 * cx * cy + 8 * j6 * cxyz
    S[..., 1, 2] = S[..., 2, 1] = -4 * j5 * sy * sz - 8 * j7 * sy * sz * cx
    S[..., 2, 0] = S[..., 0, 2] = -4 * j5 * sz * sx - 8 * j7 * sz * sx * cy
    S[..., 0, 1] = S[..., 1, 0] = -4 * j5 * sx * sy - 8 * j7 * sx * sy * cz
    return S


def dipole_kernel(p: Param, long_range: np.ndarray) -> np.ndarray:
    """A(k) = S(k) + Z*^2 L(k) on the rfftn grid, the kernel of the quadratic
    dipole energy E = (1/2N) sum_k u^(k)^H A(k) u^(k) (dVddi-dipole-dipole.F:
    20-88, one Z* for all cells).  long_range: c%long_range in C order
    (z, y[+padding], kx, b, a), as System.get_long_range returns it."""
    L = np.asarray(long_range, float)[:, :p.Ly]
    Lsym = 0.5 * (L + np.swapaxes(L, -1, -2))
    return short_range_kernel(p) + p.Z_star[0] ** 2 * Lsym


def shaped_fluctuations(A: np.ndarray, shape, T: float, var, rng) -> np.ndarray:
    """Gaussian dipole fluctuations [z][y][x][3] with the equilibrium spectrum of
    the quadratic energy A(k) + Delta (Delta: one on-site shift that stands for
    the anharmonic and strain terms, set so that the mean per-component
    variance over the cells is mean(var)), each component then scaled to its
    own variance var[a].  With E = (1/2N) sum_k u^H (A + Delta) u^ the
    equilibrium covariance of u^(k) = sum_R u(R) e^{-ikR} is N kT (A + Delta)^-1,
    drawn here as sqrt(kT) (A + Delta)^(-1/2) w^(k) from the transform w^ of
    white unit noise."""
    Lz, Ly, Lx = shape
    N = Lx * Ly * Lz
    var = np.maximum(np.asarray(var, float), 1e-12)
    kT = T / KELVIN_EV
    lam, V = np.linalg.eigh(A)
    wk = np.full(A.shape[:3], 2.0)                     # kx and -kx on the half grid
    wk[:, :, 0] = 1.0
    if Lx % 2 == 0:
        wk[:, :, -1] = 1.0
    lmin = lam.min()
    target = var.mean()

    def mean_var(delta):
        return kT / N * np.sum(wk[..., None] / (lam + delta)) / 3

    lo = -lmin + 1e-9 * max(abs(lmin), 1.0)
    hi = max(1.0, 2 * abs(lmin))
    while mean_var(hi) > target:
        hi *= 2
    for _ in range(100):                               # var(Delta) falls monotonically
        mid = 0.5 * (lo + hi)
        if mean_var(mid) > target:
            lo = mid
        else:
            hi = mid
    scale = np.sqrt(kT / (lam + hi))                   # (A + Delta)^(-1/2) sqrt(kT) in the eigenbasis
    w = np.fft.rfftn(rng.standard_normal((3,) + tuple(shape)), axes=(1, 2, 3))
    w = np.moveaxis(w, 0, -1)                          # [kz][ky][kx][a]
    c = np.einsum("...ba,...b->...a", V, w)           # V^T w
    u = np.einsum("...ab,...b->...a", V, c * scale)   # V diag(scale) V^T w
    u = np.fft.irfftn(np.moveaxis(u, -1, 0), s=tuple(shape), axes=(1, 2, 3))
    u = np.moveaxis(u, 0, -1)
    u *= np.sqrt(var / u.reshape(-1, 3).var(axis=0))
    return u


# ---- canonical frame of an equilibrium state -------------------------------------
_SHEAR = {(1, 2): 3, (2, 1): 3, (2, 0): 4, (0, 2): 4, (0, 1): 5, (1, 0): 5}


def canonical(mean_u, var_u, strain, pe, film: bool = False):
    """Rotate (permute and reflect the axes of) an equilibrium state so that
    |<u>| ascends from x to z (films: from x to y, z stays the film normal)
    and every <u>_a >= 0; returns the 13 targets (the potential energy per
    cell pe is frame-independent)."""
    mean_u, var_u, strain = np.asarray(mean_u), np.asarray(var_u), np.asarray(strain)
    if film:
        perm = np.concatenate([np.argsort(np.abs(mean_u[:2]), kind="stable"), [2]])
    else:
        perm = np.argsort(np.abs(mean_u), kind="stable")      # new axis a <- old axis perm[a]
    sign = np.where(mean_u >= 0, 1.0, -1.0)
    m = np.abs(mean_u[perm])
    v = var_u[perm]
    e = np.empty(6)
    e[:3] = strain[perm]
    for (a, b), k in ((pair, _SHEAR[pair]) for pair in ((1, 2), (2, 0), (0, 1))):
        pa, pb = perm[a], perm[b]
        e[k] = strain[_SHEAR[(pa, pb)]] * sign[pa] * sign[pb]
    return np.concatenate([m, v, e, [float(pe)]])


def equilibrium_samples(conditions, L: int = 12, n_therm: int = 3000, n_avg: int = 1000,
                        snapshots: int = 4, precision: int = 0, seed: int = 0, base: Param | None = None,
                        batch_size: int = 256):
    """Equilibrate one replica per condition (T [K], E_z [V/A], GPa) with plain
    NP-MD from the reference's default start (init_dipo_avg 0.1, dev 0.02 per
    component, p = 0), all replicas in one batch, and return the 13 canonical
    targets per replica: <u> from the device averages over the last n_avg steps,
    the variance of u over the cells averaged over `snapshots` states, the
    averaged homogeneous strain and potential energy per cell."""
    from .params import batio3_zvr
    from .system import Batch, System
    conditions = np.asarray(conditions, dtype=float)
    out = np.zeros((len(conditions), NOUT))
    rng = np.random.default_rng(seed)
    for c0 in range(0, len(conditions), batch_size):
        chunk = conditions[c0:c0 + batch_size]
        systems = []
        for T, Ez, GPa in chunk:
            p = batio3_zvr(L, method="md", kelvin=float(T)) if base is None else base.copy(kelvin=float(T))
            p.GPa = [float(GPa)] * 3
            p.external_E_field = [0.0, 0.0, float(Ez)]
            p.E_field_always_zero = float(Ez) == 0.0
            p.Q_Nose = 0.1 * p.N / 16 ** 3
            g = System(p, precision=precision)
            g.put("dipoR", 0.1 + 0.02 * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
            g.md_prologue()
            systems.append(g)
        with Batch(systems) as bt:
            bt.step(n_therm)
            for g in systems:
                g.avg_reset()
                g.avg_auto(True)
            var = np.zeros((len(systems), 3))
            per = max(1, n_avg // snapshots)
            for _ in range(snapshots):
                bt.step(per)
                for i, g in enumerate(systems):
                    var[i] += g.get("dipoR").reshape(-1, 3).var(axis=0)
            var /= snapshots
        for i, g in enumerate(systems):
            a = g.avg_get()
            n = a["n_add"] * g.N
            mean_u = np.array(a["u"]) / n
            strain = np.array(a["strain"]) / a["n_add"]
            pe = sum(a["energy"][t] for t in POTENTIAL) / (a["n_add"] * g.N)
            out[c0 + i] = canonical(mean_u, var[i], strain, pe)
            g.close()
    return out


def chain_samples(temps, fields=(0.0, 2.5e-4, 5e-4, 1e-3), pressures=(-6.0, -5.0, -4.0), seeds: int = 2,
                  L: int = 16, n_thermalize: int = 6000, n_average: int = 3000, precision: int = 0,
                  progress=None):
    """Training data on the heating branch: one chained scan over `temps` per
    (field, pressure, seed), all chains in one batch (feram_amd.sweeps), and
    the 13 canonical targets of every (chain, temperature).  Returns
    (conditions [n][3] = T, E_z, GPa; targets [n][13])."""
    from .params import batio3_zvr
    from .sweeps import temperature_chain
    bases, conds = [], []
    for Ez in fields:
        for GPa in pressures:
            for _ in range(seeds):
                p = batio3_zvr(L, method="md", kelvin=float(temps[0]))
                p.GPa = [float(GPa)] * 3
                p.external_E_field = [0.0, 0.0, float(Ez)]
                p.E_field_always_zero = float(Ez) == 0.0
                bases.append(p)
                conds.append((float(Ez), float(GPa)))
    res = temperature_chain(temps, bases=bases, n_thermalize=n_thermalize, n_average=n_average,
                            precision=precision, variance=True, progress=progress)
    X, Y = [], []
    for i, T in enumerate(res["temps"]):
        for j, (Ez, GPa) in enumerate(conds):
            X.append((T, Ez, GPa))
            Y.append(canonical(res["u"][i, j], res["variance"][i, j], res["strain"][i, j], res["pe"][i, j]))
    return np.array(X), np.array(Y)


def epit_base(L=(32, 32, 8), epi_strain: float = -0.01, GPa: float = -5.0, kelvin: float = 350.0,
              Ez: float = 0.0) -> Param:
    """The reference's epitaxial-film example (07example-BaTiO3-epit-32x32-L004-
    D0-5.0GPa-0.01-350-500K/heating/heating.sh:22-61): an 'epit' film of Lz/2
    layers (gap_id 0) on the bulk parameter set, -5 GPa, Q_Nose 0.1."""
    from .params import batio3_zvr
    p = batio3_zvr(L, method="md", kelvin=float(kelvin))
    p.bulk_or_film = "epit"
    p.gap_id = 0
    p.epi_strain = float(epi_strain)
    p.GPa = [float(GPa)] * 3
    p.Q_Nose = 0.1
    p.external_E_field = [0.0, 0.0, float(Ez)]
    p.E_field_always_zero = float(Ez) == 0.0
    return p


# the reference's default start of its epit example (heating.sh:57-58)
EPIT_INIT_AVG = np.array([0.0, 0.0, 0.15])
EPIT_INIT_DEV = 0.04


def chain_samples_film(temps, strains=(-0.015, -0.01, -0.005), fields=(0.0, 5e-4), seeds: int = 2,
                       L=(32, 32, 8), GPa: float = -5.0, n_thermalize: int = 4000, n_average: int = 2000,
                       precision: int = 0, progress=None):
    """Training data of the film network: chained heating scans of epitaxial
    films (epit_base) per (epitaxial strain, field, seed), all chains in one
    batch, from the example's default start.  Returns (conditions [n][3] = T,
    E_z, epi_strain; targets [n][13], film canonical frame)."""
    from .sweeps import temperature_chain
    bases, conds = [], []
    for s in strains:
        for Ez in fields:
            for _ in range(seeds):
                bases.append(epit_base(L, s, GPa, float(temps[0]), Ez))
                conds.append((float(Ez), float(s)))
    res = temperature_chain(temps, bases=bases, n_thermalize=n_thermalize, n_average=n_average,
                            precision=precision, variance=True, progress=progress,
                            init_avg=EPIT_INIT_AVG, init_dev=EPIT_INIT_DEV)
    X, Y = [], []
    for i, T in enumerate(res["temps"]):
        for j, (Ez, s) in enumerate(conds):
            X.append((T, Ez, s))
            Y.append(canonical(res["u"][i, j], res["variance"][i, j], res["strain"][i, j], res["pe"][i, j],
                               film=True))
    return np.array(X), np.array(Y)


class WarmStart:
    """A trained network plus the target normalisation.  control: the third
    input, "pressure" (bulk supercells, GPa) or "epi_strain" (epitaxial
    films)."""

    def __init__(self, seed: int = 0, control: str = "pressure"):
        if control not in ("pressure", "epi_strain"):
            raise ValueError("control is 'pressure' or 'epi_strain'")
        self.control = control
        self.net = Mlp(seed)
        self.y_mean = np.zeros(NOUT)
        self.y_std = np.ones(NOUT)
        self._kernels = {}

    def fit(self, conditions, targets, iters: int = 3000, lr: float = 2e-3) -> np.ndarray:
        x = normalise_inputs(*np.asarray(conditions, float).T, control=self.control)
        targets = np.asarray(targets, float)
        self.y_mean = targets.mean(axis=0)
        self.y_std = np.maximum(targets.std(axis=0), 1e-12)
        y = (targets - self.y_mean) / self.y_std
        return self.net.train(x, y, iters=iters, lr=lr)

    def predict(self, T, Ez, s) -> np.ndarray:
        """The 13 outputs at temperature T, field E_z and strain control s."""
        x = normalise_inputs(T, Ez, s, control=self.control).reshape(-1, NIN)
        return self.net.forward(x) * self.y_std + self.y_mean

    def kernel(self, system) -> np.ndarray:
        """A(k) = S(k) + Z*^2 L(k) of the system's supercell, from its device
        L(k) table (cached per supercell and parameter set)."""
        p = system.p
        key = (p.Lx, p.Ly, p.Lz, tuple(p.j), p.P_kappa2[0], p.Z_star[0], p.epsilon_inf, p.a0, p.kappa)
        if key not in self._kernels:
            self._kernels[key] = dipole_kernel(p, system.get_long_range())
        return self._kernels[key]

    def initialise(self, system, T: float, Ez: float, s: float, seed: int = 0,
                   momenta: bool = True, shaped: bool = True, energy: bool = True) -> np.ndarray:
        """u(R) = <u> + fluctuations in the canonical frame (bulk: largest mean
        along z; films: z the film normal; all means >= 0), s the strain
        control (pressure or epitaxial strain): with `shaped`, Gaussian fluctuations with the
        spectrum of the quadratic dipole energy (shaped_fluctuations) and the
        predicted per-component variance, else white noise sqrt(var) N(0, 1).
        With momenta, p from the Maxwell-Boltzmann distribution at T
        (p_a ~ N(0, m k T)); with `energy` as well, p is then scaled so that
        the kinetic energy per cell is the predicted equilibrium potential plus
        3/2 kT minus the potential energy of the drawn u (one force
        evaluation), within [1/4, 7/4] of 3/2 kT, over the counted cells (a
        film's mirrored half is rewritten by the mirror).  Returns the
        prediction."""
        pred = self.predict(T, Ez, s)[0]
        m, v = np.maximum(pred[:3], 0.0), np.maximum(pred[3:6], 0.0)
        p = system.p
        rng = np.random.default_rng(seed)
        shape = (p.Lz, p.Ly, p.Lx, 3)
        if shaped:
            system.put("dipoR", m + shaped_fluctuations(self.kernel(system), shape[:3], T, v, rng))
        else:
            system.put("dipoR", m + np.sqrt(v) * rng.standard_normal(shape))
        if momenta:
            kT = T / KELVIN_EV
            mass = np.asarray(p.mass_dipo, float)
            mom = np.sqrt(mass * kT) * rng.standard_normal(shape)
            if energy:
                e = system.mirror_force_mirror()
                pe0 = sum(e[t] for t in POTENTIAL) / p.N
                cnt = system.counted()
                ke = float(np.sum(mom[cnt] ** 2 / mass)) / (2 * int(cnt.sum()))
                target = min(max(pred[12] + 1.5 * kT - pe0, 0.25 * 1.5 * kT), 1.75 * 1.5 * kT)
                mom *= np.sqrt(target / ke)
            system.put("dipoP", mom)
        return pred


# ---- what the warm start gains ------------------------------------------------------
def mser_index(series, batch: int = 5) -> int:
    """Truncation point of the MSER-5 rule (K. P. White, "An effective
    truncation heuristic for bias reduction in simulation output", Simulation
    69 (1997) 323-334): over batch means of `batch` records, the number d of
    leading batches (d at most half of them) whose removal minimises the
    squared standard error of the mean of the rest, sum_{i>d} (y_i - ybar_d)^2
    / (n - d)^2.  Returned in records (d * batch)."""
    s = np.asarray(series, float)
    n = len(s) // batch
    y = s[:n * batch].reshape(n, batch).mean(axis=1)
    best, best_d = None, 0
    for d in range(0, n // 2 + 1):
        r = y[d:]
        v = float(((r - r.mean()) ** 2).sum()) / len(r) ** 2
        if best is None or v < best:
            best, best_d = v, d
    return best_d * batch


def equilibration_index(series, tail: float = 0.5, window: int = 5, band: float = 4.0) -> int:
    """First record after which the smoothed series stays within `band`
    standard deviations of its mean over the last `tail` of the run."""
    s = np.asarray(series, float)
    k = np.ones(window) / window
    sm = np.convolve(s, k, mode="valid")
    t0 = int(len(sm) * (1 - tail))
    mu, sd = sm[t0:].mean(), max(sm[t0:].std(), 1e-12 * max(abs(sm[t0:].mean()), 1.0))
    out = np.abs(sm - mu) > band * sd
    bad = np.nonzero(out)[0]
    return 0 if len(bad) == 0 else int(bad[-1] + 1)


def heating_scan(ws: "WarmStart", temps, L: int = 16, n_steps: int = 4000, record: int = 10,
                 Ez: float = 0.0, GPa: float = -5.0, precision: int = 0, seed: int = 1,
                 base: Param | None = None, blocks: int = 5, Q_Nose: float | None = None,
                 shaped: bool = True, energy: bool = True, cold_avg=0.1, cold_dev=0.02) -> dict:
    """Run every temperature twice in one batch -- from the reference's default
    start (init_dipo_avg 0.1, dev 0.02, p = 0) and from the warm start -- record
    the potential energy per cell every `record` steps, and return per
    temperature the equilibration step of both runs and the sorted magnitudes
    of <u> over the second half of each run, as the mean and standard error of
    `blocks` block averages.  With `base` (e.g. an epitaxial film, epit_base)
    the runs take its geometry, and a film network gets the base's
    epi_strain as its third input; cold_avg / cold_dev give the default start
    (the epit example's is EPIT_INIT_AVG, EPIT_INIT_DEV).  The equilibration step "tau" is the MSER-5
    truncation point of the potential-energy series (mser_index; at most half
    the run, so a run that has not settled counts as n_steps / 2); "tau_band"
    is the last exit of the smoothed series from a 4-sigma band around its
    plateau (equilibration_index), which single late excursions can set."""
    from .params import batio3_zvr
    from .system import Batch, System
    rng = np.random.default_rng(seed)
    systems, kinds = [], []
    for T in temps:
        for kind in ("cold", "warm"):
            p = batio3_zvr(L, method="md", kelvin=float(T)) if base is None else base.copy(kelvin=float(T))
            p.GPa = [float(GPa)] * 3
            p.external_E_field = [0.0, 0.0, float(Ez)]
            p.E_field_always_zero = float(Ez) == 0.0
            if Q_Nose is not None:
                p.Q_Nose = Q_Nose
            g = System(p, precision=precision)
            if kind == "cold":
                g.put("dipoR", cold_avg + cold_dev * rng.standard_normal((p.Lz, p.Ly, p.Lx, 3)))
            else:
                sc = p.epi_strain if ws.control == "epi_strain" else float(GPa)
                ws.initialise(g, float(T), float(Ez), float(sc), seed=int(rng.integers(1 << 31)), shaped=shaped,
                              energy=energy)
            g.md_prologue()
            systems.append(g)
            kinds.append(kind)
    nrec = n_steps // record
    half = nrec // 2
    per_block = max(1, (nrec - half) // blocks)
    series = np.zeros((len(systems), nrec))
    usum_prev = np.zeros((len(systems), 3))
    nadd_prev = np.zeros(len(systems))
    block_u = [[] for _ in systems]
    with Batch(systems) as bt:
        for r in range(nrec):
            if r == half:
                for g in systems:
                    g.avg_reset()
                    g.avg_auto(True)
            es = bt.step(record)
            for i, e in enumerate(es):
                series[i, r] = sum(e[t] for t in POTENTIAL) / systems[i].N
            if r >= half and (r + 1 - half) % per_block == 0:
                for i, g in enumerate(systems):
                    a = g.avg_get()
                    u = np.array(a["u"])
                    dn = a["n_add"] - nadd_prev[i]
                    if dn > 0:
                        block_u[i].append(np.sort(np.abs((u - usum_prev[i]) / (dn * g.avg_u_norm()))))
                    usum_prev[i], nadd_prev[i] = u, a["n_add"]
    res = {"temps": list(map(float, temps)), "record": record, "n_steps": n_steps, "L": L,
           "cold": {"tau": [], "tau_band": [], "u": [], "u_err": []},
           "warm": {"tau": [], "tau_band": [], "u": [], "u_err": []}}
    for i, g in enumerate(systems):
        b = np.array(block_u[i])
        res[kinds[i]]["tau"].append(mser_index(series[i]) * record)
        res[kinds[i]]["tau_band"].append(equilibration_index(series[i]) * record)
        res[kinds[i]]["u"].append(b.mean(axis=0).tolist())
        res[kinds[i]]["u_err"].append((b.std(axis=0, ddof=1) / np.sqrt(len(b))).tolist() if len(b) > 1
                                      else [float("nan")] * 3)
        g.close()
    res["series"] = series
    tc, tw = sum(res["cold"]["tau"]), sum(res["warm"]["tau"])
    res["gain"] = tc / max(tw, record)
    uc, uw = np.array(res["cold"]["u"]), np.array(res["warm"]["u"])
    ec, ew = np.array(res["cold"]["u_err"]), np.array(res["warm"]["u_err"])
    res["P_within_2sigma"] = (np.abs(uc - uw) <= 2 * np.sqrt(ec ** 2 + ew ** 2) + 1e-3).all(axis=1).tolist()
    return res
