"""Seeded synthetic FBA workloads for the batched LP engine.

Every LP is in the canonical minimisation form used by the reference's
backend boundary (``_problem`` in src/cpu_repeated_lp.py:119-147)::

    minimise  c @ x
    subject to A[:neq] @ x == rhs[:neq]
               A[neq:] @ x <= rhs[neq:]
               lower <= x <= upper

Five workloads are defined:

* ``species_lp``       - one genome-scale-like FBA LP (S is m x n CSR).
* ``uptake_batch``     - B LPs sharing S that differ only in Michaelis-Menten
  exchange lower bounds, lb = -Vmax*C/(Km+C) (the uptake form of
  src/dfba_simulator.py:646-666).
* ``consortium_lp``    - a three-species joint LP: block-diagonal species S,
  shared-medium coupling rows  sum_s -biomass_s * v_ex(s,i) <= C_i/dt
  (src/community_solver.py:1120-1149) and max-min growth rows t - mu_k <= 0
  (src/community_solver.py:1114-1119).
* ``consortium_batch`` - a batch of consortium LPs with per-environment
  concentrations and biomass (coupling coefficients vary per environment).
* the dFBA rollout driver (``lpe.rollout``) evolves that batch by explicit
  Euler, C <- max(0, C + sum_s v_ex*biomass_s*dt) (src/dfba_simulator.py:1352).

The joint consortium LP is fixed to the inequality form: 3 species blocks
(3600 equality rows) + 60 medium rows + 3 growth rows = 3663 rows and
5300 fluxes + t = 5301 columns.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.sparse import csr_matrix, coo_matrix


BOUND = 1000.0


@dataclass
class SpeciesModel:
    """One synthetic species: S (m x n CSR), bounds and column roles."""

    S: csr_matrix
    lower: np.ndarray
    upper: np.ndarray
    rhs: np.ndarray
    biomass_col: int
    exchange_cols: np.ndarray
    v0: np.ndarray

    @property
    def shape(self):
        return self.S.shape


def species_model(m: int = 1000, n: int = 1500, *, n_exchange: int | None = None,
                  biomass_nnz: int | None = None, seed: int = 0,
                  exchange_v0_nonnegative: bool = False) -> SpeciesModel:
    """Generate one species' stoichiometry exactly as the workload specifies.

    Generic columns get U{2..6} nonzeros at distinct uniform rows; 85% of the
    values are +-1, the rest +-2 or +-3.  The biomass column has
    ``biomass_nnz`` log-uniform [1e-3, 50] values of random sign; each
    exchange column has a single -1.  40% of the columns are reversible
    (lb = -1000), ub = 1000, and exchange uptake lb ~ -U(0.5, 20).  A point v0
    strictly inside the box fixes rhs = S v0, so the LP is feasible.
    ``exchange_v0_nonnegative`` draws v0 >= 0 on exchange columns so that any
    tighter uptake bound (lb <= 0) keeps v0 feasible.
    """
    if n_exchange is None:        # 100 exchange columns per 1500 fluxes
        n_exchange = max(1, int(round(100 * n / 1500)))
    if biomass_nnz is None:
        biomass_nnz = min(50, m)
    rng = np.random.default_rng(seed)
    biomass_col = 0
    exchange_cols = np.arange(n - n_exchange, n)
    generic = np.arange(1, n - n_exchange)
    rows, cols, vals = [], [], []
    counts = rng.integers(2, 7, size=generic.size)
    for j, k in zip(generic, counts):
        r = rng.choice(m, size=k, replace=False)
        mag = np.where(rng.random(k) < 0.85, 1.0, rng.choice([2.0, 3.0], size=k))
        sign = np.where(rng.random(k) < 0.5, -1.0, 1.0)
        rows.append(r); cols.append(np.full(k, j)); vals.append(mag * sign)
    r = rng.choice(m, size=biomass_nnz, replace=False)
    mag = np.exp(rng.uniform(np.log(1e-3), np.log(50.0), size=biomass_nnz))
    sign = np.where(rng.random(biomass_nnz) < 0.5, -1.0, 1.0)
    rows.append(r); cols.append(np.full(biomass_nnz, biomass_col)); vals.append(mag * sign)
    r = rng.integers(0, m, size=n_exchange)
    rows.append(r); cols.append(exchange_cols); vals.append(-np.ones(n_exchange))
    S = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                   shape=(m, n)).tocsr()
    S.sum_duplicates(); S.eliminate_zeros(); S.sort_indices()

    lower = np.where(rng.random(n) < 0.4, -BOUND, 0.0)
    upper = np.full(n, BOUND)
    lower[biomass_col] = 0.0
    lower[exchange_cols] = -rng.uniform(0.5, 20.0, size=n_exchange)
    # strictly interior point
    t = rng.uniform(0.02, 0.98, size=n)
    v0 = lower + t * (upper - lower)
    if exchange_v0_nonnegative:
        v0[exchange_cols] = rng.uniform(0.02, 0.98, size=n_exchange) * upper[exchange_cols]
    rhs = S @ v0
    return SpeciesModel(S=S, lower=lower, upper=upper, rhs=rhs, biomass_col=biomass_col,
                        exchange_cols=exchange_cols, v0=v0)


def species_lp(seed: int = 0, m: int = 1000, n: int = 1500):
    """Single-species FBA LP as a ``(A, rhs, lower, upper, c, neq)`` tuple."""
    sp = species_model(m, n, seed=seed)
    c = np.zeros(n); c[sp.biomass_col] = -1.0
    return sp.S.copy(), sp.rhs.copy(), sp.lower.copy(), sp.upper.copy(), c, m


def michaelis_menten_lower(vmax, km, conc):
    """Uptake bound -Vmax*C/(Km+C) (src/dfba_simulator.py:646-647)."""
    return -vmax * conc / (km + conc)


@dataclass
class UptakeBatch:
    """B LPs sharing S; per-LP exchange lower bounds."""

    S: csr_matrix
    rhs: np.ndarray          # [m]   shared
    lower: np.ndarray        # [B, n]
    upper: np.ndarray        # [B, n]
    c: np.ndarray            # [n]   shared
    neq: int
    exchange_cols: np.ndarray

    @property
    def batch(self):
        return self.lower.shape[0]

    def problem(self, b: int):
        return (self.S, self.rhs.copy(), self.lower[b].copy(), self.upper[b].copy(),
                self.c.copy(), self.neq)


def uptake_batch(batch: int = 256, seed: int = 1, m: int = 1000, n: int = 1500) -> UptakeBatch:
    """Shared-S batch: lb_j = -Vmax_j C_j/(Km_j + C_j), Vmax~U(1,20),
    Km~U(0.01,1), C~logU[1e-3,10], per instance and exchange column."""
    sp = species_model(m, n, seed=seed, exchange_v0_nonnegative=True)
    rng = np.random.default_rng(seed + 7919)
    k = sp.exchange_cols.size
    vmax = rng.uniform(1.0, 20.0, size=(batch, k))
    km = rng.uniform(0.01, 1.0, size=(batch, k))
    conc = np.exp(rng.uniform(np.log(1e-3), np.log(10.0), size=(batch, k)))
    lower = np.tile(sp.lower, (batch, 1))
    lower[:, sp.exchange_cols] = michaelis_menten_lower(vmax, km, conc)
    upper = np.tile(sp.upper, (batch, 1))
    c = np.zeros(n); c[sp.biomass_col] = -1.0
    return UptakeBatch(S=sp.S, rhs=sp.rhs, lower=lower, upper=upper, c=c, neq=m,
                       exchange_cols=sp.exchange_cols)


SPECIES_SHAPES = ((1200, 1800), (1500, 2200), (900, 1300))
UNLIMITED_SUPPLY = 1e6      # medium-row supply of unlimited metabolites (src/community_solver.py:1146-1148)
N_MEDIUM = 60


@dataclass
class Consortium:
    """Structure of the three-species joint LP (shared by every environment).

    Row order: species equality blocks, then ``N_MEDIUM`` medium rows, then
    one growth row per species.  Column order: species fluxes, then ``t``.
    The medium-row coefficient of species s on metabolite i is
    ``-biomass_s`` (consumption is a negative exchange flux); these
    ``3*N_MEDIUM`` entries are the per-environment matrix values, located at
    ``coupling_nnz`` positions of the CSR value array.

    A member need not exchange every medium metabolite (models read from SBML,
    ``lpe.sbml``): ``medium_mask[s, i]`` marks the present (species, metabolite)
    terms, and absent ones hold -1 in ``medium_cols`` / ``coupling_nnz``.  The
    present terms, species-major, are the per-environment coefficients
    (``var_positions``); ``medium_unlimited[i]`` marks metabolites whose row
    supply is effectively unbounded (water, protons: src/community_solver.py:1146-1148).
    """

    A: csr_matrix                 # template with biomass = 1
    rhs_species: np.ndarray       # [3600]
    lower: np.ndarray             # [n] template bounds
    upper: np.ndarray
    c: np.ndarray
    neq: int
    col_offsets: np.ndarray       # start column of each species
    row_offsets: np.ndarray
    biomass_cols: np.ndarray      # [3] joint column index of each species' biomass flux
    medium_cols: np.ndarray       # [3, N_MEDIUM] joint exchange columns coupled to the medium
    coupling_nnz: np.ndarray      # [3, N_MEDIUM] CSR value positions of -biomass_s
    medium_rows: np.ndarray       # [N_MEDIUM] joint row index
    vmax: np.ndarray              # [3, N_MEDIUM]
    km: np.ndarray                # [3, N_MEDIUM]
    t_col: int
    medium_mask: np.ndarray | None = None        # [S, K] bool (None: all present)
    medium_unlimited: np.ndarray | None = None   # [K] bool (None: none)

    @property
    def shape(self):
        return self.A.shape

    @property
    def n_medium(self) -> int:
        return self.medium_cols.shape[1]

    @property
    def n_species(self) -> int:
        return self.medium_cols.shape[0]

    def terms(self):
        """(species, medium index) of the present medium terms, species-major."""
        mask = self.medium_mask if self.medium_mask is not None else np.ones(self.medium_cols.shape, bool)
        return np.nonzero(mask)

    @property
    def term_cols(self) -> np.ndarray:
        s, i = self.terms()
        return self.medium_cols[s, i]

    @property
    def var_positions(self) -> np.ndarray:
        """CSR value positions of the per-environment coefficients (-biomass_s)."""
        s, i = self.terms()
        return self.coupling_nnz[s, i]

    def unlimited(self) -> np.ndarray:
        return (self.medium_unlimited if self.medium_unlimited is not None
                else np.zeros(self.n_medium, dtype=bool))


def consortium(seed: int = 2, shapes=SPECIES_SHAPES, n_medium: int = N_MEDIUM) -> Consortium:
    species = [species_model(m, n, seed=seed * 101 + k, exchange_v0_nonnegative=True,
                             n_exchange=max(int(round(100 * n / 1500)), n_medium))
               for k, (m, n) in enumerate(shapes)]
    ms = np.array([s.shape[0] for s in species]); ns = np.array([s.shape[1] for s in species])
    row_off = np.r_[0, np.cumsum(ms)]; col_off = np.r_[0, np.cumsum(ns)]
    n_species_rows, n_flux = int(row_off[-1]), int(col_off[-1])
    t_col = n_flux
    n_cols = n_flux + 1
    medium_rows = n_species_rows + np.arange(n_medium)
    growth_rows = n_species_rows + n_medium + np.arange(len(species))
    n_rows = n_species_rows + n_medium + len(species)

    rows, cols, vals = [], [], []
    for k, sp in enumerate(species):
        coo = sp.S.tocoo()
        rows.append(coo.row + row_off[k]); cols.append(coo.col + col_off[k]); vals.append(coo.data)
    medium_cols = np.stack([col_off[k] + sp.exchange_cols[:n_medium] for k, sp in enumerate(species)])
    for k in range(len(species)):
        rows.append(medium_rows); cols.append(medium_cols[k]); vals.append(-np.ones(n_medium))
    biomass_cols = np.array([col_off[k] + sp.biomass_col for k, sp in enumerate(species)])
    for k in range(len(species)):
        rows.append(np.array([growth_rows[k], growth_rows[k]]))
        cols.append(np.array([biomass_cols[k], t_col]))
        vals.append(np.array([-1.0, 1.0]))
    A = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                   shape=(n_rows, n_cols)).tocsr()
    A.sum_duplicates(); A.sort_indices()

    coupling_nnz = np.empty((len(species), n_medium), dtype=np.int64)
    for i, r in enumerate(medium_rows):
        seg = slice(A.indptr[r], A.indptr[r + 1])
        idx = A.indices[seg]
        for k in range(len(species)):
            pos = np.flatnonzero(idx == medium_cols[k, i])
            assert pos.size == 1
            coupling_nnz[k, i] = A.indptr[r] + pos[0]

    lower = np.concatenate([sp.lower for sp in species] + [np.array([0.0])])
    upper = np.concatenate([sp.upper for sp in species] + [np.array([BOUND])])
    c = np.zeros(n_cols); c[t_col] = -1.0
    rng = np.random.default_rng(seed + 104729)
    vmax = rng.uniform(1.0, 20.0, size=(len(species), n_medium))
    km = rng.uniform(0.01, 1.0, size=(len(species), n_medium))
    return Consortium(A=A, rhs_species=np.concatenate([sp.rhs for sp in species]),
                      lower=lower, upper=upper, c=c, neq=n_species_rows,
                      col_offsets=col_off, row_offsets=row_off, biomass_cols=biomass_cols,
                      medium_cols=medium_cols, coupling_nnz=coupling_nnz,
                      medium_rows=medium_rows, vmax=vmax, km=km, t_col=t_col)


@dataclass
class ConsortiumBatch:
    """Per-environment data of a consortium batch (everything else is shared)."""

    model: Consortium
    conc: np.ndarray        # [B, N_MEDIUM] mmol/L
    biomass: np.ndarray     # [B, 3] g/L
    dt: float = 1.0

    @property
    def batch(self):
        return self.conc.shape[0]

    def lower(self) -> np.ndarray:
        """[B, n] lower bounds with Michaelis-Menten uptake on medium exchanges."""
        md = self.model
        lo = np.tile(md.lower, (self.batch, 1))
        unl = md.unlimited()
        for s, i in zip(*md.terms()):
            if not unl[i]:
                lo[:, md.medium_cols[s, i]] = np.maximum(
                    michaelis_menten_lower(md.vmax[s, i], md.km[s, i], self.conc[:, i]), -BOUND)
        return lo

    def upper(self) -> np.ndarray:
        return np.tile(self.model.upper, (self.batch, 1))

    def rhs(self) -> np.ndarray:
        """[B, m]: species rhs, medium availability C/dt, growth rows 0."""
        md = self.model
        m = md.shape[0]
        out = np.zeros((self.batch, m))
        out[:, :md.neq] = md.rhs_species
        out[:, md.medium_rows] = np.where(md.unlimited(), UNLIMITED_SUPPLY, self.conc / self.dt)
        return out

    def coupling_values(self) -> np.ndarray:
        """[B, T] values of the present medium terms' coefficients (-biomass_s),
        in ``var_positions`` order."""
        s, _ = self.model.terms()
        return -np.maximum(self.biomass, 1e-12)[:, s]

    def matrix(self, b: int) -> csr_matrix:
        A = self.model.A.copy()
        A.data[self.model.var_positions] = self.coupling_values()[b]
        return A

    def problem(self, b: int):
        md = self.model
        return (self.matrix(b), self.rhs()[b], self.lower()[b], self.upper()[b],
                md.c.copy(), md.neq)


def consortium_batch(batch: int = 64, seed: int = 2, model: Consortium | None = None,
                     conc_range=(1e-3, 10.0)) -> ConsortiumBatch:
    model = model if model is not None else consortium(seed)
    rng = np.random.default_rng(seed + 15485863)
    lo, hi = conc_range
    conc = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(batch, model.medium_cols.shape[1])))
    biomass = rng.uniform(0.05, 0.5, size=(batch, model.medium_cols.shape[0]))
    return ConsortiumBatch(model=model, conc=conc, biomass=biomass)


@dataclass
class StatusBatch:
    """LPs of one pattern with known status: ``kind[b]`` is "optimal",
    "infeasible" or "unbounded"."""

    A: csr_matrix
    rhs: np.ndarray          # [B, m]
    lower: np.ndarray        # [B, n]
    upper: np.ndarray        # [B, n]
    c: np.ndarray            # [n] shared
    neq: int
    kind: list

    @property
    def batch(self):
        return self.lower.shape[0]

    def problem(self, b: int):
        return (self.A.copy(), self.rhs[b].copy(), self.lower[b].copy(), self.upper[b].copy(), self.c.copy(),
                self.neq)


def status_batch(n_each: int = 4, seed: int = 5, m: int = 200, n: int = 300) -> StatusBatch:
    """Feasible, primal infeasible and unbounded LPs sharing one pattern.

    A species LP (``species_model``) gets two extra columns p, q with +1 and
    -1 in one row, costs -0.01 each, lower bound 0.  Feasible LPs: uptake
    bounds of the exchanges drawn per LP, p and q bounded by 1000.  Primal
    infeasible LPs: one row's right-hand side moved one unit past the largest
    activity the box allows (a single-row Farkas ray).  Unbounded LPs: p and
    q without upper bound (the ray e_p + e_q keeps A x and lowers c.x)."""
    sp = species_model(m, n, seed=seed, exchange_v0_nonnegative=True)
    rng = np.random.default_rng(seed + 31)
    S = sp.S.tocoo()
    r_pq = int(rng.integers(0, m))
    rows = np.concatenate([S.row, [r_pq, r_pq]])
    cols = np.concatenate([S.col, [n, n + 1]])
    vals = np.concatenate([S.data, [1.0, -1.0]])
    A = coo_matrix((vals, (rows, cols)), shape=(m, n + 2)).tocsr()
    A.sum_duplicates(); A.sort_indices()
    c = np.zeros(n + 2); c[sp.biomass_col] = -1.0; c[n] = c[n + 1] = -0.01
    B = 3 * n_each
    lower = np.tile(np.r_[sp.lower, 0.0, 0.0], (B, 1))
    upper = np.tile(np.r_[sp.upper, BOUND, BOUND], (B, 1))
    rhs = np.tile(sp.rhs, (B, 1))
    k = sp.exchange_cols.size
    lower[:, sp.exchange_cols] = -rng.uniform(0.5, 20.0, size=(B, k))
    kind = []
    bad_rows = rng.choice(m, size=n_each, replace=False)
    for b in range(B):
        if b % 3 == 0:
            kind.append("optimal")
        elif b % 3 == 1:
            kind.append("infeasible")
            i = int(bad_rows[(b // 3) % n_each])
            seg = slice(A.indptr[i], A.indptr[i + 1])
            a_i, j = A.data[seg], A.indices[seg]
            top = np.sum(np.maximum(a_i * lower[b, j], a_i * upper[b, j]))
            rhs[b, i] = top + 1.0
        else:
            kind.append("unbounded")
            upper[b, n] = upper[b, n + 1] = np.inf
    return StatusBatch(A=A, rhs=rhs, lower=lower, upper=upper, c=c, neq=m, kind=kind)


@dataclass
class KnownOptimumBatch:
    """LPs of one pattern, each built around a chosen unique optimum: x_star
    (basic columns strictly inside their bounds, the others at a bound with
    a nonzero reduced cost of the right sign) and y_star (nonzero on the
    equality rows and the active inequality rows, zero on inactive ones)."""

    A: csr_matrix
    rhs: np.ndarray          # [B, m]
    lower: np.ndarray        # [B, n]
    upper: np.ndarray        # [B, n]
    c: np.ndarray            # [B, n]
    neq: int
    x_star: np.ndarray       # [B, n]
    y_star: np.ndarray       # [B, m]
    at_lower: np.ndarray     # [B, n] bool
    at_upper: np.ndarray     # [B, n] bool
    active_rows: np.ndarray  # [B, m - neq] bool

    @property
    def batch(self):
        return self.lower.shape[0]

    def problem(self, b: int):
        return (self.A.copy(), self.rhs[b].copy(), self.lower[b].copy(), self.upper[b].copy(), self.c[b].copy(),
                self.neq)


def known_optimum_batch(batch: int = 8, seed: int = 7, m: int = 120, n: int = 200, n_ineq: int = 30,
                        nnz_per_col: int = 6) -> KnownOptimumBatch:
    """Unique, primal- and dual-nondegenerate optima by construction: the
    active constraints' matrix restricted to the basic columns is square and
    nonsingular (columns chosen by a pivoted QR), basic columns sit at least
    20 % of their width inside the box, nonbasic reduced costs are at least
    0.5 in magnitude and active inequality duals at most -0.5."""
    from scipy.linalg import qr
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for j in range(n):
        r = rng.choice(m, size=nnz_per_col, replace=False)
        rows.append(r); cols.append(np.full(nnz_per_col, j))
        vals.append(rng.choice([-3.0, -2.0, -1.0, 1.0, 2.0, 3.0], size=nnz_per_col))
    A = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m, n)).tocsr()
    A.sum_duplicates(); A.eliminate_zeros(); A.sort_indices()
    Ad = A.toarray()
    neq = m - n_ineq
    out = {k: [] for k in ("rhs", "lower", "upper", "c", "x", "y", "lo", "up", "act")}
    for _ in range(batch):
        act = rng.random(n_ineq) < 0.5
        rows_act = np.r_[np.arange(neq), neq + np.flatnonzero(act)]
        _, R, piv = qr(Ad[rows_act], mode="economic", pivoting=True)
        k = rows_act.size
        assert abs(R[k - 1, k - 1]) > 1e-8 * abs(R[0, 0]), "active rows are rank deficient"
        basic = np.zeros(n, dtype=bool)
        basic[piv[:k]] = True
        lower = -rng.uniform(0.0, 10.0, size=n)
        upper = lower + rng.uniform(2.0, 20.0, size=n)
        at_up = ~basic & (rng.random(n) < 0.3)
        at_lo = ~basic & ~at_up
        x = np.where(at_lo, lower, np.where(at_up, upper, lower + rng.uniform(0.2, 0.8, size=n) * (upper - lower)))
        ax = Ad @ x
        rhs = ax.copy()
        inactive = neq + np.flatnonzero(~act)
        rhs[inactive] += rng.uniform(1.0, 5.0, size=inactive.size)
        y = np.zeros(m)
        y[:neq] = rng.normal(size=neq)
        y[neq + np.flatnonzero(act)] = -rng.uniform(0.5, 2.0, size=int(act.sum()))
        r = np.zeros(n)
        r[at_lo] = rng.uniform(0.5, 2.0, size=int(at_lo.sum()))
        r[at_up] = -rng.uniform(0.5, 2.0, size=int(at_up.sum()))
        c = Ad.T @ y + r
        for key, val in (("rhs", rhs), ("lower", lower), ("upper", upper), ("c", c), ("x", x), ("y", y),
                         ("lo", at_lo), ("up", at_up), ("act", act)):
            out[key].append(val)
    st = {k: np.array(v) for k, v in out.items()}
    return KnownOptimumBatch(A=A, rhs=st["rhs"], lower=st["lower"], upper=st["upper"], c=st["c"], neq=neq,
                             x_star=st["x"], y_star=st["y"], at_lower=st["lo"], at_upper=st["up"],
                             active_rows=st["act"])
