"""SBML (level 3, fbc v2) genome-scale models -> CSR LP data, and the joint
cooperative community LP built from several of them.

The reference loads its GEMs with COBRApy (models/sbml/final_consortium/*.xml,
models/sbml/helper_candidates/*.xml; src/dfba_simulator.py:1326-1435) and
assembles the cooperative community LP in src/community_solver.py:1084-1149.
COBRApy is not available here, so this module reads the SBML XML directly:
species (metabolites; boundary species dropped), reactions with their
reactant/product stoichiometry, fbc flux-bound parameters and the active fbc
objective.  Exchange reactions are the ``R_EX_`` reactions; the extracellular
metabolite they exchange gives the medium key shared across species.
"""

from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix

from .workloads import BOUND, Consortium

_CORE = "{http://www.sbml.org/sbml/level3/version1/core}"
_FBC = "{http://www.sbml.org/sbml/level3/version1/fbc/version2}"


@dataclass
class GemModel:
    name: str
    S: csr_matrix                 # [metabolites, reactions]
    lower: np.ndarray
    upper: np.ndarray
    objective: np.ndarray         # [reactions]
    reaction_ids: list
    metabolite_ids: list
    exchanges: dict               # medium key -> reaction column

    @property
    def biomass_col(self) -> int:
        nz = np.flatnonzero(self.objective)
        if nz.size != 1:
            raise ValueError(f"{self.name}: expected one objective reaction, found {nz.size}")
        return int(nz[0])


def medium_key(species_id: str) -> str:
    """Compartment-free metabolite key: 'M_glc__D_e' and 'M_glc__D__40__e__41__'
    (the '(e)' encoding) both give 'glc__D'."""
    s = species_id[2:] if species_id.startswith("M_") else species_id
    s = re.sub(r"(__40__e__41__|\(e\)|\[e\]|_e)$", "", s)
    return s


def read_sbml(path, *, bound: float = BOUND) -> GemModel:
    """Parse one SBML L3 fbc-v2 model; infinite or missing bounds become +-bound."""
    root = ET.parse(path).getroot()
    model = root.find(_CORE + "model")
    if model is None:
        raise ValueError(f"{path}: no SBML model element")
    params = {}
    lp = model.find(_CORE + "listOfParameters")
    if lp is not None:
        for p in lp:
            params[p.get("id")] = float(p.get("value"))
    mets, boundary = [], set()
    for s in model.find(_CORE + "listOfSpecies"):
        if s.get("boundaryCondition") == "true":
            boundary.add(s.get("id"))
        else:
            mets.append(s.get("id"))
    met_index = {m: i for i, m in enumerate(mets)}
    rows, cols, vals = [], [], []
    rxn_ids, lower, upper, exchanges = [], [], [], {}
    for j, r in enumerate(model.find(_CORE + "listOfReactions")):
        rid = r.get("id")
        rxn_ids.append(rid)
        lo = params.get(r.get(_FBC + "lowerFluxBound"), -bound)
        hi = params.get(r.get(_FBC + "upperFluxBound"), bound)
        lower.append(max(lo, -bound) if np.isfinite(lo) else -bound)
        upper.append(min(hi, bound) if np.isfinite(hi) else bound)
        reactants = []
        for tag, sign in (("listOfReactants", -1.0), ("listOfProducts", 1.0)):
            lst = r.find(_CORE + tag)
            if lst is None:
                continue
            for ref in lst:
                sid = ref.get("species")
                st = float(ref.get("stoichiometry", "1"))
                if sign < 0:
                    reactants.append(sid)
                if sid in met_index:
                    rows.append(met_index[sid]); cols.append(j); vals.append(sign * st)
        if rid.startswith("R_EX_") and len(reactants) == 1:
            exchanges[medium_key(reactants[0])] = j
    objective = np.zeros(len(rxn_ids))
    lo_obj = model.find(_FBC + "listOfObjectives")
    if lo_obj is not None:
        active = lo_obj.get(_FBC + "activeObjective")
        for ob in lo_obj:
            if ob.get(_FBC + "id") != active:
                continue
            sense = -1.0 if ob.get(_FBC + "type") == "minimize" else 1.0
            for fo in ob.find(_FBC + "listOfFluxObjectives"):
                objective[rxn_ids.index(fo.get(_FBC + "reaction"))] = sense * float(fo.get(_FBC + "coefficient"))
    S = coo_matrix((vals, (rows, cols)), shape=(len(mets), len(rxn_ids))).tocsr()
    S.sum_duplicates()
    S.eliminate_zeros()
    S.sort_indices()
    return GemModel(name=model.get("id"), S=S, lower=np.array(lower), upper=np.array(upper),
                    objective=objective, reaction_ids=rxn_ids, metabolite_ids=mets, exchanges=exchanges)


UNLIMITED_KEYS = frozenset({"h2o", "h", "oh1"})   # src/community_solver.py:1146-1148


def common_medium(gems) -> list:
    """Medium keys exchanged by every member (sorted)."""
    keys = set(gems[0].exchanges)
    for g in gems[1:]:
        keys &= set(g.exchanges)
    return sorted(keys)


def union_medium(gems) -> list:
    """Medium keys exchanged by any member (sorted): one shared row each, as
    the reference writes one row per extracellular metabolite
    (src/community_solver.py:1120-1149)."""
    keys = set()
    for g in gems:
        keys |= set(g.exchanges)
    return sorted(keys)


def consortium_from_gems(gems, medium=None, *, seed: int = 0) -> Consortium:
    """Joint cooperative LP (src/community_solver.py:1084-1149): species blocks
    S_k v_k = 0, one shared-medium row per medium key (default: every key any
    member exchanges) with coefficient -biomass_s on each member's exchange of
    it (template -1), one growth row t - v_bio,k <= 0 per member, objective
    max t.  Water, protons and hydroxide rows get an unlimited supply.
    Michaelis-Menten uptake parameters are drawn from ``seed``."""
    medium = list(union_medium(gems) if medium is None else medium)
    ms = np.array([g.S.shape[0] for g in gems]); ns = np.array([g.S.shape[1] for g in gems])
    row_off = np.r_[0, np.cumsum(ms)]; col_off = np.r_[0, np.cumsum(ns)]
    n_flux = int(col_off[-1]); t_col = n_flux
    n_species_rows = int(row_off[-1]); K = len(medium); S_n = len(gems)
    medium_rows = n_species_rows + np.arange(K)
    growth_rows = n_species_rows + K + np.arange(S_n)
    rows, cols, vals = [], [], []
    for k, g in enumerate(gems):
        coo = g.S.tocoo()
        rows.append(coo.row + row_off[k]); cols.append(coo.col + col_off[k]); vals.append(coo.data)
    mask = np.array([[key in g.exchanges for key in medium] for g in gems], dtype=bool).reshape(S_n, K)
    medium_cols = np.array([[col_off[k] + g.exchanges[key] if key in g.exchanges else -1 for key in medium]
                            for k, g in enumerate(gems)], dtype=np.int64).reshape(S_n, K)
    for k in range(S_n):
        present = np.flatnonzero(mask[k])
        rows.append(medium_rows[present]); cols.append(medium_cols[k, present]); vals.append(-np.ones(present.size))
    biomass_cols = np.array([col_off[k] + g.biomass_col for k, g in enumerate(gems)])
    for k in range(S_n):
        rows.append(np.array([growth_rows[k], growth_rows[k]]))
        cols.append(np.array([biomass_cols[k], t_col]))
        vals.append(np.array([-1.0, 1.0]))
    n_rows = n_species_rows + K + S_n
    A = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                   shape=(n_rows, n_flux + 1)).tocsr()
    A.sum_duplicates(); A.sort_indices()
    coupling = np.full((S_n, K), -1, dtype=np.int64)
    for i, r in enumerate(medium_rows):
        seg = A.indices[A.indptr[r]:A.indptr[r + 1]]
        for k in range(S_n):
            if mask[k, i]:
                coupling[k, i] = A.indptr[r] + int(np.flatnonzero(seg == medium_cols[k, i])[0])
    lower = np.concatenate([g.lower for g in gems] + [np.array([0.0])])
    upper = np.concatenate([g.upper for g in gems] + [np.array([BOUND])])
    c = np.zeros(n_flux + 1); c[t_col] = -1.0
    rng = np.random.default_rng(seed + 104729)
    return Consortium(A=A, rhs_species=np.zeros(n_species_rows), lower=lower, upper=upper, c=c,
                      neq=n_species_rows, col_offsets=col_off, row_offsets=row_off, biomass_cols=biomass_cols,
                      medium_cols=medium_cols, coupling_nnz=coupling, medium_rows=medium_rows,
                      vmax=rng.uniform(1.0, 20.0, size=(S_n, K)), km=rng.uniform(0.01, 1.0, size=(S_n, K)),
                      t_col=t_col, medium_mask=mask,
                      medium_unlimited=np.array([key in UNLIMITED_KEYS for key in medium], dtype=bool))


def save_gems(path, gems) -> None:
    """Store parsed models as one .npz (no pickles): S in CSR, bounds,
    objective, exchange keys and columns per model."""
    arrays = {"n_models": np.array(len(gems))}
    for k, g in enumerate(gems):
        keys = sorted(g.exchanges)
        arrays.update({
            f"m{k}_name": np.array(g.name), f"m{k}_S_data": g.S.data, f"m{k}_S_indices": g.S.indices,
            f"m{k}_S_indptr": g.S.indptr, f"m{k}_S_shape": np.array(g.S.shape), f"m{k}_lower": g.lower,
            f"m{k}_upper": g.upper, f"m{k}_objective": g.objective,
            f"m{k}_exchange_keys": np.array(keys, dtype=str),
            f"m{k}_exchange_cols": np.array([g.exchanges[key] for key in keys], dtype=np.int64),
            f"m{k}_reaction_ids": np.array(g.reaction_ids, dtype=str),
            f"m{k}_metabolite_ids": np.array(g.metabolite_ids, dtype=str)})
    np.savez_compressed(path, **arrays)


def load_gems(path) -> list:
    """Inverse of ``save_gems`` (``allow_pickle=False``)."""
    with np.load(path, allow_pickle=False) as d:
        gems = []
        for k in range(int(d["n_models"])):
            S = csr_matrix((d[f"m{k}_S_data"], d[f"m{k}_S_indices"], d[f"m{k}_S_indptr"]),
                           shape=tuple(d[f"m{k}_S_shape"]))
            keys = [str(s) for s in d[f"m{k}_exchange_keys"]]
            cols = [int(c) for c in d[f"m{k}_exchange_cols"]]
            gems.append(GemModel(name=str(d[f"m{k}_name"]), S=S, lower=np.array(d[f"m{k}_lower"]),
                                 upper=np.array(d[f"m{k}_upper"]), objective=np.array(d[f"m{k}_objective"]),
                                 reaction_ids=[str(s) for s in d[f"m{k}_reaction_ids"]],
                                 metabolite_ids=[str(s) for s in d[f"m{k}_metabolite_ids"]],
                                 exchanges=dict(zip(keys, cols))))
    return gems
