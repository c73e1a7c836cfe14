"""Molecules flattened to the per-atom arrays the hot path reads (System.expand of the reference).

Mirrors pgm_jax/system.py: a `Molecule` template (elements, charges, Gaussian radii,
polarizabilities, covalent dipoles (i, j, c), LJ R*/2 and sqrt(eps), masses) and a `System` of
molecules in order, flattened to (N,) arrays, the molecule index of every atom and the covalent
dipole index arrays cov_i, cov_j, cov_c.
"""

from __future__ import annotations

import dataclasses

import numpy as np

MASSES = {"H": 1.008, "Li": 6.94, "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998, "Na": 22.990, "P": 30.974,
          "S": 32.06, "Cl": 35.45, "K": 39.098, "Br": 79.904, "Rb": 85.468, "I": 126.90, "Cs": 132.91}  # standard atomic weights
# the parameters of the hot path (system.py:84-94; the Gaussian vdW and quadrupole quantities of the
# reference are outside it) and their default tying
QUANTITIES = ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "cov")
BY_TYPE = ("radius", "alpha", "lj_rmin_half", "lj_sqrt_eps")


@dataclasses.dataclass
class Molecule:
    """A molecule template (units nm, e, e nm, nm^3, amu)."""

    name: str
    elements: list
    q: np.ndarray
    radius: np.ndarray
    alpha: np.ndarray
    cov: list
    lj_rmin_half: np.ndarray
    lj_sqrt_eps: np.ndarray
    masses: np.ndarray | None = None
    types: list | None = None  # atom types (tying keys of the by-type quantities); default: the elements
    bonds: list | None = None  # (i, j) bonds of the symmetry classes (besides the covalent-dipole pairs)
    keys: dict | None = None   # {quantity: key per atom / per covalent dipole}: overrides of the tying

    def __post_init__(self) -> None:
        """Arrays as float; element masses and element types by default."""
        for f in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps"):
            setattr(self, f, np.asarray(getattr(self, f), float))
        self.masses = np.array([MASSES[e] for e in self.elements]) if self.masses is None else np.asarray(self.masses, float)
        self.types = list(self.elements) if self.types is None else list(self.types)
        self.bonds = [] if self.bonds is None else [tuple(b) for b in self.bonds]
        self.keys = {} if self.keys is None else dict(self.keys)

    @property
    def n(self) -> int:
        """Number of atoms."""
        return len(self.elements)

    def values(self, quantity: str) -> np.ndarray:
        """The template's values of a quantity (per atom, or per covalent dipole for "cov")."""
        return np.array([c for _, _, c in self.cov], float) if quantity == "cov" else getattr(self, quantity)

    def symmetry_classes(self) -> list:
        """Class id per atom by colour refinement on the graph of bonds and covalent-dipole pairs
        (system.py:265-298): initial colours element|type, refined by the sorted colours of the
        neighbours until the class count stops growing; ids from sorted signatures, so independent
        of the atom order."""
        adj = [set() for _ in range(self.n)]
        for i, j in list(self.bonds) + [(i, j) for i, j, _ in self.cov]:
            if i != j:
                adj[i].add(j)
                adj[j].add(i)

        def relabel(sig: list) -> list:
            rank = {x: k for k, x in enumerate(sorted(set(sig)))}
            return [rank[x] for x in sig]

        colour = relabel([f"{e}|{t}" for e, t in zip(self.elements, self.types)])
        while True:
            new = relabel([(colour[i], tuple(sorted(colour[k] for k in adj[i]))) for i in range(self.n)])
            if len(set(new)) == len(set(colour)):
                return colour
            colour = new

    def tying_keys(self) -> dict:
        """{quantity: key per atom (per covalent dipole for "cov")} (system.py:300-343): radius,
        alpha and LJ by atom type; q "<name>:<label>" and cov "<name>:<label_i>><label_j>" with label
        the type, or "<type>.<k>" when a type has several symmetry classes; `keys` overrides."""
        cls = self.symmetry_classes()
        label = {}
        for t in set(self.types):
            cs = sorted({c for c, tt in zip(cls, self.types) if tt == t})
            for k, c in enumerate(cs):
                label[c] = t if len(cs) == 1 else f"{t}.{k}"
        atom = [label[c] for c in cls]
        out = {qn: list(self.types) for qn in BY_TYPE}
        out["q"] = [f"{self.name}:{a}" for a in atom]
        out["cov"] = [f"{self.name}:{atom[i]}>{atom[j]}" for i, j, _ in self.cov]
        for qn, ks in self.keys.items():
            if qn not in QUANTITIES:
                raise KeyError(f"unknown quantity {qn!r}")
            if len(ks) != len(self.values(qn)):
                raise ValueError(f"{self.name}: {len(ks)} keys for {qn}")
            out[qn] = list(ks)
        return out


class System:
    """Flattened system: per-atom arrays q, radius, alpha, lj_rmin_half, lj_sqrt_eps, masses, mol."""

    def __init__(self, molecules: list | None = None, **arrays) -> None:
        """Flatten `molecules`, or take the arrays directly (q, radius, alpha, lj_rmin_half,
        lj_sqrt_eps, masses, mol, cov_i, cov_j, cov_c)."""
        if molecules is not None:
            offs = np.cumsum([0] + [m.n for m in molecules])
            cat = lambda f: np.concatenate([getattr(m, f) for m in molecules])  # noqa: E731
            ci, cj, cc = [], [], []
            for k, m in enumerate(molecules):
                for i, j, c in m.cov:
                    ci.append(offs[k] + i)
                    cj.append(offs[k] + j)
                    cc.append(c)
            arrays = dict(
                q=cat("q"), radius=cat("radius"), alpha=cat("alpha"), lj_rmin_half=cat("lj_rmin_half"),
                lj_sqrt_eps=cat("lj_sqrt_eps"), masses=cat("masses"),
                mol=np.concatenate([np.full(m.n, k) for k, m in enumerate(molecules)]),
                cov_i=np.array(ci), cov_j=np.array(cj), cov_c=np.array(cc, float),
            )
        self._table(molecules)
        self.q = np.ascontiguousarray(arrays["q"], float)
        self.radius = np.ascontiguousarray(arrays["radius"], float)
        self.alpha = np.ascontiguousarray(arrays["alpha"], float)
        self.lj_rmin_half = np.ascontiguousarray(arrays["lj_rmin_half"], float)
        self.lj_sqrt_eps = np.ascontiguousarray(arrays["lj_sqrt_eps"], float)
        self.masses = np.ascontiguousarray(arrays.get("masses", np.ones(len(self.q))), float)
        self.mol = np.ascontiguousarray(arrays["mol"], np.int32)
        self.cov_i = np.ascontiguousarray(arrays["cov_i"], np.int32)
        self.cov_j = np.ascontiguousarray(arrays["cov_j"], np.int32)
        self.cov_c = np.ascontiguousarray(arrays["cov_c"], float)
        n = len(self.q)
        for k in ("radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "masses", "mol"):
            if len(getattr(self, k)) != n:
                raise ValueError(f"System: {k} has {len(getattr(self, k))} entries, expected {n}")

    @property
    def n(self) -> int:
        """Number of atoms N."""
        return len(self.q)

    @property
    def nmol(self) -> int:
        """Number of molecules."""
        return int(self.mol.max()) + 1 if self.n else 0

    # ------------------------------------------------------------------ tied parameter table
    def _table(self, molecules) -> None:
        """The tied table (ParamTable, system.py:345-447): keys per quantity in first-occurrence order
        over the templates (each template counted once), initial value = mean over the tied members,
        and the key index of every atom / covalent dipole.  Systems given as arrays tie nothing: one
        key per atom (per covalent dipole)."""
        if molecules is None:
            self.param_keys = None
            return
        members = {qn: {} for qn in QUANTITIES}
        seen = set()
        for m in molecules:
            if id(m) in seen:
                continue
            seen.add(id(m))
            keys = m.tying_keys()
            for qn in QUANTITIES:
                for k, v in zip(keys[qn], m.values(qn)):
                    members[qn].setdefault(k, []).append(float(v))
        self.param_keys = {qn: list(members[qn]) for qn in QUANTITIES}
        self.params0 = {qn: np.array([np.mean(members[qn][k]) for k in self.param_keys[qn]], float) for qn in QUANTITIES}
        pos = {qn: {k: i for i, k in enumerate(self.param_keys[qn])} for qn in QUANTITIES}
        idx = {qn: [] for qn in QUANTITIES}
        for m in molecules:
            keys = m.tying_keys()
            for qn in QUANTITIES:
                idx[qn] += [pos[qn][k] for k in keys[qn]]
        self.param_index = {qn: np.array(v, np.int64) for qn, v in idx.items()}

    def initial_params(self) -> dict:
        """The initial parameter pytree {quantity: (K,) array} (ParamTable.initial)."""
        if self.param_keys is None:
            return {qn: self._per_atom(qn).copy() for qn in QUANTITIES}
        return {qn: v.copy() for qn, v in self.params0.items()}

    def _per_atom(self, qn: str) -> np.ndarray:
        return self.cov_c if qn == "cov" else getattr(self, qn)

    def expand(self, params: dict | None = None) -> dict:
        """{quantity: per-atom array (per covalent dipole for "cov")} of a parameter pytree
        (System.expand, system.py:561-577); None: the system's own values."""
        if params is None:
            return {qn: self._per_atom(qn) for qn in QUANTITIES}
        out = {}
        for qn in QUANTITIES:
            v = np.asarray(params[qn], float) if qn in params else None
            if v is None:
                out[qn] = self._per_atom(qn)
            elif self.param_keys is None:
                out[qn] = v
            else:
                out[qn] = v[self.param_index[qn]]
        return out

    def reduce_params(self, cot: dict) -> dict:
        """The gradient over the parameter table from per-atom cotangents: sums over the atoms (covalent
        dipoles) of each key -- the transpose of `expand`."""
        out = {}
        for qn in QUANTITIES:
            v = np.asarray(cot[qn].cpu() if hasattr(cot[qn], "cpu") else cot[qn], float)
            if self.param_keys is None:
                out[qn] = v
            else:
                g = np.zeros(len(self.param_keys[qn]))
                np.add.at(g, self.param_index[qn], v)
                out[qn] = g
        return out
