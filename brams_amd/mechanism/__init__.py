"""Chemical mechanism tables (species, reactions, stoichiometry).

The tables live in <mechanism>.json (relacs_tuv, relacs_mx, cb07_tuv,
racm_tuv), extracted once from the reference's generated SPACK kernels by
extract_relacs_tuv.py; they are data with the reference's species order
(e.g. RELACS_TUV/chem1_list.f90:63-109) and reaction numbering.

A process works with one mechanism, as a BRAMS build does: the environment
variable BRAMS_MECH names it (relacs_tuv by default), and the library, the
generated kernels and the oracle of that mechanism are the ones loaded.
"""
import json
import os
from dataclasses import dataclass, field
from functools import lru_cache
from typing import List

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MECHANISMS = ("relacs_tuv", "relacs_mx", "cb07_tuv", "racm_tuv")


def default_name():
    """The process's mechanism: BRAMS_MECH, relacs_tuv by default."""
    name = os.environ.get("BRAMS_MECH", "relacs_tuv").lower()
    if name not in MECHANISMS:
        raise ValueError("BRAMS_MECH=%s is not one of %s" % (name, ", ".join(MECHANISMS)))
    return name


JSON_PATH = os.path.join(HERE, "relacs_tuv.json")


@dataclass
class Species:
    name: str
    weight: float          # molar weight as the reference's REAL (float32) value
    transported: bool


@dataclass
class Reaction:
    reactants: List[int]   # 0-based species, in the order the rate multiplies them
    family: str
    params: List[float]
    post: list


@dataclass
class Mechanism:
    name: str
    species: List[Species]
    reactions: List[Reaction]
    production: dict = field(default_factory=dict)   # species -> [(coef, reaction)] 0-based
    nr_photo: int = 17

    @property
    def nspecies(self):
        return len(self.species)

    @property
    def nr(self):
        return len(self.reactions)

    def index(self, name):
        return [s.name for s in self.species].index(name)

    def stoichiometry(self):
        """Dense net stoichiometric matrix S (nspecies x nr)."""
        S = np.zeros((self.nspecies, self.nr))
        for sp, terms in self.production.items():
            for c, r in terms:
                S[sp, r] = c
        return S

    def jacobian_terms(self):
        """{(i,j): [(coef, r, other)]}: J_ij = sum coef*rk_r*y_other (other=-1: none),
        reactions ascending, one term per reactant slot equal to j
        (the order of chem_spack_jacdchemdc.f90)."""
        S = self.stoichiometry()
        present = {(sp, r) for sp, terms in self.production.items() for _, r in terms}
        out = {}
        for i in range(self.nspecies):
            for j in range(self.nspecies):
                terms = []
                for r, rx in enumerate(self.reactions):
                    if (i, r) not in present:
                        continue
                    for slot, sp in enumerate(rx.reactants):
                        if sp == j:
                            other = rx.reactants[1 - slot] if len(rx.reactants) == 2 else -1
                            terms.append((S[i, r], r, other))
                if terms:
                    out[(i, j)] = terms
        return out


def load_mechanism(name=None):
    """The tables of mechanism `name` (the process's mechanism by default)."""
    return _load(os.path.join(HERE, (name or default_name()) + ".json"))


@lru_cache(maxsize=None)
def _load(path):
    with open(path) as fh:
        d = json.load(fh)
    species = [Species(s["name"], float(np.float32(float(s["weight"]))), bool(s["transported"]))
               for s in d["species"]]
    reactions = [Reaction([x - 1 for x in r["reactants"]], r["rate"]["family"],
                          [float(x) for x in r["rate"]["params"]], r["rate"]["post"])
                 for r in d["reactions"]]
    production = {sp - 1: [(float(c), r - 1) for c, r in terms] for sp, terms in d["production"]}
    return Mechanism(d["name"], species, reactions, production, d["nr_photo"])
