"""RELACS_TUV mechanism tables (species, reactions, stoichiometry).

The tables live in relacs_tuv.json, extracted once from the reference's
generated SPACK kernels by extract_relacs_tuv.py; they are data with the
reference's species order (chem1_list.f90:63-109) and reaction numbering.
"""
import json
import os
from dataclasses import dataclass, field
from functools import lru_cache
from typing import List

import numpy as np

JSON_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "relacs_tuv.json")


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


@lru_cache(maxsize=None)
def load_mechanism(path=JSON_PATH):
    with open(path) as fh:
        d = json.load(fh)
    species = [Species(s["name"], float(np.float32(float(s["weight"]))), bool(s["transported"]))
               for s in d["species"]]
    reactions = [Reaction([x - 1 for x in r["reactants"]], r["rate"]["family"],
                          [float(x) for x in r["rate"]["params"]], r["rate"]["post"])
                 for r in d["reactions"]]
    production = {sp - 1: [(float(c), r - 1) for c, r in terms] for sp, terms in d["production"]}
    return Mechanism(d["name"], species, reactions, production, d["nr_photo"])
