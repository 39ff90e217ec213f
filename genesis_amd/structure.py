"""Bond-graph stoichiometry lens: BondStructureSystem
(src/systems/bond_structure.rs:24-250) on the resident GPU state.

Every ``interval`` ticks it takes the persistent-bond graph (the Bonds
adjacency the observation writes), its connected components with their
formulas (``gs_bond_components``), the observed binding energy of each
component (``gs_bond_binding``: raw pair potentials summed over its edges)
and, against the previous sample, the lifecycle events: aggregates that
appeared or disappeared, fusions (a new aggregate whose members are exactly
the union of two or more previous ones) and scissions (the reverse), each
recorded as a reaction between formulas.  The result mirrors the
ChemicalStructure resource (src/stats/mod.rs:93-147).

The component search and the binding energies run on the device; the
lifecycle bookkeeping is integer set logic over a few hundred components per
sample, done here as the reference does it on its host.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

Members = Tuple[int, ...]


def chemical_structure(components: Sequence[dict], binding: Sequence[float], n_entities: int,
                       prev: Sequence[Tuple[Members, str]], tick: int) -> Tuple[dict, List[Tuple[Members, str]]]:
    """One BondStructureSystem sample (src/systems/bond_structure.rs:84-249).

    components: bond_components records (size, formula, members ascending),
    in their order (size descending, then formula); binding[k]: the binding
    energy of components[k]; prev: the (members, formula) list of the
    previous sample.  Returns the ChemicalStructure record and the list to
    pass as ``prev`` next time."""
    bound_entities = int(sum(c["size"] for c in components))
    largest = max((int(c["size"]) for c in components), default=0)
    # per-formula count and binding sum, in component order (:98-123)
    per_formula: Dict[str, List[float]] = {}
    for c, e in zip(components, binding):
        entry = per_formula.setdefault(c["formula"], [0.0, 0])
        entry[0] += float(e)
        entry[1] += 1
    compositions = [{"formula": f, "count": int(cnt), "mean_binding": (s / cnt) if cnt > 0 else 0.0}
                    for f, (s, cnt) in per_formula.items()]
    compositions.sort(key=lambda c: (-c["count"], c["formula"]))
    # lifecycle against the previous sample (:138-223)
    current = [(tuple(int(m) for m in c["members"]), c["formula"]) for c in components]
    current_keys = {m for m, _ in current}
    prev_keys = {m for m, _ in prev}
    current_by = dict(current)
    prev_by = dict(prev)
    appeared = len(current_keys - prev_keys)
    disappeared = len(prev_keys - current_keys)
    reactions: Dict[Tuple[Tuple[str, ...], Tuple[str, ...]], int] = {}
    fusions = 0
    for new_key in current_keys - prev_keys:
        new_set = set(new_key)
        contained = [pk for pk in prev_keys if set(pk) <= new_set]
        if len(contained) >= 2 and set().union(*map(set, contained)) == new_set:
            fusions += 1
            reactants = tuple(sorted(prev_by.get(k, "?") for k in contained))
            key = (reactants, (current_by.get(new_key, "?"),))
            reactions[key] = reactions.get(key, 0) + 1
    scissions = 0
    for old_key in prev_keys - current_keys:
        old_set = set(old_key)
        contained = [ck for ck in current_keys if set(ck) <= old_set]
        if len(contained) >= 2 and set().union(*map(set, contained)) == old_set:
            scissions += 1
            products = tuple(sorted(current_by.get(k, "?") for k in contained))
            key = ((prev_by.get(old_key, "?"),), products)
            reactions[key] = reactions.get(key, 0) + 1
    # count descending, then the reactants joined by '+' (:225-233); the
    # products break the ties the reference leaves to its hash order
    reaction_list = [{"reactants": list(r), "products": list(p), "count": c} for (r, p), c in reactions.items()]
    reaction_list.sort(key=lambda r: (-r["count"], "+".join(r["reactants"]), "+".join(r["products"])))
    rec = {"tick": int(tick), "aggregates": len(components), "bound_entities": bound_entities,
           "monomers": max(0, int(n_entities) - bound_entities), "largest": largest,
           "compositions": compositions, "appeared": appeared, "disappeared": disappeared,
           "fusions": fusions, "scissions": scissions, "reactions": reaction_list}
    return rec, current


class BondStructureSystem:
    """BondStructureSystem (src/systems/bond_structure.rs:24-250) on a GPU universe:
    samples every ``interval`` ticks while bond observation is enabled."""

    def __init__(self, interval: int = 1):
        self.interval = max(1, int(interval))
        self.prev: List[Tuple[Members, str]] = []

    def name(self) -> str:
        return "bond-structure"

    def run(self, universe, tick: int, species) -> Optional[dict]:
        if tick % self.interval != 0 or not universe.physics.enable_bond_observation:
            return None
        species = np.asarray(species)
        components = universe.bond_components(species)
        per_entity = universe.bond_binding()
        # a component's binding energy: its members' partial sums (each edge
        # is counted at its smaller id), in member order
        binding = [float(np.sum(per_entity[np.asarray(c["members"], dtype=np.int64)])) for c in components]
        rec, self.prev = chemical_structure(components, binding, universe.n, self.prev, tick)
        return rec
