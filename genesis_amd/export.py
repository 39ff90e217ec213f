"""Observability exports: the metrics CSV and XYZ position frames
(src/export/mod.rs), written from a StatsSnapshot whose figures come from
device reductions (genesis_amd.stats) and from positions downloaded from the
resident state.

Numbers are formatted as Rust's ``format!`` does: ``{}`` of an integer is
its decimal digits, ``{:.N}`` of an f64 is the correctly rounded fixed-point
decimal with N digits (Python's ``f`` format gives the same digits for every
finite value, ties included, since both round the exact binary value),
``NaN`` and ``inf`` as Rust spells them.
"""
from __future__ import annotations

import math
import os
from typing import Sequence

import numpy as np

from .params import SYMBOLS
from .stats import StatsSnapshot

# src/export/mod.rs:14
CSV_HEADER = ("tick,time_s,entities,energy_total,energy_kinetic,energy_potential,energy_avg,temperature_avg,"
              "mean_speed,density,collisions,systems_run,fps,memory_kb,aggregates,largest,monomers,bound_pairs,"
              "bonded_pairs,bonds_formed,bond_lifetime_ticks,bond_energy,bond_species,chemical_aggregates,"
              "chemical_appeared,chemical_disappeared,chemical_fusions,chemical_scissions,chemical_compositions,"
              "chemical_reactions")


def fixed(x: float, digits: int) -> str:
    """Rust's ``{:.digits}`` of an f64."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return format(x, f".{digits}f")


def bond_species(matrix: Sequence[float]) -> str:
    """The non-zero species pairs of the (symmetric, flattened 10 x 10) bond
    matrix, lower triangle with the diagonal, as ``A-B:count`` sorted by count
    descending then pair name, joined by ``;`` (src/export/mod.rs:23-47)."""
    n = len(SYMBOLS)
    pairs = []
    for ia in range(n):
        for ib in range(ia + 1):
            k = ia * n + ib
            c = float(matrix[k]) if k < len(matrix) else 0.0
            if c > 0.0:
                pairs.append((f"{SYMBOLS[ia]}-{SYMBOLS[ib]}", int(c)))
    pairs.sort(key=lambda p: (-p[1], p[0]))
    return ";".join(f"{s}:{c}" for s, c in pairs)


def chemical_compositions(s: StatsSnapshot) -> str:
    """``formula:count@mean_binding`` per composition, ``;``-joined
    (src/export/mod.rs:52-60)."""
    if s.chemical is None:
        return ""
    return ";".join(f"{e['formula']}:{int(e['count'])}@{fixed(e['mean_binding'], 2)}"
                    for e in s.chemical["compositions"])


def chemical_reactions(s: StatsSnapshot) -> str:
    """``reactants->products`` (``+``-joined), ``xN`` when seen N > 1 times,
    ``;``-joined (src/export/mod.rs:67-88)."""
    if s.chemical is None:
        return ""
    out = []
    for r in s.chemical["reactions"]:
        part = "+".join(r["reactants"]) + "->" + "+".join(r["products"])
        if int(r["count"]) > 1:
            part += f"x{int(r['count'])}"
        out.append(part)
    return ";".join(out)


def csv_row(s: StatsSnapshot) -> str:
    """One metrics row in CSV_HEADER's column order (src/export/mod.rs:91-134)."""
    st = s.structure
    aggregates, monomers, largest, bound = ((st["aggregates"], st["monomers"], st["largest"], st["bound_pairs"])
                                            if st is not None else (0, 0, 0, 0))
    ch = s.chemical
    cagg = ch["aggregates"] if ch is not None else 0
    appeared = ch["appeared"] if ch is not None else 0
    disappeared = ch["disappeared"] if ch is not None else 0
    fusions = ch["fusions"] if ch is not None else 0
    scissions = ch["scissions"] if ch is not None else 0
    cols = [
        str(int(s.tick)), fixed(s.time, 6), str(int(s.entities)), fixed(s.energy_total, 6),
        fixed(s.energy_total - s.energy_potential, 6), fixed(s.energy_potential, 6), fixed(s.energy_avg, 6),
        fixed(s.temperature_avg, 6), fixed(s.mean_speed, 6), fixed(s.density, 6), str(int(s.collisions)),
        str(int(s.systems_run)), fixed(s.fps, 3), fixed(s.memory_bytes / 1024.0, 1), str(int(aggregates)),
        str(int(largest)), str(int(monomers)), str(int(bound)), str(int(s.bonded_pairs)),
        str(int(s.bonds_formed)), fixed(s.bond_lifetime_ticks, 1), fixed(s.bond_energy, 6),
        bond_species(s.bond_matrix), str(int(cagg)), str(int(appeared)), str(int(disappeared)),
        str(int(fusions)), str(int(scissions)), chemical_compositions(s), chemical_reactions(s),
    ]
    return ",".join(cols)


def append_csv(path: str, s: StatsSnapshot) -> None:
    """Appends one row, writing the header first when the file does not exist
    (src/export/mod.rs:138-150)."""
    fresh = not os.path.exists(path)
    with open(path, "a") as f:
        if fresh:
            f.write(CSV_HEADER + "\n")
        f.write(csv_row(s) + "\n")


def frame_text(tick: int, species: Sequence[int], x, y, z) -> str:
    """An XYZ frame: the atom count, ``frame {tick}``, then ``symbol x y z``
    with six decimals per atom, in entity order (src/export/mod.rs:155-173)."""
    species = np.asarray(species)
    lines = [str(len(species)), f"frame {int(tick)}"]
    for k in range(len(species)):
        lines.append(f"{SYMBOLS[int(species[k])]} {fixed(x[k], 6)} {fixed(y[k], 6)} {fixed(z[k], 6)}")
    return "\n".join(lines) + "\n"


def write_frame(prefix: str, tick: int, universe) -> str:
    """Writes ``{prefix}_{tick:08}.xyz`` from the resident positions (downloaded
    in entity order) and returns the path."""
    st = universe.download()
    path = f"{prefix}_{int(tick):08d}.xyz"
    with open(path, "w") as f:
        f.write(frame_text(tick, st.species, st.x, st.y, st.z))
    return path
