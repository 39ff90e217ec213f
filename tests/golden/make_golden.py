#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ (inputs + oracle outputs).

The reference ships no golden vectors for this path (its tests are closed
forms and conservation checks), so these are the oracle's outputs on seeded
inputs, cross-checked when written: forces between the oracle's grid path and
its brute-force path, g(r) counts against a numpy brute force, and cluster
labels against a brute-force union-find (tests/test_oracle_lenses.py does the
same on other inputs).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from genesis_amd.inputs import mixture, plain_lj  # noqa: E402
from genesis_amd.params import DEFAULT_SIGMA, GRID_MINIMAGE, Physics  # noqa: E402
from oracle import oracle as O  # noqa: E402


def physics_vector(ph: Physics) -> np.ndarray:
    """Flat record of the physics parameters the fixture used."""
    return np.array(list(ph.box) + [ph.dt] + list(ph.sigma) + list(ph.epsilon_k) + list(ph.charge) + list(ph.mass)
                    + [ph.thermal_constant, ph.lj_cutoff_factor, ph.coulomb_constant, ph.coulomb_cutoff,
                       ph.thermostat_temperature, ph.thermostat_tau, float(ph.enable_electrostatics),
                       float(ph.enable_thermostat), float(ph.grid_semantics)])


def physics_from_vector(v: np.ndarray) -> Physics:
    v = list(v)
    return Physics(box=tuple(v[0:3]), dt=v[3], sigma=tuple(v[4:14]), epsilon_k=tuple(v[14:24]),
                   charge=tuple(v[24:34]), mass=tuple(v[34:44]), thermal_constant=v[44], lj_cutoff_factor=v[45],
                   coulomb_constant=v[46], coulomb_cutoff=v[47], thermostat_temperature=v[48],
                   thermostat_tau=v[49], enable_electrostatics=bool(v[50]), enable_thermostat=bool(v[51]),
                   grid_semantics=int(v[52]))


def write(name, ph, st, nsteps=10, gr=None):
    f = O.forces(ph, st)
    if ph.grid_dims() and min(ph.grid_dims()) >= 3:
        b = O.forces(Physics(**{**ph.__dict__, "grid_semantics": GRID_MINIMAGE}), st)
        for a in ("ax", "ay", "az"):
            assert np.allclose(f[a], b[a], rtol=0, atol=1e-12 * np.abs(f[a]).max())
    fin, obs = O.run(ph, st, nsteps)
    out = dict(physics=physics_vector(ph), x=st.x, y=st.y, z=st.z, vx=st.vx, vy=st.vy, vz=st.vz,
               species=st.species, charge=st.charge, ax=f["ax"], ay=f["ay"], az=f["az"], V=np.array(f["V"]),
               pe=f["pe"], sabs=f["sabs"], nsteps=np.array(nsteps), obs=obs, fin_x=fin.x, fin_y=fin.y, fin_z=fin.z)
    if gr is not None:
        r_max, nb = gr
        c, g, rm, dr = O.radial_distribution(st.x, st.y, st.z, np.asarray(ph.box), r_max, nb)
        lab, cs = O.clusters(st.x, st.y, st.z, st.species, DEFAULT_SIGMA, np.asarray(ph.box))
        out.update(gr_rmax=np.array(r_max), gr_counts=c, gr_g=g, labels=lab,
                   cluster_stats=np.array([cs["monomers"], cs["aggregates"], cs["largest"], cs["mean_size"],
                                           cs["bound_pairs"]]))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, st.n, "atoms, dims", ph.grid_dims(), "V", f["V"])


def main():
    w = plain_lj(seed=11, n_side=4)
    write("lj64_one_cell", w.physics, w.state)
    w = mixture(seed=12, n=512)
    write("mix512_two_cells", w.physics, w.state, gr=(0.5 * w.physics.box[0], 50))
    w = mixture(seed=13, n=1728)
    write("mix1728", w.physics, w.state, gr=(0.25 * w.physics.box[0], 100))


if __name__ == "__main__":
    main()
