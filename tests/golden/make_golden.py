"""Writes the committed golden fixtures of the hot path from the FP64 oracle.

The reference ships no numeric vectors for this path (its only statement of
tolerance is chargeagent.cpp:368-370), so the fixtures are the oracle's own
outputs on small seeded inputs, kept as data: inputs and expected outputs.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import pyoracle as O  # noqa: E402
from _util import oracle_state  # noqa: E402

TRAJ = {
    "transistor_48x48": dict(grid_x=48, grid_y=48, grid_z=1, coulomb_carriers=1, electron_percentage=0.0,
                             hole_percentage=0.02, seed_charges=1.0, voltage_left=1.0,
                             e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9,
                             random_seed=101),
    "solar_12x12x6": dict(simulation_type=1, grid_x=12, grid_y=12, grid_z=6, coulomb_carriers=1,
                          electron_percentage=0.02, hole_percentage=0.02, seed_charges=0.5,
                          generation_rate=0.5, recombination_rate=0.3, recombination_range=1,
                          exciton_binding=0.1, electrostatic_cutoff=20, random_seed=102),
}
STEPS = (0, 1, 5, 20, 60)


def main():
    out = {}
    for name, kw in TRAJ.items():
        w = O.World(O.OrcParams.make(**kw))
        st = oracle_state(w)
        out[f"{name}/init_e"] = st["electrons"]
        out[f"{name}/init_h"] = st["holes"]
        done = 0
        for s in STEPS:
            w.step(s - done)
            done = s
            e, h = w.carriers(0), w.carriers(1)
            win, k = w.rng()
            a, su, _ = w.flux()
            out[f"{name}/s{s}/e_site"] = e["site"]
            out[f"{name}/s{s}/h_site"] = h["site"]
            out[f"{name}/s{s}/e_path"] = e["pathlength"]
            out[f"{name}/s{s}/h_path"] = h["pathlength"]
            out[f"{name}/s{s}/rng_consumed"] = np.array([k], np.uint64)
            out[f"{name}/s{s}/rng_head"] = win[:8].copy()
            out[f"{name}/s{s}/flux"] = np.stack([a, su])
        # dE of the next proposals
        w.propose()
        w.coulomb()
        out[f"{name}/next_fsite"] = np.concatenate([w.carriers(0)["fsite"], w.carriers(1)["fsite"]])
        out[f"{name}/next_de"] = np.concatenate([w.carriers(0)["de"], w.carriers(1)["de"]])
    # pair potentials and a map on a small lattice
    rng = np.random.default_rng(103)
    dims = (20, 16, 8)
    V = 20 * 16 * 8
    cs = rng.choice(V, 150, replace=False).astype(np.int32)
    cq = rng.choice([-1, 1], 150).astype(np.int32)
    ts = rng.integers(0, V, 64).astype(np.int32)
    out["pairs/dims"] = np.array(dims, np.int32)
    out["pairs/charge_site"] = cs
    out["pairs/charge_q"] = cq
    out["pairs/target_site"] = ts
    out["pairs/coulomb_c9"] = O.pair_potential(dims, 9, cs, cq, ts)
    out["pairs/image_c9"] = O.pair_potential(dims, 9, cs, cq, ts, flags=O.IMAGE)
    out["pairs/gauss_c9_s1p5"] = O.pair_potential(dims, 9, cs, cq, ts, flags=O.GAUSS, sigma=1.5)
    out["pairs/map_c9"] = O.potential_map(dims, 9, cs, cq, threads=4)
    out["pairs/map_nocut"] = O.potential_map(dims, 0, cs, cq, threads=4)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
