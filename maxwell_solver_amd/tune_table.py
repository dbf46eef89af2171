"""Writes the size-class tuning table shipped beside the library (tuning_gfx950.json): the autotuner
(fdtd_autotune: every columns-per-lane and z-chunk choice and two tile heights, then the slab part's
chunk depth, timed in a run's launch form, the automatic choice kept unless another is faster by
more than 1%) run on one
simulation of each size class of the workloads -- the 256^3 sphere grid, the 384x384x256 layered
ground, the GIC grid, the 96 grids of the batched workload and a sample of the served requests.
Run on the GPU the table is for:  python -m maxwell_solver_amd.tune_table [out.json]
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def main(out: str) -> None:
    from maxwell_solver_amd import Simulation
    from maxwell_solver_amd import scenarios as S
    tmp = os.path.join(tempfile.mkdtemp(prefix="tune_"), "table.json")
    os.environ["FDTD_TUNE_CACHE"] = tmp  # the autotuner writes every class it tunes here
    cases = [S.config2(T=64), S.config3(T=64), S.config_gic(T=64)]
    cases += [sc for sc, _ in S.config4()]
    cases += [sc for sc, _, _ in S.config5(200)[::10]]
    seen = set()
    for sc in cases:
        key = (sc.I, sc.J.bit_length() - 1, sc.K.bit_length() - 1, sc.P, sc.mat_index is not None, sc.mode != 0)
        if key in seen:
            continue
        seen.add(key)
        g = Simulation()
        S.setup(g, sc, init_line=bool(sc.mode))
        g.run(0, 2, sc.mode)
        ms = g.autotune(2, sc.mode, reps=10)
        print(f"{sc.name}: {sc.I}x{sc.J}x{sc.K} P{sc.P} -> {g.get_tuning()} {ms * 1e3:.1f} us/step", flush=True)
        g.close()
    with open(tmp) as f:
        table = json.load(f)
    table["note"] = ("autotuned on one MI355X (gfx950) by maxwell_solver_amd/tune_table.py; rows / z_chunk / "
                     "cols_per_lane / slab_chunk 0 mean the automatic choice")
    with open(out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"{len(table['classes'])} classes -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "tuning_gfx950.json"))
