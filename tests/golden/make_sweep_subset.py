"""Write tests/golden/sweep_3f2.json: a data subset of the reference's sweep
inputs (sweep_data/3F2/3F2_part00.jsonl, sweep_data/trajectories/
dim5_trajectories.json, sweep_data/shifts/dim5_shifts.json).

The files are data (pFq spec records, direction vectors, shift records); the
subset keeps every 16th spec (with its n-form companion "matrix" strings,
which pin the per-axis companion builder), 24 trajectories spread over the
list and every 8th shift; and, for the sweep benchmark, the first 64
trajectories with all 512 shifts. Run here, where the reference tree is mounted:

    python tests/golden/make_sweep_subset.py
"""

import json
import os

REF = "/root/reference/sweep_data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sweep_3f2.json")


def main() -> None:
    with open(os.path.join(REF, "3F2", "3F2_part00.jsonl")) as f:
        specs = [json.loads(ln) for ln in f if ln.strip()]
    with open(os.path.join(REF, "trajectories", "dim5_trajectories.json")) as f:
        trajs = json.load(f)
    with open(os.path.join(REF, "shifts", "dim5_shifts.json")) as f:
        shifts = json.load(f)
    tl = trajs["trajectories"]
    tsel = sorted(set(list(range(5)) + [i * (len(tl) - 1) // 18 for i in range(19)]))
    out = {
        "source": "sweep_data/3F2/3F2_part00.jsonl, sweep_data/trajectories/dim5_trajectories.json, "
                  "sweep_data/shifts/dim5_shifts.json",
        "n_specs_total": len(specs),
        "n_traj_total": trajs["count"],
        "n_shifts_total": shifts["count"],
        "specs": specs[::16],
        "traj_index": tsel,
        "trajectories": [tl[i] for i in tsel],
        "shift_index": list(range(0, len(shifts["shifts"]), 8)),
        "shifts": shifts["shifts"][::8],
        # the sweep a driver runs with --max-traj 64: the first 64
        # trajectories and every shift
        "bench_trajectories": tl[:64],
        "bench_shifts": shifts["shifts"],
    }
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
