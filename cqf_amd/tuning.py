"""Committed autotuner choices for gfx950.

The autotuner (gqf_autotune) times bulk_insert of a 2^nbits-slot table
loaded to a given fraction in one batch, for each candidate region size and
sort digit width, and keeps the fastest; results are cached per nbits and
load rounded to 0.05.  `measure` runs it over a grid on the GPU and `write`
stores the choices as JSON (tuning_gfx950.json beside this file); `load`
puts a stored table into the cache with gqf_autotune_set, so a process that
loads it tunes those sizes without timing anything.

    python -m cqf_amd.tuning --write        # on an MI355X: measure and store
"""
import argparse
import json
import os

from cqf_amd import gqf

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")
NBITS = tuple(range(16, 29))
LOADS = tuple(round(0.05 * k, 2) for k in range(10, 20))  # 0.50 .. 0.95


def measure(nbits=NBITS, loads=LOADS, log=None):
    """{(nbits, load): (region_blocks, sort_digit_bits)} from the autotuner."""
    out = {}
    for q in nbits:
        for ld in loads:
            out[(q, ld)] = gqf.autotune(q, ld)
            if log:
                log(f"nbits={q} load={ld:.2f} -> region_blocks={out[(q, ld)][0]} digit={out[(q, ld)][1]}")
    return out


def to_json(choices, device=""):
    rows = [{"nbits": q, "load": ld, "region_blocks": r, "sort_digit_bits": d}
            for (q, ld), (r, d) in sorted(choices.items())]
    return {"device": device, "note": "gqf_autotune choices: fastest bulk_insert of one batch to the given load",
            "choices": rows}


def write(choices, path=TABLE, device=""):
    with open(path, "w") as fh:
        json.dump(to_json(choices, device), fh, indent=1)
        fh.write("\n")


def read(path=TABLE):
    """{(nbits, load): (region_blocks, sort_digit_bits)} from a stored table."""
    with open(path) as fh:
        doc = json.load(fh)
    return {(int(c["nbits"]), float(c["load"])): (int(c["region_blocks"]), int(c["sort_digit_bits"]))
            for c in doc["choices"]}


def load(path=TABLE):
    """Seeds the autotuner's cache from a stored table; returns the entries."""
    L = gqf.load_library()
    choices = read(path)
    for (q, ld), (r, d) in choices.items():
        rc = L.gqf_autotune_set(q, ld, r, d)
        if rc != 0:
            raise ValueError(f"gqf_autotune_set({q}, {ld}, {r}, {d}) rc={rc}")
    return len(choices)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--write", action="store_true", help="measure the grid and store it")
    ap.add_argument("--path", default=TABLE)
    args = ap.parse_args()
    import torch
    torch.cuda.set_device(0)
    choices = measure(log=lambda s: print(s, flush=True))
    if args.write:
        write(choices, args.path, torch.cuda.get_device_name(0))
        print(f"wrote {len(choices)} choices to {args.path}", flush=True)


if __name__ == "__main__":
    main()
