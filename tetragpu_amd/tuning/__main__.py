"""python -m tetragpu_amd.tuning [--sides 6 10 ...] [--reps 5] [--out FILE]

Measure the tuning table on the GPU of this machine and save it (default:
tetragpu_amd/tuning/mi355x.json, which bench.py and the batch path load)."""
import argparse
import json

from . import DEFAULT, autotune, save


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--sides", type=int, nargs="+", default=[6, 10, 16, 24, 40, 64, 96, 128])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=str(DEFAULT))
    args = ap.parse_args()
    import tetragpu_amd as t

    t.init(0)
    table = autotune(args.sides, args.reps)
    save(table, args.out)
    print(json.dumps(table))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
