"""Writes oracle_mobley.json: the FP64 oracle's own end-to-end results for the
example_milc/mobley_3053621 case (KH, ctip3p.xvk, 256^3 in 64 A), with and
without -z, so GPU tests can compare against the oracle at 256^3 without
re-running it (about 5 minutes per solve on 8 cores).

Run from the repository root:  python tests/golden/make_oracle_vectors.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def main():
    out = {}
    for q0 in (False, True):
        o, th = O.run_inp(os.path.join(HERE, "3drism.dat"), os.path.join(HERE, "mobley_3053621"),
                          q0=q0, solvent_dir=HERE)
        r = o.result
        rec = {k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in th.items()}
        rec["steps"] = int(r.steps)
        rec["status"] = int(r.status)
        rec["rms_last"] = float(r.rms[-1])
        rec["rms_first5"] = [float(x) for x in r.rms[:5]]
        out["xmu0" if q0 else "xmu"] = rec
        print(("xmu0" if q0 else "xmu"), rec["steps"], rec["SFE_SC"], flush=True)
    with open(os.path.join(HERE, "oracle_mobley.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
