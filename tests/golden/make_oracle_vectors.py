"""Writes the FP64 oracle's own end-to-end results for the reference's
bundled example inputs, so GPU tests can compare against the oracle at full
size without re-running it:

* oracle_mobley.json -- example_milc/mobley_3053621 (KH, ctip3p.xvk, 256^3 in
  64 A), with and without -z (about 5 minutes per solve on 8 cores);
* oracle_pegasus_test1.json -- example-pegasus/test1.inp (a water molecule in
  h2o_vv.xvk, KH, 128^3 in 64 A; about half a minute on 8 cores), with the
  per-site integrals.

Run from the repository root:
  python tests/golden/make_oracle_vectors.py [mobley] [test1]
(no argument: both).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def record(o, th):
    r = o.result
    rec = {k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in th.items()}
    rec["steps"] = int(r.steps)
    rec["status"] = int(r.status)
    rec["rms_last"] = float(r.rms[-1])
    rec["rms_first5"] = [float(x) for x in r.rms[:5]]
    return rec


def mobley():
    out = {}
    for q0 in (False, True):
        o, th = O.run_inp(os.path.join(HERE, "3drism.dat"), os.path.join(HERE, "mobley_3053621"),
                          q0=q0, solvent_dir=HERE)
        rec = record(o, th)
        out["xmu0" if q0 else "xmu"] = rec
        print(("xmu0" if q0 else "xmu"), rec["steps"], rec["SFE_SC"], flush=True)
    with open(os.path.join(HERE, "oracle_mobley.json"), "w") as f:
        json.dump(out, f, indent=1)


def test1():
    o, th = O.run_inp(os.path.join(HERE, "pegasus_test1.inp"), solvent_dir=HERE)
    rec = record(o, th)
    print("test1", rec["steps"], rec["SFE_SC"], rec["SFE_GF"], rec["PMV"], flush=True)
    with open(os.path.join(HERE, "oracle_pegasus_test1.json"), "w") as f:
        json.dump({"xmu": rec}, f, indent=1)


if __name__ == "__main__":
    which = sys.argv[1:] or ["mobley", "test1"]
    if "mobley" in which:
        mobley()
    if "test1" in which:
        test1()
