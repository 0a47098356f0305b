"""Extract the H..Ne entries of the basis sets this package ships.

The basis sets are the Basis Set Exchange JSON files that OCC reads
(share/basis/<name>.json, loaded by AOBasis::load in src/gto/shell.cpp:499-571).
Only elements 1-10 are kept, in the same BSE schema, so the package carries
the data its generators need (water clusters, alkanes) without the full
library.  Run once in a checkout that has the OCC sources at
/root/reference; it writes the basis data committed under occ_amd/data/basis/
(package data, not test fixtures: the script sits with the other recipes that
read the reference).
"""
import json
import pathlib
import sys

NAMES = ["6-31g", "def2-svp", "3-21g", "sto-3g", "def2-tzvp", "6-31g(d,p)"]


def main(src="/root/reference/share/basis"):
    out = pathlib.Path(__file__).resolve().parents[2] / "occ_amd" / "data" / "basis"
    out.mkdir(exist_ok=True)
    for name in NAMES:
        d = json.load(open(f"{src}/{name}.json"))
        els = {k: {"electron_shells": v["electron_shells"]}
               for k, v in d["elements"].items() if int(k) <= 10}
        doc = {"name": d.get("name", name), "elements": els}
        (out / f"{name}.json").write_text(json.dumps(doc, indent=0))
        print(name, len(els))


if __name__ == "__main__":
    main(*sys.argv[1:])
