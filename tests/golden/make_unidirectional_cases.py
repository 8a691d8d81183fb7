"""Writes unidirectional_cases.json: the reference's own regression inputs of
the unidirectional_heat_transfer case (every tests/*.json of the case, as
data) with the norms its .mpirun=4.output prints after each Newton solve, and
the defaults that applications/mp-heat-transfer/parameters.schema.json lists
for the parameter keys the JSON front end models. Run where the reference is
available at REF; tests only read the written fixture.
"""
import glob
import json
import os
import re

REF = os.environ.get("MPDG_REFERENCE", "/root/reference")
APP = "applications/mp-heat-transfer"
CASE = f"{APP}/cases/unidirectional_heat_transfer/tests"
HERE = os.path.dirname(os.path.abspath(__file__))
SECTIONS = ("base", "time stepping", "heat", "material", "application specific")


def schema_defaults(node, path=()):
    out = {}
    for key, sub in node.get("properties", {}).items():
        p = path + (key,)
        if "properties" in sub:
            out.update(schema_defaults(sub, p))
        elif "default" in sub:
            out["/".join(p)] = sub["default"]
    return out


def main():
    schema = json.load(open(os.path.join(REF, APP, "parameters.schema.json")))
    defaults = {k: v for k, v in schema_defaults(schema).items() if k.split("/")[0] in SECTIONS}
    cases = {}
    for path in sorted(glob.glob(os.path.join(REF, CASE, "*.json"))):
        name = os.path.basename(path)[: -len(".json")]
        prm = json.load(open(path))
        prm.pop("$schema", None)
        out = os.path.join(REF, CASE, name + ".mpirun=4.output")
        norms = [float(v) for v in re.findall(r"\|\|solution\|\| = ([0-9.eE+-]+)", open(out).read())]
        cases[name] = {"input": prm, "l2_norms": norms}
    fixture = {"source": f"{CASE}/*.json, *.mpirun=4.output and {APP}/parameters.schema.json",
               "schema_defaults": defaults, "cases": cases}
    path = os.path.join(HERE, "unidirectional_cases.json")
    with open(path, "w") as f:
        json.dump(fixture, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", path, len(cases), "cases,", len(defaults), "defaults")


if __name__ == "__main__":
    main()
