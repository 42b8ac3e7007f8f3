"""Writes tests/golden/schema_fixture.json from the reference's schema document (reference
schema/lmweather.dsl.json, the file its scripts/gen-schema.py generates from the host's registrations): its sha256 and
size, and the canonical signature strings of the operators it declares ("context wx: Wx" and "method Type.name(params)
-> returns", sorted, the form of the host's --lmcode-signatures listing).  Run in the build container, where the
reference tree is present; the schema document itself is not copied into this repository."""
import hashlib
import json
import os

REF = os.environ.get("LMW_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    raw = open(os.path.join(REF, "schema", "lmweather.dsl.json"), "rb").read()
    doc = json.loads(raw)
    sigs = [f"context {k}: {v}" for k, v in doc["context"].items()]
    for t, body in doc["types"].items():
        for name, overloads in body["methods"].items():
            for o in overloads:
                sigs.append(f"method {t}.{name}({', '.join(o['params'])}) -> {o['returns']}")
    for name, overloads in doc["functions"].items():
        for o in overloads:
            sigs.append(f"fn {name}({', '.join(o['params'])}) -> {o['returns']}")
    out = {"source": "reference schema/lmweather.dsl.json", "sha256": hashlib.sha256(raw).hexdigest(), "bytes": len(raw),
           "methods": {t: len(b["methods"]) for t, b in doc["types"].items()}, "signatures": sorted(sigs)}
    with open(os.path.join(HERE, "schema_fixture.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(out["sha256"], out["bytes"], out["methods"], len(out["signatures"]))


if __name__ == "__main__":
    main()
