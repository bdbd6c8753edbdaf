"""Transcribe the reference's golden vectors into JSON fixtures.

Reads the generated header files tests/cpp/golden_metric.hpp,
tests/cpp/golden_metric_3d.hpp, tests/cpp/golden_patch.hpp and
tests/cpp/golden_geometry.hpp of the
reference checkout (numbers only: inputs and expected outputs of the
reference's own C++ oracle bindings, produced there by
tests/cpp/gen_golden.py) and writes them as data under tests/golden/.

Run once in the build container:  python tests/golden/extract_golden.py
"""

from __future__ import annotations

import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "tests" / "cpp"
OUT = Path(__file__).resolve().parent

_FIELD = re.compile(r"\.(\w+)\s*=\s*(\{[^}]*\}|[-+0-9.eE]+)")


def _parse_value(txt: str):
    txt = txt.strip()
    if txt.startswith("{"):
        body = txt[1:-1].strip()
        return [float(x) for x in body.split(",") if x.strip()]
    return float(txt)


def _records(header: Path, array_name: str) -> list[dict]:
    text = header.read_text()
    start = text.index(array_name)
    body = text[start:]
    recs = []
    # each record is a "{ .field = ..., ... }" block at depth 2
    for block in re.finditer(r"\{\s*\n\s*\.(?:.|\n)*?\n\s*\}", body):
        rec = {m.group(1): _parse_value(m.group(2)) for m in _FIELD.finditer(block.group(0))}
        if rec:
            recs.append(rec)
    return recs


def main() -> None:
    metric = _records(REF / "golden_metric.hpp", "kSamples")
    metric3d = _records(REF / "golden_metric_3d.hpp", "kSamples3D")
    patch = _records(REF / "golden_patch.hpp", "kPatchSamples")
    geometry = _records(REF / "golden_geometry.hpp", "kGeoSamples")
    assert len(metric) == 8 and len(metric3d) == 8 and len(patch) == 5 and len(geometry) == 9, (
        len(metric),
        len(metric3d),
        len(patch),
        len(geometry),
    )
    for rec in geometry:
        rec["tag"] = int(rec["tag"])
    for rec in patch:
        for k in ("n_nodes", "P", "tag"):
            rec[k] = int(rec[k])
        for k in ("gc", "gn0", "gn1", "role"):
            rec[k] = [int(v) for v in rec[k]]
    prov = {
        "source": "bezmi/egg tests/cpp/golden_{metric,metric_3d,patch,geometry}.hpp",
        "generator": "reference tests/cpp/gen_golden.py (C++ oracle bindings)",
        "note": "numbers transcribed verbatim; patch RNG seed 20240601 upstream",
    }
    (OUT / "golden_metric.json").write_text(json.dumps({"provenance": prov, "samples": metric}, indent=1))
    (OUT / "golden_metric_3d.json").write_text(
        json.dumps({"provenance": prov, "samples": metric3d}, indent=1)
    )
    (OUT / "golden_patch.json").write_text(json.dumps({"provenance": prov, "samples": patch}, indent=1))
    (OUT / "golden_geometry.json").write_text(json.dumps({"provenance": prov, "samples": geometry}, indent=1))
    print("wrote", len(metric), len(metric3d), len(patch), len(geometry))


if __name__ == "__main__":
    main()
