"""Write the GAMS frontend's fixtures from the reference checkout.

    python tests/golden/make_gams_golden.py [REFERENCE_ROOT]

* tests/golden/gams/*.gms: the reference's MINLPLib test models
  (test/data/gams/*.gms), copied as data so the solves can run where the
  reference is absent.
* tests/golden/gams_corpus.json: for each of those models, the row of the
  reference's corpus table (MINLP_STATUS.md, "Instance | Parses | Sense | Vars |
  Cons | Nodes | Ref primal | Ref dual | ..."): the variable, constraint and DAG
  node counts its frontend produced, and MINLPLib's published bounds.
"""
from __future__ import annotations

import json
import pathlib
import re
import shutil
import sys

HERE = pathlib.Path(__file__).resolve().parent


def number(cell: str):
    cell = cell.strip()
    if cell in ("", "—", "-"):
        return None
    return float(cell)


def main(argv: list[str]) -> int:
    ref = pathlib.Path(argv[1] if len(argv) > 1 else "/root/reference")
    src = ref / "test" / "data" / "gams"
    out_dir = HERE / "gams"
    out_dir.mkdir(exist_ok=True)
    names = []
    for f in sorted(src.glob("*.gms")):
        shutil.copyfile(f, out_dir / f.name)
        names.append(f.stem)

    rows = {}
    pat = re.compile(r"^\| (\S+) \| (yes|no) \| (min|max) \| (\d+) \| (\d+) \| (\d+) \| ([^|]*)\| ([^|]*)\|")
    for line in (ref / "MINLP_STATUS.md").read_text().splitlines():
        m = pat.match(line)
        if m and m.group(1) in names:
            rows[m.group(1)] = {
                "sense": m.group(3),
                "vars": int(m.group(4)),
                "constraints": int(m.group(5)),
                "nodes": int(m.group(6)),
                "ref_primal": number(m.group(7)),
                "ref_dual": number(m.group(8)),
            }
    doc = {
        "source": "reference MINLP_STATUS.md corpus table (frontend counts) and test/data/gams",
        "models": names,
        "table": rows,
    }
    (HERE / "gams_corpus.json").write_text(json.dumps(doc, indent=1, sort_keys=True) + "\n")
    print(f"{len(names)} models, {len(rows)} table rows")
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv))
