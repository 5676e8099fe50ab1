"""Write results_reference.json: the reference's published result records, trimmed.

Reads results/quick.json (a benchmark record) and results/validation.json (a validation
record) of the reference checkout and keeps every key, but cuts long per-point arrays
(centreline profiles, Poiseuille profiles, force histories) to their first three entries.
The fixture pins the result schema that pylbm_amd.report reads and that the package's
benchmark and validation runners write.  Run from the repository root:

    python tests/golden/make_results_fixture.py [REFERENCE_ROOT]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

KEEP = 3


def trim(x):
    if isinstance(x, dict):
        return {k: trim(v) for k, v in x.items()}
    if isinstance(x, list):
        if len(x) > KEEP and all(isinstance(v, (int, float)) for v in x):
            return x[:KEEP]
        return [trim(v) for v in x]
    return x


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    bench = json.loads((ref / "results" / "quick.json").read_text())
    val = json.loads((ref / "results" / "validation.json").read_text())
    out = Path(__file__).with_name("results_reference.json")
    out.write_text(json.dumps({"bench": trim(bench), "validation": trim(val)}, indent=0))
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
