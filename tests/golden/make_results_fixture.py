"""Write results_reference.json: the reference's published result records, trimmed.

Reads results/quick.json (a benchmark record), results/validation.json (a validation
record), results/rtx3070-laptop.json (the benchmark record the reference's paper quotes)
and results/counterparts.json (its head-to-head record) of the reference checkout and
keeps every key, but cuts long per-point arrays (centreline profiles, Poiseuille profiles,
force histories) to their first three entries and drops the free-text notes and the GPU
state snapshots.  It also stores, as the expected output, the macros of the reference's
generated paper/results.tex (each ``\newcommand`` name with its body, and the two
provisional flags), which the reference wrote from those same three records.  The fixture
pins the result schema that pylbm_amd.report and pylbm_amd.paper read and that the
package's benchmark and validation runners write.  Run from the repository root:

    python tests/golden/make_results_fixture.py [REFERENCE_ROOT]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

KEEP = 3
DROP = {"notes", "gpu_state_start", "gpu_state_end", "license_note"}


def trim(x, drop=frozenset()):
    if isinstance(x, dict):
        return {k: trim(v, drop) for k, v in x.items() if k not in drop}
    if isinstance(x, list):
        if len(x) > KEEP and all(isinstance(v, (int, float)) for v in x):
            return x[:KEEP]
        return [trim(v, drop) for v in x]
    return x


def tex_macros(text: str) -> dict:
    """``\\newcommand{\\name}{body}`` entries (bodies may span lines) and the flags."""
    macros, name, body = {}, None, []
    for line in text.splitlines():
        if line.startswith("\\newcommand{\\"):
            if name is not None:
                macros[name] = "\n".join(body)[:-1]
            head, _, rest = line[len("\\newcommand{\\"):].partition("}{")
            name, body = head, [rest]
        elif name is not None:
            body.append(line)
    if name is not None:
        macros[name] = "\n".join(body)[:-1]
    flags = {f: f"\\{f}true" in text for f in ("provisional", "cpprovisional")}
    return {"macros": macros, **flags}


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    bench = json.loads((ref / "results" / "quick.json").read_text())
    val = json.loads((ref / "results" / "validation.json").read_text())
    out = Path(__file__).with_name("results_reference.json")
    full = json.loads((ref / "results" / "rtx3070-laptop.json").read_text())
    cp = json.loads((ref / "results" / "counterparts.json").read_text())
    tex = tex_macros((ref / "paper" / "results.tex").read_text())
    tex["source"] = "rtx3070-laptop.json"
    out.write_text(json.dumps({"bench": trim(bench), "validation": trim(val),
                               "bench_paper": trim(full, DROP),
                               "counterparts": trim(cp, DROP), "results_tex": tex},
                              indent=0))
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
