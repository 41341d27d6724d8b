"""Write the golden proof fixtures under tests/golden/ from the Python oracle.

Inputs are regenerated from seeds (zorch_amd.workloads generators), so a
fixture holds only the instance recipe and the expected outputs: messages,
raw challenge words, reduced value, final folded values and the final
transcript stream state (SHA-256 or BLAKE3). Run:
``python tests/golden/gen_fixtures.py [name ...]``.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from tests.golden.recipes import RECIPES, build, state_json  # noqa: E402
from tests.oracle_glue import run_py  # noqa: E402


def main() -> None:
    only = set(sys.argv[1:])
    for name, recipe in RECIPES.items():
        if only and name not in only:
            continue
        inst = build(recipe)
        res, _ = run_py(inst)
        out = {
            "recipe": recipe,
            "msgs": res.msgs.tolist(),
            "point": res.point.tolist(),
            "value": res.value.tolist(),
            "final": res.final.tolist(),
            "transcript": state_json(res.state, inst.transcript_kind),
        }
        (HERE / f"{name}.json").write_text(json.dumps(out, separators=(",", ":")) + "\n")
        print("wrote", name)


if __name__ == "__main__":
    main()
