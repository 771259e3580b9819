"""Every `file:line` citation of the reference in the boundary header, the oracle and the library
sources names a file of the reference and lines that exist in it. A bare file name must name one
file of the reference tree; a path is taken from the reference's root. Skipped where the reference
tree is absent (it is not shipped with this repository)."""
import os
import re
from pathlib import Path

import pytest

REF = Path("/root/reference")
ROOT = Path(__file__).resolve().parents[1]
CITE = re.compile(r"((?:[A-Za-z0-9_]+/)*[A-Za-z0-9_]+\.(?:cpp|h|hpp|txt|json)):(\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)")
# files of this repository that carry citations of the reference
FILES = (["include/sbs_gpu.h", "soft-body-simulator_amd/host/sbs_amd/physics.h"] + sorted(str(p.relative_to(ROOT)) for p in (ROOT / "oracle").glob("*.[ch]*"))
         + sorted(str(p.relative_to(ROOT)) for p in (ROOT / "oracle").glob("*.py"))
         + sorted(str(p.relative_to(ROOT)) for p in (ROOT / "soft-body-simulator_amd" / "csrc").glob("*.*"))
         + sorted(str(p.relative_to(ROOT)) for p in (ROOT / "soft-body-simulator_amd" / "sbs_amd").glob("*.py")))

needs_ref = pytest.mark.skipif(not (REF / "src").is_dir(), reason="the reference tree is not present")


def _index():
    by_name = {}
    for dirpath, dirnames, files in os.walk(REF):
        dirnames[:] = [d for d in dirnames if not d.startswith(".") and d != "build"]
        for f in files:
            by_name.setdefault(f, []).append(Path(dirpath) / f)
    return by_name


def _resolve(name, by_name):
    if "/" in name:
        for root in (REF, REF / "include" / "sbs", REF / "src"):
            if (root / name).is_file():
                return [root / name]
        base = name.rsplit("/", 1)[1]
        return [p for p in by_name.get(base, []) if str(p).endswith("/" + name)]
    return by_name.get(name, [])


@needs_ref
@pytest.mark.parametrize("rel", FILES)
def test_citations_name_existing_reference_lines(rel):
    text = (ROOT / rel).read_text(errors="replace")
    by_name = _index()
    bad = []
    for m in CITE.finditer(text):
        name, spans = m.group(1), m.group(2)
        if (ROOT / name).exists() or name.startswith(("tests/", "oracle/", "include/sbs_gpu")):
            continue  # a file of this repository
        cands = _resolve(name, by_name)
        if not cands:
            if any(name.endswith(s) for s in (".json",)) or name in ("sbs_gpu.h", "kernels.h"):
                continue
            bad.append(f"{name}: no such file in the reference")
            continue
        if len(cands) > 1:
            bad.append(f"{name}: names {len(cands)} files of the reference")
            continue
        nlines = len(cands[0].read_text(errors="replace").splitlines())
        for span in spans.split(","):
            a, _, b = span.partition("-")
            hi = int(b or a)
            if int(a) < 1 or hi > nlines or (b and int(b) < int(a)):
                bad.append(f"{name}:{span} outside its {nlines} lines")
    assert not bad, f"{rel}: " + "; ".join(bad)
