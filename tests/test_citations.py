"""Every reference citation (file:line or file:a-b) in the C ABI header, the HIP
sources, the package and the oracle points inside an existing reference file.
Runs where the reference checkout is present (skipped elsewhere)."""

import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
SCOPES = [ROOT / "include", ROOT / "oracle", ROOT / "warp-preconditioners_amd" / "csrc",
          ROOT / "warp-preconditioners_amd" / "warp_preconditioners", ROOT / "bench.py",
          ROOT / "__graft_entry__.py", ROOT / "tests"]
SUFFIXES = {".h", ".hip", ".cpp", ".py", ".md"}
CITE = re.compile(r"(?<![\w/.])((?:[\w-]+/)*[\w-]+\.(?:py|md|json|txt|toml)):(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")


def _resolve(name):
    # paths relative to the reference root, its package, benchmarks/ or tests/
    for base in (REF, REF / "warp_preconditioners", REF / "benchmarks", REF / "tests"):
        p = base / name
        if p.is_file():
            return p
    return None


def citations():
    out = []
    for scope in SCOPES:
        files = [scope] if scope.is_file() else sorted(p for p in scope.rglob("*")
                                                         if p.suffix in SUFFIXES)
        for f in files:
            if "golden" in f.parts or f.name == "test_citations.py":
                continue
            for lineno, line in enumerate(f.read_text(errors="replace").splitlines(), 1):
                for m in CITE.finditer(line):
                    out.append((f.relative_to(ROOT), lineno, m.group(1), m.group(2)))
    return out


@pytest.mark.skipif(not REF.is_dir(), reason="reference checkout not present")
def test_reference_citations_point_inside_files():
    found = citations()
    assert len(found) > 100
    bad = []
    for where, lineno, name, spans in found:
        path = _resolve(name)
        if path is None:
            if (ROOT / name).exists() or any((ROOT / d / name).exists() for d in
                                             ("warp-preconditioners_amd", "tests", "oracle")):
                continue  # a file of this repository
            bad.append(f"{where}:{lineno}: {name} not in the reference")
            continue
        count = len(path.read_text(errors="replace").splitlines())
        for span in spans.replace(" ", "").split(","):
            lo, _, hi = span.partition("-")
            lo, hi = int(lo), int(hi or lo)
            if not (1 <= lo <= hi <= count):
                bad.append(f"{where}:{lineno}: {name}:{span} outside 1-{count}")
    assert not bad, "\n".join(bad)
