"""The Fortran binding (include/impact_hip_c.f90) against the C header: the
field ids and the ABI version it declares are the C enum's, and it compiles,
as do the drop-in subroutines on top of it (include/impact_hip_host.f90)."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def c_fields():
    h = (ROOT / "include" / "impact_hip.h").read_text()
    body = re.search(r"typedef enum \{(.*?)\} ih_field;", h, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = [t.split("=")[0].strip() for t in body.replace("\n", " ").split(",") if t.strip()]
    return {n: q for q, n in enumerate(names) if n != "IH_NFIELDS"}


def f_parameters():
    f = (ROOT / "include" / "impact_hip_c.f90").read_text()
    out = {}
    for line in f.splitlines():
        if "parameter ::" in line and not line.lstrip().startswith("!"):
            for name, val in re.findall(r"(IH_\w+)\s*=\s*(-?\d+)", line.split("::", 1)[1]):
                out[name] = int(val)
    return out


def test_field_ids_match_the_header():
    c = c_fields()
    f = f_parameters()
    assert len(c) == 29
    for name, val in c.items():
        assert f.get(name) == val, name


def test_abi_version_matches_the_header():
    h = (ROOT / "include" / "impact_hip.h").read_text()
    v = int(re.search(r"#define IH_ABI_VERSION (\d+)", h).group(1))
    assert f_parameters()["IH_ABI_VERSION"] == v


def test_binding_compiles(tmp_path):
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if Path("/opt/rocm/bin/amdflang").exists() else None)
    if not fc:
        pytest.skip("no Fortran compiler")
    for src in ("impact_hip_c.f90", "impact_hip_host.f90"):
        r = subprocess.run([fc, "-c", f"-J{tmp_path}", str(ROOT / "include" / src), "-o",
                            str(tmp_path / (src[:-4] + ".o"))], capture_output=True, text=True, timeout=300,
                           cwd=tmp_path)
        assert r.returncode == 0, r.stderr
