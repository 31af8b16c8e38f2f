"""Builds the component-test library (tests/csrc/amx_components.hip) for a
network into tests/_build/<name>/libamxt_<name>.so."""
from __future__ import annotations

import subprocess
from pathlib import Path

from amici_amd import codegen

HERE = Path(__file__).resolve().parent


def build(net, team=None, force=False) -> Path:
    d = HERE / "_build" / net.name
    d.mkdir(parents=True, exist_ok=True)
    traits = d / "traits.h"
    src = codegen.emit_model_traits(net, team)
    so = d / f"libamxt_{net.name}.so"
    stamp = d / "stamp"
    digest = codegen._source_digest(src + (HERE / "csrc" / "amx_components.hip").read_text())
    if so.exists() and stamp.exists() and stamp.read_text() == digest and not force:
        return so
    traits.write_text(src)
    cmd = [codegen.HIPCC, *codegen.HIP_FLAGS, "-I", str(codegen.CSRC),
           f'-DAMX_MODEL_TRAITS="{traits}"', str(HERE / "csrc" / "amx_components.hip"),
           "-o", str(so)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr[-6000:])
    stamp.write_text(digest)
    return so
