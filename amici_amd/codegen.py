"""Model code generator: reaction network -> gfx950 HIP model module.

Plays the role of AMICI's ``DEExporter`` (python/sdist/amici/exporters/sundials/
de_export.py:90, _write_function_file :374-508), which writes one C++ function
per symbolic quantity (xdot, dwdx, dxdotdw, dxdotdp_explicit, y, Jy, ...) with
CSC sparsity tables (cxx_functions.py:101-...).  Here the quantities are
derived with sympy from the mass-action network (:func:`symbolic_model`) and
emitted as compile-time sparsity tables plus the team size, consumed by the
device functions of ``csrc/amx_massaction.h`` (flux / RHS gather, Jacobian
nonzeros, J^T v through the flux structure, df/dp columns, observables).
Every table entry is checked against the sympy derivatives by the tests.

``compile_model`` writes ``<build>/<name>/amx_model.hip`` and compiles it with
hipcc for gfx950 into ``libamx_<name>.so``, which exports the C-ABI of
``include/amx.h``.
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np

from .networks import ReactionNetwork

CSRC = Path(__file__).resolve().parent / "csrc"
BUILD_ROOT = Path(__file__).resolve().parent / "_models"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]


def default_team(nx: int) -> int:
    """One instance per 64-lane wavefront (team = wavefront) for nx <= 64; one
    instance per workgroup of 2 or 4 wavefronts (team = workgroup) up to 256
    states."""
    if nx <= 64:
        return 64
    if nx <= 128:
        return 128
    if nx <= 256:
        return 256
    raise ValueError("models with more than 256 states are not supported (one lane per state)")


def _csr(rows: list, n: int):
    ptr = [0]
    flat = []
    for i in range(n):
        flat.extend(rows[i])
        ptr.append(len(flat))
    return ptr, flat


def model_tables(net: ReactionNetwork) -> dict:
    """All sparsity tables of the emitted model (pure numpy / python)."""
    net.validate()
    nx, nr, npar, ny = net.nx, net.nr, net.np, net.ny
    S = net.stoichiometry()
    ra, rb = [], []
    for rx in net.reactions:
        re = sorted(rx.reactants)
        ra.append(re[0] if len(re) >= 1 else -1)
        rb.append(re[1] if len(re) == 2 else -1)
    # rows of S (species -> reactions) and columns (reaction -> species)
    sp_rows = [[(r, S[i, r]) for r in range(nr) if S[i, r] != 0.0] for i in range(nx)]
    rs_cols = [[(i, S[i, r]) for i in range(nx) if S[i, r] != 0.0] for r in range(nr)]
    # dw_r/dx_s = mult * k_r * x_other
    dwdx = []  # per reaction: list of (s, other, mult)
    for r in range(nr):
        a, b = ra[r], rb[r]
        if a < 0:
            dwdx.append([])
        elif b < 0:
            dwdx.append([(a, -1, 1.0)])
        elif a == b:
            dwdx.append([(a, a, 2.0)])
        else:
            dwdx.append([(a, b, 1.0), (b, a, 1.0)])
    jterms: dict = {}
    for r in range(nr):
        for (s, o, mult) in dwdx[r]:
            for i in range(nx):
                if S[i, r] != 0.0:
                    jterms.setdefault((i, s), []).append((r, o, S[i, r] * mult))
    jkeys = sorted(jterms)
    dx_rows = [[] for _ in range(nx)]
    for r in range(nr):
        for (s, o, mult) in dwdx[r]:
            dx_rows[s].append((r, o, mult))
    pr_rows = [[r for r in range(nr) if net.reactions[r].rate == j] for j in range(npar)]
    C = net.observable_matrix
    ob_rows = [[(i, C[k, i]) for i in range(nx) if C[k, i] != 0.0] for k in range(ny)]
    ot_rows = [[(k, C[k, i]) for k in range(ny) if C[k, i] != 0.0] for i in range(nx)]

    t = {}
    t["RA"], t["RB"] = ra, rb
    t["KI"] = [rx.rate for rx in net.reactions]
    p, f = _csr([[e[0] for e in row] for row in sp_rows], nx)
    t["SP_PTR"], t["SP_R"] = p, f
    t["SP_C"] = [e[1] for row in sp_rows for e in row]
    p, f = _csr([[e[0] for e in row] for row in rs_cols], nr)
    t["RS_PTR"], t["RS_S"] = p, f
    t["RS_C"] = [e[1] for row in rs_cols for e in row]
    t["J_I"] = [k[0] for k in jkeys]
    t["J_J"] = [k[1] for k in jkeys]
    p, f = _csr([[e[0] for e in jterms[k]] for k in jkeys], len(jkeys))
    t["JT_PTR"], t["JT_R"] = p, f
    t["JT_O"] = [e[1] for k in jkeys for e in jterms[k]]
    t["JT_C"] = [e[2] for k in jkeys for e in jterms[k]]
    p, f = _csr([[e[0] for e in row] for row in dx_rows], nx)
    t["DX_PTR"], t["DX_R"] = p, f
    t["DX_O"] = [e[1] for row in dx_rows for e in row]
    t["DX_C"] = [e[2] for row in dx_rows for e in row]
    p, f = _csr(pr_rows, npar)
    t["PR_PTR"], t["PR_R"] = p, f
    p, f = _csr([[e[0] for e in row] for row in ob_rows], ny)
    t["OB_PTR"], t["OB_S"] = p, f
    t["OB_C"] = [e[1] for row in ob_rows for e in row]
    p, f = _csr([[e[0] for e in row] for row in ot_rows], nx)
    t["OT_PTR"], t["OT_Y"] = p, f
    t["OT_C"] = [e[1] for row in ot_rows for e in row]
    t["X0C"] = [float(v) for v in net.x0_const]
    t["X0K"] = [int(v) for v in net.x0_fixed]
    t["JNZ"] = len(jkeys)
    return t


_INT_TABLES = ["RA", "RB", "KI", "SP_PTR", "SP_R", "RS_PTR", "RS_S", "J_I", "J_J", "JT_PTR",
               "JT_R", "JT_O", "DX_PTR", "DX_R", "DX_O", "PR_PTR", "PR_R", "OB_PTR", "OB_S",
               "OT_PTR", "OT_Y", "X0K"]
_DBL_TABLES = ["SP_C", "RS_C", "JT_C", "DX_C", "OB_C", "OT_C", "X0C"]


def _lane_major(rows, team: int, pad):
    """Pads the entry lists ``rows`` (row l is handled by team lane l % team in
    pass l // team) to the longest row and lays them out lane-major: entry q
    of lane i in pass e at [(e * n + q) * team + i], so that one fixed-length,
    fully unrolled loop over q reads each entry of all lanes with one
    coalesced load.  Returns n and the flat list."""
    passes = max(1, -(-len(rows) // team))
    n = max([len(r) for r in rows] + [1])
    flat = [pad] * (passes * n * team)
    for l, row in enumerate(rows):
        e, i = divmod(l, team)
        for q, v in enumerate(row):
            flat[(e * n + q) * team + i] = v
    return n, flat


def lane_tables(net: ReactionNetwork, team: int, t: dict | None = None) -> dict:
    """Lane-major padded forms of the per-lane sparse sums (model_tables):
    DXL / DXLC (J^T v, lane = species: entry r | o' << 16 with o' = NX for a
    unimolecular entry, whose factor x[NX] is 1, and the coefficient), SPL /
    SPLC (S w, lane = species), RSL / RSLC (S^T v, lane = reaction) and PRL /
    PRLC (df/dp^T, lane = parameter).  Padding entries have coefficient 0 and
    indices inside the arrays they address."""
    t = t or model_tables(net)
    nx = net.nx

    def rows(ptr, *cols):
        return [list(zip(*(c[ptr[i]:ptr[i + 1]] for c in cols))) for i in range(len(ptr) - 1)]

    out = {}
    dx = rows(t["DX_PTR"], t["DX_R"], t["DX_O"], t["DX_C"])
    n, idx = _lane_major([[r | ((o if o >= 0 else nx) << 16) for r, o, _ in row] for row in dx],
                         team, nx << 16)
    out["DXL_N"], out["DXL"] = n, idx
    out["DXLC"] = _lane_major([[c for _, _, c in row] for row in dx], team, 0.0)[1]
    sp = rows(t["SP_PTR"], t["SP_R"], t["SP_C"])
    out["SPL_N"], out["SPL"] = _lane_major([[r for r, _ in row] for row in sp], team, 0)
    out["SPLC"] = _lane_major([[c for _, c in row] for row in sp], team, 0.0)[1]
    rs = rows(t["RS_PTR"], t["RS_S"], t["RS_C"])
    out["RSL_N"], out["RSL"] = _lane_major([[i for i, _ in row] for row in rs], team, 0)
    out["RSLC"] = _lane_major([[c for _, c in row] for row in rs], team, 0.0)[1]
    pr = rows(t["PR_PTR"], t["PR_R"])
    out["PRL_N"], out["PRL"] = _lane_major([[r for (r,) in row] for row in pr], team, 0)
    out["PRLC"] = _lane_major([[1.0 for _ in row] for row in pr], team, 0.0)[1]
    return out


def _fmt_array(ctype: str, name: str, vals) -> str:
    vals = list(vals) if len(vals) else [0]
    if ctype == "double":
        body = ", ".join(repr(float(v)) for v in vals)
    else:
        body = ", ".join(str(int(v)) for v in vals)
    return f"  static constexpr {ctype} {name}[{len(vals)}] = {{{body}}};\n"


def emit_model_source(net: ReactionNetwork, team: int | None = None,
                      waves: int | None = None) -> str:
    """Model module source: traits plus the C ABI (amx_capi.inc)."""
    return emit_model_traits(net, team, waves) + '#include "amx_capi.inc"\n'


WAVES_CHOICES = (1, 2, 4)


def default_waves(nx: int) -> int:
    """Register budget of the default module (Launch<M>::WAVES_PER_EU)."""
    return 2 if nx <= 16 else 1


def emit_model_traits(net: ReactionNetwork, team: int | None = None,
                      waves: int | None = None) -> str:
    """The AmxModel traits struct (dimensions, team size, sparsity tables).
    ``waves``: register budget of the kernels as resident waves per SIMD (1:
    512 registers per lane, 2: 256, 4: 128); None keeps the default (2 up to
    16 states, else 1)."""
    t = model_tables(net)
    team = team or default_team(net.nx)
    if team < net.nx or team not in (64, 128, 256):
        raise ValueError(f"team size {team} does not fit nx={net.nx} (one lane per state; "
                         "a team is one 64-lane wavefront or a workgroup of 128 / 256 lanes)")
    if waves is not None and waves not in WAVES_CHOICES:
        raise ValueError(f"waves per SIMD must be one of {WAVES_CHOICES}")
    lines = [
        f"// Generated model module for network '{net.name}' (amici_amd.codegen).\n",
        "// Sparsity tables of xdot = S w, J, J^T v, df/dp and y = C x; see\n",
        "// amici_amd/codegen.py:model_tables for the meaning of every table.\n",
        f"#define AMX_TEAM_LANES {team if team > 64 else 64}\n",
        *([f"#define AMX_WAVES_PER_EU {waves}\n"] if waves is not None else []),
        '#include "amx_massaction.h"\n\n',
        "struct AmxModel {\n",
        f"  static constexpr int NX = {net.nx}, NR = {net.nr}, NP = {net.np}, NK = {net.nk},"
        f" NY = {net.ny};\n",
        f"  static constexpr int T = {team};\n",
        f"  static constexpr int NOISE = {int(net.noise)};\n",
        f"  static constexpr int JNZ = {t['JNZ']};\n",
    ]
    for name in _INT_TABLES:
        lines.append(_fmt_array("int", name, t[name]))
    for name in _DBL_TABLES:
        lines.append(_fmt_array("double", name, t[name]))
    lt = lane_tables(net, team, t)
    for name, ctype in (("DXL", "unsigned"), ("DXLC", "double"), ("SPL", "int"), ("SPLC", "double"),
                        ("RSL", "int"), ("RSLC", "double"), ("PRL", "int"), ("PRLC", "double")):
        if name + "_N" in lt:
            lines.append(f"  static constexpr int {name}_N = {lt[name + '_N']};\n")
        lines.append(_fmt_array(ctype, name, lt[name]))
    lines.append("};\n\n")
    return "".join(lines)


def model_manifest(net: ReactionNetwork, team: int, waves: int | None = None) -> dict:
    return dict(name=net.name, nx=net.nx, nr=net.nr, np=net.np, nk=net.nk, ny=net.ny,
                team=team, waves=waves, noise=int(net.noise))


def variant_name(net: ReactionNetwork, team: int | None = None, waves: int | None = None) -> str:
    """Module name of a launch-shape variant: the network's name for the
    default shape, with _t<team> / _w<waves> for the others."""
    team = team or default_team(net.nx)
    name = net.name if team == default_team(net.nx) else f"{net.name}_t{team}"
    return name if waves in (None, default_waves(net.nx)) else f"{name}_w{waves}"


def _source_digest(src: str) -> str:
    h = hashlib.sha256(src.encode())
    for f in sorted(CSRC.glob("*")):
        if f.is_file():
            h.update(f.read_bytes())
    return h.hexdigest()[:16]


def model_dir(name: str) -> Path:
    return BUILD_ROOT / name


def compile_model(net: ReactionNetwork, team: int | None = None, force: bool = False,
                  name: str | None = None, waves: int | None = None) -> Path:
    """Generate and compile the model module; returns the path of the .so.
    ``team`` and ``waves`` select a launch-shape variant (lanes per instance,
    register budget; see emit_model_traits).

    Reuses an existing library when its recorded source digest matches
    (unless ``force``).
    """
    team = team or default_team(net.nx)
    if waves == default_waves(net.nx):
        waves = None  # the default module
    name = name or variant_name(net, team, waves)
    src = emit_model_source(net, team, waves)
    d = model_dir(name)
    d.mkdir(parents=True, exist_ok=True)
    so = d / f"libamx_{name}.so"
    digest = _source_digest(src)
    man_path = d / "manifest.json"
    if so.exists() and man_path.exists() and not force:
        man = json.loads(man_path.read_text())
        if man.get("digest") == digest:
            return so
    (d / "amx_model.hip").write_text(src)
    tmp = d / f"libamx_{name}.so.tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", str(CSRC), str(d / "amx_model.hip"), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for model {name}:\n{' '.join(cmd)}\n{res.stderr[-8000:]}")
    shutil.move(str(tmp), str(so))
    man = model_manifest(net, team, waves)
    man["digest"] = digest
    man_path.write_text(json.dumps(man, indent=1))
    return so


# ---------------------------------------------------------------------------
# symbolic model (sympy) — the derivation the tables must agree with

def symbolic_model(net: ReactionNetwork):
    """sympy expressions for w, xdot, J = dxdot/dx, dxdot/dp and y.

    Mirrors DEModel's equations (python/sdist/amici/_symbolic/de_model.py:
    xdot :1385-1410, derivatives :1975-2056) for the mass-action case.
    """
    import sympy as sp

    x = sp.symbols(f"x0:{net.nx}")
    p = sp.symbols(f"p0:{net.np}")
    w = []
    for rx in net.reactions:
        e = p[rx.rate]
        for a in rx.reactants:
            e = e * x[a]
        w.append(e)
    S = net.stoichiometry()
    xdot = [sp.expand(sum(int(S[i, r]) * w[r] for r in range(net.nr) if S[i, r] != 0))
            for i in range(net.nx)]
    J = sp.Matrix(xdot).jacobian(sp.Matrix(x))
    dfdp = sp.Matrix(xdot).jacobian(sp.Matrix(p))
    y = [sum(sp.Float(net.observable_matrix[k, i]) * x[i]
             for i in range(net.nx) if net.observable_matrix[k, i] != 0) for k in range(net.ny)]
    return dict(x=x, p=p, w=w, xdot=xdot, J=J, dfdp=dfdp, y=y)


def op_counts(net: ReactionNetwork) -> dict:
    """Operation counts of the emitted quantities (sympy.count_ops), used by
    the algorithmic-FLOP model of the benchmark."""
    import sympy as sp

    sm = symbolic_model(net)
    c_f = sum(sp.count_ops(e) for e in sm["w"]) + sum(sp.count_ops(e) for e in sm["xdot"])
    c_J = sum(sp.count_ops(e) for e in sm["J"] if e != 0)
    c_fp = sum(1 for e in sm["dfdp"] if e != 0)
    c_y = sum(sp.count_ops(e) for e in sm["y"])
    return dict(c_f=int(c_f), c_J=int(c_J), c_fp=int(c_fp), c_y=int(c_y),
                nnz_J=int(sum(1 for e in sm["J"] if e != 0)))
