"""SolverFactory for the device PCG-MG path: the reference's INI solver configuration
(kernel/solver/solver_factory.hpp:376-640, data/ini/solver_example.ini) mapped onto the fused
solver object.

A configuration is a set of ``[section]`` blocks of ``key = value`` lines (``#`` starts a
comment), as read by the reference's PropertyMap.  ``create_scalar_solver(hierarchy, config,
"linsolver")`` follows the tree from the named root the way the reference's factory does and
accepts the shape this path implements::

    [linsolver]            type = pcg        (tolerances and iteration limits of the outer PCG)
      precon = mgv  ->     type = mg         cycle = v | f | w, hierarchy = h1, lvl_min = -1, lvl_max = 0,
                                             memorytype = hip, datatype = float | double
                                             (the cycle's precision; the extension key
                                             valuetype = half stores its matrices in binary16)
        hierarchy = h1 ->  type = hierarchy  smoother = rich, coarse = coarse
          smoother ->      type = richardson max_iter = min_iter = steps (pre = post), omega,
                                             precon = jacobi (type = jacobi, omega)
                           or type = jacobi  (one damped Jacobi step per smoothing)
          coarse ->        type = umfpack | cusolver | direct   (solved exactly in FP64)

``precon = none`` gives plain CG.  The keys the reference's solvers read from their sections
(IterativeSolver: tol_rel, tol_abs, tol_abs_low, div_rel, div_abs, stag_rate, max_iter,
min_iter, min_stag_iter, iterative.hpp:252-301; Richardson/Jacobi: omega, richardson.hpp:113,
jacobi_precond.hpp:106) keep their meaning.  A tree outside this shape (another Krylov solver, an
iterative coarse solver, a smoother with a stopping tolerance) raises ``FactoryError`` naming the
section: the fused path does not approximate it.  The extension keys
``format = csr | sell | banded`` (root or mg section) and ``graph = 0 | 1`` select the level-matrix
format and graph replay.
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

from .problems import Hierarchy
from .solver import CYCLE, MGPCG


class FactoryError(ValueError):
    """A configuration this solver path does not implement (names the section and key)."""


PropertyMap = Dict[str, Dict[str, str]]

_EXACT_COARSE = ("umfpack", "cusolver", "direct")
_INNER = {("double", None): "f64", ("float", None): "f32", ("float", "half"): "f16", ("double", "half"): "f16"}


def read_ini(text: str) -> PropertyMap:
    """Sections of ``key = value`` lines; ``#`` comments, surrounding blanks removed; a key
    repeated in a section keeps its last value; lines before the first section go to ""."""
    out: PropertyMap = {}
    cur = ""
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("[") and line.endswith("]"):
            cur = line[1:-1].strip()
            out.setdefault(cur, {})
            continue
        if "=" not in line:
            raise FactoryError(f"line {raw!r}: expected 'key = value' or '[section]'")
        k, v = line.split("=", 1)
        out.setdefault(cur, {})[k.strip()] = v.strip()
    return out


def _section(base: PropertyMap, name: str, parent: str) -> Dict[str, str]:
    if name not in base:
        raise FactoryError(f"section not found in property map: {name} (referenced by {parent})")
    return base[name]


def _num(sec: Dict[str, str], where: str, key: str, kind, default=None):
    if key not in sec:
        return default
    try:
        v = kind(float(sec[key])) if kind is int else kind(sec[key])
    except ValueError:
        raise FactoryError(f"{where}.{key}: cannot parse {sec[key]!r}")
    if kind is int and (float(sec[key]) != int(v) or v < 0):
        raise FactoryError(f"{where}.{key}: an integer >= 0 expected, got {sec[key]!r}")
    return v


def _smoother(base: PropertyMap, name: str) -> Tuple[float, float, int]:
    """(omega_jacobi, omega_richardson, steps) of a smoother section."""
    sec = _section(base, name, "hierarchy.smoother")
    t = sec.get("type")
    if t == "jacobi":
        return _num(sec, name, "omega", float, 1.0), 1.0, 1
    if t != "richardson":
        raise FactoryError(f"{name}.type = {t}: the device multigrid smooths with richardson(jacobi) or jacobi")
    mx, mn = _num(sec, name, "max_iter", int, 100), _num(sec, name, "min_iter", int, 0)
    if mx != mn:
        raise FactoryError(f"{name}: a smoother runs a fixed number of steps here (min_iter = max_iter), "
                           f"got min_iter = {mn}, max_iter = {mx}")
    if mx < 1:
        raise FactoryError(f"{name}.max_iter: at least one smoothing step")
    if "min_stag_iter" in sec and _num(sec, name, "min_stag_iter", int) != 0:
        raise FactoryError(f"{name}.min_stag_iter: no stagnation test in the smoother")
    omega_r = _num(sec, name, "omega", float, 1.0)
    pre = sec.get("precon", "none")
    if pre == "none":
        raise FactoryError(f"{name}.precon: the device smoother is richardson with a jacobi preconditioner")
    psec = _section(base, pre, f"{name}.precon")
    if psec.get("type") != "jacobi":
        raise FactoryError(f"{pre}.type = {psec.get('type')}: the smoother's preconditioner must be jacobi")
    omega_j = _num(psec, pre, "omega", float, 1.0)
    return omega_j, omega_r, mx


def solver_kwargs(base: PropertyMap, section: str = "linsolver") -> dict:
    """MGPCG keyword arguments for the tree rooted at ``section`` (see the module docstring)."""
    root = _section(base, section, "the caller")
    if root.get("type") != "pcg":
        raise FactoryError(f"{section}.type = {root.get('type')}: this path is PCG preconditioned by multigrid (type = pcg)")
    if root.get("datatype", "double") != "double":
        raise FactoryError(f"{section}.datatype: the outer PCG runs in double")
    kw: dict = {}
    for key, kind in (("tol_rel", float), ("tol_abs", float), ("tol_abs_low", float), ("max_iter", int),
                      ("min_iter", int)):
        v = _num(root, section, key, kind)
        if v is not None:
            kw[key] = v
    for key in ("div_rel", "div_abs", "stag_rate", "min_stag_iter"):
        if key in root:
            raise FactoryError(f"{section}.{key}: not configurable on the device PCG (the reference's defaults apply)")
    fmt = root.get("format")
    pre = root.get("precon", "none")
    if pre == "none":
        kw["precond"] = "none"
    else:
        mg = _section(base, pre, f"{section}.precon")
        if mg.get("type") != "mg":
            raise FactoryError(f"{pre}.type = {mg.get('type')}: the preconditioner must be a multigrid (type = mg)")
        if "cycle" not in mg:
            raise FactoryError(f"{pre}: mg section without cycle key is not allowed")
        cyc = mg["cycle"].lower()
        if cyc not in CYCLE:
            raise FactoryError(f"{pre}.cycle = {mg['cycle']}: mg cycle unknown (v, f or w)")
        if cyc != "v":
            kw["cycle"] = cyc
        if int(_num(mg, pre, "lvl_max", float, 0)) != 0 or int(_num(mg, pre, "lvl_min", float, -1)) != -1:
            raise FactoryError(f"{pre}: the cycle spans the whole hierarchy (lvl_max = 0, lvl_min = -1)")
        if mg.get("adapt_cgc", "fixed") != "fixed":
            raise FactoryError(f"{pre}.adapt_cgc: the coarse-grid correction is fixed (omega = 1)")
        mem = mg.get("memorytype", "hip")
        if mem not in ("hip", "main"):
            raise FactoryError(f"{pre}.memorytype = {mem}: the cycle runs on the device (memorytype = hip)")
        key = (mg.get("datatype", "double"), mg.get("valuetype"))
        if key not in _INNER:
            raise FactoryError(f"{pre}: datatype = {key[0]}, valuetype = {key[1]} is not one of double, float, float+half")
        kw["inner"] = _INNER[key]
        fmt = mg.get("format", fmt)
        hname = mg.get("hierarchy")
        if hname is None:
            raise FactoryError(f"{pre}: no hierarchy key found in mg section")
        h = _section(base, hname, f"{pre}.hierarchy")
        if h.get("type") != "hierarchy":
            raise FactoryError(f"{hname}.type must be 'hierarchy'")
        if "smoother" not in h or "coarse" not in h:
            raise FactoryError(f"{hname}: a hierarchy section needs smoother and coarse keys")
        oj, orr, steps = _smoother(base, h["smoother"])
        kw.update(omega_jacobi=oj, omega_richardson=orr, pre=steps, post=steps)
        c = _section(base, h["coarse"], f"{hname}.coarse")
        if c.get("type") not in _EXACT_COARSE:
            raise FactoryError(f"{h['coarse']}.type = {c.get('type')}: the coarse level is solved exactly "
                               f"(type = {' | '.join(_EXACT_COARSE)})")
    if fmt is not None:
        if fmt not in ("csr", "sell", "banded"):
            raise FactoryError(f"format = {fmt}: one of csr, sell, banded")
        kw["fmt"] = fmt
    if "graph" in root:
        kw["graph"] = bool(_num(root, section, "graph", int))
    return kw


def create_scalar_solver(hierarchy: Hierarchy, config, section: str = "linsolver",
                         options: Optional[dict] = None) -> MGPCG:
    """SolverFactory::create_scalar_solver for the device path: ``config`` is INI text or a
    parsed property map; returns the (not yet initialised) solver object."""
    base = read_ini(config) if isinstance(config, str) else config
    return MGPCG(hierarchy, options=options, **solver_kwargs(base, section))
