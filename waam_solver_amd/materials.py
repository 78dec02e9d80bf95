"""Material properties (reference materials.py:23-247, material_tables.py, physics/surfactant.py).

A material is the ``constants`` block of a YAML file plus optional surface, electrical and
surfactant sections and optional piecewise-linear tables. The surfactant model adjusts dγ/dT
at load time (Heiple static sulphur scale by default, reference physics/surfactant.py:28-50,
or the Sahoo/DebRoy Fe-S isotherm, :122-174).
"""

from __future__ import annotations

import math
import pathlib
from dataclasses import dataclass, field

PKG_MATERIALS = pathlib.Path(__file__).resolve().parent / "materials"


@dataclass
class MaterialTables:
    cp: list[tuple[float, float]] = field(default_factory=list)
    k: list[tuple[float, float]] = field(default_factory=list)
    mu: list[tuple[float, float]] = field(default_factory=list)
    dgamma_dT: list[tuple[float, float]] = field(default_factory=list)

    def has_any(self) -> bool:
        return bool(self.cp or self.k or self.mu or self.dgamma_dT)


def interp_linear(T: float, knots, fallback: float) -> float:
    if not knots:
        return fallback
    if T <= knots[0][0]:
        return knots[0][1]
    if T >= knots[-1][0]:
        return knots[-1][1]
    for (t0, v0), (t1, v1) in zip(knots[:-1], knots[1:]):
        if t0 <= T <= t1:
            if t1 == t0:
                return v0
            return v0 + (T - t0) / (t1 - t0) * (v1 - v0)
    return fallback


@dataclass
class MaterialProps:
    name: str
    k: float
    rho: float
    cp: float
    alpha: float
    T_solidus: float
    T_liquidus: float
    L_fusion: float
    gamma_0: float
    dgamma_dT: float
    T_ref_gamma: float
    mu: float
    beta_T: float = 1.2e-4
    contact_angle_deg: float = 80.0
    rho_e_ohm_m: float = 1.5e-7
    eta_stick: float = 0.85
    sulphur_ppm: float = 30.0
    status: str = "placeholder"
    source: str = ""
    high_sulphur: bool = False
    tables: MaterialTables = field(default_factory=MaterialTables)

    def cp_at(self, T: float) -> float:
        return interp_linear(T, self.tables.cp, self.cp)


def props_from_constants(name: str, status: str, source: str, c: dict) -> MaterialProps:
    rho, cp, k = float(c["rho"]), float(c["cp"]), float(c["k"])
    return MaterialProps(
        name=name, k=k, rho=rho, cp=cp, alpha=float(c.get("alpha", k / (rho * cp))),
        T_solidus=float(c["T_solidus"]), T_liquidus=float(c["T_liquidus"]),
        L_fusion=float(c["L_fusion"]), gamma_0=float(c["gamma_0"]),
        dgamma_dT=float(c["dgamma_dT"]), T_ref_gamma=float(c.get("T_ref_gamma", c["T_liquidus"])),
        mu=float(c["mu"]), beta_T=float(c.get("beta_T", 1.2e-4)), status=status,
        source=source or "")


# ---- surfactant models ---------------------------------------------------------------------
def heiple_scale(sulphur_ppm: float) -> float:
    s = max(0.0, sulphur_ppm)
    if s <= 30.0:
        return 1.0
    if s >= 60.0:
        return -0.35
    return 1.0 - 1.35 * (s - 30.0) / 30.0


def heiple_dgamma_dT(base: float, sulphur_ppm: float) -> float:
    s = heiple_scale(sulphur_ppm)
    return -abs(base) * s if s < 0 else base * s


_SAHOO = dict(gamma_m=1.943, A=4.3e-4, T_m=1809.0, Gamma_s=1.3e-5, k1=0.00318, dH0=-1.66e5,
              R=8.314462618)


def dgamma_dT_sahoo(T: float, sulphur_ppm: float) -> float:
    p = _SAHOO
    T = max(float(T), 300.0)
    a_s = max(0.0, float(sulphur_ppm)) / 10000.0
    Ka = p["k1"] * math.exp(-p["dH0"] / (p["R"] * T)) * a_s
    frac = Ka / (1.0 + Ka) if Ka > 0.0 else 0.0
    return -p["A"] - p["R"] * p["Gamma_s"] * math.log(1.0 + Ka) - frac * p["Gamma_s"] * p["dH0"] / T


def gamma_sahoo(T: float, sulphur_ppm: float, T_m: float) -> float:
    p = _SAHOO
    T = max(float(T), 300.0)
    a_s = max(0.0, float(sulphur_ppm)) / 10000.0
    Ka = p["k1"] * math.exp(-p["dH0"] / (p["R"] * T)) * a_s
    return p["gamma_m"] - p["A"] * (T - T_m) - p["R"] * T * p["Gamma_s"] * math.log(1.0 + Ka)


def apply_surfactant_model(props: MaterialProps, model: str = "heiple") -> None:
    key = (model or "heiple").lower().replace("-", "_")
    if key in ("sahoo", "sahoo_debroy", "debroy"):
        T_liq = float(props.T_liquidus)
        T_hi = max(T_liq + 800.0, 2600.0)
        n = 8
        props.tables.dgamma_dT = [
            (T_liq + (T_hi - T_liq) * i / (n - 1),
             dgamma_dT_sahoo(T_liq + (T_hi - T_liq) * i / (n - 1), props.sulphur_ppm))
            for i in range(n)]
        props.dgamma_dT = dgamma_dT_sahoo(T_liq + 50.0, props.sulphur_ppm)
        props.gamma_0 = gamma_sahoo(T_liq, props.sulphur_ppm, T_liq)
        props.T_ref_gamma = T_liq
        return
    props.dgamma_dT = heiple_dgamma_dT(props.dgamma_dT, props.sulphur_ppm)
    if props.tables.dgamma_dT:
        props.tables.dgamma_dT = [(t, heiple_dgamma_dT(v, props.sulphur_ppm))
                                  for t, v in props.tables.dgamma_dT]


def _knots(raw) -> list[tuple[float, float]]:
    out = [(float(r[0]), float(r[1])) for r in (raw or []) if len(r) == 2]
    out.sort(key=lambda p: p[0])
    return out


def material_from_dict(data: dict) -> MaterialProps:
    for key in ("name", "constants"):
        if key not in data:
            raise ValueError(f"material definition missing '{key}'")
    c = data["constants"]
    for key in ("rho", "cp", "k", "T_solidus", "T_liquidus", "L_fusion"):
        if key not in c:
            raise ValueError(f"material constants missing '{key}'")
    props = props_from_constants(data["name"], data.get("status", "placeholder"),
                                 data.get("source", ""), c)
    surface = data.get("surface", {}) or {}
    elec = data.get("electrical", {}) or {}
    surf = data.get("surfactant", {}) or {}
    props.contact_angle_deg = float(surface.get("contact_angle_deg", 80.0))
    props.rho_e_ohm_m = float(elec.get("rho_e_ohm_m", 1.5e-7))
    props.eta_stick = float(elec.get("eta_stick", 0.85))
    props.sulphur_ppm = float(surf.get("sulphur_ppm", 30.0))
    tbl = data.get("tables") or {}
    props.tables = MaterialTables(cp=_knots(tbl.get("cp")), k=_knots(tbl.get("k")),
                                  mu=_knots(tbl.get("mu")), dgamma_dT=_knots(tbl.get("dgamma_dT")))
    apply_surfactant_model(props, str(surf.get("model", "heiple")))
    return props


def resolve_material_path(name: str | pathlib.Path) -> pathlib.Path | None:
    """A YAML path (contains '/' or ends in .yaml/.yml; relative paths also resolve against the
    package directory), else ``materials/validated/<name>.yaml`` then ``materials/<name>.yaml``
    in the package (reference materials.py:153-167, :199-200)."""
    s = str(name)
    if "/" in s or s.endswith((".yaml", ".yml")):
        # as given, else relative to the package (job files name "materials/validated/...")
        for p in (pathlib.Path(s), PKG_MATERIALS.parent / s):
            if p.exists():
                return p
        return None
    for base in (PKG_MATERIALS / "validated", PKG_MATERIALS):
        for cand in (base / f"{s}.yaml", base / f"{s}.yml"):
            if cand.exists():
                return cand
    return None


def load_material(name: str | pathlib.Path, high_sulphur: bool = False) -> MaterialProps:
    """Load a material by name (package materials/ directory) or by YAML path."""
    import yaml

    p = resolve_material_path(name)
    if p is None:
        raise KeyError(f"Material '{name}' not found")
    with open(p) as f:
        props = material_from_dict(yaml.safe_load(f))
    if high_sulphur:
        props.sulphur_ppm = max(float(props.sulphur_ppm), 150.0)
        props.high_sulphur = True
        props.dgamma_dT = abs(float(props.dgamma_dT))
    return props
