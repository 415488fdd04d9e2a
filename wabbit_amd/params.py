"""Run parameters: WABBIT's INI format and the subset of type_params /
type_params_acm the per-block time step reads.

INI syntax follows LIB/PARAMS/module_ini_files_parser.f90: ``[Section]``
headers, ``key=value;`` lines, everything after the first ``;`` is a comment,
lines starting with ``;`` are skipped, an empty value selects the default,
vectors are blank separated, matrices are written ``key=(/ row; row /)``
over several lines.  Defaults are the reference's own, cited per key.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

# order_discretization -> scheme code used by the C ABI (wabbit_hip.h)
SCHEMES = {
    "FD_2nd_central": 2,
    "FD_4th_central": 4,
    "FD_4th_central_optimized": 40,
    "FD_6th_central": 6,
}
# stencil reach of each scheme: the ghost width the RHS needs
FD_REACH = {2: 1, 4: 2, 40: 3, 6: 3}


class IniFile:
    """Parsed INI file: {section: {key: raw value string}}."""

    def __init__(self, text: str = ""):
        self.data: Dict[str, Dict[str, str]] = {}
        self._parse(text)

    @classmethod
    def read(cls, path: str) -> "IniFile":
        with open(path, "r", errors="replace") as f:
            return cls(f.read())

    def _parse(self, text: str) -> None:
        sec = None
        lines = text.splitlines()
        i = 0
        while i < len(lines):
            line = lines[i].strip()
            i += 1
            if not line or line.startswith(";") or line.startswith("#"):
                continue
            m = re.match(r"^\[(.+)\]", line)
            if m:
                sec = m.group(1).strip()
                self.data.setdefault(sec, {})
                continue
            if "=" not in line or sec is None:
                continue
            key, val = line.split("=", 1)
            key = key.strip()
            if val.lstrip().startswith("(/"):
                # matrix spanning lines until "/)"
                buf = val.lstrip()[2:]
                while "/)" not in buf and i < len(lines):
                    buf += " ; " + lines[i].strip()
                    i += 1
                buf = buf.split("/)")[0]
                self.data[sec][key] = "(/" + buf + "/)"
                continue
            if ";" in val:
                val = val.split(";", 1)[0]
            self.data[sec][key] = val.strip()

    def raw(self, section: str, key: str) -> Optional[str]:
        v = self.data.get(section, {}).get(key)
        if v is None or v == "":
            return None
        return v

    def get_float(self, section, key, default):
        v = self.raw(section, key)
        return float(v.split()[0].replace("d", "e").replace("D", "e")) if v else default

    def get_int(self, section, key, default):
        v = self.raw(section, key)
        return int(float(v.split()[0])) if v else default

    def get_bool(self, section, key, default):
        v = self.raw(section, key)
        if v is None:
            return default
        t = v.split()[0].lower()
        return t in ("1", "yes", "true", ".true.", "y", "t")

    def get_str(self, section, key, default):
        v = self.raw(section, key)
        return v.split()[0] if v else default

    def get_vec(self, section, key, default: List[float]):
        v = self.raw(section, key)
        if not v:
            return list(default)
        return [float(x) for x in v.replace(",", " ").split()]

    def get_matrix(self, section, key, default):
        v = self.raw(section, key)
        if not v or not v.startswith("(/"):
            return np.asarray(default, dtype=np.float64)
        rows = [r for r in v[2:-2].split(";") if r.strip()]
        return np.asarray([[float(x) for x in r.replace(",", " ").split()] for r in rows], np.float64)


def butcher_rk4() -> np.ndarray:
    """Default Butcher tableau (LIB/MESH/ini_file_to_params.f90:640-646)."""
    return np.array([[0.0, 0.0, 0.0, 0.0, 0.0],
                     [0.5, 0.5, 0.0, 0.0, 0.0],
                     [0.5, 0.0, 0.5, 0.0, 0.0],
                     [1.0, 0.0, 0.0, 1.0, 0.0],
                     [0.0, 1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]], dtype=np.float64)


def cfl_nu_default(scheme: int, dim: int) -> float:
    """module_ACM.f90:367-377: default diffusive CFL by scheme order."""
    denom = {2: 4.000, 4: 5.333, 40: 5.333, 6: 6.0444}[scheme]
    return 0.95 * 2.79 / (denom * float(dim))


@dataclass
class ACMParams:
    """type_params_acm subset (LIB/EQUATION/ACMnew/module_ACM.f90:56-149)."""
    dim: int = 2
    scheme: int = 4                       # order_discretization
    skew_symmetry: bool = False           # module_ACM.f90 'skew_symmetry'
    penalization: bool = False            # [VPM] penalization (default .true. in the reference)
    use_sponge: bool = False              # [Sponge] use_sponge
    c_0: float = 20.0
    nu: float = 1.0e-3
    gamma_p: float = 1.0                  # module_ACM.f90:237
    C_eta: float = 1.0                    # module_ACM.f90:272
    C_sponge: float = 1.0e-2              # module_ACM.f90:360
    u_mean_set: List[float] = field(default_factory=lambda: [1.0, 0.0, 0.0])  # :349
    CFL: float = 1.0                      # module_ACM.f90:364
    CFL_nu: float = 0.0                   # 0 -> scheme default (cfl_nu_default)
    CFL_eta: float = 0.99                 # module_ACM.f90:365
    domain_size: List[float] = field(default_factory=lambda: [1.0, 1.0, 1.0])

    def __post_init__(self):
        if not self.CFL_nu:
            self.CFL_nu = cfl_nu_default(self.scheme, self.dim)
        self.u_mean_set = list(self.u_mean_set) + [0.0] * (3 - len(self.u_mean_set))

    @property
    def neq(self) -> int:
        return self.dim + 1


@dataclass
class TimeParams:
    """The dt clamps of calculate_time_step (LIB/TIME/calculate_time_step.f90:58-118)
    and the stop conditions of the main loop (LIB/MAIN/main.f90:293)."""
    time_max: float = 1.0                 # ini_file_to_params.f90:586
    nt: int = 99999999                    # :591
    dt_fixed: float = 0.0
    dt_max: float = 0.0
    write_method: str = "fixed_freq"      # :597
    write_time: float = 1.0               # :601
    write_time_first: float = 0.0         # :606
    write_backup_n: int = 0
    write_backup_time: float = 1.0
    tsave_stats: float = 9999999.9        # :189
    butcher: np.ndarray = field(default_factory=butcher_rk4)


def params_from_ini(ini: IniFile):
    """Read the ACM and time parameters the hot path needs from an INI file."""
    dim = ini.get_int("Domain", "dim", 2)
    dom = ini.get_vec("Domain", "domain_size", [1.0, 1.0, 1.0])
    disc = ini.get_str("Discretization", "order_discretization", "FD_4th_central")
    scheme = SCHEMES[disc]
    acm = ACMParams(
        dim=dim, scheme=scheme,
        skew_symmetry=ini.get_bool("ACM-new", "skew_symmetry", False),
        penalization=ini.get_bool("VPM", "penalization", True),
        use_sponge=ini.get_bool("Sponge", "use_sponge", False),
        c_0=ini.get_float("ACM-new", "c_0", 10.0),
        nu=ini.get_float("ACM-new", "nu", 1.0e-1),
        gamma_p=ini.get_float("ACM-new", "gamma_p", 1.0),
        C_eta=ini.get_float("VPM", "C_eta", 1.0),
        C_sponge=ini.get_float("Sponge", "C_sponge", 1.0e-2),
        u_mean_set=ini.get_vec("ACM-new", "u_mean_set", [1.0, 0.0, 0.0]),
        CFL=ini.get_float("Time", "CFL", 1.0),
        CFL_nu=ini.get_float("Time", "CFL_nu", 0.0),
        CFL_eta=ini.get_float("Time", "CFL_eta", 0.99),
        domain_size=dom + [1.0] * (3 - len(dom)),
    )
    tp = TimeParams(
        time_max=ini.get_float("Time", "time_max", 1.0),
        nt=ini.get_int("Time", "nt", 99999999),
        dt_fixed=ini.get_float("Time", "dt_fixed", 0.0),
        dt_max=ini.get_float("Time", "dt_max", 0.0),
        write_method=ini.get_str("Time", "write_method", "fixed_freq"),
        write_time=ini.get_float("Time", "write_time", 1.0),
        write_time_first=ini.get_float("Time", "write_time_first", 0.0),
        write_backup_n=ini.get_int("Time", "write_backup_n", 0),
        tsave_stats=ini.get_float("Statistics", "tsave_stats", 9999999.9),
        butcher=ini.get_matrix("Time", "butcher_tableau", butcher_rk4()),
    )
    return acm, tp


def clamp_dt(dt_phys: float, time: float, tp: TimeParams) -> float:
    """Host part of calculate_time_step (calculate_time_step.f90:22-118): apply
    dt_fixed / dt_max and truncate dt so that output, backup and statistics
    times and time_max are hit exactly.  ``dt_phys`` is the minimum over blocks
    of the physics dt (the device reduction)."""
    import math

    def fmod(a, b):  # Fortran MOD for reals: a - int(a/b)*b
        return math.fmod(a, b)

    if tp.dt_fixed > 0.0:
        dt = tp.dt_fixed
    else:
        dt = dt_phys
        if tp.dt_max > 0.0:
            dt = min(tp.dt_max, dt)
    if tp.write_method == "fixed_time":
        if (fmod(time + dt, tp.write_time) < fmod(time + 1e-12, tp.write_time)
                and not abs(fmod(time, tp.write_time)) < 1e-12
                and time + 1e-12 > tp.write_time_first):
            dt = tp.write_time - fmod(time, tp.write_time)
    if tp.write_backup_n > 0:
        wb = tp.write_backup_time
        if fmod(time + dt, wb) < fmod(time + 1e-12, wb) and not abs(fmod(time, wb)) < 1e-12:
            dt = wb - fmod(time, wb)
    if abs(tp.tsave_stats - 9999999.9) > 1e-3:
        ts = tp.tsave_stats
        if fmod(time + dt, ts) < fmod(time + 1e-12, ts) and not abs(fmod(time, ts)) < 1e-12:
            dt = ts - fmod(time, ts)
    if time + dt > tp.time_max and time <= tp.time_max:
        dt = tp.time_max - time
    return dt
