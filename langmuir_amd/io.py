"""Langmuir's own input and output files, read and written unchanged (host side).

* Input / checkpoint files (``*.inp``, ``*.chk``, optionally gzipped): sections
  ``[Electrons] [Holes] [Defects] [Traps] [TrapPotentials] [FluxState]
  [RandomState] [Parameters]`` as CheckPointer::load/save read and write them
  (checkpointer.cpp:22-170, 205-623), with the ``key = value`` lines of
  KeyValueParser::parse (keyvalueparser.cpp:120-175) and the ``key = value``
  layout of TypedVariable::keyValue (variable.h:195-240).
* ``out.dat``: the flux table of FluxWriter (writer.cpp:18-43, 360-379).
* ``*.coulomb``: the full-lattice map of Logger::saveCoulombEnergy
  (writer.cpp:545-581), every site written (the reference's read-back wrote
  only n values; see lmg_potential_map).

``[RandomState]`` is what ``operator<<(std::ostream&, Random&)`` writes
(rand.cpp:359-364; read back in that order by rand.cpp:366-381): the seed
``m_seed``, then the text state of Boost's mt19937.  That text is the C++
standard's representation of the engine, the last 624 raw words X(i-624) ..
X(i-1), oldest first (Boost >= 1.47 rewinds its in-place block to print them;
1.44's ``compute(j)`` prints the same words with a trailing blank).  Those words
are exactly the engine window ``rng_window`` of this package's simulation state,
so a checkpoint resumes the same random stream.
"""
from __future__ import annotations

import gzip
import re
import io as _io
import os
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Deque, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .host import SimulationParameters

# (key, type, default, constant) in registration order (keyvalueparser.cpp:15-105;
# defaults parameters.h:285-390).  Types: "str", "i32", "u32", "u64", "real", "bool", "datetime".
PARAMETER_TABLE: Tuple[Tuple[str, str, object, bool], ...] = (
    ("simulation.type", "str", "transistor", False),
    ("current.step", "u32", 0, False),
    ("iterations.real", "i32", 1000, False),
    ("random.seed", "u64", 0, False),
    ("grid.z", "i32", 1, False),
    ("grid.y", "i32", 128, False),
    ("grid.x", "i32", 128, False),
    ("hopping.range", "i32", 1, False),
    ("output.is.on", "bool", True, False),
    ("iterations.print", "i32", 10, False),
    ("output.precision", "i32", 15, False),
    ("output.width", "i32", 23, False),
    ("output.stub", "str", "out", False),
    ("output.ids.on.delete", "bool", False, False),
    ("output.ids.on.encounter", "bool", False, False),
    ("output.coulomb", "i32", 0, False),
    ("output.step.chk", "i32", 1, False),
    ("output.chk.trap.potential", "bool", False, False),
    ("output.potential", "bool", False, False),
    ("output.xyz", "i32", 0, False),
    ("output.xyz.e", "bool", True, False),
    ("output.xyz.h", "bool", True, False),
    ("output.xyz.d", "bool", True, False),
    ("output.xyz.t", "bool", True, False),
    ("output.xyz.mode", "i32", 0, False),
    ("image.traps", "bool", False, False),
    ("image.defects", "bool", False, False),
    ("image.carriers", "i32", 0, False),
    ("electron.percentage", "real", 0.01, False),
    ("hole.percentage", "real", 0.0, False),
    ("seed.charges", "real", 0.0, False),
    ("defect.percentage", "real", 0.0, False),
    ("trap.percentage", "real", 0.0, False),
    ("trap.potential", "real", 0.10, False),
    ("gaussian.stdev", "real", 0.0, False),
    ("seed.percentage", "real", 1.0, False),
    ("voltage.right", "real", 0.0, False),
    ("voltage.left", "real", 0.0, False),
    ("slope.z", "real", 0.0, False),
    ("coulomb.carriers", "bool", False, False),
    ("coulomb.gaussian.sigma", "real", 0.0, False),
    ("defects.charge", "i32", 0, False),
    ("exciton.binding", "real", 0.0, False),
    ("temperature.kelvin", "real", 300.0, False),
    ("source.rate", "real", 0.90, False),
    ("e.source.l.rate", "real", -1.0, False),
    ("e.source.r.rate", "real", -1.0, False),
    ("h.source.l.rate", "real", -1.0, False),
    ("h.source.r.rate", "real", -1.0, False),
    ("generation.rate", "real", 0.001, False),
    ("source.metropolis", "bool", False, False),
    ("source.coulomb", "bool", False, False),
    ("source.scale.area", "real", 65536.0, False),
    ("balance.charges", "bool", False, False),
    ("drain.rate", "real", 0.90, False),
    ("e.drain.l.rate", "real", -1.0, False),
    ("e.drain.r.rate", "real", -1.0, False),
    ("h.drain.l.rate", "real", -1.0, False),
    ("h.drain.r.rate", "real", -1.0, False),
    ("recombination.rate", "real", 0.0, False),
    ("recombination.range", "i32", 0, False),
    ("use.opencl", "bool", False, False),
    ("work.x", "i32", 4, False),
    ("work.y", "i32", 4, False),
    ("work.z", "i32", 4, False),
    ("work.size", "i32", 256, False),
    ("opencl.threshold", "i32", 256, False),
    ("opencl.device.id", "i32", 0, False),
    ("max.threads", "i32", -1, False),
    ("boltzmann.constant", "real", 1.3806504e-23, True),
    ("dielectric.constant", "real", 3.5, True),
    ("elementary.charge", "real", 1.60217646e-19, True),
    ("permittivity.space", "real", 8.854187817e-12, True),
    ("grid.factor", "real", 1e-9, True),
    ("electrostatic.cutoff", "i32", 50, True),
    ("electrostatic.prefactor", "real", 0.0, True),
    ("inverse.kt", "real", 0.0, True),
    ("ok.cl", "bool", False, True),
    ("simulation.start", "datetime", 0, True),
)
_TABLE = {k: (t, d, c) for k, t, d, c in PARAMETER_TABLE}

# flux-agent object names in World order (sourceagent.cpp:20-70, drainagent.cpp:20-70,
# world.cpp:961-1122); they head the out.dat columns
FLUX_AGENT_NAMES = ("eSourceL", "eSourceR", "hSourceL", "hSourceR", "xSource",
                    "eDrainL", "eDrainR", "hDrainL", "hDrainR", "xDrain")

SECTIONS = ("Parameters", "Electrons", "Holes", "Defects", "Traps", "TrapPotentials", "RandomState",
            "FluxState")


class ParseError(ValueError):
    pass


_INT_RANGE = {"i32": (-(1 << 31), (1 << 31) - 1), "u32": (0, (1 << 32) - 1), "u64": (0, (1 << 64) - 1)}
_INT_TOKEN = re.compile(r"^\s*[+-]?\d+\s*$")
_REAL_TOKEN = re.compile(r"^\s*[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?\s*$")


def _convert(key: str, kind: str, token: str):
    """TypedVariable<T>::convert (variable.h:309-395): QString::toDouble / toInt / toUInt /
    toULongLong (decimal, whole token, in range); bool is true/false (any case) or an
    integer."""
    if kind == "str":
        return token.strip()
    if kind == "real":
        if not _REAL_TOKEN.match(token):
            raise ParseError(f"can not convert to double: {token}")
        return float(token)
    if kind in _INT_RANGE:
        lo, hi = _INT_RANGE[kind]
        if not _INT_TOKEN.match(token) or not lo <= int(token) <= hi:
            raise ParseError(f"can not convert to {kind}: {token}")
        return int(token)
    if kind == "bool":
        low = token.strip().lower()
        if low in ("true", "false"):
            return low == "true"
        if not _INT_TOKEN.match(token) or not -(1 << 31) <= int(token) < (1 << 31):
            raise ParseError(f"can not convert to bool: {token}")
        return int(token) != 0
    raise AssertionError(kind)


def _format_value(kind: str, v) -> str:
    """TypedVariable<T>::value (variable.h:153-190): reals through QTextStream in scientific
    notation with 15 digits after the point (a two-digit exponent at least), booleans
    True/False, integers and strings as they are."""
    if kind == "real":
        return f"{float(v):.15e}"
    if kind == "bool":
        return "True" if v else "False"
    return str(v)


class KeyValues:
    """The parameter set a KeyValueParser holds: every registered key with its typed value
    (defaults until set), in registration order."""

    def __init__(self):
        self.values: Dict[str, object] = {k: d for k, _, d, _ in PARAMETER_TABLE}
        self.ignored_constants: List[str] = []

    def parse_line(self, line: str) -> None:
        """KeyValueParser::parse: '#' comments, blank lines, exactly one '=', key case
        folded, value trimmed of surrounding quotes; an empty value is skipped; constant
        keys are ignored (TypedVariable::read)."""
        parsed = line.strip()
        hash_at = parsed.find("#")
        if hash_at >= 0:
            parsed = parsed[:hash_at].rstrip()
        if not parsed:
            return
        if parsed.count("=") != 1:
            raise ParseError(f"parse error: missing equal sign\n\tline: {line.strip()}")
        key, value = parsed.split("=")
        key = key.strip().lower()
        if not key:
            raise ParseError(f"parse error: key is empty\n\tline: {line.strip()}")
        if key not in _TABLE:
            raise ParseError(f"parse error: invalid key\n\tline: {line.strip()}")
        value = value.strip().lstrip('"').rstrip('"')
        if not value:
            return
        kind, _, constant = _TABLE[key]
        if constant:
            self.ignored_constants.append(key)
            return
        self.values[key] = _convert(key, kind, value)

    def __getitem__(self, key: str):
        return self.values[key]

    def __setitem__(self, key: str, value) -> None:
        kind, _, _ = _TABLE[key]
        self.values[key] = _convert(key, kind, str(value)) if isinstance(value, str) else value

    def format(self) -> str:
        """operator<<(std::ostream&, KeyValueParser) (keyvalueparser.cpp:230-241): '\\n' +
        'key = value' for every non-constant key, the key left-justified in 30 columns,
        then '=' in 3 columns (2 before a negative value)."""
        out = []
        for key, kind, _, constant in PARAMETER_TABLE:
            if constant:
                continue
            v = _format_value(kind, self.values[key])
            if kind == "str":    # TypedVariable<QString>::keyValue (variable.h:221-240)
                out.append("\n" + key.ljust(30) + ("=" if v == "" else "=  " + v))
            else:
                out.append("\n" + key.ljust(30) + ("= " if v.startswith("-") else "=  ") + v)
        return "".join(out)

    def save(self, path: str) -> None:
        """KeyValueParser::save (keyvalueparser.cpp:176-196): the %stub.parm file."""
        with open(path, "w", encoding="latin-1", newline="\n") as f:
            f.write("[Parameters]" + self.format())

    def check(self) -> None:
        """checkSimulationParameters (parameters.h:393-646): every fatal check."""
        v = self.values
        if v["simulation.type"] not in ("transistor", "solarcell"):
            raise ParseError(f"simulation.type({v['simulation.type']}) must be transistor or solarcell")
        for k in ("grid.x", "grid.y", "grid.z"):
            if v[k] < 1:
                raise ParseError(f"{k}({v[k]}) >= 1")
        if v["iterations.print"] <= 0:
            raise ParseError(f"iterations.print({v['iterations.print']}) <= 0")
        if v["iterations.real"] < 0:
            raise ParseError(f"iterations.real({v['iterations.real']}) < 0")
        if v["iterations.real"] % v["iterations.print"] != 0:
            raise ParseError("iterations.real % iterations.print != 0")
        for k in ("electron.percentage", "hole.percentage", "defect.percentage", "trap.percentage",
                  "seed.percentage", "source.rate", "drain.rate", "seed.charges", "recombination.rate"):
            if not 0.0 <= v[k] <= 1.0:
                raise ParseError(f"{k}({v[k]}) < 0 || > 1.0")
        if v["defect.percentage"] > 1.0 - v["trap.percentage"]:
            raise ParseError("defect.percentage > 1.0 - trap.percentage")
        for k in ("e.source.l.rate", "e.source.r.rate", "h.source.l.rate", "h.source.r.rate",
                  "generation.rate", "e.drain.l.rate", "e.drain.r.rate", "h.drain.l.rate", "h.drain.r.rate"):
            if v[k] > 1.0:
                raise ParseError(f"{k}({v[k]}) < 0 || > 1.0")
        if v["gaussian.stdev"] < 0.0:
            raise ParseError("gaussian.stdev < 0.0")
        if v["temperature.kelvin"] < 0.0:
            raise ParseError("temperature.kelvin < 0.0")
        if v["defects.charge"] != 0 and not v["coulomb.carriers"]:
            raise ParseError("defects.charge != 0 && coulomb.carriers = false")
        if not 0 <= v["hopping.range"] <= 2:
            raise ParseError(f"hopping.range({v['hopping.range']}) < 0 || > 2")
        if not v["source.metropolis"] and v["source.coulomb"]:
            raise ParseError("source.metropolis = false, yet source.coulomb = true")
        solar = v["simulation.type"] == "solarcell"
        both = v["electron.percentage"] > 0 and v["hole.percentage"] > 0
        if v["recombination.rate"] > 0 and not solar:
            raise ParseError("recombination.rate > 0, yet simulation.type != solarcell")
        if v["recombination.rate"] and not both:
            raise ParseError("recombination.rate > 0, yet electron.percentage or hole.percentage is 0")
        if v["output.ids.on.encounter"] and not solar:
            raise ParseError("output.ids.on.encounter == true, yet simulation.type != solarcell")
        if v["output.ids.on.encounter"] and not both:
            raise ParseError("output.ids.on.encounter == true, yet electron.percentage or hole.percentage is 0")
        if (v["output.xyz"] > 0 and not (v["output.xyz.e"] or v["output.xyz.h"] or v["output.xyz.d"]
                                         or v["output.xyz.t"])):
            raise ParseError("output.xyz > 0, yet output.xyz.e/h/d/t are all false")
        if v["output.xyz.mode"] not in (0, 1):
            raise ParseError("output.xyz.mode must be 0 or 1")
        if v["opencl.threshold"] <= 0:
            raise ParseError("opencl.threshold must be >= 0")
        if v["opencl.device.id"] < 0:
            raise ParseError("opencl.device.id must be >= 0")
        if v["balance.charges"] and not solar:
            raise ParseError("balance.charges == true, yet simulation.type != solarcell")
        if v["coulomb.gaussian.sigma"] > 0 and 1 in (v["grid.x"], v["grid.y"], v["grid.z"]):
            raise ParseError("coulomb.gaussian.sigma > 0, yet the simulation is not 3D")
        if v["output.ids.on.encounter"]:
            raise ParseError("output.ids.on.encouter is deprecated")

    def to_parameters(self) -> SimulationParameters:
        """The SimulationParameters of this set (the keys the hot path reads)."""
        keys = {k: v for k, v in self.values.items() if k in SimulationParameters.KEYS}
        keys.pop("electrostatic.cutoff", None)   # constant in files (the default stands)
        return SimulationParameters.from_keys(keys)


@dataclass
class Checkpoint:
    """One input or checkpoint file.  Absent sections are None."""
    parameters: KeyValues = field(default_factory=KeyValues)
    electrons: Optional[List[int]] = None
    holes: Optional[List[int]] = None
    defects: Optional[List[int]] = None
    traps: Optional[List[int]] = None
    trap_potentials: Optional[List[float]] = None
    flux: Optional[List[int]] = None          # attempts and successes, interleaved per agent
    random_state: Optional[List[int]] = None  # 624 raw mt19937 words, oldest first
    random_seed: Optional[int] = None         # Random::m_seed, written before the words


def _open_text(path: str) -> str:
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:2] == b"\x1f\x8b":   # gzipped (gzipper.cpp)
        raw = gzip.decompress(raw)
    return raw.decode("latin-1")


def fix_traps(ck: Checkpoint) -> None:
    """LangmuirPython's CheckPoint.fix_traps (LangmuirPython/langmuir/checkpoint.py:337-388):
    trap.percentage = (n + 0.25) / V, so that int(trap.percentage * V) is the listed n and
    World seeds no more traps; seed.percentage = 0; a single shared trap potential moves
    into trap.potential (a mixed list stays, with trap.potential = 0)."""
    if not ck.traps:
        return
    size = len(ck.traps)
    if ck.trap_potentials:
        if len(ck.trap_potentials) != size:
            raise ValueError("number of traps and potentials do not agree")
        first = ck.trap_potentials[0]
        value = first if all(v == first for v in ck.trap_potentials) else None
    else:
        value = ck.parameters["trap.potential"]
    volume = ck.parameters["grid.x"] * ck.parameters["grid.y"] * ck.parameters["grid.z"]
    if size > volume:
        raise ValueError(f"number of traps exceeds volume: {size} > {volume}")
    ck.parameters["trap.percentage"] = (size + 0.25) / volume
    ck.parameters["seed.percentage"] = 0.0
    if value is None:
        ck.parameters["trap.potential"] = 0.0
    else:
        ck.trap_potentials = None
        ck.parameters["trap.potential"] = float(value)


def read_checkpoint(path: str) -> Checkpoint:
    """CheckPointer::load (checkpointer.cpp:22-170) of a file, gzipped or not."""
    return parse_checkpoint(_open_text(path))


def parse_checkpoint(text: str) -> Checkpoint:
    """CheckPointer::load on text: whitespace-separated section headers ('[Name]' or
    'Name'), each followed by a count and that many values; [RandomState] is the
    Boost engine's 624 words; [Parameters] runs line by line up to a line 'end' or
    the end of the file.  As std::getline there, a last parameter line with no newline
    after it is not read (checkpointer.cpp:403-428)."""
    ck = Checkpoint()
    stream = _io.StringIO(text)
    buf: Deque[str] = deque()

    def next_token() -> Optional[str]:
        while not buf:
            line = stream.readline()
            if not line:
                return None
            buf.extend(line.split())
        return buf.popleft()

    def uint(tok: str) -> int:
        if not _INT_TOKEN.match(tok) or not 0 <= int(tok) < (1 << 32):
            raise ValueError
        return int(tok)

    def u64(tok: str) -> int:
        if not _INT_TOKEN.match(tok) or not 0 <= int(tok) < (1 << 64):
            raise ValueError
        return int(tok)

    def real(tok: str) -> float:
        if not _REAL_TOKEN.match(tok):
            raise ValueError
        return float(tok)

    def read_values(what: str, conv, n: int) -> list:
        out = []
        for i in range(n):
            tok = next_token()
            if tok is None:
                raise ParseError(f"stream error: end of file\n\texpected {what} {i + 1} of {n}")
            try:
                out.append(conv(tok))
            except ValueError:
                raise ParseError(f"stream error: can not convert {tok}\n\texpected {what} {i + 1} of {n}") from None
        return out

    def read_count(what: str) -> int:
        return read_values(f"number of {what}", uint, 1)[0]

    while True:
        tok = next_token()
        if tok is None:
            break
        section = re.sub(r"[\[\]\s]", "", tok)
        if section not in SECTIONS:
            raise ParseError(f"invalid section encountered: {section}")
        if section == "Parameters":
            rest = " ".join(buf)      # the remainder of the header line
            buf.clear()
            ck.parameters.parse_line(rest)
            while True:
                line = stream.readline()
                if not line.endswith("\n"):
                    break
                if line.strip().lower() == "end":
                    break
                ck.parameters.parse_line(line)
        elif section == "Electrons":
            ck.electrons = read_values("electron", uint, read_count("electrons"))
        elif section == "Holes":
            ck.holes = read_values("hole", uint, read_count("holes"))
        elif section == "Defects":
            ck.defects = read_values("defect", uint, read_count("defects"))
        elif section == "Traps":
            ck.traps = read_values("trap", uint, read_count("traps"))
        elif section == "TrapPotentials":
            ck.trap_potentials = read_values("trap potential", real, read_count("trap potentials"))
        elif section == "FluxState":
            ck.flux = read_values("flux", u64, read_count("fluxes"))
        elif section == "RandomState":
            # stream >> m_seed, then stream >> *twister (rand.cpp:366-381)
            ck.random_seed = read_values("random number generator seed", u64, 1)[0]
            ck.random_state = read_values("random number generator state word", uint, 624)
    return ck


def format_checkpoint(ck: Checkpoint, trap_potentials: Optional[bool] = None) -> str:
    """CheckPointer::save (checkpointer.cpp:172-195, 473-623): Electrons, Holes, Defects,
    Traps, TrapPotentials (when output.chk.trap.potential; values with 17 significant
    decimals), FluxState, RandomState, then Parameters; a newline after each section
    but the last."""
    def ints(name, values):
        values = [] if values is None else list(values)
        return f"[{name}]\n{len(values)}" + "".join(f"\n{int(v)}" for v in values)

    parts = [ints("Electrons", ck.electrons), ints("Holes", ck.holes), ints("Defects", ck.defects),
             ints("Traps", ck.traps)]
    if ck.parameters["output.chk.trap.potential"] if trap_potentials is None else trap_potentials:
        tp = [] if ck.trap_potentials is None else list(ck.trap_potentials)
        parts.append(f"[TrapPotentials]\n{len(tp)}" + "".join(f"\n{float(v):.17e}" for v in tp))
    parts.append(ints("FluxState", [0] * (2 * len(FLUX_AGENT_NAMES)) if ck.flux is None else ck.flux))
    words = ck.random_state
    if words is None or len(words) != 624:
        raise ValueError("a checkpoint carries the engine state ([RandomState], 624 words)")
    if ck.random_seed is None:
        raise ValueError("a checkpoint carries the engine's seed (Random::m_seed) before its words")
    # stream << m_seed << ' ' << *twister (rand.cpp:359-364)
    parts.append(f"[RandomState]\n{int(ck.random_seed)} " + " ".join(str(int(w)) for w in words))
    return "\n".join(parts) + "\n[Parameters]" + ck.parameters.format()


def write_checkpoint(path: str, ck: Checkpoint, trap_potentials: Optional[bool] = None) -> None:
    with open(path, "w", encoding="latin-1", newline="\n") as f:
        f.write(format_checkpoint(ck, trap_potentials))


def output_name(template: str, parameters: KeyValues, directory: Optional[str] = None) -> str:
    """OutputInfo (output.cpp:10-60): %stub and %step (any case) replaced by output.stub
    and current.step; in the file's base name (up to its first '.') leading and trailing
    '-' removed and runs of '-' folded to one."""
    name = re.sub("%stub", lambda _: str(parameters["output.stub"]), template, flags=re.IGNORECASE)
    name = re.sub("%step", lambda _: str(parameters["current.step"]), name, flags=re.IGNORECASE)
    head, tail = os.path.split(name)
    base, dot, suffix = tail.partition(".")
    base = re.sub("-+", "-", base.strip("-"))
    if not base:
        raise ValueError(f"can not generate file name: the file name is empty: {name}")
    out = os.path.join(head, base + (("." + suffix) if suffix else ""))
    return out if directory is None else os.path.join(directory, out)


# ---------------------------------------------------------------------------
# simulation <-> checkpoint
# ---------------------------------------------------------------------------
def setup_of(ck: Checkpoint):
    """World::initialize for a loaded file (world.cpp:501-660): the listed carriers,
    defects and traps, the rest seeded from the engine -- the [RandomState] when the
    file has one, else random.seed (checkpointer.cpp:142-156).  Returns a
    world.WorldSetup."""
    from .world import initialize
    ck.parameters.check()
    return initialize(ck.parameters.to_parameters(), electrons=ck.electrons or (), holes=ck.holes or (),
                      defects=ck.defects or (), traps=ck.traps or (),
                      trap_potentials=ck.trap_potentials or None, rng_window=ck.random_state,
                      rng_seed=ck.random_seed)


def simulation_kwargs(ck: Checkpoint, setup) -> dict:
    """Simulation(...) keyword arguments of a loaded file and its set-up: sites in list
    order, the background potential, the engine window after set-up, the flux counters
    (setFluxInfo, world.cpp:626) and current.step.  Lifetimes and path lengths start at
    zero, as a load does in the reference (only sites are saved)."""
    kw = dict(electrons=setup.electrons, holes=setup.holes, defects=setup.defects,
              potential=setup.potential, rng_window=setup.rng_window, rng_consumed=setup.rng_consumed,
              current_step=int(ck.parameters["current.step"]))
    if ck.flux:
        f = np.asarray(ck.flux, dtype=np.uint64)
        if f.size != 2 * len(FLUX_AGENT_NAMES):
            raise ParseError(f"[FluxState] holds {f.size} values, expected {2 * len(FLUX_AGENT_NAMES)}")
        kw["flux_attempts"], kw["flux_successes"] = f[0::2].copy(), f[1::2].copy()
    return kw


def checkpoint_of(state: dict, parameters: KeyValues, setup) -> Checkpoint:
    """The checkpoint of a simulation state (Simulation.state()) with the defects and
    traps of its set-up: sites in list order, flux counters, the engine window and
    current.step."""
    kv = KeyValues()
    kv.values.update(parameters.values)
    kv.values["current.step"] = int(state["current_step"])
    flux: List[int] = []
    for a, s in zip(state["flux_attempts"], state["flux_successes"]):
        flux += [int(a), int(s)]
    return Checkpoint(parameters=kv, electrons=[int(v) for v in state["e_site"]],
                      holes=[int(v) for v in state["h_site"]], defects=list(setup.defects),
                      traps=list(setup.traps), trap_potentials=list(setup.trap_potentials),
                      flux=flux, random_state=[int(w) for w in np.asarray(state["rng_window"], np.uint64)],
                      random_seed=int(setup.rng_seed))


# ---------------------------------------------------------------------------
# out.dat and .coulomb
# ---------------------------------------------------------------------------
class FluxWriter:
    """out.dat (writer.cpp:18-43, 360-379): a header row, then one row per print interval:
    current.step, every flux agent's attempts and successes, the carrier counts and the
    milliseconds since the start; every field right-aligned in output.width columns, no
    separator, a bare newline."""

    def __init__(self, path: str, parameters: KeyValues, start: Optional[float] = None):
        self.width = int(parameters["output.width"])
        self.start = time.time() if start is None else start
        self.f = open(path, "w", encoding="latin-1", newline="\n")
        head = ["simulation:time"]
        for name in FLUX_AGENT_NAMES:
            head += [f"{name}:attempt", f"{name}:success"]
        head += ["electron:count", "hole:count", "real:time"]
        self._row(head)

    def _row(self, fields) -> None:
        self.f.write("".join(str(v).rjust(self.width) for v in fields) + "\n")
        self.f.flush()

    def write(self, state: dict, real_ms: Optional[int] = None) -> None:
        fields = [int(state["current_step"])]
        for a, s in zip(state["flux_attempts"], state["flux_successes"]):
            fields += [int(a), int(s)]
        ms = int((time.time() - self.start) * 1e3) if real_ms is None else int(real_ms)
        fields += [len(state["e_site"]), len(state["h_site"]), ms]
        self._row(fields)

    def close(self) -> None:
        self.f.close()


def read_flux_table(path: str) -> Tuple[List[str], np.ndarray]:
    """Column names and the integer rows of an out.dat file."""
    with open(path, encoding="latin-1") as f:
        head = f.readline().split()
        rows = [list(map(int, line.split())) for line in f if line.strip()]
    return head, np.asarray(rows, dtype=np.int64).reshape(-1, len(head))


def format_coulomb(dims: Sequence[int], values: np.ndarray, precision: int = 15) -> str:
    """Logger::saveCoulombEnergy (writer.cpp:545-581): 's x y z v', then one line per
    site, x outermost and z innermost, the value in scientific notation with
    output.precision digits after the point."""
    X, Y, Z = (int(n) for n in dims)
    v = np.asarray(values, dtype=np.float64).reshape(Z, Y, X)
    out = ["s x y z v\n"]
    for i in range(X):
        for j in range(Y):
            for k in range(Z):
                out.append(f"{i + X * (j + Y * k)} {i} {j} {k} {v[k, j, i]:.{precision}e}\n")
    return "".join(out)


def write_coulomb(path: str, dims: Sequence[int], values: np.ndarray, precision: int = 15) -> None:
    with open(path, "w", encoding="latin-1", newline="\n") as f:
        f.write(format_coulomb(dims, values, precision))


def read_coulomb(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(site ids, values) of a .coulomb file, in file order."""
    data = np.loadtxt(path, skiprows=1, ndmin=2)
    return data[:, 0].astype(np.int64), data[:, 4]


# ---------------------------------------------------------------------------
# the run loop of the langmuir program
# ---------------------------------------------------------------------------
def run_input(path: str, directory: Optional[str] = None, device_id: int = 0, de_path: str = "auto"):
    """langmuir.cpp:40-100 with Simulation::performIterations's output block
    (simulation.cpp:166-216): load the file, write %stub.parm, then for each print
    interval run iterations.print steps on the device and write the out.dat row, a
    %stub-%step.coulomb map every output.coulomb intervals (when carriers exist) and a
    %stub.chk checkpoint every output.step.chk intervals; a last checkpoint at the end.
    Returns (final state, parameters, set-up)."""
    from .host import GpuHelper, Simulation
    ck = read_checkpoint(path)
    kv = ck.parameters
    setup = setup_of(ck)
    kv.values["random.seed"] = setup.params.random_seed
    sim = Simulation(setup.params, device_id=device_id, de_path=de_path, **simulation_kwargs(ck, setup))
    on = bool(kv["output.is.on"])
    if directory is not None:
        os.makedirs(directory, exist_ok=True)
    if on:
        kv.save(output_name("%stub.parm", kv, directory))
    flux = FluxWriter(output_name("%stub.dat", kv, directory), kv) if on else None
    every = int(kv["iterations.print"])
    state = sim.state()
    try:
        for _ in range(int(kv["current.step"]), int(kv["iterations.real"]), every):
            sim.perform_iterations(every)
            state = sim.state()
            kv.values["current.step"] = step = int(state["current_step"])
            if not on:
                continue
            flux.write(state)
            n = len(state["e_site"]) + len(state["h_site"])
            if n > 0 and kv["output.coulomb"] > 0 and step % (every * kv["output.coulomb"]) == 0:
                helper = GpuHelper(sim.ctx)
                vmap = helper.launch_coulomb_kernel1(state["e_site"], state["h_site"], setup.defects,
                                                     setup.params.defects_charge)
                p = setup.params
                write_coulomb(output_name("%stub-%step.coulomb", kv, directory), (p.grid_x, p.grid_y, p.grid_z),
                              vmap, int(kv["output.precision"]))
            if kv["output.step.chk"] > 0 and step % (every * kv["output.step.chk"]) == 0:
                write_checkpoint(output_name("%stub.chk", kv, directory), checkpoint_of(state, kv, setup))
        if on:
            write_checkpoint(output_name("%stub.chk", kv, directory), checkpoint_of(state, kv, setup))
    finally:
        if flux is not None:
            flux.close()
    return state, kv, setup
