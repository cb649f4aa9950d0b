"""Ad hoc parameter changes of molflowCLI (--setParams, --setParamsByFile):
the ParameterParser of the reference (src/ParameterParser.cpp:43-255).

    changes = parse(["facet.3.sticking=0.01", "facet.\\"pumps\\".sticking=0.2;simulation.mass=2"],
                    selections)
    n_errors = apply(model, changes)

Lines are "facet.<ids>.<name>=<value>" or "simulation.<name>=<value>";
several may be joined with ';' in one argument. <ids> is a 1-based facet
number, a range "a-b", a comma list of those, or a quoted selection group
name of the XML file. Facet names: opacity, temperature, sticking,
outgassing (mbar l/s), specificOutgassing (mbar l/s/cm^2); simulation
names: mass, enableDecay, halfLife. Values are read like strtod (an
unreadable value is 0). Out-of-range opacity or sticking values are applied
and counted as errors, as in ChangeFacetParams, and so are facet ids
beyond the model; setting a positive
outgassing on a facet that does not desorb makes it desorb with a cosine
law. Unknown lines, names and facet ids are reported and skipped.
"""
from __future__ import annotations

import re
import sys
from dataclasses import dataclass, field

from . import _abi as A

MBARLS_TO_PAM3S = 0.1
FACET_NAMES = ("opacity", "temperature", "sticking", "outgassing", "specificOutgassing")
SIMU_NAMES = ("mass", "enableDecay", "halfLife")


@dataclass
class ParamChanges:
    facet: list = field(default_factory=list)     # (facet index 0-based, name, value)
    simulation: list = field(default_factory=list)  # (name, value)
    messages: list = field(default_factory=list)


def _strtod(s: str) -> float:
    m = re.match(r"\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?|inf|nan)", s, re.IGNORECASE)
    return float(m.group(0)) if m else 0.0


def split_facet_list(text: str, n_max: int = 10_000_000) -> list[int]:
    """1-based ids and inclusive ranges "a-b", comma separated, to 0-based
    facet indices."""
    out = []
    for part in text.split(","):
        part = part.strip()
        if not part:
            continue
        if "-" in part[1:]:
            a, b = part.split("-", 1)
            a, b = int(a), int(b)
            if a < 1 or b < a or b > n_max:
                raise ValueError(f"bad facet range {part!r}")
            out.extend(range(a - 1, b))
        else:
            a = int(part)
            if a < 1 or a > n_max:
                raise ValueError(f"bad facet id {part!r}")
            out.append(a - 1)
    return out


def parse(lines, selections: dict | None = None) -> ParamChanges:
    """Parse --setParams arguments (each may hold several ';'-joined
    changes) or the lines of a --setParamsByFile file."""
    selections = selections or {}
    ch = ParamChanges()
    flat = []
    for item in lines:
        for line in str(item).splitlines():
            flat.extend(x for x in line.split(";"))
    for i, line in enumerate(flat, 1):
        line = line.strip()
        if not line:
            continue
        kind, _, rest = line.partition(".")
        if kind == "facet":
            # the id part ends at the first '.' outside quotes
            if rest.startswith('"'):
                end = rest.find('"', 1)
                ids_txt, rest2 = rest[: end + 1], rest[end + 2:]
            else:
                ids_txt, _, rest2 = rest.partition(".")
            name, _, val = rest2.partition("=")
            if ids_txt.startswith('"'):
                grp = ids_txt.strip('"')
                if grp not in selections:
                    ch.messages.append(f"selection \"{grp}\" not found in file, ignoring")
                    continue
                ids = list(selections[grp])
            else:
                try:
                    ids = split_facet_list(ids_txt)
                except ValueError as e:
                    ch.messages.append(f"could not parse facet id or range: {ids_txt} ({e})")
                    continue
            if name not in FACET_NAMES:
                ch.messages.append(f"invalid option: {name}")
                continue
            v = _strtod(val)
            ch.facet.extend((f, name, v) for f in ids)
        elif kind == "simulation":
            name, _, val = rest.partition("=")
            if name not in SIMU_NAMES:
                ch.messages.append(f"invalid option: {name}")
                continue
            ch.simulation.append((name, _strtod(val)))
        else:
            ch.messages.append(f"[line #{i}] unknown input {line}")
    return ch


def parse_file(path: str, selections: dict | None = None) -> ParamChanges:
    with open(path) as fh:
        return parse(fh.read().splitlines(), selections)


def selections_of(root) -> dict:
    """Selection groups of a MolFlow XML document (Interface/Selections):
    name -> 0-based facet indices."""
    out = {}
    intf = root.find("Interface") if root is not None else None
    sel = intf.find("Selections") if intf is not None else None
    for s in (sel.findall("Selection") if sel is not None else []):
        out[s.attrib.get("name", "")] = [int(float(i.attrib["facet"])) for i in s.findall("selItem")]
    return out


def apply(model, changes: ParamChanges, log=None) -> int:
    """Apply parsed changes to a Model (ChangeSimuParams,
    ChangeFacetParams); returns the number of invalid values."""
    log = log or (lambda msg: print(msg, file=sys.stderr))
    for m in changes.messages:
        log(f"[ParameterChange] {m}")
    s = model.settings
    for name, v in changes.simulation:
        if name == "enableDecay":
            s.enable_decay = int(v > 0.0)
        elif name == "mass":
            s.gas_mass = v
        elif name == "halfLife":
            s.half_life = v
    F = model.facets
    areas = None
    errors = 0
    for f, name, v in changes.facet:
        if f >= model.n_facets:
            log(f"[ParameterChange][Facet][ID: {f + 1}] Facet ID out of range")
            errors += 1
            continue
        if name == "opacity":
            F["opacity"][f] = v
            if v < 0.0 or v > 1.0:
                errors += 1
                log(f"[ParameterChange][Facet][ID: {f}] Invalid opacity on facet: {v}")
        elif name in ("outgassing", "specificOutgassing"):
            if name == "outgassing":
                F["outgassing"][f] = v * MBARLS_TO_PAM3S
            else:
                if areas is None:
                    areas = model.facet_areas()
                F["outgassing"][f] = v * MBARLS_TO_PAM3S * areas[f]
            if F["outgassing"][f] > 0.0 and F["desorb_type"][f] == A.DES_NONE:
                F["desorb_type"][f] = A.DES_COSINE
        elif name == "sticking":
            F["sticking"][f] = v
            if v < 0.0 or v > 1.0:
                errors += 1
                log(f"[ParameterChange][Facet][ID: {f + 1}] Invalid sticking coefficient on facet: {v}")
        elif name == "temperature":
            F["temperature"][f] = v
    return errors
