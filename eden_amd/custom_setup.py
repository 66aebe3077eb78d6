"""EdenCustomSetup files: per-instance parameter settings of a simulation.

The reference reads them from <Simulation><EdenCustomSetup filename="..."> (the file name
relative to the file holding the Simulation element) with ParseSimulation_CustomSetupFile
(eden/NeuroML.cpp:7270-7980) and applies the statements, in order, to the instantiated model's
tables (eden/Eden.cpp:9437-9660). The format is line-oriented: tokens split on whitespace,
blank lines and lines whose first token starts with '#' skipped, and

    set cell <population> <ids | all> <segments> <parameter> <value | multi> [unit]
    set input <inputList> <ids | all> <parameter> <value | multi> [unit]
    set synapse <projection> <ids | all> <pre | post> <parameter> <value | multi> [unit]

where <ids> is a comma-separated list of population instance ids, input ids or connection
ids (each once), <segments> is "all" for point neurons, the unit follows the value when the
parameter has a dimension, and "multi" takes its values from the next line,
"values v1 v2 ...", one per target (every instance of the group for "all"), each in the
statement's unit.

This module tokenises and checks the grammar; eden_amd.neuroml resolves the statements
against the network and fills the per-instance parameter tables. The reference's cable modes
("cable", "multicable") and values that are references to other state variables are refused
(the reference itself leaves the cable modes unimplemented, eden/Eden.cpp:9561-9566).
"""
import re
from dataclasses import dataclass


class CustomSetupError(ValueError):
    pass


_INT = re.compile(r"^[-+]?\d+$")
_REAL = re.compile(r"^[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?$")


@dataclass
class Statement:
    kind: str            # "cell", "input" or "synapse"
    group: str           # population, input list or projection id
    targets: list        # instance / input / connection ids, or None for all
    where: str           # cell: the segment specifier; synapse: "pre" or "post"; input: ""
    param: str
    tokens: list         # the value and unit tokens after the parameter ("multi" included)
    values: list         # multi: the tokens of the values line; else None
    line: int            # 1-based line of the statement in the file

    @property
    def multi(self):
        return self.tokens[0] == "multi"

    def what(self):
        return f"EdenCustomSetup line {self.line} (set {self.kind} {self.group})"


def tokenize(text):
    """(line number, tokens) of each non-blank, non-comment line."""
    out = []
    for n, line in enumerate(text.split("\n"), 1):
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        out.append((n, tok))
    return out


def parse(text):
    """The statements of a custom setup file, in order."""
    lines = tokenize(text)
    out = []
    i = 0
    while i < len(lines):
        n, tok = lines[i]
        i += 1
        if tok[0] != "set":
            raise CustomSetupError(f"EdenCustomSetup line {n}: unknown statement type '{tok[0]}'")
        if len(tok) < 3:
            raise CustomSetupError(f"EdenCustomSetup line {n}: 'set' statement needs at least an item type and locator")
        kind = tok[1]
        need = {"cell": 7, "input": 6, "synapse": 7}.get(kind)
        if need is None:
            raise CustomSetupError(f"EdenCustomSetup line {n}: 'set' statement: unknown type '{kind}'")
        if len(tok) < need:
            raise CustomSetupError(f"EdenCustomSetup line {n}: 'set {kind}' statement is incomplete")
        if kind == "input":
            where, param, rest = "", tok[4], tok[5:]
        else:
            where, param, rest = tok[4], tok[5], tok[6:]
        targets = None
        if tok[3] != "all":
            targets = []
            for t in tok[3].split(","):
                if not _INT.match(t):
                    raise CustomSetupError(f"EdenCustomSetup line {n}: could not resolve id '{t}'")
                targets.append(int(t))
            if len(set(targets)) != len(targets):
                raise CustomSetupError(f"EdenCustomSetup line {n}: an id is specified more than once")
        if kind == "synapse" and where not in ("pre", "post"):
            raise CustomSetupError(f"EdenCustomSetup line {n}: unknown synapse location '{where}'")
        if rest[0] in ("cable", "multicable"):
            raise CustomSetupError(f"EdenCustomSetup line {n}: value specifier '{rest[0]}' is not supported")
        values = None
        if rest[0] == "multi":
            if i >= len(lines) or lines[i][1][0] != "values":
                raise CustomSetupError(f"EdenCustomSetup line {n}: 'set' line has no 'values' line following")
            values = lines[i][1][1:]
            for v in values:
                if not _REAL.match(v):
                    raise CustomSetupError(f"EdenCustomSetup line {lines[i][0]}: could not parse as number: {v}")
            i += 1
        elif not _REAL.match(rest[0]):
            raise CustomSetupError(f"EdenCustomSetup line {n}: could not parse '{rest[0]}' as a numerical value or "
                                   "'multi' (references to state variables are not supported)")
        out.append(Statement(kind, tok[2], targets, where, param, rest, values, n))
    return out


def values_of(st, n_targets, convert_one):
    """Engine values of a statement for its n_targets targets (in target order):
    convert_one(number text, unit or None) converts one value. A single value applies to every
    target; multi values must number n_targets."""
    unit = st.tokens[1] if len(st.tokens) > 1 else None
    if st.multi:
        if len(st.values) != n_targets:
            raise CustomSetupError(f"{st.what()}: {n_targets} values were expected, but {len(st.values)} were "
                                   "provided instead")
        return [convert_one(v, unit) for v in st.values]
    return [convert_one(st.tokens[0], unit)] * n_targets
