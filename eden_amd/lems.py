"""Generic LEMS synapse mechanisms: a restatement of the NeuroML core synapse ComponentTypes,
a LEMS expression parser, and an emitter of their per-step update as C / HIP functions.

The reference turns every LEMS ComponentType it instantiates into C text (DescribeLems,
eden/Eden.cpp:2339-2995), compiles it with gcc and loads it with dlopen
(eden/Eden.cpp:6992-7216). Here the same per-step semantics are emitted as device functions
that the library compiles with hiprtc into a kernel template at context creation
(eden_amd/csrc/lems_kernel.hip, include/edenhip.h `edenhip_lems_synapses`); the oracle
compiles the C dialect of the same text with gcc.

Semantics of one step of one instance (eden/Eden.cpp:2700-2995):
  - parameters, derived parameters (from parameters), the `weight` property;
  - derived variables from the present state, in dependency order; a conditional derived
    variable takes the value of its first case whose condition holds, else its default case,
    else 0 (eden/Eden.cpp:2749-2776);
  - on an initialisation step: the OnStart assignments in order;
  - otherwise: next = state + dt * derivative for every state variable with a time
    derivative (each derivative from the present state), next = state for the others; then
    the OnCondition blocks in order, each whose test holds running its assignments, then the
    OnEvent assignments when the instance receives its event this step;
  - an assignment sets the variable's present value and its next value (Emit_AssignState,
    eden/Eden.cpp:2865-2887), so later tests and assignments of the step read it, while
    derived variables keep the values computed at the top;
  - exposures i (current) and g (conductance; 0 for current-based synapses) last, a state
    variable's as assigned.
Expressions are printed fully parenthesised, left to right as parsed, with every literal
rounded to float, `expf`/`logf`/`powf` and the trigonometric and hyperbolic functions as
EXP/LOG/POW/SIN/... (the dialect's own: on the device and in the oracle, each rounded once
from its double evaluation), H as stepf (0 below zero, else 1, as
the reference defines it, eden/Eden.cpp:3197), so both compilers perform the same float
operations in the same order. Engine units (us, mV, nA, uS) make every unit conversion
factor of these types 1 (the reference's Convert::Suffix then emits nothing).

The ComponentType table restates the dynamics of the NeuroML 2 core synapse types (the
NeuroML standard's definitions, as embedded by the reference in
eden/neuroml/LEMS_CoreComponents.inc.xml:7-372, 1834-1942); a blockingPlasticSynapse is
composed with its block and plasticity mechanisms into one type (blocking_plastic_type).
"""
import re
from dataclasses import dataclass, field

import numpy as np


class LemsError(ValueError):
    pass


# ---------------------------------------------------------------- component types

@dataclass
class SynapseType:
    """A synapse ComponentType: `params` (name, NeuroML dimension or None for plain numbers),
    `derived_params` (name, expression), `constants` (name, engine value), `state` (names),
    `derived` (name, expression, or name, [(condition or None for the default, expression),
    ...] for a conditional one) in dependency order, `derivatives` (state name ->
    expression), `on_start` and `on_event` (state name -> expression, in order),
    `on_conditions` [(test, [(state name, expression), ...]), ...], exposures `i` and
    (optionally) `g` (names of derived or state variables)."""
    params: list
    state: list
    derived: list
    derivatives: dict
    on_start: dict
    on_event: dict
    expose_i: str
    expose_g: str = None
    derived_params: list = field(default_factory=list)
    constants: list = field(default_factory=list)
    on_conditions: list = field(default_factory=list)
    is_input: bool = False     # an input (its current joins the compartment's input total)


_COND = [("gbase", "conductance"), ("erev", "voltage")]
_PYNN = [("tau_syn", None)]
_PYNN_CONST = [("MVOLT", 1.0), ("MSEC", 1000.0), ("NAMP", 1.0)]   # 1 mV, 1 ms, 1 nA in engine units

SYNAPSE_TYPES = {
    # the reference runs expOneSynapse natively, exposing the conductance of the step's start
    # also on an event step (eden/Eden.cpp:3424-3460): hence the exposure through a derived
    # variable, not the state variable an event assignment changes
    "expOneSynapse": SynapseType(
        params=_COND + [("tauDecay", "time")], state=["g"], derived=[("i", "g * (erev - v)"), ("g0", "g")],
        derivatives={"g": "-g / tauDecay"}, on_start={"g": "0"}, on_event={"g": "g + (gbase * weight)"},
        expose_i="i", expose_g="g0"),
    "alphaCurrentSynapse": SynapseType(
        params=[("tau", "time"), ("ibase", "current")], state=["I", "J"], derived=[("i", "I")],
        derivatives={"I": "(2.7182818284590451*J - I)/tau", "J": "-J/tau"},
        on_start={"I": "0", "J": "0"}, on_event={"J": "J + weight * ibase"}, expose_i="i"),
    "alphaSynapse": SynapseType(
        params=_COND + [("tau", "time")], state=["g", "A"], derived=[("i", "g * (erev - v)")],
        derivatives={"g": "(2.7182818284590451 * A - g)/tau", "A": "-A / tau"},
        on_start={"g": "0", "A": "0"}, on_event={"A": "A + (gbase*weight)"}, expose_i="i", expose_g="g"),
    "expTwoSynapse": SynapseType(
        params=_COND + [("tauRise", "time"), ("tauDecay", "time")],
        derived_params=[("peakTime", "log(tauDecay / tauRise) * (tauRise * tauDecay)/(tauDecay - tauRise)"),
                        ("waveformFactor", "1 / (-exp(-peakTime / tauRise) + exp(-peakTime / tauDecay))")],
        state=["A", "B"], derived=[("g", "gbase * (B - A)"), ("i", "g * (erev - v)")],
        derivatives={"A": "-A / tauRise", "B": "-B / tauDecay"}, on_start={"A": "0", "B": "0"},
        on_event={"A": "A + (weight * waveformFactor)", "B": "B + (weight * waveformFactor)"},
        expose_i="i", expose_g="g"),
    "expThreeSynapse": SynapseType(
        params=[("gbase1", "conductance"), ("gbase2", "conductance"), ("erev", "voltage"), ("tauRise", "time"),
                ("tauDecay1", "time"), ("tauDecay2", "time")],
        derived_params=[
            ("peakTime1", "log(tauDecay1 / tauRise) * (tauRise * tauDecay1)/(tauDecay1 - tauRise)"),
            ("waveformFactor1", "1 / (-exp(-peakTime1 / tauRise) + exp(-peakTime1 / tauDecay1))"),
            ("peakTime2", "log(tauDecay2 / tauRise) * (tauRise * tauDecay2)/(tauDecay2 - tauRise)"),
            ("waveformFactor2", "1 / (-exp(-peakTime2 / tauRise) + exp(-peakTime2 / tauDecay2))")],
        state=["A", "B", "C"], derived=[("g", "gbase1*(B - A) + gbase2*(C-A)"), ("i", "g * (erev - v)")],
        derivatives={"A": "-A / tauRise", "B": "-B / tauDecay1", "C": "-C / tauDecay2"},
        on_start={"A": "0", "B": "0", "C": "0"},
        on_event={"A": "A + (gbase1*weight * waveformFactor1 + gbase2*weight*waveformFactor2 )/(gbase1+gbase2)",
                  "B": "B + (weight * waveformFactor1)", "C": "C + (weight * waveformFactor2)"},
        expose_i="i", expose_g="g"),
    "expCondSynapse": SynapseType(
        params=_PYNN + [("e_rev", None)], constants=_PYNN_CONST, state=["g"],
        derived=[("i", "g * (e_rev - (v/MVOLT)) * NAMP")], derivatives={"g": "-g / (tau_syn*MSEC)"},
        on_start={}, on_event={"g": "g+weight"}, expose_i="i"),
    "expCurrSynapse": SynapseType(
        params=_PYNN, constants=_PYNN_CONST, state=["I"], derived=[("i", "I * NAMP")],
        derivatives={"I": "-I / (tau_syn*MSEC)"}, on_start={}, on_event={"I": "I + weight"}, expose_i="i"),
    "alphaCondSynapse": SynapseType(
        params=_PYNN + [("e_rev", None)], constants=_PYNN_CONST, state=["g", "A"],
        derived=[("i", "g * (e_rev - (v/MVOLT)) * NAMP")],
        derivatives={"g": "(2.7182818*A - g)/(tau_syn*MSEC)", "A": "-A /(tau_syn*MSEC)"},
        on_start={}, on_event={"A": "A + weight"}, expose_i="i"),
    "alphaCurrSynapse": SynapseType(
        params=_PYNN, constants=_PYNN_CONST, state=["I", "A"], derived=[("i", "I * NAMP")],
        derivatives={"I": "(2.7182818*A - I)/(tau_syn*MSEC)", "A": "-A /(tau_syn*MSEC)"},
        on_start={}, on_event={"A": "A + weight"}, expose_i="i"),
    # a linear gap junction: current conductance (vpeer - v), no conductance exposure (the
    # reference's Exposure_g = Gsyn * 0, eden/Eden.cpp:3407-3421), vpeer the voltage of the
    # compartment at the other end at the start of the step
    "gapJunction": SynapseType(
        params=[("conductance", "conductance")], state=[], derived=[("i", "conductance * (vpeer - v)")],
        derivatives={}, on_start={}, on_event={}, expose_i="i"),
    # the two ends of a continuous (graded) connection; vpeer the voltage at the other end
    "silentSynapse": SynapseType(
        params=[], state=[], derived=[("i", "0")], derivatives={}, on_start={}, on_event={}, expose_i="i"),
    "linearGradedSynapse": SynapseType(
        params=[("conductance", "conductance")], state=[], derived=[("i", "conductance * (vpeer - v)")],
        derivatives={}, on_start={}, on_event={}, expose_i="i"),
    "gradedSynapse": SynapseType(
        params=[("conductance", "conductance"), ("delta", "voltage"), ("k", "frequency"), ("Vth", "voltage"),
                ("erev", "voltage")],
        state=["s"],
        derived=[("inf", "1/(1 + exp((Vth - vpeer)/delta))"), ("tau", "(1-inf)/k"),
                 ("i", "conductance * s * (erev-v)"),
                 ("s_rate", [("(1-inf) .gt. 1e-4", "(inf - s)/tau"), (None, "0")])],
        derivatives={"s": "s_rate"}, on_start={"s": "0"}, on_event={},
        on_conditions=[("(1-inf) .lt. 1e-4", [("s", "inf")])], expose_i="i"),
}

# LEMS current inputs (LEMS_CoreComponents.inc.xml:402-622; the pulseGenerator and
# poissonFiringSynapse are native): their current, or dimensionless current I, joins the
# compartment's input total as i * weight per instance (eden/Eden.cpp:3914-3931)
_WINDOW = [("t .lt. delay", "0"), ("t .geq. delay .and. t .lt. duration+delay", None), ("t .geq. duration+delay", "0")]


def _windowed(var, active, params):
    conds = [(test, [(var, active if val is None else val)]) for test, val in _WINDOW]
    return SynapseType(params=params, state=[var], derived=[], derivatives={}, on_start={}, on_event={},
                       on_conditions=conds, expose_i=var, is_input=True)


def _ramp(var, dim):
    base, ramp = "weight * baselineAmplitude", \
        "weight * (startAmplitude + (finishAmplitude - startAmplitude) * (t - delay) / (duration))"
    return SynapseType(
        params=[("delay", "time"), ("duration", "time"), ("startAmplitude", dim), ("finishAmplitude", dim),
                ("baselineAmplitude", dim)],
        state=[var], derived=[], derivatives={}, on_start={var: "baselineAmplitude"}, on_event={},
        on_conditions=[("t .lt. delay", [(var, base)]), ("t .geq. delay .and. t .lt. duration+delay", [(var, ramp)]),
                       ("t .geq. duration+delay", [(var, base)])],
        expose_i=var, is_input=True)


_SINE = "weight * amplitude * sin(phase + (2 * 3.14159265 * (t-delay)/period) )"
_CLAMP = [("delay", "time"), ("duration", "time"), ("simpleSeriesResistance", "resistance")]
INPUT_TYPES = {
    "pulseGeneratorDL": _windowed("I", "weight * amplitude", [("delay", "time"), ("duration", "time"),
                                                             ("amplitude", None)]),
    "sineGenerator": _windowed("i", _SINE, [("phase", None), ("delay", "time"), ("duration", "time"),
                                            ("amplitude", "current"), ("period", "time")]),
    "sineGeneratorDL": _windowed("I", _SINE, [("phase", None), ("delay", "time"), ("duration", "time"),
                                              ("amplitude", None), ("period", "time")]),
    "rampGenerator": _ramp("i", "current"),
    "rampGeneratorDL": _ramp("I", None),
    "voltageClamp": SynapseType(
        params=_CLAMP + [("targetVoltage", "voltage")], state=["i"], derived=[], derivatives={}, on_start={},
        on_event={}, on_conditions=[("t .lt. delay", [("i", "0")]),
                                    ("t .geq. delay", [("i", "weight * (targetVoltage - v) / simpleSeriesResistance")]),
                                    ("t .gt. duration + delay", [("i", "0")])],
        expose_i="i", is_input=True),
    "voltageClampTriple": SynapseType(
        params=_CLAMP + [("active", None), ("conditioningVoltage", "voltage"), ("testingVoltage", "voltage"),
                         ("returnVoltage", "voltage")],
        state=["i"], derived=[], derivatives={}, on_start={}, on_event={},
        on_conditions=[("active .eq. 1 .and. t .lt. delay",
                        [("i", "weight * (conditioningVoltage - v) / simpleSeriesResistance")]),
                       ("active .eq. 1 .and. t .geq. delay",
                        [("i", "weight * (testingVoltage - v) / simpleSeriesResistance")]),
                       ("active .eq. 1 .and. t .gt. duration + delay",
                        [("i", "weight * (returnVoltage - v) / simpleSeriesResistance")])],
        expose_i="i", is_input=True),
}
# the NeuroML base input types a user input ComponentType may extend
BASE_INPUTS = {"basePointCurrent": [], "basePointCurrentDL": [], "baseVoltageDepPointCurrent": [],
               "baseVoltageDepPointCurrentDL": []}

# blockingPlasticSynapse's child mechanisms (LEMS_CoreComponents.inc.xml:231-312): parameters,
# state, derived variables (exposing blockFactor / plasticityFactor), derivatives, OnStart and
# OnEvent (port in: the synapse's own event)
BLOCK_MECHANISMS = {
    "voltageConcDepBlockMechanism": dict(
        params=[("blockConcentration", "concentration"), ("scalingConc", "concentration"), ("scalingVolt", "voltage")],
        derived=[("blockFactor", "1/(1 + (blockConcentration / scalingConc)* exp(-1 * (v / scalingVolt)))")]),
}
PLASTICITY_MECHANISMS = {
    "tsodyksMarkramDepMechanism": dict(
        params=[("initReleaseProb", None), ("tauRec", "time")], state=["R"],
        derived=[("U", "initReleaseProb"), ("plasticityFactor", "R * U")], derivatives={"R": "(1 - R) / tauRec"},
        on_start={"R": "1"}, on_event={"R": "R * (1 - U)"}),
    "tsodyksMarkramDepFacMechanism": dict(
        params=[("initReleaseProb", None), ("tauRec", "time"), ("tauFac", "time")], state=["R", "U"],
        derived=[("plasticityFactor", "R * U")],
        derivatives={"R": "(1 - R) / tauRec", "U": "(initReleaseProb - U) / tauFac"},
        on_start={"R": "1", "U": "initReleaseProb"},
        on_event={"R": "R * (1 - U)", "U": "U + initReleaseProb * (1 - U)"}),
}


def blocking_plastic_type(block=None, plasticity=None):
    """One SynapseType for a blockingPlasticSynapse with the given block and plasticity
    mechanism types (None: absent, factor 1). The reference runs the three components one
    after the other for each instance, block first, then plasticity, then the synapse, which
    reads the mechanisms' exposures as computed from their present state
    (eden/Eden.cpp:3491-3532, LEMS_CoreComponents.inc.xml:315-372); their variables do not
    share names, so the composition is one namespace in that order."""
    b = BLOCK_MECHANISMS[block] if block else dict(params=[], derived=[("blockFactor", "1")])
    p = PLASTICITY_MECHANISMS[plasticity] if plasticity else dict(
        params=[], state=[], derived=[("plasticityFactor", "1")], derivatives={}, on_start={}, on_event={})
    return SynapseType(
        params=_COND + [("tauRise", "time"), ("tauDecay", "time")] + b["params"] + p["params"],
        derived_params=SYNAPSE_TYPES["expTwoSynapse"].derived_params,
        state=p["state"] + ["A", "B"],
        derived=b["derived"] + p["derived"] + [("g", "blockFactor * gbase * (B - A)"), ("i", "g * (erev - v)")],
        derivatives={**p["derivatives"], "A": "-A / tauRise", "B": "-B / tauDecay"},
        on_start={**p["on_start"], "A": "0", "B": "0"},
        on_event={**p["on_event"], "A": "A + (weight * plasticityFactor * waveformFactor)",
                  "B": "B + (weight * plasticityFactor * waveformFactor)"},
        expose_i="i", expose_g="g")

# PyNN synapses' numbers are plain (ms, mV); their base type gives tau_syn no default here

# ---------------------------------------------------------------- user ComponentTypes

# the NeuroML 2 base synapse types (its Synapses.xml) and the parameters they contribute
BASE_SYNAPSES = {
    "baseSynapse": [], "baseVoltageDepSynapse": [], "baseCurrentBasedSynapse": [], "baseGradedSynapse": [],
    "baseConductanceBasedSynapse": [("gbase", "conductance"), ("erev", "voltage")],
    "baseConductanceBasedSynapseTwo": [("gbase1", "conductance"), ("gbase2", "conductance"), ("erev", "voltage")],
}
# LEMS dimension names -> the reader's unit dimensions (None: dimensionless)
LEMS_DIMS = {"time": "time", "per_time": "frequency", "voltage": "voltage", "current": "current",
             "conductance": "conductance", "capacitance": "capacitance", "length": "length",
             "concentration": "concentration", "conductance_per_voltage": "conductance_per_voltage",
             "voltage_per_time": "voltage_per_time", "currentDensity": "currentDensity", "per_voltage": "per_voltage",
             "resistance": "resistance", "conductanceDensity": "conductivity", "temperature": "temperature",
             "specificCapacitance": "specific_capacitance",
             "none": None, "": None}


def _ltag(el):
    t = el.tag
    return t.rsplit("}", 1)[-1] if isinstance(t, str) else ""


def _base_of(name, component_types):
    seen = set()
    while name in component_types and name not in seen:
        seen.add(name)
        name = component_types[name].get("extends", "")
    return name


def is_synapse_component_type(name, component_types):
    """Whether a ComponentType (by name) extends, through any chain, a NeuroML base synapse."""
    return _base_of(name, component_types) in BASE_SYNAPSES


def is_input_component_type(name, component_types):
    """Whether a ComponentType (by name) extends, through any chain, a NeuroML base current input."""
    return _base_of(name, component_types) in BASE_INPUTS


def synapse_type_from_lems(name, component_types, convert):
    """SynapseType of a user-defined synapse <ComponentType> (the reference turns any lemsified
    ComponentType into C, eden/Eden.cpp:2339-2995): Parameters and DerivedParameters, Constants
    (converted to engine units by `convert(text, dimension)`), the `weight` Property, and the
    nearest Dynamics along the extends chain: StateVariables, DerivedVariables and
    ConditionalDerivedVariables (ordered by their dependencies), TimeDerivatives, OnStart,
    OnCondition and OnEvent (port `in`) StateAssignments; exposures `i` and `g`. Returns
    (SynapseType, default weight)."""
    chain, n = [], name
    while n in component_types:
        chain.append(component_types[n])
        n = component_types[n].get("extends", "")
    if n not in BASE_SYNAPSES and n not in BASE_INPUTS:
        raise LemsError(f"ComponentType {name} does not extend a base synapse or input type")
    is_input = n in BASE_INPUTS
    params = list(BASE_SYNAPSES.get(n, BASE_INPUTS.get(n, [])))
    dparams, consts, weight, dyn = [], [], 1.0, None
    for ct in reversed(chain):                     # base first, so derived types override
        for el in ct:
            t = _ltag(el)
            if t == "Parameter":
                dim = el.get("dimension", "none")
                if dim not in LEMS_DIMS:
                    raise LemsError(f"{name}: dimension '{dim}' of parameter {el.get('name')} is not supported")
                params = [p for p in params if p[0] != el.get("name")] + [(el.get("name"), LEMS_DIMS[dim])]
            elif t == "DerivedParameter":
                dparams.append((el.get("name"), el.get("value")))
            elif t == "Constant":
                dim = LEMS_DIMS.get(el.get("dimension", "none"), "unsupported")
                if dim == "unsupported":
                    raise LemsError(f"{name}: dimension of constant {el.get('name')} is not supported")
                consts.append((el.get("name"), convert(el.get("value"), dim)))
            elif t == "Property" and el.get("name") == "weight":
                weight = float(el.get("defaultValue", "1"))
            elif t == "Dynamics":
                dyn = el
    if dyn is None:
        raise LemsError(f"ComponentType {name} has no Dynamics")
    state, derived, ders, on_start, on_event, expose, on_cond = [], {}, {}, {}, {}, {}, []
    for el in dyn:
        t = _ltag(el)
        if t == "StateVariable":
            state.append(el.get("name"))
            if el.get("exposure"):
                expose[el.get("exposure")] = el.get("name")
        elif t == "DerivedVariable":
            if el.get("value") is None:
                raise LemsError(f"{name}: derived variable {el.get('name')} without a value (select) is not supported")
            derived[el.get("name")] = el.get("value")
            if el.get("exposure"):
                expose[el.get("exposure")] = el.get("name")
        elif t == "ConditionalDerivedVariable":
            cases = [(c.get("condition"), c.get("value")) for c in el if _ltag(c) == "Case"]
            if not cases:
                raise LemsError(f"{name}: conditional derived variable {el.get('name')} has no cases")
            derived[el.get("name")] = cases
            if el.get("exposure"):
                expose[el.get("exposure")] = el.get("name")
        elif t == "TimeDerivative":
            ders[el.get("variable")] = el.get("value")
        elif t == "OnCondition":
            assigns = []
            for a in el:
                if _ltag(a) == "StateAssignment":
                    assigns.append((a.get("variable"), a.get("value")))
                else:
                    raise LemsError(f"{name}: <{_ltag(a)}> in <OnCondition> is not supported in a synapse")
            on_cond.append((el.get("test"), assigns))
        elif t in ("OnStart", "OnEvent"):
            if t == "OnEvent" and el.get("port", "in") != "in":
                raise LemsError(f"{name}: OnEvent port '{el.get('port')}' is not supported")
            tgt = on_start if t == "OnStart" else on_event
            for a in el:
                if _ltag(a) == "StateAssignment":
                    tgt[a.get("variable")] = a.get("value")
                elif _ltag(a) != "EventOut":
                    raise LemsError(f"{name}: <{_ltag(a)}> in <{t}> is not supported")
        elif t in ("Regime", "KineticScheme"):
            raise LemsError(f"{name}: <{t}> is not supported")
    if is_input and "i" not in expose and "I" in expose:
        expose["i"] = expose["I"]          # a dimensionless input's current
    if "i" not in expose:
        raise LemsError(f"ComponentType {name} exposes no current i")

    def syms(expr):
        texts = [expr] if isinstance(expr, str) else [x for case in expr for x in case if x is not None]
        out, stack = set(), [parse(x) for x in texts]
        while stack:
            x = stack.pop()
            if x[0] == "sym":
                out.add(x[1])
            elif x[0] != "num":
                stack.extend(y for y in x[1:] if isinstance(y, tuple))
        return out

    order, done = [], set()

    def visit(v, path=()):
        if v in done:
            return
        if v in path:
            raise LemsError(f"{name}: derived variables depend on each other in a cycle ({v})")
        for u in syms(derived[v]):
            if u in derived:
                visit(u, path + (v,))
        done.add(v)
        order.append((v, derived[v]))

    for v in derived:
        visit(v)
    st = SynapseType(params=params, state=state, derived=order, derivatives=ders, on_start=on_start,
                     on_event=on_event, expose_i=expose["i"], expose_g=expose.get("g"), derived_params=dparams,
                     constants=consts, on_conditions=on_cond, is_input=is_input)
    return st, weight


# ---------------------------------------------------------------- expressions

_TOKEN = re.compile(r"\s*(?:(\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)|([A-Za-z_]\w*)|"
                    r"(\*\*|[-+*/^(),]|\.(?:gt|lt|geq|leq|ge|le|eq|neq|ne|and|or|not)\.))")
# LEMS relational operators and their C forms
_REL = {".gt.": ">", ".lt.": "<", ".geq.": ">=", ".ge.": ">=", ".leq.": "<=", ".le.": "<=", ".eq.": "==",
        ".neq.": "!=", ".ne.": "!="}
_FUNCS = {"exp": "EXP", "log": "LOG", "ln": "LOG", "sqrt": "sqrtf", "abs": "fabsf", "sin": "SIN", "cos": "COS",
          "tan": "TAN", "sinh": "SINH", "cosh": "COSH", "tanh": "TANH", "ceil": "ceilf", "floor": "floorf",
          "H": "stepf", "random": "random"}


def _tokens(text):
    pos, out = 0, []
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise LemsError(f"cannot read expression '{text}' at '{text[pos:]}'")
        num, name, op = m.groups()
        out.append(("num", num) if num else ("name", name) if name else ("op", op))
        pos = m.end()
    return out


def parse(text):
    """Expression tree: ('num', value) | ('sym', name) | ('neg', x) | (op, l, r) | ('fn', f, x)
    | ('cmp', c_op, l, r) | ('and', l, r) | ('or', l, r) | ('not', x), with the usual
    precedence: unary minus, ^ (right-associative), * /, + - (left), relations (.gt. .lt.
    .geq. .leq. .eq. .neq.), .not., .and., .or. (left)."""
    toks = _tokens(text)
    pos = [0]

    def peek():
        return toks[pos[0]] if pos[0] < len(toks) else ("end", None)

    def take():
        t = peek()
        pos[0] += 1
        return t

    def primary():
        t = take()
        if t[0] == "num":
            return ("num", float(t[1]))
        if t[0] == "name":
            if peek() == ("op", "("):
                take()
                arg = logical_or()
                if take() != ("op", ")"):
                    raise LemsError(f"missing ')' in '{text}'")
                if t[1] not in _FUNCS:
                    raise LemsError(f"function {t[1]} is not supported")
                return ("fn", t[1], arg)
            return ("sym", t[1])
        if t == ("op", "("):
            e = logical_or()
            if take() != ("op", ")"):
                raise LemsError(f"missing ')' in '{text}'")
            return e
        raise LemsError(f"unexpected token {t[1]!r} in '{text}'")

    def unary():
        if peek() == ("op", "-"):
            take()
            return ("neg", unary())
        if peek() == ("op", "+"):
            take()
            return unary()
        return power()

    def power():
        b = primary()
        if peek() in (("op", "^"), ("op", "**")):
            take()
            return ("pow", b, unary())
        return b

    def term():
        e = unary()
        while peek() in (("op", "*"), ("op", "/")):
            op = take()[1]
            e = ("*" if op == "*" else "/", e, unary())
        return e

    def additive():
        e = term()
        while peek() in (("op", "+"), ("op", "-")):
            op = take()[1]
            e = (op, e, term())
        return e

    def relation():
        e = additive()
        if peek()[0] == "op" and peek()[1] in _REL:
            op = _REL[take()[1]]
            e = ("cmp", op, e, additive())
        return e

    def negation():
        if peek() == ("op", ".not."):
            take()
            return ("not", negation())
        return relation()

    def logical_and():
        e = negation()
        while peek() == ("op", ".and."):
            take()
            e = ("and", e, negation())
        return e

    def logical_or():
        e = logical_and()
        while peek() == ("op", ".or."):
            take()
            e = ("or", e, logical_and())
        return e

    tree = logical_or()
    if pos[0] != len(toks):
        raise LemsError(f"trailing tokens in '{text}'")
    return tree


def _lit(v):
    return "( (float)%s)" % repr(float(v))


def infix(tree, names):
    """Fully parenthesised float C text of an expression tree; `names` maps LEMS symbols to
    C variable names."""
    k = tree[0]
    if k == "num":
        return _lit(tree[1])
    if k == "sym":
        if tree[1] not in names:
            raise LemsError(f"unknown symbol {tree[1]}")
        return names[tree[1]]
    if k == "neg":
        return "( -" + infix(tree[1], names) + " )"
    if k == "fn":
        if tree[1] == "random":
            raise LemsError("random() in synapse and input types is not supported")
        return _FUNCS[tree[1]] + "( " + infix(tree[2], names) + " )"
    if k == "pow":
        return "POW( " + infix(tree[1], names) + " , " + infix(tree[2], names) + " )"
    if k == "cmp":
        return "( ( " + infix(tree[2], names) + " " + tree[1] + " " + infix(tree[3], names) + " ) )"
    if k == "not":
        return "( !" + infix(tree[1], names) + " )"
    if k in ("and", "or"):
        op = "&&" if k == "and" else "||"
        return "( ( " + infix(tree[1], names) + " " + op + " " + infix(tree[2], names) + " ) )"
    return "( ( " + infix(tree[1], names) + " " + k + " " + infix(tree[2], names) + " ) )"


_NP_FUNCS = {"exp": np.exp, "log": np.log, "ln": np.log, "sqrt": np.sqrt, "abs": np.abs, "sin": np.sin,
             "cos": np.cos, "tan": np.tan, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh, "ceil": np.ceil,
             "floor": np.floor, "H": lambda x: np.float32(0.0 if x < 0 else 1.0)}


def evaluate(tree, values):
    """Value of an expression tree in float arithmetic (literals and symbols rounded to float,
    every operation in float), the evaluation the reader needs at instantiation time (e.g.
    inhomogeneous channel densities); `values` maps symbols to numbers."""
    F = np.float32
    k = tree[0]
    if k == "num":
        return F(tree[1])
    if k == "sym":
        if tree[1] not in values:
            raise LemsError(f"unknown symbol {tree[1]}")
        return F(values[tree[1]])
    if k == "neg":
        return F(-evaluate(tree[1], values))
    if k == "fn":
        if tree[1] not in _NP_FUNCS:
            raise LemsError(f"{tree[1]}() cannot be evaluated at setup")
        return F(_NP_FUNCS[tree[1]](evaluate(tree[2], values)))
    if k == "cmp":
        a, b = evaluate(tree[2], values), evaluate(tree[3], values)
        return F({">": a > b, "<": a < b, ">=": a >= b, "<=": a <= b, "==": a == b, "!=": a != b}[tree[1]])
    if k == "not":
        return F(not evaluate(tree[1], values))
    if k in ("and", "or"):
        a, b = bool(evaluate(tree[1], values)), bool(evaluate(tree[2], values))
        return F((a and b) if k == "and" else (a or b))
    a, b = evaluate(tree[1], values), evaluate(tree[2], values)
    with np.errstate(all="ignore"):
        if k == "pow":
            return F(np.power(a, b))
        return F({"+": a + b, "-": a - b, "*": a * b, "/": a / b}[k])


# ---------------------------------------------------------------- emitter

def _emit_derived(L, st, names, indent="    "):
    """Derived variables in order: plain ones as constants, conditional ones as an if / else
    if chain over their cases with the default case last, 0 when none applies
    (eden/Eden.cpp:2739-2776)."""
    for d, expr in st.derived:
        if isinstance(expr, str):
            L.append(f"{indent}const float d_{d} = {infix(parse(expr), names)};")
        else:
            L.append(f"{indent}float d_{d} = 0.0f;")
            L.append(f"{indent}if (0) {{ }}")
            for cond, val in expr:
                if cond is not None:
                    L.append(f"{indent}else if ({infix(parse(cond), names)}) {{ d_{d} = {infix(parse(val), names)}; }}")
            default = [val for cond, val in expr if cond is None]
            if default:
                L.append(f"{indent}else {{ d_{d} = {infix(parse(default[0]), names)}; }}")
        names[d] = f"d_{d}"


def _emit_assign(L, assigns, names, indent):
    """State assignments in order, each setting the present and the next value (later
    expressions of the step read the assigned value)."""
    for s_, expr in assigns:
        if f"n_{s_}" not in {f"n_{x}" for x in names.get("__state__", ())}:
            raise LemsError(f"assignment to {s_}, which is not a state variable")
        L.append(f"{indent}s_{s_} = {infix(parse(expr), names)}; n_{s_} = s_{s_};")


def emit(name, st, fn_name, dialect):
    """C / HIP text of one type's step function:

        void fn(const float *P, const float *S, float *Sn, long stride, float v, float vpeer,
                float t, int init, int event, float dt, float *i_out, float *g_out)

    P: the type's parameter vector (params in table order, then the `weight` property); vpeer:
    the voltage at the other end of a gap junction or graded connection (v for other types); t:
    the step's time (time_f32);
    S / Sn: the instance's present / next state, variable k at k * stride (they may alias:
    every present value is read before anything is written)."""
    q = "__device__ __forceinline__ " if dialect == "hip" else "static "
    L = [f"{q}void {fn_name}(const float *P, const float *S, float *Sn, long stride, float v, float vpeer, float t,"
         f" int init, int event, float dt, float *i_out, float *g_out) {{",
         f"    /* {name} */",
         "    (void)S; (void)Sn; (void)stride; (void)vpeer; (void)t; (void)init; (void)event; (void)dt;"]
    names = {"v": "v", "vpeer": "vpeer", "dt": "dt", "t": "t"}
    for k, (p, _) in enumerate(st.params):
        L.append(f"    const float p_{p} = P[{k}];")
        names[p] = f"p_{p}"
    L.append(f"    const float p_weight = P[{len(st.params)}];")
    names["weight"] = "p_weight"
    for c, val in st.constants:
        L.append(f"    const float c_{c} = {_lit(val)};")
        names[c] = f"c_{c}"
    for d, expr in st.derived_params:
        L.append(f"    const float dp_{d} = {infix(parse(expr), names)};")
        names[d] = f"dp_{d}"
    for k, s_ in enumerate(st.state):
        L.append(f"    float s_{s_} = S[{k} * stride];")
        names[s_] = f"s_{s_}"
    names["__state__"] = tuple(st.state)
    _emit_derived(L, st, names)
    for s_ in st.state:
        L.append(f"    float n_{s_} = s_{s_};")
    L.append("    if (init) {")
    _emit_assign(L, st.on_start.items(), names, "        ")
    L.append("    } else {")
    for k, s_ in enumerate(st.state):
        if s_ in st.derivatives:
            L.append(f"        const float der_{k} = {infix(parse(st.derivatives[s_]), names)};")
            L.append(f"        n_{s_} = s_{s_} + dt * der_{k};")
    for test, assigns in st.on_conditions:
        L.append(f"        if ({infix(parse(test), names)}) {{")
        _emit_assign(L, assigns, names, "            ")
        L.append("        }")
    L.append("        if (event) {")
    _emit_assign(L, st.on_event.items(), names, "            ")
    L.append("        }")
    L.append("    }")
    for k, s_ in enumerate(st.state):
        L.append(f"    Sn[{k} * stride] = n_{s_};")
    L.append(f"    *i_out = {names[st.expose_i]};")
    L.append(f"    *g_out = {names[st.expose_g] if st.expose_g else '0.0f'};")
    L.append("}")
    return "\n".join(L)


@dataclass
class LemsSynapses:
    """The generic synapse types of a model: per type its component type name, parameter
    vector and state size, and the emitted source in both dialects."""
    type_names: list
    params: list          # per type: np.float32 vector
    n_state: list
    source_hip: str
    source_c: str
    type_flags: list = field(default_factory=list)   # per type: bit 0 an input type


def build(types, custom=None):
    """types: list of (component type name, {param: engine value}, weight property); custom:
    user ComponentTypes by name (SynapseType). Returns LemsSynapses whose functions are
    lems_syn_<k> and a dispatcher lems_syn(type, ...)."""
    hip, c, params, nst, names, flags = [], [], [], [], [], []
    custom = custom or {}
    for k, (tname, vals, weight) in enumerate(types):
        if tname not in SYNAPSE_TYPES and tname not in INPUT_TYPES and tname not in custom:
            raise LemsError(f"synapse type {tname} is not supported")
        st = custom[tname] if tname in custom else SYNAPSE_TYPES.get(tname) or INPUT_TYPES[tname]
        flags.append(1 if st.is_input else 0)
        missing = [p for p, _ in st.params if p not in vals]
        if missing:
            raise LemsError(f"{tname}: parameters {missing} missing")
        params.append(np.array([vals[p] for p, _ in st.params] + [weight], np.float32))
        nst.append(len(st.state))
        names.append(tname)
        hip.append(emit(tname, st, f"lems_syn_{k}", "hip"))
        c.append(emit(tname, st, f"lems_syn_{k}", "c"))

    def dispatcher(q):
        L = [f"{q}void lems_syn(int type, const float *P, const float *S, float *Sn, long stride, float v, float vpeer,"
             " float t,"
             " int init, int event, float dt, float *i_out, float *g_out) {", "    switch (type) {"]
        for k in range(len(types)):
            L.append(f"    case {k}: lems_syn_{k}(P, S, Sn, stride, v, vpeer, t, init, event, dt, i_out, g_out); break;")
        L.append("    default: *i_out = 0.0f; *g_out = 0.0f; break;")
        L.append("    }")
        L.append("}")
        return "\n".join(L)

    src_hip = "\n\n".join(hip + [dispatcher("__device__ __forceinline__ ")]) + "\n"
    src_c = "\n\n".join(c + [dispatcher("")]) + "\n"
    return LemsSynapses(names, params, nst, src_hip, src_c, flags)
