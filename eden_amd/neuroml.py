"""NeuroML 2 / LEMS reader for the GPU path: a LEMS or NeuroML file becomes the instantiated
model (`eden_amd.model.Model`) and the description of its output files.

This is the subset of EDEN's front end (ReadNeuroML, eden/NeuroML.cpp:9572-10068, and the
instantiation in GenerateModel, eden/Eden.cpp:677-10377) that the per-step loop's mechanism
set covers:

  includes          <Include file> (LEMS) and <include href> (NeuroML), resolved relative to
                    the including file; NeuroML's core type files are skipped
                    (eden/NeuroML.cpp:9813-9891)
  ion channels      ionChannelHH / ionChannel with gateHHrates (HHExpRate, HHExpLinearRate,
                    HHSigmoidRate), gateHHtauInf, gateHHratesTauInf, gateHHratesTau,
                    gateHHratesInf, gateHHInstantaneous (fixedTimeCourse,
                    HHExpVariable, HHExpLinearVariable, HHSigmoidVariable), user rate, time
                    course and steady state ComponentTypes (eden_amd.lems_rates), q10Settings
                    (q10Fixed, q10ExpTemp); channels without gates
                    are passive (eden/NeuroML.cpp:4380-4598)
  cells             <cell>: morphology (segments, parent / fractionAlong, implicit proximal
                    points, segmentGroups with member / include, inhomogeneousParameter),
                    biophysicalProperties (channelDensity, channelDensityNonUniform with a
                    condDensity expression of the path length from the root,
                    specificCapacitance, initMembPotential, spikeThresh, resistivity), each over a segment group, later specs overriding earlier
                    ones (eden/NeuroML.cpp:2530-3566)
  point neurons     the NeuroML core artificial cells (iafTauCell, iafTauRefCell, iafCell,
                    iafRefCell, izhikevichCell, izhikevich2007Cell, adExIaFCell,
                    fitzHughNagumoCell, fitzHughNagumo1969Cell, hindmarshRose1984Cell,
                    pinskyRinzelCA3Cell, the PyNN
                    IF_* / EIF_* / HH_cond_exp cells, the spike sources spikeArray,
                    spikeGenerator, spikeGeneratorRandom, spikeGeneratorPoisson,
                    spikeGeneratorRefPoisson, SpikeSourcePoisson) and user <ComponentType>s extending a base
                    cell type, each a one-compartment cell run by its generated step
                    (eden_amd.lems_cells; eden/Eden.cpp:6603-6910), recorded by state variable
  discretisation    a cable segment group (neuroLexId sao864921383) becomes
                    numberInternalDivisions compartments of equal path length, any other
                    segment one compartment; mechanisms and V0 apply to every compartment a
                    spec's segments overlap; inputs, connections and recordings go to the
                    compartment holding their segment's fractionAlong (0.5 when absent)
                    (eden/Eden.cpp:1387-1772, 1015-1038, 4511-4560)
  synapses          expTwoSynapse natively; expOneSynapse, alphaSynapse, alphaCurrentSynapse,
                    expThreeSynapse, blockingPlasticSynapse (voltageConcDepBlockMechanism,
                    tsodyksMarkramDepMechanism, tsodyksMarkramDepFacMechanism), gapJunction,
                    silentSynapse, linearGradedSynapse, gradedSynapse and the PyNN synapses
                    (expCond, expCurr, alphaCond, alphaCurr) as generic LEMS types
                    (eden_amd.lems), compiled by the library at run time
                    (eden/NeuroML.cpp:6506-6829); user-defined
                    <ComponentType>s that extend a base synapse type, from their LEMS
                    Dynamics (eden/Eden.cpp:2339-2995)
  inputs            pulseGenerator, poissonFiringSynapse natively (eden/NeuroML.cpp:6968-7107);
                    timedSynapticInput, poissonFiringSynapse of a synapse other than
                    expTwoSynapse and transientPoissonFiringSynapse as input instances of their
                    synapse's generic type with a spike list or a Poisson spiker
                    (eden/Eden.cpp:3759-3791, 3859-3912);
                    pulseGeneratorDL, sine and ramp generators (and their DL forms), voltageClamp,
                    voltageClampTriple and user <ComponentType>s extending a base current input
                    as generic LEMS types whose instances sit after a compartment's synapses
                    (eden/Eden.cpp:3914-3931)
  networks          population (size or instances), projection with connection /
                    connectionWD, synapticConnection, electricalProjection (gapJunction:
                    an instance on each end, each with the other end's voltage as its peer),
                    continuousProjection (continuousConnection[Instance[W]]: the post
                    component on the post end, the pre component on the pre end, peers as
                    for gap junctions), explicitInput, inputList with input / inputW
                    (eden/NeuroML.cpp:6136-6420, eden/Eden.cpp:9225-9270)
  simulation        <Simulation length step seed target> with OutputFile / OutputColumn and
                    EventOutputFile / EventSelection, <Target> (eden/NeuroML.cpp:7979-8311)
  custom setup      <EdenCustomSetup filename>: per-instance parameters of point neurons,
                    input-list inputs and projection connections (eden_amd.custom_setup;
                    eden/NeuroML.cpp:7270-7980, eden/Eden.cpp:9437-9660)

Instantiation follows the reference's order: cells numbered over populations in order (the
work-item / RNG id, eden/Eden.cpp:7857-7911); per compartment, channel distributions in the
order of their channelDensity elements; pulse and Poisson inputs in network input order
(eden/Eden.cpp:8791-8919); synapse instances per (compartment, synapse type) in projection
then connection order, the types of a compartment in the order their synapse elements were
read (IdListRle is a sorted set of component sequence numbers, eden/Eden.cpp:1861-1930,
8931-9281); spike recipients per source compartment in connection order
(AddSpikeTarget, eden/Eden.cpp:8029-8046). Quantities convert to engine units as
ParseQuantity does: the number read as a double, times 10^(unit power - native power), then
stored as float (eden/NeuroML.cpp:3046-3077, eden/NeuroML.h:530-532).

Constructs outside this set (other channel and synapse types, artificial cells, other
inputs) raise NeuroMLError naming the element.
"""
import dataclasses
import os
import re
import time
import xml.etree.ElementTree as ET
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import custom_setup as setup_mod
from . import lems as lems_mod
from . import lems_cells as cells_mod
from . import lems_rates as rates_mod
from .model import Model, cell_rng_seeds, channel_type


class NeuroMLError(ValueError):
    pass


# ---------------------------------------------------------------- units (eden/NeuroML.cpp:305-425)
# dimension -> (native power of ten, {unit name: power of ten}); every scale factor is 1
_DIMS = {
    "time": (-6, {"s": 0, "ms": -3, "msec": -3, "us": -6, "usec": -6, "sec": 0, "second": 0, "seconds": 0,
                  "hour": None}),
    "frequency": (6, {"per_s": 0, "Hz": 0, "per_ms": 3, "kHz": 3, "per_us": 6, "MHz": 6}),
    "length": (-6, {"m": 0, "cm": -2, "mm": -3, "um": -6}),
    "voltage": (-3, {"V": 0, "mV": -3}),
    "current": (-9, {"A": 0, "mA": -3, "uA": -6, "nA": -9, "pA": -12}),
    "resistance": (6, {"ohm": 0, "kohm": 3, "Mohm": 6}),
    "conductance": (-6, {"S": 0, "mS": -3, "uS": -6, "nS": -9, "pS": -12}),
    "resistivity": (1, {"kohm_cm": 1, "ohm_m": 0, "ohm_cm": -2}),
    "conductivity": (1, {"S_per_m2": 0, "mS_per_cm2": 1, "S_per_cm2": 4, "S_per_mm2": 6}),
    "capacitance": (-12, {"F": 0, "mF": -3, "uF": -6, "nF": -9, "pF": -12}),
    "specific_capacitance": (-2, {"uF_per_cm2": -2, "F_per_m2": 0, "uF_per_mm2": 0}),
    "temperature": (0, {"K": 0, "degK": 0, "degC": "degC"}),
    "concentration": (-3, {"mol_per_m3": 0, "mol_per_cm3": 6, "M": 3, "mM": 0, "uM": -3, "nM": -6}),
    "conductance_per_voltage": (-6, {"S_per_V": 0, "nS_per_mV": -6}),
    "voltage_per_time": (3, {"V_per_sec": 0, "V_per_s": 0, "mV_per_msec": 0, "mV_per_ms": 0, "mV_per_usec": 3}),
    "currentDensity": (-2, {"A_per_m2": 0, "uA_per_cm2": -2, "mA_per_cm2": 1}),
    "per_voltage": (3, {"per_V": 0, "per_mV": 3}),
}

_NUM = re.compile(r"^\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)\s*(\S*)\s*$")


def convert(text, dim, what="quantity"):
    """Engine-unit value (a Python float, the double EDEN computes before storing it) of a
    NeuroML quantity such as '-65mV' or '3.0 S_per_m2'."""
    m = _NUM.match(text or "")
    if not m:
        raise NeuroMLError(f"{what}: '{text}' is not a number followed by a unit")
    value, unit = float(m.group(1)), m.group(2)
    if dim is None:
        if unit:
            raise NeuroMLError(f"{what}: '{text}' should have no unit")
        return value
    native, units = _DIMS[dim]
    if unit not in units:
        raise NeuroMLError(f"{what}: unknown {dim} unit '{unit}' (supported: {' '.join(units)})")
    p = units[unit]
    if p == "degC":
        return value + 273.15
    if p is None:
        return value * 3600.0 * 10.0 ** (0 - native)
    # ScaleEntry::ConvertTo: value * (scale / to.scale) * pow10(pow - to.pow) + offset term 0
    return value * (1.0 / 1.0) * 10.0 ** (p - native) + 0.0


def _f32(x):
    return float(np.float32(x))


def _tag(el):
    t = el.tag
    return t.rsplit("}", 1)[-1] if "}" in t else t


# NeuroML's own type-definition files, included by LEMS files for jLEMS; EDEN embeds its own
# core components instead and skips them (eden/NeuroML.cpp:9837-9862)
_CORE_INCLUDES = {"Cells.xml", "Networks.xml", "Simulation.xml", "PyNN.xml", "Synapses.xml", "Inputs.xml",
                  "Channels.xml", "NeuroMLCoreDimensions.xml", "NeuroMLCoreCompTypes.xml"}

# element names of the synapse and input types this reader instantiates, by role; a LEMS
# <Component type="..."> of these types is read as the element itself
_SYNAPSES = set(lems_mod.SYNAPSE_TYPES) | {"blockingPlasticSynapse"}
_INPUTS = {"pulseGenerator", "poissonFiringSynapse", "transientPoissonFiringSynapse", "timedSynapticInput"} | \
    set(lems_mod.INPUT_TYPES)
# inputs that drive their own instance of a synapse (eden/Eden.cpp:3859-3912)
_FIRING_INPUTS = {"poissonFiringSynapse", "transientPoissonFiringSynapse", "timedSynapticInput"}
# synapse types that may sit on either end of a continuous connection (no spike input)
_GRADED = {"silentSynapse", "linearGradedSynapse", "gradedSynapse", "gapJunction"}
_UNSUPPORTED_SYNAPSES = {"doubleSynapse"}
_UNSUPPORTED_INPUTS = set()
_UNSUPPORTED_CELLS = set()


@dataclass
class OutputFileSpec:
    id: str
    file_name: str
    columns: list = field(default_factory=list)      # (column id, quantity path, compartment)


@dataclass
class EventFileSpec:
    id: str
    file_name: str
    format: str
    selections: list = field(default_factory=list)   # (selection id int, select path, compartment)


@dataclass
class Outputs:
    trajectories: list
    events: list
    base_dir: str


class _Doc:
    """Elements of a LEMS / NeuroML file and of the files it includes, in reading order."""

    def __init__(self, path):
        self.items = OrderedDict()       # id -> (role, element), role in cell/channel/synapse/input/...
        self.order = []                  # (role, id) in document order
        self.networks = OrderedDict()
        self.simulations = OrderedDict()
        self.sim_base = {}               # simulation id -> directory of the file holding it
        self.morphologies = {}
        self.biophysics = {}
        self.component_types = {}        # user <ComponentType> elements by name
        self.target = None
        self.seen = set()
        self._read(os.path.abspath(path))

    def _read(self, path):
        if path in self.seen:
            return
        self.seen.add(path)
        try:
            root = ET.parse(path).getroot()
        except (OSError, ET.ParseError) as e:
            raise NeuroMLError(f"cannot read {path}: {e}") from None
        base = os.path.dirname(path)
        self._walk(root, base)

    def _include(self, name, base):
        p = os.path.join(base, name)
        if not os.path.exists(p):
            if os.path.basename(name) in _CORE_INCLUDES:
                return
            raise NeuroMLError(f"included file {name} not found (relative to {base})")
        if os.path.basename(name) in _CORE_INCLUDES and not p.endswith(".nml"):
            return
        self._read(p)

    def _add(self, role, el):
        i = el.get("id")
        if not i:
            raise NeuroMLError(f"<{_tag(el)}> without an id")
        if i in self.items:
            raise NeuroMLError(f"id '{i}' defined twice")
        self.items[i] = (role, el)
        self.order.append((role, i))

    def _walk(self, root, base):
        for el in root:
            t = _tag(el)
            if t == "Component" and el.get("type"):
                # a LEMS component of a NeuroML type: read as that element
                el = _as_element(el)
                t = _tag(el)
            if t == "Include":
                self._include(el.get("file", ""), base)
            elif t == "include":
                self._include(el.get("href", ""), base)
            elif t == "Target":
                self.target = el.get("component")
            elif t in ("ionChannelHH", "ionChannel", "ionChannelPassive"):
                self._add("channel", el)
            elif t == "cell":
                self._add("cell", el)
            elif t == "morphology":
                self.morphologies[el.get("id")] = el
            elif t == "biophysicalProperties":
                self.biophysics[el.get("id")] = el
            elif t == "ComponentType":
                self.component_types[el.get("name")] = el
            elif t in _SYNAPSES or lems_mod.is_synapse_component_type(t, self.component_types):
                self._add("synapse", el)
            elif t in _INPUTS or lems_mod.is_input_component_type(t, self.component_types):
                self._add("input", el)
            elif t in cells_mod.CELL_TYPES or t == "spikeArray" or cells_mod.base_cell_of(t, self.component_types):
                self._add("acell", el)
            elif t == "network":
                self.networks[el.get("id")] = el
            elif t == "Simulation":
                self.simulations[el.get("id")] = el
                self.sim_base[el.get("id")] = base
            elif t in _UNSUPPORTED_SYNAPSES or t in _UNSUPPORTED_CELLS or t in _UNSUPPORTED_INPUTS:
                # read, and refused only when the model uses it
                self._add("unsupported", el)
            elif t == "neuroml":
                self._walk(el, base)


def _as_element(comp):
    """<Component type="X" id=... a=...> -> an element <X id=... a=...> with the same children."""
    el = ET.Element(comp.get("type"), {k: v for k, v in comp.attrib.items() if k != "type"})
    el.extend(list(comp))
    return el


# ---------------------------------------------------------------- ion channels

_RATE_FORMS = {"HHExpRate": N.RATE_EXP, "HHExpLinearRate": N.RATE_EXPLINEAR, "HHSigmoidRate": N.RATE_SIGMOID}


def _rate(el, what, user=None):
    t = el.get("type")
    if t not in _RATE_FORMS and user is not None:
        return user(el, "baseVoltageDepRate", what)
    if t not in _RATE_FORMS:
        raise NeuroMLError(f"{what}: rate type '{t}' is not supported (HHExpRate, HHExpLinearRate, HHSigmoidRate)")
    r = convert(el.get("rate"), "frequency", what + " rate")
    mid = convert(el.get("midpoint"), "voltage", what + " midpoint")
    sc = convert(el.get("scale"), "voltage", what + " scale")
    return N.Rate(_RATE_FORMS[t], _f32(r), _f32(mid), _f32(sc))


_VARIABLE_FORMS = {"HHExpVariable": N.RATE_EXP, "HHExpLinearVariable": N.RATE_EXPLINEAR,
                   "HHSigmoidVariable": N.RATE_SIGMOID}
# gate elements by kind: alpha/beta rates, or tau and inf given (a gateHHratesTauInf's rates
# play no part in its dynamics, eden/Eden.cpp:5282-5383)
_GATE_KINDS = {"gateHHrates": N.GATE_RATES, "gateHHtauInf": N.GATE_TAU_INF, "gateHHratesTauInf": N.GATE_TAU_INF,
               "gateHHInstantaneous": N.GATE_INSTANTANEOUS, "gateHHratesTau": N.GATE_RATES_TAU,
               "gateHHratesInf": N.GATE_RATES_INF}


def _steady_state(el, what, user=None):
    """A steadyState element: HHExpVariable / HHExpLinearVariable / HHSigmoidVariable, its
    rate a plain number (eden/NeuroML.cpp:4028-4057), or a user ComponentType."""
    t = el.get("type")
    if t not in _VARIABLE_FORMS and user is not None:
        return user(el, "baseVoltageDepVariable", what)
    if t not in _VARIABLE_FORMS:
        raise NeuroMLError(f"{what}: steadyState type '{t}' is not supported ({', '.join(_VARIABLE_FORMS)})")
    r = convert(el.get("rate"), None, what + " rate")
    mid = convert(el.get("midpoint"), "voltage", what + " midpoint")
    sc = convert(el.get("scale"), "voltage", what + " scale")
    return N.Rate(_VARIABLE_FORMS[t], _f32(r), _f32(mid), _f32(sc))


def _time_course(el, what, user=None):
    """A timeCourse element: fixedTimeCourse, tau in us (eden/NeuroML.cpp:4087-4108), or a
    user ComponentType."""
    if el.get("type") != "fixedTimeCourse" and user is not None:
        return user(el, "baseVoltageDepTime", what)
    if el.get("type") != "fixedTimeCourse":
        raise NeuroMLError(f"{what}: timeCourse type '{el.get('type')}' is not supported (fixedTimeCourse)")
    return N.Rate(N.RATE_FIXED, _f32(convert(el.get("tau"), "time", what + " tau")), 0.0, 0.0)


def _channel(el, temperature, user=None):
    """(channel type, gate kinds) of an ion channel element; `user(element, base type, what)`
    turns a user rate ComponentType into an EDENHIP_RATE_LEMS rate."""
    cid = el.get("id")
    gates, kinds, extras = [], [], []
    scaling = 1.0
    for g in el:
        t = _tag(g)
        if t in ("notes", "annotation"):
            continue
        if t == "q10ConductanceScaling":
            # conductance_scaling = q10Factor^((T - experimentalTemp) / 10), the start of the
            # open fraction (eden/Eden.cpp:5700-5721)
            if temperature is None:
                raise NeuroMLError(f"channel {cid}: q10ConductanceScaling needs the network's temperature")
            f = convert(g.get("q10Factor"), None, "q10Factor")
            t0 = convert(g.get("experimentalTemp"), "temperature", "experimentalTemp")
            scaling = f ** ((temperature - t0) / 10.0)
            continue
        gtype = g.get("type") if t == "gate" else t
        if gtype not in _GATE_KINDS:
            raise NeuroMLError(f"channel {cid}: <{t}> gates of type {gtype} are not supported "
                               f"({', '.join(_GATE_KINDS)})")
        kind = _GATE_KINDS[gtype]
        inst = int(g.get("instances", "1"))
        fw = rv = tau = inf = None
        q10 = 1.0
        for c in g:
            ct = _tag(c)
            with_rates = kind in (N.GATE_RATES, N.GATE_RATES_TAU, N.GATE_RATES_INF)
            if ct == "forwardRate" and with_rates:
                fw = _rate(c, f"channel {cid} gate {g.get('id')} forward", user)
            elif ct == "reverseRate" and with_rates:
                rv = _rate(c, f"channel {cid} gate {g.get('id')} reverse", user)
            elif ct == "timeCourse":
                tau = _time_course(c, f"channel {cid} gate {g.get('id')} timeCourse", user)
            elif ct == "steadyState":
                inf = _steady_state(c, f"channel {cid} gate {g.get('id')} steadyState", user)
            elif ct == "q10Settings":
                qt = c.get("type")
                if qt == "q10Fixed":
                    q10 = convert(c.get("fixedQ10"), None, "fixedQ10")
                elif qt == "q10ExpTemp":
                    if temperature is None:
                        raise NeuroMLError(f"channel {cid}: q10ExpTemp needs the network's temperature")
                    f = convert(c.get("q10Factor"), None, "q10Factor")
                    t0 = convert(c.get("experimentalTemp"), "temperature", "experimentalTemp")
                    q10 = f ** ((temperature - t0) / 10.0)
                else:
                    raise NeuroMLError(f"channel {cid}: q10Settings type '{qt}' is not supported")
        extra = None
        if kind in (N.GATE_RATES, N.GATE_RATES_TAU, N.GATE_RATES_INF):
            if fw is None or rv is None:
                raise NeuroMLError(f"channel {cid} gate {g.get('id')}: forwardRate and reverseRate are required")
            gates.append((inst, fw, rv, _f32(q10)))
            extra = tau if kind == N.GATE_RATES_TAU else inf if kind == N.GATE_RATES_INF else None
            if kind != N.GATE_RATES and extra is None:
                raise NeuroMLError(f"channel {cid} gate {g.get('id')}: "
                                   f"{'timeCourse' if kind == N.GATE_RATES_TAU else 'steadyState'} is required")
        elif kind == N.GATE_INSTANTANEOUS:
            if inf is None:
                raise NeuroMLError(f"channel {cid} gate {g.get('id')}: steadyState is required")
            gates.append((inst, N.Rate(N.RATE_FIXED, 0.0, 0.0, 0.0), inf, _f32(q10)))
        else:
            if tau is None or inf is None:
                raise NeuroMLError(f"channel {cid} gate {g.get('id')}: timeCourse and steadyState are required")
            gates.append((inst, tau, inf, _f32(q10)))
        kinds.append(kind)
        extras.append(extra)
    if len(gates) > N.MAX_GATES:
        raise NeuroMLError(f"channel {cid}: more than {N.MAX_GATES} gates")
    pad = N.MAX_GATES - len(kinds)
    return channel_type(gates), kinds + [N.GATE_RATES] * pad, extras + [None] * pad, _f32(scaling)


# ---------------------------------------------------------------- cells

@dataclass
class _Cell:
    id: str
    seg_ids: list            # segment ids in file order (parents first)
    parent: np.ndarray       # compartment parent (cell-local), -1 for the root
    seg_map: list            # per segment: [(compartment, fractionAlong where it ends on the segment)]
    channels: list           # per compartment: [(channel id, gbase mS/cm^2, erev mV, distribution id)]
    v0: np.ndarray           # per compartment
    vthr: np.ndarray
    bioph: str = ""          # id of the biophysicalProperties
    point: tuple = None      # a LEMS point neuron: (type name, {parameter: engine value}, CellType)
    area: np.ndarray = None
    cap: np.ndarray = None
    r_axial: np.ndarray = None
    inv_rc: np.ndarray = None

    @property
    def n_comp(self):
        return len(self.parent)

    def comp_of_segment(self, seg_id, what, fraction_along=0.5):
        """The compartment that holds a location on a segment: the first whose span on the
        segment ends at or beyond it (GetCompartmentForSegmentLocation, eden/Eden.cpp:1015-1038)."""
        try:
            k = self.seg_ids.index(int(seg_id))
        except ValueError:
            raise NeuroMLError(f"{what}: segment {seg_id} not in cell {self.id}") from None
        return _comp_at(self.seg_map, k, fraction_along)


def _comp_at(seg_map, k, fraction_along):
    f = np.float32(fraction_along)
    for comp, f_end in seg_map[k]:
        if f <= f_end:
            return comp
    if not seg_map[k]:
        raise NeuroMLError(f"segment index {k} is in no compartment")
    return seg_map[k][-1][0]


def _point(el):
    return [convert(el.get(k), None, f"point {k}") for k in ("x", "y", "z", "diameter")]


def _morphology(cid, m):
    segs = [s for s in m if _tag(s) == "segment"]
    if not segs:
        raise NeuroMLError(f"cell {cid}: morphology without segments")
    ids = [int(s.get("id")) for s in segs]
    index = {sid: k for k, sid in enumerate(ids)}
    n = len(segs)
    parent = np.full(n, -1, np.int32)
    frac = np.ones(n, np.float32)
    prox = np.zeros((n, 4))
    dist = np.zeros((n, 4))
    for k, s in enumerate(segs):
        pe = next((c for c in s if _tag(c) == "parent"), None)
        pr = next((c for c in s if _tag(c) == "proximal"), None)
        di = next((c for c in s if _tag(c) == "distal"), None)
        if di is None:
            raise NeuroMLError(f"cell {cid} segment {ids[k]}: no distal point")
        dist[k] = _point(di)
        if pe is not None:
            p = index.get(int(pe.get("segment")))
            if p is None or p >= k:
                raise NeuroMLError(f"cell {cid} segment {ids[k]}: parent must be a segment listed before it")
            parent[k] = p
            f = convert(pe.get("fractionAlong", "1"), None, "fractionAlong")
            if not 0 <= f <= 1:
                raise NeuroMLError(f"cell {cid} segment {ids[k]}: fractionAlong {f} outside [0, 1]")
            frac[k] = f
            if pr is None:
                # the implicit proximal point lies fractionAlong along the parent
                prox[k] = prox[p] + (dist[p] - prox[p]) * f
        elif k != 0:
            raise NeuroMLError(f"cell {cid}: segment {ids[k]} has no parent (only the first segment may be the root)")
        if pr is not None:
            prox[k] = _point(pr)
        elif pe is None:
            raise NeuroMLError(f"cell {cid} segment {ids[k]}: a root segment needs a proximal point")
    groups = {"all": set(range(n))}
    raw = {}
    cable_ids = []
    inho = {}    # inhomogeneous parameter id -> (variable, group, subtract the minimum, divide by the maximum)
    for g in m:
        if _tag(g) != "segmentGroup":
            continue
        gid = g.get("id")
        raw[gid] = g
        if g.get("neuroLexId") == "sao864921383":
            cable_ids.append(gid)
        for ip in g:
            if _tag(ip) != "inhomogeneousParameter":
                continue
            if (ip.get("metric") or "").lower() != "path length from root":
                raise NeuroMLError(f"cell {cid} group {gid}: inhomogeneous parameter metric "
                                   f"'{ip.get('metric')}' is not supported (Path Length from root)")
            prox_el = next((c for c in ip if _tag(c) == "proximal"), None)
            dist_el = next((c for c in ip if _tag(c) == "distal"), None)
            ts = prox_el.get("translationStart") if prox_el is not None else None
            ne = dist_el.get("normalizationEnd") if dist_el is not None else None
            if ts is not None and float(ts) != 0:
                raise NeuroMLError(f"cell {cid}: translationStart must be 0")
            if ne is not None and float(ne) != 1:
                raise NeuroMLError(f"cell {cid}: normalizationEnd must be 1")
            inho[ip.get("id")] = (ip.get("variable"), gid, ts is not None, ne is not None)

    def resolve(gid, stack=()):
        if gid in groups:
            return groups[gid]
        if gid not in raw or gid in stack:
            raise NeuroMLError(f"cell {cid}: segment group {gid} not found")
        out = set()
        for c in raw[gid]:
            t = _tag(c)
            if t == "member":
                s = index.get(int(c.get("segment")))
                if s is None:
                    raise NeuroMLError(f"cell {cid} group {gid}: segment {c.get('segment')} not found")
                out.add(s)
            elif t == "include":
                out |= resolve(c.get("segmentGroup"), stack + (gid,))
            elif t in ("path", "subTree"):
                raise NeuroMLError(f"cell {cid} group {gid}: <{t}> is not supported")
        groups[gid] = out
        return out

    for gid in raw:
        resolve(gid)
    # cables: the group's segments in file order, and its number of compartments (the
    # numberInternalDivisions property, 1 when absent; eden/NeuroML.cpp:2709-2860)
    cables = []
    for gid in cable_ids:
        nseg = 1
        for c in raw[gid]:
            if _tag(c) == "property" and c.get("tag") == "numberInternalDivisions":
                nseg = int(c.get("value"))
                if nseg < 1:
                    raise NeuroMLError(f"cell {cid} group {gid}: numberInternalDivisions must be at least 1")
        cables.append((gid, sorted(groups[gid]), nseg))
    for g in raw.values():
        if g.get("neuroLexId") != "sao864921383":
            for c in g:
                if _tag(c) == "property" and c.get("tag") == "numberInternalDivisions":
                    raise NeuroMLError(f"cell {cid}: group {g.get('id')} has numberInternalDivisions but is not a "
                                       "cable (neuroLexId sao864921383)")
    return ids, parent, frac, prox.astype(np.float32), dist.astype(np.float32), groups, cables, inho


def _lerp_len(p, q, f0, f1):
    """Float length of the slice [f0, f1] of the segment p -> q (LerpPt3d, DistancePt3d)."""
    F = np.float32
    a = [F(p[i] + F(F(q[i] - p[i]) * F(f0))) for i in range(3)]
    b = [F(p[i] + F(F(q[i] - p[i]) * F(f1))) for i in range(3)]
    d = [F(b[i] - a[i]) for i in range(3)]
    return F(np.sqrt(F(F(F(d[0] * d[0]) + F(d[1] * d[1])) + F(d[2] * d[2]))))


def _path_length_from_root(cparent, slices, prox, dist):
    """Path length of each compartment's midpoint from the root's proximal end: the root's
    half length, then the parent's plus half of each of the two lengths (eden/Eden.cpp:4119-4134)."""
    F = np.float32
    L = np.zeros(len(cparent), np.float32)
    for c, sl in enumerate(slices):
        acc = F(0)
        for s_, f0, f1 in sl:
            acc = F(acc + _lerp_len(prox[s_], dist[s_], f0, f1))
        L[c] = acc
    out = np.zeros(len(cparent), np.float32)
    for c in range(len(cparent)):
        p = cparent[c]
        out[c] = F(L[c] / F(2)) if p < 0 else F(out[p] + F(F(L[p] + L[c]) / F(2)))
    return out


def _inhomogeneous_values(expr, var, vals, subtract_min, divide_max, cid):
    """Channel densities (mS/cm^2) of a channelDensityNonUniform over its compartments: the
    parameter (path length from root, um) optionally less its minimum and over its maximum,
    with NEURON's guard values (minimum from 1e9 down, maximum from 1 up), put into the value
    expression, whose result is S/m^2 (eden/Eden.cpp:4560-4657)."""
    F = np.float32
    vals = np.asarray(vals, np.float32).copy()
    p0 = F(min([F(1e9)] + list(vals)))
    p1 = F(max([F(1)] + list(vals)))
    if subtract_min:
        vals = (vals - p0).astype(np.float32)
        p1 = F(p1 - p0)
    if divide_max:
        vals = (vals / p1).astype(np.float32)
    try:
        tree = lems_mod.parse(expr)
        out = [lems_mod.evaluate(tree, {var: v}) for v in vals]
    except lems_mod.LemsError as e:
        raise NeuroMLError(f"cell {cid}: inhomogeneousValue '{expr}': {e}") from None
    return [F(F(g) * F(0.1)) for g in out]      # S/m^2 -> mS/cm^2


def _discretise(cid, ids, parent, frac, prox, dist, cables):
    """Compartments of a cell (eden/Eden.cpp:1387-1772): every cable group becomes nseg
    compartments of equal path length, allotted when its first segment comes up in segment
    order; every other segment is one compartment. Float arithmetic as the reference's (Real
    is float). Returns (compartment parents, per compartment its slices [(segment, from, to)],
    per segment [(compartment, fractionAlong where it ends)])."""
    F = np.float32
    n = len(ids)
    cable_of = [-1] * n
    for ci, (gid, segs, nseg) in enumerate(cables):
        for s_ in segs:
            if cable_of[s_] >= 0:
                raise NeuroMLError(f"cell {cid}: cables {cables[cable_of[s_]][0]} and {gid} overlap on segment {ids[s_]}")
            cable_of[s_] = ci
    free_comp = [-1] * n
    cable_off = [-1] * len(cables)
    ncomp = 0
    for s_ in range(n):
        ci = cable_of[s_]
        if ci >= 0:
            if cable_off[ci] < 0:
                cable_off[ci] = ncomp
                ncomp += cables[ci][2]
        else:
            free_comp[s_] = ncomp
            ncomp += 1
    slices = [[] for _ in range(ncomp)]
    seg_map = [[] for _ in range(n)]

    def add(comp, s_, f0, f1):
        slices[comp].append((s_, F(f0), F(f1)))
        seg_map[s_].append((comp, F(f1)))

    def seg_len(s_):
        d = dist[s_, :3] - prox[s_, :3]
        return F(np.sqrt(F(F(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))

    for ci, (gid, segs, nseg) in enumerate(cables):
        cum = []
        total = F(0)
        for i, s_ in enumerate(segs):
            if i > 0:
                if parent[s_] != segs[i - 1]:
                    raise NeuroMLError(f"cell {cid} cable {gid}: not an unbranched cable (segment {ids[s_]})")
                if parent[s_] >= 0 and frac[s_] != 1:
                    raise NeuroMLError(f"cell {cid} cable {gid}: segment {ids[s_]} is not attached to the end of "
                                       "its parent")
            total = F(total + seg_len(s_))
            cum.append(total)

        def at(si, q):
            a = cum[si - 1] if si > 0 else F(0)
            b = cum[si]
            f = F(F(q - a) / F(b - a)) if b != a else F(np.nan)
            if abs(float(F(b - a))) < 0.001:
                f = F(0)
            return f

        under = 0
        for k in range(nseg):
            start = F(F(F(k) / F(nseg)) * total)
            end = F(F(F(k + 1) / F(nseg)) * total)
            comp = cable_off[ci] + k
            idx = start
            while under < len(cum) and cum[under] <= end:
                fa = at(under, idx)
                if not abs(float(F(fa - F(1)))) < 0.0001:
                    add(comp, segs[under], fa, 1.0)
                idx = cum[under]
                under += 1
            if under < len(cum) and end < cum[under]:
                f0, f1 = at(under, idx), at(under, end)
                if not abs(float(F(f1 - f0))) < 0.0001:
                    add(comp, segs[under], f0, f1)
    for s_ in range(n):
        if free_comp[s_] >= 0:
            add(free_comp[s_], s_, 0.0, 1.0)
    cparent = np.full(ncomp, -1, np.int32)
    for s_ in range(n):
        if free_comp[s_] >= 0 and parent[s_] >= 0:
            cparent[free_comp[s_]] = _comp_at(seg_map, parent[s_], frac[s_])
    for ci, (gid, segs, nseg) in enumerate(cables):
        c0 = cable_off[ci]
        for c in range(c0 + 1, c0 + nseg):
            cparent[c] = c - 1
        if not slices[c0]:
            raise NeuroMLError(f"cell {cid} cable {gid}: its first compartment holds no segment")
        first = slices[c0][0][0]
        if parent[first] >= 0:
            cparent[c0] = _comp_at(seg_map, parent[first], frac[first])
    for c in range(ncomp):
        if not slices[c]:
            raise NeuroMLError(f"cell {cid}: compartment {c} holds no segment")
        if (c == 0) != (cparent[c] < 0) or cparent[c] >= c:
            raise NeuroMLError(f"cell {cid}: the compartments do not list parents before children")
    return cparent, slices, seg_map


def _lems_value(text, lems_dim, what):
    """Engine value of a LEMS parameter or constant of the given LEMS dimension."""
    d = lems_mod.LEMS_DIMS.get(lems_dim, lems_dim if lems_dim in _DIMS else "unsupported")
    if d == "unsupported":
        raise NeuroMLError(f"{what}: dimension '{lems_dim}' is not supported")
    return convert(text, d, what)


def _point_cell(doc, el):
    """A LEMS point neuron (artificial cell) element as a one-compartment cell: its type, its
    parameters in engine units (eden/NeuroML.cpp:8688-8745 for the base types' parameters)."""
    cid, t = el.get("id"), _tag(el)
    if t == "spikeArray":
        times = [_f32(convert(sp.get("time"), "time", f"spikeArray {cid} spike")) for sp in el if _tag(sp) == "spike"]
        c = _Cell(cid, [0], np.array([-1], np.int32), [[(0, 1.0)]], [[]], np.zeros(1, np.float32),
                  np.full(1, np.nan, np.float32), "", point=(t, {"times": times}, cells_mod.SpikeList()))
        c.area, c.cap = np.ones(1, np.float32), np.ones(1, np.float32)
        c.r_axial, c.inv_rc = np.zeros(1, np.float32), np.zeros(1, np.float32)
        return c
    try:
        ct = cells_mod.CELL_TYPES.get(t) or cells_mod.cell_type_from_lems(
            t, doc.component_types, lambda text, d: _lems_value(text, d, f"ComponentType {t} constant"))
    except lems_mod.LemsError as e:
        raise NeuroMLError(str(e)) from None
    vals = {}
    for p_, d_ in ct.params:
        if el.get(p_) is None:
            raise NeuroMLError(f"cell {cid} ({t}): parameter {p_} missing")
        vals[p_] = _lems_value(el.get(p_), d_, f"cell {cid} {p_}")
    c = _Cell(cid, [0], np.array([-1], np.int32), [[(0, 1.0)]], [[]], np.zeros(1, np.float32),
              np.full(1, np.nan, np.float32), "", point=(t, vals, ct))
    c.area = np.ones(1, np.float32)
    c.cap = np.ones(1, np.float32)
    c.r_axial = np.zeros(1, np.float32)
    c.inv_rc = np.zeros(1, np.float32)
    return c


def _cell(doc, el, temperature, channel_index):
    cid = el.get("id")
    m = next((c for c in el if _tag(c) == "morphology"), None)
    if m is None:
        m = doc.morphologies.get(el.get("morphology"))
    if m is None:
        raise NeuroMLError(f"cell {cid}: no morphology")
    b = next((c for c in el if _tag(c) == "biophysicalProperties"), None)
    if b is None:
        b = doc.biophysics.get(el.get("biophysicalProperties"))
    if b is None:
        raise NeuroMLError(f"cell {cid}: no biophysicalProperties")
    ids, parent, frac, prox, dist, groups, cables, inho = _morphology(cid, m)
    n = len(ids)
    cparent, slices, seg_map = _discretise(cid, ids, parent, frac, prox, dist, cables)
    ncomp = len(cparent)
    path_len = _path_length_from_root(cparent, slices, prox, dist)
    # capacitance and resistivity per segment; channels, V0 and threshold per compartment,
    # each spec over every compartment its segments overlap (SpecToCompList,
    # eden/Eden.cpp:4511-4560)
    chans = [[] for _ in range(ncomp)]
    cm = np.full(n, np.nan)
    ra = np.full(n, np.nan)
    v0 = np.full(ncomp, np.nan)
    vthr = np.full(ncomp, 0.0)    # NeuroML's default spike threshold, 0 mV

    def where(spec):
        if spec.get("segment") is not None:
            k = ids.index(int(spec.get("segment"))) if int(spec.get("segment")) in ids else None
            if k is None:
                raise NeuroMLError(f"cell {cid}: segment {spec.get('segment')} not found")
            return [k]
        g = spec.get("segmentGroup", "all")
        if g not in groups:
            raise NeuroMLError(f"cell {cid}: segment group {g} not found")
        return sorted(groups[g])

    def comps(spec):
        return sorted({c for k in where(spec) for c, _ in seg_map[k]})

    for part in b:
        pt = _tag(part)
        if pt == "membraneProperties":
            for s in part:
                st = _tag(s)
                if st == "channelDensity":
                    ch = s.get("ionChannel")
                    if ch not in channel_index:
                        raise NeuroMLError(f"cell {cid}: ion channel {ch} not found or not supported")
                    g = convert(s.get("condDensity"), "conductivity", "condDensity")
                    e = convert(s.get("erev"), "voltage", "erev")
                    for k in comps(s):
                        chans[k].append((ch, _f32(g), _f32(e), s.get("id")))
                elif st == "channelDensityNonUniform":
                    ch = s.get("ionChannel")
                    if ch not in channel_index:
                        raise NeuroMLError(f"cell {cid}: ion channel {ch} not found or not supported")
                    e = convert(s.get("erev"), "voltage", "erev")
                    for vp in s:
                        if _tag(vp) != "variableParameter":
                            continue
                        if vp.get("parameter") != "condDensity":
                            raise NeuroMLError(f"cell {cid}: variableParameter '{vp.get('parameter')}' is not "
                                               "supported (condDensity)")
                        iv = next((c for c in vp if _tag(c) == "inhomogeneousValue"), None)
                        pid = iv.get("inhomogeneousParameter") if iv is not None else None
                        if pid not in inho or inho[pid][1] != vp.get("segmentGroup"):
                            raise NeuroMLError(f"cell {cid}: inhomogeneous parameter {pid} not found in group "
                                               f"{vp.get('segmentGroup')}")
                        var, _, sub_min, div_max = inho[pid]
                        cl = comps(vp)
                        for k, g in zip(cl, _inhomogeneous_values(iv.get("value"), var, path_len[cl], sub_min,
                                                                  div_max, cid)):
                            chans[k].append((ch, g, _f32(e), s.get("id")))
                elif st == "specificCapacitance":
                    for k in where(s):
                        cm[k] = convert(s.get("value"), "specific_capacitance", "specificCapacitance")
                elif st == "initMembPotential":
                    for k in comps(s):
                        v0[k] = convert(s.get("value"), "voltage", "initMembPotential")
                elif st == "spikeThresh":
                    for k in comps(s):
                        vthr[k] = convert(s.get("value"), "voltage", "spikeThresh")
                elif st in ("notes", "annotation"):
                    pass
                else:
                    raise NeuroMLError(f"cell {cid}: <{st}> in membraneProperties is not supported")
        elif pt == "intracellularProperties":
            for s in part:
                st = _tag(s)
                if st == "resistivity":
                    for k in where(s):
                        ra[k] = convert(s.get("value"), "resistivity", "resistivity")
                elif st in ("notes", "annotation"):
                    pass
                else:
                    raise NeuroMLError(f"cell {cid}: <{st}> in intracellularProperties is not supported")
    if np.isnan(cm).any():
        raise NeuroMLError(f"cell {cid}: specificCapacitance missing on some segments")
    if np.isnan(v0).any():
        raise NeuroMLError(f"cell {cid}: initMembPotential missing on some segments")
    if n > 1 and np.isnan(ra).any():
        raise NeuroMLError(f"cell {cid}: resistivity missing on some segments")
    ra = np.where(np.isnan(ra), 0.0, ra).astype(np.float32)
    cm = cm.astype(np.float32)
    c = _Cell(cid, ids, cparent, seg_map, chans, v0.astype(np.float32), vthr.astype(np.float32), b.get("id", ""))
    try:
        if not cables:
            c.area, c.cap, c.r_axial, c.inv_rc = N.cable_setup(parent, prox, dist, cm, ra)
        else:
            sb = np.zeros(ncomp + 1, np.int32)
            np.cumsum([len(x) for x in slices], out=sb[1:])
            flat = [x for lst in slices for x in lst]
            c.area, c.cap, c.r_axial, c.inv_rc = N.cable_setup_slices(
                prox, dist, cm, ra, cparent, sb, [x[0] for x in flat], [x[1] for x in flat], [x[2] for x in flat])
    except ValueError as e:
        raise NeuroMLError(f"cell {cid}: {e}") from None
    return c


# ---------------------------------------------------------------- paths

_POP_IDX = re.compile(r"^([^\[/]+)\[(\d+)\]$")


def _cell_ref(text):
    """Population name and cell instance id of a cell reference ('../pop/3/cell', 'pop/3/cell',
    'pop[3]', '/3' with the population known), as ParseSynapsePopulationRef /
    ParseSynapseCellRef read them (eden/NeuroML.cpp:4603-4649)."""
    s = text[3:] if text.startswith("../") else text
    b, sl = s.find("["), s.find("/")
    ends = [x for x in (b, sl) if x >= 0]
    pop = s[: min(ends)] if ends else s
    if b >= 0:
        rest = s[b + 1:]
    elif sl >= 0:
        rest = s[sl + 1:]
    else:
        rest = s
    m = re.match(r"\d+", rest)
    if not m:
        raise NeuroMLError(f"cell reference '{text}' has no instance id")
    return pop, int(m.group(0))


def _quantity_path(text):
    """(population, instance id, segment id, gate) of an OutputColumn quantity: the membrane
    voltage ('pop/3/cell/5/v', 'pop/3/cell/v', 'pop[3]/v'; gate None) or a channel gate's
    state, 'pop/3/cell/5/<biophysicalProperties id>/membraneProperties/<channel distribution
    id>/<ion channel id>/<gate id>/q' (gate = (biophysics, distribution, channel, gate); the
    path grammar of eden/NeuroML.cpp:1370-1371, 1820-1905)."""
    parts = [p for p in text.split("/") if p]
    m = _POP_IDX.match(parts[0])
    if m:
        pop, inst, rest = m.group(1), int(m.group(2)), parts[1:]
    else:
        if len(parts) < 3 or not parts[1].isdigit():
            raise NeuroMLError(f"quantity path '{text}' not understood")
        pop, inst, rest = parts[0], int(parts[1]), parts[3:] if not parts[2].isdigit() else parts[2:]
    seg = 0
    if len(rest) == 2 and rest[0].isdigit():
        seg = int(rest[0])
        rest = rest[1:]
    if len(rest) == 7 and rest[0].isdigit() and rest[2] == "membraneProperties" and rest[6] == "q":
        return pop, inst, int(rest[0]), (rest[1], rest[3], rest[4], rest[5])
    if len(rest) == 1 and rest[0] != "v" and re.match(r"^[A-Za-z_]\w*$", rest[0]):
        return pop, inst, seg, ("state", rest[0])     # a point neuron's state variable
    if rest != ["v"]:
        raise NeuroMLError(f"quantity path '{text}': only the membrane voltage 'v' of a segment, a channel "
                           "gate's 'q' or a point neuron's state variable can be recorded")
    return pop, inst, seg, None


def _select_path(text):
    """(population, instance id, segment id) of an EventSelection: 'pop/3/cell', 'pop[3]'."""
    parts = [p for p in text.split("/") if p]
    m = _POP_IDX.match(parts[0])
    if m:
        return m.group(1), int(m.group(2)), 0
    if len(parts) >= 2 and parts[1].isdigit():
        seg = int(parts[3]) if len(parts) >= 4 and parts[3].isdigit() else 0
        return parts[0], int(parts[1]), seg
    raise NeuroMLError(f"event selection '{text}' not understood")


# ---------------------------------------------------------------- model

@dataclass
class _Pop:
    id: str
    cell: _Cell
    inst_ids: list           # instance id per sequential instance
    first_cell: int          # global cell index of its first instance

    def seq(self, inst_id, what):
        try:
            return self.inst_ids.index(inst_id)
        except ValueError:
            raise NeuroMLError(f"{what}: instance {inst_id} not in population {self.id}") from None


def _rates_or_none(entries, temperature):
    if not entries:
        return None
    try:
        return rates_mod.build(entries, temperature)
    except lems_mod.LemsError as e:
        raise NeuroMLError(str(e)) from None


def _custom_setup(doc, sim):
    """The statements of the simulation's EdenCustomSetup file, if it names one (its file name
    relative to the file holding the Simulation, eden/NeuroML.cpp:8302-8313, 7281-7283)."""
    els = [el for el in sim if _tag(el) == "EdenCustomSetup"]
    if not els:
        return []
    if len(els) > 1:
        raise NeuroMLError("EdenCustomSetup already defined previously")
    name = els[0].get("filename")
    if not name:
        raise NeuroMLError("EdenCustomSetup: attribute filename missing")
    p = os.path.join(doc.sim_base.get(sim.get("id"), "."), name)
    try:
        with open(p) as f:
            text = f.read()
    except OSError as e:
        raise NeuroMLError(f"EdenCustomSetup: could not open file \"{p}\": {e.strerror}") from None
    try:
        return setup_mod.parse(text)
    except setup_mod.CustomSetupError as e:
        raise NeuroMLError(str(e)) from None


def _names(items):
    return [x if isinstance(x, str) else x[0] for x in items]


def _value_converter(dim, what):
    """convert_one(number text, unit) for a parameter of the given engine dimension (None:
    dimensionless, any unit token ignored as the reference does)."""
    def one(v, unit):
        if dim is None:
            return float(v)
        if unit is None:
            raise NeuroMLError(f"{what}: statement needs units for a property of dimension {dim}")
        return convert(f"{v} {unit}", dim, what)
    return one


def _apply_setting(st, doc, pops, input_lists, projections, input_where, pulses, poisson, generic_index, generic_st,
                   syn_over, cell_over, spiker_over):
    """One EdenCustomSetup statement (eden/Eden.cpp:9501-9650): point-neuron parameters per
    cell, input parameters per input of a list (native pulse and Poisson tables directly,
    generic inputs per instance), synapse parameters per connection end."""
    what = st.what()

    def seqs_of(ids, n, label):
        if ids is None:
            return list(range(n))
        pos = {x: i for i, x in enumerate(label)}
        out = []
        for x in ids:
            if x not in pos:
                raise NeuroMLError(f"{what}: could not resolve id {x}")
            out.append(pos[x])
        return out

    def settable(names_dims, state_names, owner, name=None):
        name = st.param if name is None else name
        if name not in names_dims:
            if name in state_names:
                raise NeuroMLError(f"{what}: {name} is a state variable of {owner}; only parameters can be set")
            raise NeuroMLError(f"{what}: {owner} has no parameter {name}")
        return names_dims[name]

    try:
        if st.kind == "cell":
            p = pops.get(st.group)
            if p is None:
                raise NeuroMLError(f"{what}: population {st.group} not found")
            if p.cell.point is None:
                raise NeuroMLError(f"{what}: population {st.group} holds physical cells; 'set cell' sets point-neuron "
                                   "parameters")
            if st.where != "all":
                raise NeuroMLError(f"{what}: segment specifier '{st.where}' must be 'all' for point neurons")
            ct = p.cell.point[2]
            if isinstance(ct, cells_mod.SpikeList):
                raise NeuroMLError(f"{what}: a spikeArray has no settable parameters")
            ldim = settable(dict(ct.params), [s for s, _ in ct.state], f"cell type {p.cell.id}")
            dim = lems_mod.LEMS_DIMS.get(ldim, ldim if ldim in _DIMS else "unsupported")
            seqs = seqs_of(st.targets, len(p.inst_ids), p.inst_ids)
            vals = setup_mod.values_of(st, len(seqs), _value_converter(dim, what))
            for s, v in zip(seqs, vals):
                cell_over.setdefault(p.first_cell + s, {})[st.param] = v
        elif st.kind == "input":
            if st.group not in input_lists:
                raise NeuroMLError(f"{what}: input list {st.group} not found")
            comp_id, items = input_lists[st.group]
            seqs = seqs_of(st.targets, len(items), [x[0] for x in items])
            el = doc.items[comp_id][1]
            tag = _tag(el)
            # paths into an input (eden/NeuroML.cpp:1700-1752): "synapse/<parameter>" for the
            # inputs that drive a synapse, the spiker's parameters, a LEMS input's parameters,
            # a pulse's delay, duration and amplitude; none into a spike list
            key = st.param
            if tag in _FIRING_INPUTS and st.param.startswith("synapse/"):
                key = st.param[len("synapse/"):]
                if comp_id not in generic_index:
                    raise NeuroMLError(f"{what}: the synapse of a natively run poissonFiringSynapse (an expTwoSynapse) "
                                       "has no per-instance parameters here")
                sty = generic_st[generic_index[comp_id]]
                dim = settable(dict(sty.params), _names(sty.state), f"the synapse of input {comp_id}", key)
                where = "syn"
            elif tag == "timedSynapticInput":
                raise NeuroMLError(f"{what}: timedSynapticInput paths other than synapse/... are not supported (nor "
                                   "by the reference, eden/NeuroML.cpp:1746-1749)")
            elif tag in _FIRING_INPUTS and comp_id in generic_index:
                table = {"averageRate": "frequency"}
                if tag == "transientPoissonFiringSynapse":
                    table.update(delay="time", duration="time")
                dim = settable(table, ["tsince", "isi", "nextIsi"], f"input {comp_id}")
                where = "spiker"
            elif comp_id in generic_index:
                sty = generic_st[generic_index[comp_id]]
                dim = settable(dict(sty.params), _names(sty.state), f"input {comp_id}")
                where = "syn"
            elif tag == "pulseGenerator":
                dim = settable({"delay": "time", "duration": "time", "amplitude": "current"}, [], f"input {comp_id}")
            else:
                dim = settable({"averageRate": "frequency"}, ["tsince", "isi"], f"input {comp_id}")
            vals = setup_mod.values_of(st, len(seqs), _value_converter(dim, what))
            for s, v in zip(seqs, vals):
                w = input_where[items[s][1]]
                if w[0] == "conn":
                    (syn_over if where == "syn" else spiker_over).setdefault(w[1], {})[key] = v
                elif w[0] == "pulse":
                    g, j = w[1], w[2]
                    k = ("delay", "duration", "amplitude").index(st.param)
                    rec = list(pulses[g][j])
                    rec[k] = _f32(v)
                    pulses[g][j] = tuple(rec)
                else:
                    g, j = w[1], w[2]
                    rec = list(poisson[g][j])
                    rec[0] = _f32(v)
                    poisson[g][j] = tuple(rec)
        else:
            if st.group not in projections:
                raise NeuroMLError(f"{what}: projection {st.group} not found")
            kind, items = projections[st.group]
            if st.where == "pre" and kind == "spiking":
                raise NeuroMLError(f"{what}: 'pre' specifier can only be used for graded synapses")
            seqs = seqs_of(st.targets, len(items), [x[0] for x in items])
            if not seqs:
                raise NeuroMLError(f"{what}: item set is empty")
            end = 1 if st.where == "post" else 2
            comps = {(items[s][3], items[s][4]) for s in seqs}
            if len(comps) > 1:
                raise NeuroMLError(f"{what}: target set contains different connection types")
            syn_id = items[seqs[0]][end + 2]
            sty = generic_st[generic_index[syn_id]]
            dim = settable(dict(sty.params), _names(sty.state), f"synapse {syn_id}")
            vals = setup_mod.values_of(st, len(seqs), _value_converter(dim, what))
            for s, v in zip(seqs, vals):
                syn_over.setdefault(items[s][end], {})[st.param] = v
    except setup_mod.CustomSetupError as e:
        raise NeuroMLError(str(e)) from None


def read(path, rng_seed=None):
    """Read a LEMS or NeuroML file; return (Model, Outputs). rng_seed overrides the
    simulation's seed (EDEN's rng_seed token)."""
    doc = _Doc(path)
    if not doc.simulations:
        raise NeuroMLError("no <Simulation> found")
    sim = doc.simulations.get(doc.target) if doc.target else None
    if sim is None:
        if doc.target:
            raise NeuroMLError(f"<Target component='{doc.target}'> is not a Simulation")
        sim = next(iter(doc.simulations.values()))
    net_id = sim.get("target")
    net = doc.networks.get(net_id)
    if net is None:
        raise NeuroMLError(f"simulation target network '{net_id}' not found")
    dt = convert(sim.get("step"), "time", "Simulation step")
    length = convert(sim.get("length"), "time", "Simulation length")
    # the rng_seed option, else <Simulation seed>, else the wall clock's seconds
    # (eden/Eden.cpp:776-792; only random inputs depend on it)
    seed = rng_seed
    if seed is None:
        s = sim.get("seed")
        seed = int(s) if s is not None else int(time.time())
    temperature = convert(net.get("temperature"), "temperature", "temperature") if net.get("temperature") else None
    statements = _custom_setup(doc, sim)

    # channel types in document order, then the cells the network uses
    channel_index, channel_types, gate_ids, gate_kinds, gate_extras, chan_scaling = {}, [], {}, [], [], []
    rate_entries, rate_types = [], {}

    def user_rate(el, base, what):
        """A user rate / time course / steady state ComponentType instance: its rate of form
        EDENHIP_RATE_LEMS (eden/NeuroML.cpp:4028-4108)."""
        t = el.get("type")
        if rates_mod.base_rate_of(t, doc.component_types) != base:
            raise NeuroMLError(f"{what}: type '{t}' is not supported (a built-in form or a user ComponentType "
                               f"extending {base})")
        if t not in rate_types:
            try:
                rate_types[t] = rates_mod.rate_type_from_lems(
                    t, doc.component_types, lambda text, d: _lems_value(text, d, f"ComponentType {t} constant"))
            except lems_mod.LemsError as e:
                raise NeuroMLError(str(e)) from None
        rt = rate_types[t]
        vals = {}
        for p, d in rt.params:
            if el.get(p) is None:
                raise NeuroMLError(f"{what}: parameter {p} of {t} missing")
            vals[p] = _lems_value(el.get(p), d, f"{what} {p}")
        rate_entries.append((rt, vals))
        return N.Rate(N.RATE_LEMS, float(len(rate_entries) - 1), 0.0, 0.0)

    for role, i in doc.order:
        if role == "channel":
            channel_index[i] = len(channel_types)
            ct_, kinds_, extras_, scaling_ = _channel(doc.items[i][1], temperature, user_rate)
            chan_scaling.append(scaling_)
            channel_types.append(ct_)
            gate_kinds += kinds_
            gate_extras += extras_
            gate_ids[i] = [g.get("id") for g in doc.items[i][1] if _tag(g) not in ("notes", "annotation", "q10ConductanceScaling")]
    cells = {}

    def get_cell(cid, what):
        if cid in cells:
            return cells[cid]
        role, el = doc.items.get(cid, (None, None))
        if role == "unsupported":
            raise NeuroMLError(f"{what}: cell type <{_tag(el)}> is not supported")
        if role == "acell":
            cells[cid] = _point_cell(doc, el)
            return cells[cid]
        if role != "cell":
            raise NeuroMLError(f"{what}: cell '{cid}' not found")
        cells[cid] = _cell(doc, el, temperature, channel_index)
        return cells[cid]

    pops = OrderedDict()
    n_cells = 0
    for el in net:
        if _tag(el) != "population":
            continue
        pid = el.get("id")
        c = get_cell(el.get("component"), f"population {pid}")
        inst = [int(i.get("id")) for i in el if _tag(i) == "instance"]
        if not inst:
            inst = list(range(int(el.get("size", "0"))))
        pops[pid] = _Pop(pid, c, inst, n_cells)
        n_cells += len(inst)
    if n_cells == 0:
        raise NeuroMLError(f"network {net_id} has no cells")

    # compartments: cells in population order, compartments in segment order
    comp0 = []
    pop_of_cell = []
    nc = 0
    for p in pops.values():
        for _ in p.inst_ids:
            comp0.append(nc)
            pop_of_cell.append(p)
            nc += p.cell.n_comp
    comp0.append(nc)
    cell_begin = np.array(comp0, np.int64)

    def comp(pop_id, inst_id, seg_id, what, fraction_along=0.5):
        if pop_id not in pops:
            raise NeuroMLError(f"{what}: population {pop_id} not found")
        p = pops[pop_id]
        c = p.first_cell + p.seq(inst_id, what)
        return int(cell_begin[c]) + p.cell.comp_of_segment(seg_id, what, fraction_along)

    def frac(el, attr):
        f = convert(el.get(attr, "0.5"), None, attr)
        if not 0 <= f <= 1:
            raise NeuroMLError(f"{attr} {f} outside [0, 1]")
        return f

    # synapse types in reading order (their component sequence numbers). Models whose
    # projections use only expTwoSynapse run it natively; otherwise every synapse element
    # becomes a generic LEMS type (in the same order, so per-compartment type order is the
    # reference's), and expTwo types stay native for the Poisson inputs that drive them.
    syn_elems = [(i, doc.items[i][1]) for role, i in doc.order if role == "synapse"]
    native_index, syn_types = {}, []
    for i, el in syn_elems:
        if _tag(el) == "expTwoSynapse":
            native_index[i] = len(syn_types)
            syn_types.append(N.ExpTwoType(
                _f32(convert(el.get("gbase"), "conductance", f"{i} gbase")),
                _f32(convert(el.get("erev"), "voltage", f"{i} erev")),
                _f32(convert(el.get("tauRise"), "time", f"{i} tauRise")),
                _f32(convert(el.get("tauDecay"), "time", f"{i} tauDecay")), 1.0))
    used = set()
    for el in net:
        if _tag(el) in ("projection", "synapticConnection"):
            used.add(el.get("synapse"))
        elif _tag(el) == "electricalProjection":
            used |= {c.get("synapse") for c in el if c.get("synapse")}
        elif _tag(el) == "continuousProjection":
            used |= {c.get(k) for c in el for k in ("preComponent", "postComponent") if c.get(k)}
    # (point neurons take their synapses through the generic path, and so do LEMS inputs)
    points = any(p.cell.point is not None for p in pops.values())
    def native_poisson(el):
        """A poissonFiringSynapse of an expTwoSynapse runs natively; others take the generic path."""
        if _tag(el) != "poissonFiringSynapse":
            return False
        return _tag(doc.items.get(el.get("synapse"), (None, ET.Element("none")))[1]) == "expTwoSynapse"

    input_elems = [(i, doc.items[i][1]) for role, i in doc.order
                   if role == "input" and _tag(doc.items[i][1]) != "pulseGenerator" and
                   not native_poisson(doc.items[i][1])]
    used_inputs = {el.get("input") for el in net if _tag(el) == "explicitInput"} | \
                  {el.get("component") for el in net if _tag(el) == "inputList"}
    lems_inputs = any(i in used_inputs for i, _ in input_elems)
    # (per-instance synapse settings take the generic path too: its parameters are per instance)
    generic = lems_inputs or any(s.kind == "synapse" for s in statements) or \
        any(doc.items.get(u, (None, None))[0] == "synapse" and
            (_tag(doc.items[u][1]) != "expTwoSynapse" or points) for u in used)
    generic_index, generic_types, generic_st, custom = {}, [], [], {}

    def generic_synapse(i, el):
        """(type name, SynapseType, engine values, weight property) of a synapse element."""
        t = _tag(el)
        weight = 1.0
        extra = {}
        if t == "blockingPlasticSynapse":
            # composed with its block and plasticity mechanisms (their parameters are the
            # child elements' attributes), one type per combination
            mech = {}
            for ch in el:
                ct = _tag(ch)
                if ct not in ("blockMechanism", "plasticityMechanism"):
                    continue
                table = lems_mod.BLOCK_MECHANISMS if ct == "blockMechanism" else lems_mod.PLASTICITY_MECHANISMS
                if ch.get("type") not in table:
                    raise NeuroMLError(f"synapse {i}: {ct} type '{ch.get('type')}' is not supported")
                mech[ct] = ch.get("type")
                for p_, d_ in table[ch.get("type")]["params"]:
                    if ch.get(p_) is None:
                        raise NeuroMLError(f"synapse {i}: {ct} parameter {p_} missing")
                    extra[p_] = convert(ch.get(p_), d_, f"{i} {p_}")
            t = f"blockingPlasticSynapse/{mech.get('blockMechanism', '')}/{mech.get('plasticityMechanism', '')}"
            if t not in custom:
                custom[t] = (lems_mod.blocking_plastic_type(mech.get("blockMechanism"),
                                                            mech.get("plasticityMechanism")), 1.0)
            st, weight = custom[t]
        elif t in lems_mod.SYNAPSE_TYPES:
            st = lems_mod.SYNAPSE_TYPES[t]
        else:
            if t not in custom:
                try:
                    custom[t] = lems_mod.synapse_type_from_lems(t, doc.component_types, convert)
                except lems_mod.LemsError as e:
                    raise NeuroMLError(str(e)) from None
            st, weight = custom[t]
        vals = {p: convert(el.get(p), d, f"{i} {p}") for p, d in st.params if el.get(p) is not None}
        vals.update(extra)
        return t, st, vals, weight

    timed_times = {}   # timedSynapticInput id -> its spike times (us, listed order)
    spikers = {}       # generic Poisson input id -> (averageRate 1/us, transient, delay us, duration us)
    if generic:
        for i, el in syn_elems:
            t, st, vals, weight = generic_synapse(i, el)
            generic_index[i] = len(generic_types)
            generic_types.append((t, vals, weight))
            generic_st.append(st)
        # LEMS current inputs after the synapse types (their instances come last in a
        # compartment's list)
        for i, el in input_elems:
            t = _tag(el)
            if t in _FIRING_INPUTS:
                # a spiker driving its own instance of a synapse type, whose current joins the
                # input total (eden/Eden.cpp:3859-3912): the synapse's type as an input type
                sid = el.get("synapse")
                role, sel = doc.items.get(sid, (None, None))
                if role != "synapse":
                    raise NeuroMLError(f"{t} {i}: synapse '{sid}' not found or not supported")
                st_name, st, vals, weight = generic_synapse(sid, sel)
                key = f"{t}/{st_name}"
                if key not in custom:
                    custom[key] = (dataclasses.replace(st, is_input=True), weight)
                if t == "timedSynapticInput":
                    timed_times[i] = [_f32(convert(sp.get("time"), "time", f"timedSynapticInput {i} spike"))
                                      for sp in el if _tag(sp) == "spike"]
                else:
                    tr_ = t == "transientPoissonFiringSynapse"
                    spikers[i] = (_f32(convert(el.get("averageRate"), "frequency", f"{i} averageRate")), int(tr_),
                                  _f32(convert(el.get("delay"), "time", f"{i} delay")) if tr_ else 0.0,
                                  _f32(convert(el.get("duration"), "time", f"{i} duration")) if tr_ else 0.0)
                generic_index[i] = len(generic_types)
                generic_types.append((key, vals, weight))
                generic_st.append(custom[key][0])
                continue
            if t in lems_mod.INPUT_TYPES:
                st, weight = lems_mod.INPUT_TYPES[t], 1.0
            else:
                if t not in custom:
                    try:
                        custom[t] = lems_mod.synapse_type_from_lems(t, doc.component_types, convert)
                    except lems_mod.LemsError as e:
                        raise NeuroMLError(str(e)) from None
                st, weight = custom[t]
            vals = {p: convert(el.get(p), d, f"{i} {p}") for p, d in st.params if el.get(p) is not None}
            generic_index[i] = len(generic_types)
            generic_types.append((t, vals, weight))
            generic_st.append(st)

    def synapse(sid, what, native=False):
        role, el = doc.items.get(sid, (None, None))
        if role == "unsupported":
            raise NeuroMLError(f"{what}: synapse type <{_tag(el)}> is not supported")
        if role != "synapse":
            raise NeuroMLError(f"{what}: synapse '{sid}' not found")
        if native or not generic:
            if sid not in native_index:
                raise NeuroMLError(f"{what}: synapse '{sid}' ({_tag(el)}) needs the generic synapse path; "
                                   "Poisson inputs drive expTwoSynapse only")
            return native_index[sid]
        return generic_index[sid]

    # network inputs in order (explicitInput and inputList entries as they appear)
    inputs = []    # (component id, compartment, weight)
    conns = []     # (pre compartment, post compartment, synapse type, weight, delay us)
    # what per-instance settings address: input lists' inputs (id, index into inputs) and
    # projections' connections (id, post end's and pre end's index into conns or -1, the
    # synapse component on each end)
    input_lists, projections = {}, {}
    for el in net:
        t = _tag(el)
        if t == "explicitInput":
            pop, inst = _cell_ref(el.get("target"))
            inputs.append((el.get("input"), comp(pop, inst, int(el.get("segmentId", "0")), "explicitInput",
                                                 frac(el, "fractionAlong")), 1.0))
        elif t == "inputList":
            pop = el.get("population")
            items = input_lists.setdefault(el.get("id"), (el.get("component"), []))[1]
            for i in el:
                it = _tag(i)
                if it not in ("input", "inputW"):
                    continue
                items.append((int(i.get("id", len(items))), len(inputs)))
                _, inst = _cell_ref(i.get("target"))
                w = convert(i.get("weight"), None, "inputW weight") if it == "inputW" else 1.0
                inputs.append((el.get("component"), comp(pop, inst, int(i.get("segmentId", "0")), "input",
                                                         frac(i, "fractionAlong")), w))
        elif t == "projection":
            pre, post = el.get("presynapticPopulation"), el.get("postsynapticPopulation")
            st = synapse(el.get("synapse"), f"projection {el.get('id')}")
            items = projections.setdefault(el.get("id"), ("spiking", []))[1]
            for c in el:
                ct = _tag(c)
                if ct not in ("connection", "connectionWD"):
                    continue
                items.append((int(c.get("id", len(items))), len(conns), -1, el.get("synapse"), None))
                _, pi = _cell_ref(c.get("preCellId"))
                _, qi = _cell_ref(c.get("postCellId"))
                a = comp(pre, pi, int(c.get("preSegmentId", "0")), "connection", frac(c, "preFractionAlong"))
                b = comp(post, qi, int(c.get("postSegmentId", "0")), "connection", frac(c, "postFractionAlong"))
                w = convert(c.get("weight"), None, "weight") if ct == "connectionWD" else 1.0
                d = convert(c.get("delay"), "time", "delay") if ct == "connectionWD" else 0.0
                conns.append((a, b, st, w, d, -1))
        elif t == "synapticConnection":
            st = synapse(el.get("synapse"), "synapticConnection")
            fp, fi = _cell_ref(el.get("from"))
            tp, ti = _cell_ref(el.get("to"))
            conns.append((comp(fp, fi, 0, "synapticConnection"), comp(tp, ti, 0, "synapticConnection"), st, 1.0, 0.0,
                          -1))
        elif t == "electricalProjection":
            # gap junctions: an instance on each end, the postsynaptic end's first, each with
            # the other end as its peer (eden/Eden.cpp:9225-9248)
            pre, post = el.get("presynapticPopulation"), el.get("postsynapticPopulation")
            items = projections.setdefault(el.get("id"), ("electrical", []))[1]
            for c in el:
                ct = _tag(c)
                if ct not in ("electricalConnection", "electricalConnectionInstance", "electricalConnectionInstanceW"):
                    continue
                items.append((int(c.get("id", len(items))), len(conns), len(conns) + 1, c.get("synapse"), c.get("synapse")))
                what = f"electricalProjection {el.get('id')}"
                st = synapse(c.get("synapse"), what)
                if ct == "electricalConnection":
                    pi, qi = int(c.get("preCell")), int(c.get("postCell"))
                else:
                    _, pi = _cell_ref(c.get("preCell"))
                    _, qi = _cell_ref(c.get("postCell"))
                a = comp(pre, pi, int(c.get("preSegment", "0")), what, frac(c, "preFractionAlong"))
                b = comp(post, qi, int(c.get("postSegment", "0")), what, frac(c, "postFractionAlong"))
                w = convert(c.get("weight"), None, "weight") if ct == "electricalConnectionInstanceW" else 1.0
                conns.append((-1, b, st, w, 0.0, a))
                conns.append((-1, a, st, w, 0.0, b))
        elif t == "continuousProjection":
            # graded connections: the postsynaptic component on the post end, then the
            # presynaptic one on the pre end, each with the other end as its peer
            # (eden/Eden.cpp:9254-9270)
            pre, post = el.get("presynapticPopulation"), el.get("postsynapticPopulation")
            items = projections.setdefault(el.get("id"), ("continuous", []))[1]
            for c in el:
                ct = _tag(c)
                if ct not in ("continuousConnection", "continuousConnectionInstance", "continuousConnectionInstanceW"):
                    continue
                items.append((int(c.get("id", len(items))), len(conns), len(conns) + 1, c.get("postComponent"),
                              c.get("preComponent")))
                what = f"continuousProjection {el.get('id')}"
                s_pre = synapse(c.get("preComponent"), what)
                s_post = synapse(c.get("postComponent"), what)
                for sid in (c.get("preComponent"), c.get("postComponent")):
                    if _tag(doc.items[sid][1]) not in _GRADED:
                        raise NeuroMLError(f"{what}: component {sid} ({_tag(doc.items[sid][1])}) is not a graded "
                                           f"synapse ({', '.join(sorted(_GRADED))})")
                if ct == "continuousConnection":
                    pi, qi = int(c.get("preCell")), int(c.get("postCell"))
                else:
                    _, pi = _cell_ref(c.get("preCell"))
                    _, qi = _cell_ref(c.get("postCell"))
                a = comp(pre, pi, int(c.get("preSegment", "0")), what, frac(c, "preFractionAlong"))
                b = comp(post, qi, int(c.get("postSegment", "0")), what, frac(c, "postFractionAlong"))
                w = convert(c.get("weight"), None, "weight") if ct == "continuousConnectionInstanceW" else 1.0
                conns.append((-1, b, s_post, w, 0.0, a))
                conns.append((-1, a, s_pre, w, 0.0, b))

    # LEMS point neurons: one type per cell component, each cell's type (-1: physical)
    point_types, cell_point_type = None, None
    if points:
        order, cell_point_type = {}, np.full(n_cells, -1, np.int32)
        for c in range(n_cells):
            cell = pop_of_cell[c].cell
            if cell.point is not None:
                cell_point_type[c] = order.setdefault(cell.id, len(order))
        by_id = {p.cell.id: p.cell for p in pops.values()}
        ids = sorted(order, key=order.get)
        try:
            point_types = cells_mod.build([(by_id[i].id, by_id[i].point[1]) for i in ids],
                                          {by_id[i].id: by_id[i].point[2] for i in ids})
        except lems_mod.LemsError as e:
            raise NeuroMLError(str(e)) from None

    # per-compartment tables
    v0 = np.concatenate([pop_of_cell[c].cell.v0 for c in range(n_cells)])
    cap = np.concatenate([pop_of_cell[c].cell.cap for c in range(n_cells)])
    area = np.concatenate([pop_of_cell[c].cell.area for c in range(n_cells)])
    inv = np.concatenate([pop_of_cell[c].cell.inv_rc for c in range(n_cells)])
    rax = np.concatenate([pop_of_cell[c].cell.r_axial for c in range(n_cells)])
    parent = np.concatenate([pop_of_cell[c].cell.parent for c in range(n_cells)])
    vthr_all = np.concatenate([pop_of_cell[c].cell.vthr for c in range(n_cells)])
    ch_t, ch_g, ch_e, ch_n = [], [], [], []
    for c in range(n_cells):
        for lst in pop_of_cell[c].cell.channels:
            ch_n.append(len(lst))
            for (ch, g, e, _) in lst:
                ch_t.append(channel_index[ch])
                ch_g.append(g)
                ch_e.append(e)
    ch_b = np.zeros(nc + 1, np.int64)
    np.cumsum(ch_n, out=ch_b[1:])

    # inputs
    pulses = [[] for _ in range(nc)]
    poisson = [[] for _ in range(nc)]
    poi_count = {}    # (cell, input component) -> instances so far (the RNG stamp's instance)
    input_where = []  # per input: ("conn", index into conns), ("pulse", g, j) or ("poisson", g, j)
    timed_conn = {}   # index into conns of a timedSynapticInput instance -> its component
    spiker_conn = {}  # index into conns of a generic Poisson input instance -> (component, RNG instance)
    for (iid, g, w) in inputs:
        role, el = doc.items.get(iid, (None, None))
        if role != "input":
            raise NeuroMLError(f"input '{iid}' not found or not supported (pulseGenerator, poissonFiringSynapse, "
                               f"{', '.join(sorted(lems_mod.INPUT_TYPES))}, user input ComponentTypes)")
        t = _tag(el)
        if iid in generic_index:
            # a LEMS input: an instance in the compartment's synapse list
            if iid in timed_times:
                timed_conn[len(conns)] = iid
            if iid in spikers:
                # its RNG instance: the earlier instances of the same input on the cell
                cell = int(np.searchsorted(cell_begin, g, side="right") - 1)
                k = poi_count.get((cell, iid), 0)
                poi_count[(cell, iid)] = k + 1
                spiker_conn[len(conns)] = (iid, k)
            input_where.append(("conn", len(conns)))
            conns.append((-1, g, generic_index[iid], _f32(w), 0.0, -1))
        elif t == "pulseGenerator":
            input_where.append(("pulse", g, len(pulses[g])))
            pulses[g].append((_f32(convert(el.get("delay"), "time", "delay")),
                              _f32(convert(el.get("duration"), "time", "duration")),
                              _f32(convert(el.get("amplitude"), "current", "amplitude")), _f32(w)))
        else:
            cell = int(np.searchsorted(cell_begin, g, side="right") - 1)
            k = poi_count.get((cell, iid), 0)
            poi_count[(cell, iid)] = k + 1
            rate = _f32(convert(el.get("averageRate"), "frequency", "averageRate"))
            input_where.append(("poisson", g, len(poisson[g])))
            poisson[g].append((rate, synapse(el.get("synapse"), f"poissonFiringSynapse {iid}", native=True), _f32(w), k))
    # per-instance settings, in statement order (native inputs' tables directly; generic
    # synapse instances and point neurons get their own parameter vectors below)
    syn_over, cell_over, spiker_over = {}, {}, {}
    for st_ in statements:
        _apply_setting(st_, doc, pops, input_lists, projections, input_where, pulses, poisson, generic_index,
                       generic_st, syn_over, cell_over, spiker_over)

    pu_b = np.zeros(nc + 1, np.int64)
    np.cumsum([len(x) for x in pulses], out=pu_b[1:])
    pu = [x for lst in pulses for x in lst]
    po_b = np.zeros(nc + 1, np.int64)
    np.cumsum([len(x) for x in poisson], out=po_b[1:])
    po = [x for lst in poisson for x in lst]

    # synapse instances: per compartment, by synapse type (reading order), then connection order
    per_comp = [[] for _ in range(nc)]
    for k, (a, b, st, w, d, peer) in enumerate(conns):
        per_comp[b].append((st, k))
    syn_b = np.zeros(nc + 1, np.int64)
    inst_of_conn = np.zeros(len(conns), np.int64)
    syn_list = []
    for g in range(nc):
        for (st, k) in sorted(per_comp[g], key=lambda x: x[0]):   # stable: connection order within a type
            inst_of_conn[k] = len(syn_list)
            syn_list.append(k)
        syn_b[g + 1] = len(syn_list)
    spk = [[] for _ in range(nc)]
    for k, (a, b, st, w, d, peer) in enumerate(conns):
        if a >= 0:
            spk[a].append(inst_of_conn[k])
    spk_b = np.zeros(nc + 1, np.int64)
    np.cumsum([len(x) for x in spk], out=spk_b[1:])

    # outputs
    base_dir = os.path.dirname(os.path.abspath(path))
    trajectories, events = [], []
    rec, evc, rec_gate = [], [], []

    def gate_variable(cell, g, gate, what):
        """The compartment's gate variable a gate path names: the first channel distribution
        of that id on the compartment (eden/NeuroML.cpp:1843-1853), the gate's place in its
        channel, counted after the gates of the compartment's earlier distributions."""
        bioph, dist, chan, gid = gate
        if bioph not in (cell.bioph, "biophysicalProperties"):
            raise NeuroMLError(f"{what}: cell {cell.id} has no biophysicalProperties '{bioph}'")
        lst = cell.channels[g - int(cell_begin[np.searchsorted(cell_begin, g, side="right") - 1])]
        j = 0
        for (ch, _, _, did) in lst:
            if did == dist:
                if ch != chan:
                    raise NeuroMLError(f"{what}: distribution {dist} holds ion channel {ch}, not {chan}")
                if gid not in gate_ids[ch]:
                    raise NeuroMLError(f"{what}: ion channel {ch} has no gate {gid}")
                return j + gate_ids[ch].index(gid)
            j += len(gate_ids[ch])
        raise NeuroMLError(f"{what}: no channel distribution {dist} on that segment of cell {cell.id}")

    for el in sim:
        t = _tag(el)
        if t == "OutputFile":
            f = OutputFileSpec(el.get("id"), el.get("fileName"))
            for c in el:
                if _tag(c) == "OutputColumn":
                    pop, inst, seg, gate = _quantity_path(c.get("quantity"))
                    what = f"OutputColumn {c.get('id')}"
                    g = comp(pop, inst, seg, what)
                    f.columns.append((c.get("id"), c.get("quantity"), len(rec)))
                    rec.append(g)
                    pc = pops[pop].cell
                    if pc.point is not None:
                        ct = pc.point[2]
                        names = [x for x, _ in ct.state]
                        var = "v" if gate is None else gate[1] if gate[0] == "state" else None
                        if var == "v" and ct.v_state is None:
                            raise NeuroMLError(f"{what}: cell type {ct.name} has no membrane voltage v")
                        if var != "v" and var not in names:
                            raise NeuroMLError(f"{what}: cell type {ct.name} has no state variable {var or gate}")
                        rec_gate.append(-1 if var == "v" else names.index(var))
                    elif gate is not None and gate[0] == "state":
                        raise NeuroMLError(f"{what}: '{gate[1]}' is not a quantity of a physical cell")
                    else:
                        rec_gate.append(-1 if gate is None else gate_variable(pc, g, gate, what))
            trajectories.append(f)
        elif t == "EventOutputFile":
            fmt = el.get("format", "TIME_ID")
            if fmt not in ("TIME_ID", "ID_TIME"):
                raise NeuroMLError(f"EventOutputFile {el.get('id')}: format '{fmt}' is not supported")
            f = EventFileSpec(el.get("id"), el.get("fileName"), fmt)
            for c in el:
                if _tag(c) == "EventSelection":
                    port = c.get("eventPort", "spike")
                    if port != "spike":
                        raise NeuroMLError(f"EventSelection {c.get('id')}: only the 'spike' port is supported")
                    pop, inst, seg = _select_path(c.get("select"))
                    g = comp(pop, inst, seg, f"EventSelection {c.get('id')}")
                    f.selections.append((int(c.get("id")), c.get("select"), len(evc)))
                    evc.append(g)
            events.append(f)

    # spike thresholds only where spikes are read: event columns and spike sources
    vthr = np.full(nc, np.nan, np.float32)
    used = np.zeros(nc, bool)
    used[np.asarray(evc, np.int64)] = True
    used[np.diff(spk_b) > 0] = True
    vthr[used] = vthr_all[used]

    lems_tables = lems_mod.build(generic_types, {k: v[0] for k, v in custom.items()}) if generic else None
    syn_spike_list, spike_times = None, None
    if timed_conn:
        # each timedSynapticInput's spike list once, ended by +inf; its instances start there
        start, times = {}, []
        for iid in sorted(set(timed_conn.values()), key=list(timed_times).index):
            start[iid] = len(times)
            times += timed_times[iid] + [np.inf]
        syn_spike_list = np.full(len(syn_list), -1, np.int64)
        for i, k in enumerate(syn_list):
            if k in timed_conn:
                syn_spike_list[i] = start[timed_conn[k]]
        spike_times = np.array(times, np.float32)
    syn_poisson_spiker, poisson_spikers = None, None
    if spiker_conn:
        syn_poisson_spiker = np.full(len(syn_list), -1, np.int64)
        rows = []
        for i, k in enumerate(syn_list):
            if k in spiker_conn:
                iid, inst = spiker_conn[k]
                rate, tr_, delay, duration = spikers[iid]
                over = spiker_over.get(k, {})
                rate = _f32(over.get("averageRate", rate))
                delay = _f32(over.get("delay", delay))
                duration = _f32(over.get("duration", duration))
                syn_poisson_spiker[i] = len(rows)
                rows.append((rate, tr_, delay, duration, inst, 0))
        poisson_spikers = {name: np.array([r[j] for r in rows], dt) for j, (name, dt) in enumerate(
            (("rate", np.float32), ("transient", np.int32), ("delay", np.float32), ("duration", np.float32),
             ("instance", np.int32), ("rng_offset", np.int32)))}
    syn_pidx, syn_vecs = None, None
    if syn_over:
        syn_pidx, syn_vecs = np.full(len(syn_list), -1, np.int64), []
        for i, k in enumerate(syn_list):
            if k in syn_over:
                ty = conns[k][2]
                names = [p for p, _ in generic_st[ty].params]
                vec = np.array(lems_tables.params[ty], np.float32)
                for p, v in syn_over[k].items():
                    vec[names.index(p)] = _f32(v)
                syn_pidx[i] = len(syn_vecs)
                syn_vecs.append(vec)
    cell_pidx, cell_vecs = None, None
    if cell_over:
        cell_pidx, cell_vecs = np.full(n_cells, -1, np.int64), []
        for c in sorted(cell_over):
            ty = int(cell_point_type[c])
            names = [p for p, _ in pop_of_cell[c].cell.point[2].params]
            vec = np.array(point_types.params[ty], np.float32)
            for p, v in cell_over[c].items():
                vec[names.index(p)] = _f32(v)
            cell_pidx[c] = len(cell_vecs)
            cell_vecs.append(vec)

    f32 = lambda xs, i: np.array([x[i] for x in xs], np.float32)
    model = Model(
        dt=_f32(dt), t_final=float(length), cell_comp_begin=cell_begin,
        v0=v0, capacitance=cap, area=area, inv_rc_axial=inv, r_axial=rax, v_threshold=vthr, parent=parent,
        channel_types=channel_types, comp_chan_begin=ch_b, chan_type=np.array(ch_t, np.int32),
        chan_gbase=np.array(ch_g, np.float32), chan_erev=np.array(ch_e, np.float32),
        cell_rng_seed=cell_rng_seeds(seed, n_cells),
        comp_pulse_begin=pu_b, pulse_start=f32(pu, 0), pulse_duration=f32(pu, 1), pulse_amplitude=f32(pu, 2),
        pulse_weight=f32(pu, 3),
        syn_types=syn_types, comp_syn_begin=syn_b,
        syn_type=np.array([conns[k][2] for k in syn_list], np.int32),
        syn_weight=np.array([_f32(conns[k][3]) for k in syn_list], np.float32),
        syn_delay=np.array([_f32(conns[k][4]) for k in syn_list], np.float32),
        syn_peer=(np.array([conns[k][5] for k in syn_list], np.int64) if any(c[5] >= 0 for c in conns) else None),
        comp_spike_begin=spk_b, spike_target=np.array([x for lst in spk for x in lst], np.int64),
        comp_poisson_begin=po_b, poisson_rate=f32(po, 0), poisson_syn_type=np.array([x[1] for x in po], np.int32),
        poisson_weight=f32(po, 2), poisson_instance=np.array([x[3] for x in po], np.int32),
        poisson_rng_offset=np.zeros(len(po), np.int32),
        record_comp=np.array(rec, np.int64), event_comp=np.array(evc, np.int64),
        record_gate=np.array(rec_gate, np.int32) if any(j >= 0 for j in rec_gate) else None,
        lems=lems_tables, lems_cells=point_types, cell_lems_type=cell_point_type,
        gate_kind=np.array(gate_kinds, np.int32) if any(gate_kinds) else None,
        gate_extra_rate=gate_extras if any(x is not None for x in gate_extras) else None,
        chan_scaling=np.array(chan_scaling, np.float32) if any(x != 1.0 for x in chan_scaling) else None,
        lems_rates=_rates_or_none(rate_entries, temperature),
        syn_param_index=syn_pidx, syn_param_vecs=syn_vecs, cell_param_index=cell_pidx, cell_param_vecs=cell_vecs,
        syn_spike_list=syn_spike_list, spike_times=spike_times,
        syn_poisson_spiker=syn_poisson_spiker, poisson_spikers=poisson_spikers,
        name=os.path.splitext(os.path.basename(path))[0])
    model.sim_seed = seed
    return model, Outputs(trajectories, events, base_dir)
