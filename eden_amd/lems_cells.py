"""LEMS point neurons (the reference's artificial cells): a restatement of the NeuroML core
cell ComponentTypes, user-defined cell <ComponentType>s, and an emitter of their per-step
update as C / HIP functions that the library compiles with hiprtc beside the generic synapse
functions (include/edenhip.h `edenhip_lems_cells`); the oracle compiles the C dialect.

One step of one cell (the reference's artificial-cell work item, eden/Eden.cpp:6603-6910,
with the LEMS emission of eden/Eden.cpp:2600-2995):
  - the external current: the cell's synapse currents, then its input currents, summed
    (I_synapses_total + I_input_total), is the requirement iSyn (or ISyn);
  - derived parameters, derived variables from the present state in dependency order
    (conditional ones: first case whose condition holds, else the default case, else 0);
  - on an initialisation step the OnStart assignments; otherwise next = state + dt * derivative
    for every state variable with a time derivative, then each OnCondition whose test holds
    runs its assignments (each sets the present and the next value, so later tests and
    assignments read it) and raises the spike output if it has an EventOut;
  - a spike reaches the cell's recipients as a physical cell's does.

Engine units follow the reference's native unit per dimension (eden/NeuroML.cpp:300-487: us,
mV, nA, uS, pF, ...). A product or quotient whose operands' native units do not multiply to
the result's native unit is scaled by the power of ten between them, as the reference's
ExpressionInfix does (eden/Eden.cpp:2341-2515), and so is a time derivative whose unit times
the native time unit is not its variable's (eden/Eden.cpp:2932-2944); dimensions without a
native unit of their own are held in SI units.

The core table restates the dynamics of the NeuroML 2 standard's cell types (its Cells.xml and
PyNN.xml definitions, as embedded by the reference in
eden/neuroml/LEMS_CoreComponents.inc.xml:957-1833), with the magic parameters, constants and
requirements the reference adds for their base types (eden/NeuroML.cpp:8456-8505, 8688-8745).
"""
from dataclasses import dataclass, field

from .lems import _FUNCS, LemsError, _lit, _ltag, parse

# ---------------------------------------------------------------- dimensions

# LEMS dimension name -> (M, L, T, I, K, N, J exponents, power of ten of the native unit);
# the first name of a vector is its canonical one
DIMENSIONS = {
    "none": ((0, 0, 0, 0, 0, 0, 0), 0),
    "time": ((0, 0, 1, 0, 0, 0, 0), -6),
    "per_time": ((0, 0, -1, 0, 0, 0, 0), 6),
    "voltage": ((1, 2, -3, -1, 0, 0, 0), -3),
    "conductance": ((-1, -2, 3, 2, 0, 0, 0), -6),
    "conductanceDensity": ((-1, -4, 3, 2, 0, 0, 0), 1),
    "capacitance": ((-1, -2, 4, 2, 0, 0, 0), -12),
    "specificCapacitance": ((-1, -4, 4, 2, 0, 0, 0), -2),
    "resistance": ((1, 2, -3, -2, 0, 0, 0), 6),
    "resistivity": ((1, 3, -3, -2, 0, 0, 0), 1),
    "current": ((0, 0, 0, 1, 0, 0, 0), -9),
    "length": ((0, 1, 0, 0, 0, 0, 0), -6),
    "concentration": ((0, -3, 0, 0, 0, 1, 0), -3),
    "permeability": ((0, 1, -1, 0, 0, 0, 0), 0),
    "temperature": ((0, 0, 0, 0, 1, 0, 0), 0),
    "rho_factor": ((0, -1, -1, -1, 0, 1, 0), 0),
    "per_voltage": ((-1, -3, 3, 1, 0, 0, 0), 3),
    "charge": ((0, 0, 1, 1, 0, 0, 0), -12),
    "charge_per_mole": ((0, 0, 1, 1, 0, -1, 0), 0),
    "currentDensity": ((0, -2, 0, 1, 0, 0, 0), -2),
    "area": ((0, 2, 0, 0, 0, 0, 0), -12),
    "volume": ((0, 3, 0, 0, 0, 0, 0), -18),
    "substance": ((0, 0, 0, 0, 0, 1, 0), -18),
    "idealGasConstantDims": ((1, 2, -2, 0, -1, -1, 0), 0),
    "conductance_per_voltage": ((-2, -4, 6, 3, 0, 0, 0), -6),
    "voltage_per_time": ((1, 2, -4, -1, 0, 0, 0), 3),
}
DIMENSIONS["unitless"] = DIMENSIONS["none"]
_NATIVE = {}
for _n, (_v, _p) in DIMENSIONS.items():
    _NATIVE.setdefault(_v, _p)
UNITY = DIMENSIONS["none"][0]


def dim(name):
    if name not in DIMENSIONS:
        raise LemsError(f"dimension '{name}' is not supported")
    return DIMENSIONS[name][0]


def native(vec):
    """Power of ten of the native unit of a dimension vector (SI for unnamed ones)."""
    return _NATIVE.get(vec, 0)


def _mul(a, b, sign=1):
    return tuple(x + sign * y for x, y in zip(a, b))


def infix_dim(tree, names, dims):
    """(C text, dimension vector) of an expression tree: `names` maps symbols to C names,
    `dims` symbols to dimension vectors. Products and quotients carry the power-of-ten
    factor between their operands' native units and the result's."""
    k = tree[0]
    if k == "num":
        return _lit(tree[1]), UNITY
    if k == "sym":
        s = tree[1]
        if s not in names:
            raise LemsError(f"unknown symbol {s}")
        return names[s], dims[s]
    if k == "neg":
        t, d = infix_dim(tree[1], names, dims)
        return "( -" + t + " )", d
    if k == "fn" and tree[1] == "random":
        # random(x): the reference's randof(x, rng_object_id, instance, step, rng_offset + n)
        # with n counting the calls in emission order (eden/Eden.cpp:2450-2459)
        if "__rng__" not in names:
            raise LemsError("random() is not supported here")
        t, d = infix_dim(tree[2], names, dims)
        n = names["__rng__"][0]
        names["__rng__"][0] += 1
        return f"lems_randof( {t} , rng_seed, rng_step, {n} )", d
    if k == "fn":
        t, d = infix_dim(tree[2], names, dims)
        fn = "stepf" if tree[1] == "H" else _FUNCS[tree[1]]
        return fn + "( " + t + " )", d
    if k == "pow":
        a, _ = infix_dim(tree[1], names, dims)
        b, _ = infix_dim(tree[2], names, dims)
        return "POW( " + a + " , " + b + " )", UNITY
    if k == "cmp":
        a, _ = infix_dim(tree[2], names, dims)
        b, _ = infix_dim(tree[3], names, dims)
        return "( ( " + a + " " + tree[1] + " " + b + " ) )", UNITY
    if k == "not":
        a, _ = infix_dim(tree[1], names, dims)
        return "( !" + a + " )", UNITY
    if k in ("and", "or"):
        a, _ = infix_dim(tree[1], names, dims)
        b, _ = infix_dim(tree[2], names, dims)
        return "( ( " + a + (" && " if k == "and" else " || ") + b + " ) )", UNITY
    a, da = infix_dim(tree[1], names, dims)
    b, db = infix_dim(tree[2], names, dims)
    if k in ("+", "-"):
        return "( ( " + a + " " + k + " " + b + " ) )", db
    out = _mul(da, db, 1 if k == "*" else -1)
    f = native(da) + (native(db) if k == "*" else -native(db)) - native(out)
    suffix = f" * 1e{f}f" if f else ""
    return "( ( " + a + " " + k + " " + b + " )" + suffix + " )", out


# ---------------------------------------------------------------- cell types

@dataclass
class CellType:
    """A point-neuron ComponentType: `params` [(name, dimension)], `constants` [(name,
    dimension, engine value)], `derived_params` [(name, expression)], `state` [(name,
    dimension)], `derived` [(name, dimension, expression or [(condition or None, expression),
    ...])] in dependency order, `derivatives` {state: expression}, `on_start` [(state,
    expression)], `on_conditions` [(test, [(state, expression), ...], spike)], `requirements`
    {name: dimension} (t, iSyn / ISyn), `v_state` the state variable exposed as the membrane
    voltage v (or None), `spikes` whether it has a spike output."""
    name: str
    params: list
    state: list
    derived: list
    derivatives: dict
    on_start: list
    on_conditions: list
    requirements: dict
    constants: list = field(default_factory=list)
    derived_params: list = field(default_factory=list)
    v_state: str = None
    spikes: bool = True

    @property
    def v_index(self):
        names = [s for s, _ in self.state]
        return names.index(self.v_state) if self.v_state in names else -1


_V = "voltage"
# the reference's additions per base cell type: (parameters, constants, requirements)
_PYNN = ([("cm", "none"), ("i_offset", "none"), ("v_init", "none"), ("tau_syn_E", "none"), ("tau_syn_I", "none")],
         [("MSEC", "time", 1000.0), ("MVOLT", _V, 1.0), ("NFARAD", "capacitance", 1000.0)],
         {"t": "time", "iSyn": "current"})
_PYNN_IAF = ([*_PYNN[0], ("tau_refrac", "none"), ("v_thresh", "none"), ("tau_m", "none"), ("v_rest", "none"),
              ("v_reset", "none")], _PYNN[1], _PYNN[2])
BASE_CELLS = {
    "baseCell": ([], [], {"t": "time"}),
    "baseSpikeSource": ([], [], {"t": "time"}),
    "baseSpikingCell": ([], [], {"t": "time"}),
    "baseCellMembPot": ([], [], {"t": "time"}),
    "baseCellMembPotDL": ([], [], {"t": "time"}),
    "baseCellMembPotCap": ([("C", "capacitance")], [], {"t": "time", "iSyn": "current"}),
    "baseIaf": ([("thresh", _V), ("reset", _V)], [], {"t": "time"}),
    "baseIafCapCell": ([("thresh", _V), ("reset", _V), ("C", "capacitance")], [], {"t": "time", "iSyn": "current"}),
    "basePyNNCell": _PYNN,
    "basePyNNIaFCell": _PYNN_IAF,
    "basePyNNIaFCondCell": ([*_PYNN_IAF[0], ("e_rev_E", "none"), ("e_rev_I", "none")], _PYNN[1], _PYNN[2]),
}


def _cell(name, base, params=(), state=(), derived=(), ders=None, on_start=(), on_cond=(), consts=(), dparams=(),
          reqs=None, v="v", spikes=True):
    bp, bc, br = BASE_CELLS[base]
    r = dict(br)
    r.update(reqs or {})
    return CellType(name=name, params=list(bp) + list(params), state=list(state), derived=list(derived),
                    derivatives=dict(ders or {}), on_start=list(on_start), on_conditions=list(on_cond),
                    requirements=r, constants=list(bc) + list(consts), derived_params=list(dparams),
                    v_state=v if v in [s for s, _ in state] else None, spikes=spikes)


_REFRACT = "t .gt. lastSpikeTime + refract"
# the Poisson spike sources' state and the time their next spike is used at: the ideal time,
# or just after the present time when that has passed (dt permitting)
_POISSON_SOURCE_STATE = [("tsince", "time"), ("tnextIdeal", "time"), ("tnextUsed", "time"), ("isi", "time")]
_POISSON_SOURCE_DERS = {"tsince": "1", "tnextUsed": "0", "tnextIdeal": "0"}
_TNEXT_USED = "tnextIdeal*H( (tnextIdeal-t)/t ) + (t+SMALL_TIME)*H( (t-tnextIdeal)/t )"
_PYNN_REFRACT = "t .gt. lastSpikeTime + (tau_refrac*MSEC)"
_PYNN_RATE = "(MVOLT * ((i_offset/cm) +  ((v_rest - (v/MVOLT)) / tau_m))/MSEC) + (iSyn / (cm * NFARAD))"


def _pynn_iaf(name, base):
    return _cell(
        name, base, state=[("v", _V), ("lastSpikeTime", "time")],
        derived=[("voltage_rate_active", "current", _PYNN_RATE),
                 ("voltage_rate", "voltage_per_time", [(_PYNN_REFRACT, "voltage_rate_active"), (None, "0")])],
        ders={"v": "voltage_rate"},
        on_start=[("v", "v_init * MVOLT"), ("lastSpikeTime", "t - (tau_refrac*MSEC)")],
        on_cond=[(f"(v .gt. (v_thresh * MVOLT)) .and. ({_PYNN_REFRACT})",
                  [("lastSpikeTime", "t"), ("v", "v_reset * MVOLT")], True)])


def _pynn_eif(name):
    return _cell(
        name, "basePyNNIaFCondCell",
        params=[("v_spike", "none"), ("delta_T", "none"), ("tau_w", "none"), ("a", "none"), ("b", "none")],
        dparams=[("eif_threshold", "v_spike * H(delta_T-1e-12) + v_thresh * H(-1*delta_T+1e-9)")],
        state=[("v", _V), ("w", "none"), ("lastSpikeTime", "time")],
        derived=[("delta_I", "none", [("delta_T .gt. 0", "delta_T * exp(((v / MVOLT) - v_thresh) / delta_T)"),
                                      ("delta_T .eq. 0", "0")]),
                 ("voltage_rate_active", "current",
                  "(MVOLT * ((-1 * ((v / MVOLT) - v_rest) + delta_I) / tau_m + (i_offset - w) / cm) / MSEC)"
                  " + (iSyn / (cm * NFARAD))"),
                 ("voltage_rate", "voltage_per_time", [(_PYNN_REFRACT, "voltage_rate_active"), (None, "0")])],
        ders={"v": "voltage_rate", "w": "(1 / tau_w) * (a * ((v / MVOLT) - v_rest) - w) / MSEC"},
        on_start=[("v", "v_init * MVOLT"), ("w", "0"), ("lastSpikeTime", "t - (tau_refrac*MSEC)")],
        on_cond=[(f"(v .gt. eif_threshold * MVOLT) .and. ({_PYNN_REFRACT})",
                  [("lastSpikeTime", "t"), ("v", "v_reset * MVOLT"), ("w", "w+b")], True)])


CELL_TYPES = {
    "iafTauCell": _cell(
        "iafTauCell", "baseIaf", params=[("leakReversal", _V), ("tau", "time")], state=[("v", _V)],
        ders={"v": "(leakReversal - v) / tau"}, on_start=[("v", "leakReversal")],
        on_cond=[("v .gt. thresh", [("v", "reset")], True)]),
    "iafTauRefCell": _cell(
        "iafTauRefCell", "baseIaf", params=[("leakReversal", _V), ("tau", "time"), ("refract", "time")],
        state=[("v", _V), ("lastSpikeTime", "time")],
        derived=[("voltage_rate", "voltage_per_time", [(_REFRACT, "(leakReversal - v) / tau"), (None, "0")])],
        ders={"v": "voltage_rate"}, on_start=[("v", "leakReversal"), ("lastSpikeTime", "t - refract")],
        on_cond=[(f"(v .gt. thresh) .and. ({_REFRACT})", [("lastSpikeTime", "t"), ("v", "reset")], True)]),
    "iafCell": _cell(
        "iafCell", "baseIafCapCell", params=[("leakConductance", "conductance"), ("leakReversal", _V)],
        state=[("v", _V)], derived=[("iMemb", "current", "leakConductance * (leakReversal - v) + iSyn")],
        ders={"v": "iMemb / C"}, on_start=[("v", "leakReversal")],
        on_cond=[("v .gt. thresh", [("v", "reset")], True)]),
    "iafRefCell": _cell(
        "iafRefCell", "baseIafCapCell",
        params=[("leakConductance", "conductance"), ("leakReversal", _V), ("refract", "time")],
        state=[("v", _V), ("lastSpikeTime", "time")],
        derived=[("iMemb", "current", "leakConductance * (leakReversal - v) + iSyn"),
                 ("voltage_rate", "voltage_per_time", [(_REFRACT, "iMemb / C"), (None, "0")])],
        ders={"v": "voltage_rate"}, on_start=[("v", "leakReversal"), ("lastSpikeTime", "t - refract")],
        on_cond=[(f"(v .gt. thresh) .and. ({_REFRACT})", [("lastSpikeTime", "t"), ("v", "reset")], True)]),
    "izhikevichCell": _cell(
        "izhikevichCell", "baseCellMembPot",
        params=[("v0", _V), ("a", "none"), ("b", "none"), ("c", "none"), ("d", "none"), ("thresh", _V)],
        consts=[("MSEC", "time", 1000.0), ("MVOLT", _V, 1.0)], reqs={"ISyn": "none"},
        state=[("v", _V), ("U", "none")],
        ders={"v": "(0.04 * (v*v / MVOLT) + 5 * v + (140.0 - U + ISyn) * MVOLT)/MSEC",
              "U": "a * (b * v / MVOLT - U) / MSEC"},
        on_start=[("v", "v0"), ("U", "v0 * b / MVOLT")],
        on_cond=[("v .gt. thresh", [("v", "c * MVOLT"), ("U", "U + d")], True)]),
    "izhikevich2007Cell": _cell(
        "izhikevich2007Cell", "baseCellMembPotCap",
        params=[("v0", _V), ("k", "conductance_per_voltage"), ("vr", _V), ("vt", _V), ("vpeak", _V),
                ("a", "per_time"), ("b", "conductance"), ("c", _V), ("d", "current")],
        state=[("v", _V), ("u", "current")],
        derived=[("iMemb", "current", "k * (v-vr) * (v-vt) + iSyn - u")],
        ders={"v": "iMemb / C", "u": "a * (b * (v-vr) - u)"}, on_start=[("v", "v0"), ("u", "0")],
        on_cond=[("v .gt. vpeak", [("v", "c"), ("u", "u + d")], True)]),
    "adExIaFCell": _cell(
        "adExIaFCell", "baseCellMembPotCap",
        params=[("gL", "conductance"), ("EL", _V), ("VT", _V), ("thresh", _V), ("reset", _V), ("delT", _V),
                ("tauw", "time"), ("refract", "time"), ("a", "conductance"), ("b", "current")],
        state=[("v", _V), ("w", "current"), ("lastSpikeTime", "time")],
        derived=[("iMemb", "current", "-1 * gL * (v - EL) + gL * delT * exp((v - VT) / delT) - w + iSyn"),
                 ("voltage_rate", "voltage_per_time", [(_REFRACT, "iMemb / C"), (None, "0")])],
        ders={"v": "voltage_rate", "w": "(a * (v - EL) - w) / tauw"},
        on_start=[("v", "EL"), ("w", "0"), ("lastSpikeTime", "t - refract")],
        on_cond=[(f"(v .gt. thresh) .and. ({_REFRACT})",
                  [("lastSpikeTime", "t"), ("v", "reset"), ("w", "w + b")], True)]),
    "fitzHughNagumoCell": _cell(
        "fitzHughNagumoCell", "baseCellMembPotDL", params=[("I", "none")], consts=[("SEC", "time", 1e6)],
        state=[("V", "none"), ("W", "none")],
        ders={"V": "( (V - ((V^3) / 3)) - W + I) / SEC", "W": "(0.08 * (V + 0.7 - 0.8 * W)) / SEC"}),
    "fitzHughNagumo1969Cell": _cell(
        "fitzHughNagumo1969Cell", "baseCellMembPotDL",
        params=[("a", "none"), ("b", "none"), ("I", "none"), ("phi", "none"), ("V0", "none"), ("W0", "none")],
        consts=[("TS", "time", 1000.0)], state=[("V", "none"), ("W", "none")],
        derived=[("F", "none", "V - V^3 / 3")],
        ders={"V": "(F - W + I) / TS", "W": "phi * (V + a - b * W) / TS"},
        on_start=[("V", "V0"), ("W", "W0")]),
    "hindmarshRose1984Cell": _cell(
        "hindmarshRose1984Cell", "baseCellMembPotCap",
        params=[("a", "none"), ("b", "none"), ("c", "none"), ("d", "none"), ("r", "none"), ("s", "none"),
                ("x1", "none"), ("v_scaling", _V), ("x0", "none"), ("y0", "none"), ("z0", "none")],
        consts=[("MSEC", "time", 1000.0)],
        state=[("v", _V), ("y", "none"), ("z", "none"), ("spiking", "none")],
        derived=[("x", "none", "v / v_scaling"), ("phi", "none", "y - a * x^3 + b * x^2"),
                 ("chi", "none", "c - d * x^2 - y"), ("rho", "none", "s * ( x - x1 ) - z"),
                 ("iMemb", "current", "(C * (v_scaling * (phi - z) / MSEC)) + iSyn")],
        ders={"v": "iMemb/C", "y": "chi / MSEC", "z": "r * rho / MSEC"},
        on_start=[("v", "x0 * v_scaling"), ("y", "y0"), ("z", "z0")],
        on_cond=[("v .gt. 0 .and. spiking .lt. 0.5", [("spiking", "1")], True),
                 ("v .lt. 0", [("spiking", "0")], False)]),
    "spikeGenerator": _cell(
        "spikeGenerator", "baseSpikeSource", params=[("period", "time")], consts=[("SMALL_TIME", "time", 1e-6)],
        state=[("tsince", "time"), ("tnext", "time")], ders={"tsince": "1", "tnext": "0"},
        on_start=[("tsince", "0"), ("tnext", "period")],
        on_cond=[("tnext-t .lt. SMALL_TIME", [("tsince", "0"), ("tnext", "tnext+period")], True)]),
    # Pinsky and Rinzel's two-compartment CA3 cell in one component
    # (eden/neuroml/LEMS_CoreComponents.inc.xml:1306-1442); its v is a derived variable, so the
    # reference gives it no membrane-voltage state (its Vs and Vd are recorded by name)
    "pinskyRinzelCA3Cell": _cell(
        "pinskyRinzelCA3Cell", "baseCellMembPot",
        params=[("iSoma", "currentDensity"), ("iDend", "currentDensity"), ("gLs", "conductanceDensity"),
                ("gLd", "conductanceDensity"), ("gNa", "conductanceDensity"), ("gKdr", "conductanceDensity"),
                ("gCa", "conductanceDensity"), ("gKahp", "conductanceDensity"), ("gKC", "conductanceDensity"),
                ("gc", "conductanceDensity"), ("eNa", _V), ("eCa", _V), ("eK", _V), ("eL", _V), ("pp", "none"),
                ("cm", "specificCapacitance"), ("alphac", "none"), ("betac", "none"), ("gNmda", "conductanceDensity"),
                ("gAmpa", "conductanceDensity"), ("qd0", "none")],
        consts=[("MSEC", "time", 1000.0), ("MVOLT", _V, 1.0), ("UAMP_PER_CM2", "currentDensity", 1.0),
                ("Smax", "none", 125.0), ("Vsyn", _V, 60.0), ("betaqd", "none", 0.001)],
        state=[("Vs", _V), ("Vd", _V), ("Cad", "none"), ("hs", "none"), ("ns", "none"), ("sd", "none"),
               ("cd", "none"), ("qd", "none"), ("Si", "none"), ("Wi", "none")],
        derived=[
            ("ICad", "currentDensity", "gCa*sd*sd*(Vd-eCa)"),
            ("alphams_Vs", "none", "0.32*(-46.9-Vs/MVOLT)/(exp((-46.9-Vs/MVOLT)/4.0)-1.0)"),
            ("betams_Vs", "none", "0.28*(Vs/MVOLT+19.9)/(exp((Vs/MVOLT+19.9)/5.0)-1.0)"),
            ("Minfs_Vs", "none", "alphams_Vs/(alphams_Vs+betams_Vs)"),
            ("alphans_Vs", "none", "0.016*(-24.9-Vs/MVOLT)/(exp((-24.9-Vs/MVOLT)/5.0)-1.0)"),
            ("betans_Vs", "none", "0.25*exp(-1.0-0.025*Vs/MVOLT)"),
            ("alphahs_Vs", "none", "0.128*exp((-43.0-Vs/MVOLT)/18.0)"),
            ("betahs_Vs", "none", "4.0/(1.0+exp((-20.0-Vs/MVOLT)/5.0))"),
            ("alphasd_Vd", "none", "1.6/(1.0+exp(-0.072*(Vd/MVOLT-5.0)))"),
            ("betasd_Vd", "none", "0.02*(Vd/MVOLT+8.9)/(exp((Vd/MVOLT+8.9)/5.0)-1.0)"),
            ("Iampa", "currentDensity", "gAmpa*Wi*(Vd-Vsyn)"),
            ("Sisat", "none", [("Si .gt. Smax", "Smax"), (None, "Si")]),
            ("Inmda", "currentDensity", "gNmda*Sisat*(Vd-Vsyn)/(1.0+0.28*exp(-0.062*(Vd/MVOLT-60.0)))"),
            ("Isyn", "currentDensity", "Iampa+Inmda"),
            ("alphaqd", "none", [("0.00002*Cad .gt. 0.01", "0.01"), (None, "0.00002*Cad")]),
            ("chid", "none", [("Cad/250 .gt. 1", "1"), (None, "Cad/250")]),
            ("alphacd_Vd", "none", [("Vd .lt. -10*MVOLT", "exp((Vd/MVOLT+50.0)/11-(Vd/MVOLT+53.5)/27)/18.975"),
                                    (None, "2.0*exp((-53.5-Vd/MVOLT)/27.0)")]),
            ("betacd_Vd", "none", [("Vd .lt. -10*MVOLT", "(2.0*exp((-53.5-Vd/MVOLT)/27.0)-alphacd_Vd)"),
                                   (None, "0")])],
        ders={"Vs": "(-gLs*(Vs-eL)-gNa*(Minfs_Vs^2)*hs*(Vs-eNa)-gKdr*ns*(Vs-eK)+(gc/pp)*(Vd-Vs)+iSoma/pp) / cm",
              "Vd": "(iDend/(1.0-pp)-Isyn/(1.0-pp)-gLd*(Vd-eL)-ICad-gKahp*qd*(Vd-eK)-gKC*cd*chid*(Vd-eK)"
                    "+(gc*(Vs-Vd))/(1.0-pp)) / cm",
              "Cad": "(-0.13*ICad/UAMP_PER_CM2-0.075*Cad) / MSEC",
              "hs": "(alphahs_Vs-(alphahs_Vs+betahs_Vs)*hs) / MSEC",
              "ns": "(alphans_Vs-(alphans_Vs+betans_Vs)*ns) / MSEC",
              "sd": "(alphasd_Vd-(alphasd_Vd+betasd_Vd)*sd) / MSEC",
              "cd": "(alphacd_Vd-(alphacd_Vd+betacd_Vd)*cd) / MSEC",
              "qd": "(alphaqd-(alphaqd+betaqd)*qd) / MSEC",
              "Si": "-Si/150.0", "Wi": "-Wi/2.0"},
        on_start=[("Vs", "eL"), ("Vd", "eL"), ("qd", "qd0")]),
    # random spike sources (eden/neuroml/LEMS_CoreComponents.inc.xml:778-937, 1944-2000)
    "spikeGeneratorRandom": _cell(
        "spikeGeneratorRandom", "baseSpikeSource", params=[("maxISI", "time"), ("minISI", "time")],
        consts=[("MSEC", "time", 1000.0)], state=[("tsince", "time"), ("tnext", "time"), ("isi", "time")],
        ders={"tsince": "1", "tnext": "0"},
        on_start=[("tsince", "0"), ("isi", "minISI + MSEC * random((maxISI - minISI) / MSEC)"), ("tnext", "isi")],
        on_cond=[("t .gt. tnext", [("isi", "minISI + MSEC * random((maxISI - minISI) / MSEC)"), ("tsince", "0"),
                                   ("tnext", "tnext+isi")], True)]),
    "spikeGeneratorPoisson": _cell(
        "spikeGeneratorPoisson", "baseSpikeSource", params=[("averageRate", "per_time")],
        consts=[("SMALL_TIME", "time", 1e-6)], state=_POISSON_SOURCE_STATE, ders=_POISSON_SOURCE_DERS,
        on_start=[("tsince", "0"), ("isi", "-1 * log(random(1)) / averageRate"), ("tnextIdeal", "isi"),
                  ("tnextUsed", "isi")],
        on_cond=[("t .gt. tnextUsed", [("tsince", "0"), ("isi", "-1 * log(random(1)) / averageRate"),
                                       ("tnextIdeal", "(tnextIdeal+isi)"), ("tnextUsed", _TNEXT_USED)], True)]),
    "spikeGeneratorRefPoisson": _cell(
        "spikeGeneratorRefPoisson", "baseSpikeSource", params=[("averageRate", "per_time"), ("minimumISI", "time")],
        consts=[("SMALL_TIME", "time", 1e-6)], dparams=[("averageIsi", "1 / averageRate")],
        state=_POISSON_SOURCE_STATE, ders=_POISSON_SOURCE_DERS,
        on_start=[("tsince", "0"), ("isi", "minimumISI - (averageIsi-minimumISI) * log(random(1))"),
                  ("tnextIdeal", "isi"), ("tnextUsed", "isi")],
        on_cond=[("t .gt. tnextUsed", [("tsince", "0"),
                                       ("isi", "minimumISI - (averageIsi-minimumISI) * log(random(1))"),
                                       ("tnextIdeal", "(tnextIdeal+isi)"), ("tnextUsed", _TNEXT_USED)], True)]),
    "SpikeSourcePoisson": _cell(
        "SpikeSourcePoisson", "baseSpikeSource", params=[("start", "time"), ("duration", "time"), ("rate", "per_time")],
        consts=[("LONG_TIME", "time", 3.6e18), ("SMALL_TIME", "time", 1e-6)],
        dparams=[("end", "start + duration")], state=_POISSON_SOURCE_STATE, ders=_POISSON_SOURCE_DERS,
        on_start=[("isi", "start - log(random(1))/rate"), ("tsince", "0"),
                  ("tnextIdeal", "isi + H(((isi) - (start+duration))/duration)*LONG_TIME"),
                  ("tnextUsed", "tnextIdeal")],
        on_cond=[("t .gt. tnextUsed", [
            ("isi", "-1 * log(random(1))/rate"),
            ("tnextIdeal", "(tnextIdeal+isi) + H(((tnextIdeal+isi) - (start+duration))/duration)*LONG_TIME"),
            ("tnextUsed", _TNEXT_USED), ("tsince", "0")], True)]),
    "IF_curr_alpha": _pynn_iaf("IF_curr_alpha", "basePyNNIaFCell"),
    "IF_curr_exp": _pynn_iaf("IF_curr_exp", "basePyNNIaFCell"),
    "IF_cond_alpha": _pynn_iaf("IF_cond_alpha", "basePyNNIaFCondCell"),
    "IF_cond_exp": _pynn_iaf("IF_cond_exp", "basePyNNIaFCondCell"),
    "EIF_cond_exp_isfa_ista": _pynn_eif("EIF_cond_exp_isfa_ista"),
    "EIF_cond_alpha_isfa_ista": _pynn_eif("EIF_cond_alpha_isfa_ista"),
    "HH_cond_exp": _cell(
        "HH_cond_exp", "basePyNNCell",
        params=[("gbar_K", "none"), ("gbar_Na", "none"), ("g_leak", "none"), ("e_rev_K", "none"),
                ("e_rev_Na", "none"), ("e_rev_leak", "none"), ("v_offset", "none"), ("e_rev_E", "none"),
                ("e_rev_I", "none")],
        state=[("v", _V), ("m", "none"), ("h", "none"), ("n", "none")],
        derived=[("iLeak", "none", "g_leak * (e_rev_leak - (v / MVOLT))"),
                 ("iNa", "none", "gbar_Na * (m * m * m) * h * (e_rev_Na - (v / MVOLT))"),
                 ("iK", "none", "gbar_K * (n * n * n * n) * (e_rev_K - (v / MVOLT))"),
                 ("iMemb", "none", "iLeak + iNa + iK + i_offset"),
                 ("alpham", "none", "0.32 * (13 - (v / MVOLT) + v_offset) / (exp((13 - (v / MVOLT) + v_offset) / 4.0) - 1)"),
                 ("betam", "none", "0.28 * ((v / MVOLT) - v_offset - 40) / (exp(((v / MVOLT) - v_offset - 40) / 5.0) - 1)"),
                 ("alphah", "none", "0.128 * exp((17 - (v / MVOLT) + v_offset) / 18.0)"),
                 ("betah", "none", "4.0 / (1 + exp((40 - (v / MVOLT) + v_offset) / 5))"),
                 ("alphan", "none", "0.032 * (15 - (v / MVOLT) + v_offset) / (exp((15 - (v / MVOLT) + v_offset) / 5) - 1)"),
                 ("betan", "none", "0.5 * exp((10 - (v / MVOLT) + v_offset) / 40)")],
        ders={"v": "(MVOLT * (iMemb / cm) / MSEC) + (iSyn / (cm * NFARAD))",
              "m": "(alpham * (1 - m) - betam * m) / MSEC", "h": "(alphah * (1 - h) - betah * h) / MSEC",
              "n": "(alphan * (1 - n) - betan * n) / MSEC"},
        on_start=[("v", "v_init * MVOLT"), ("m", "0"), ("h", "0"), ("n", "0")], spikes=False),
}


def _order_derived(name, derived):
    """Derived variables in dependency order (each after the ones it reads)."""
    by = {d[0]: d for d in derived}

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
        for u in syms(by[v][2]):
            if u in by:
                visit(u, path + (v,))
        done.add(v)
        order.append(by[v])

    for d in derived:
        visit(d[0])
    return order


def base_cell_of(name, component_types):
    """The NeuroML base cell type a ComponentType extends through its chain, or None."""
    seen = set()
    while name in component_types and name not in seen:
        seen.add(name)
        name = component_types[name].get("extends", "")
    return name if name in BASE_CELLS else None


def cell_type_from_lems(name, component_types, convert):
    """CellType of a user-defined cell <ComponentType> extending a NeuroML base cell type: its
    Parameters, Constants (converted by `convert(text, dimension)` to engine units),
    DerivedParameters, Requirements, and the nearest Dynamics along the extends chain
    (StateVariables, DerivedVariables, ConditionalDerivedVariables, TimeDerivatives, OnStart,
    OnCondition with StateAssignments and EventOut). OnEvent blocks (a spike input port of
    the cell itself, which no connection reaches) never run."""
    chain, n = [], name
    while n in component_types:
        chain.append(component_types[n])
        n = component_types[n].get("extends", "")
    if n not in BASE_CELLS:
        raise LemsError(f"ComponentType {name} does not extend a base cell type")
    bp, bc, br = BASE_CELLS[n]
    params, consts, reqs = list(bp), list(bc), dict(br)
    dparams, dyn = [], None
    for ct in reversed(chain):
        for el in ct:
            t = _ltag(el)
            if t == "Parameter":
                dim(el.get("dimension", "none"))
                params = [p for p in params if p[0] != el.get("name")] + [(el.get("name"), el.get("dimension", "none"))]
            elif t == "DerivedParameter":
                dparams.append((el.get("name"), el.get("value")))
            elif t == "Constant":
                d = el.get("dimension", "none")
                dim(d)
                consts.append((el.get("name"), d, convert(el.get("value"), d)))
            elif t == "Requirement":
                dim(el.get("dimension", "none"))
                reqs[el.get("name")] = el.get("dimension", "none")
            elif t == "Dynamics":
                dyn = el
    if dyn is None:
        raise LemsError(f"ComponentType {name} has no Dynamics")
    state, derived, ders, on_start, on_cond, v_state = [], [], {}, [], [], None
    for el in dyn:
        t = _ltag(el)
        if t == "StateVariable":
            state.append((el.get("name"), el.get("dimension", "none")))
            dim(state[-1][1])
            if el.get("exposure") == "v":
                v_state = el.get("name")
        elif t == "DerivedVariable":
            if el.get("value") is None:
                raise LemsError(f"{name}: derived variable {el.get('name')} without a value (select) is not supported")
            derived.append((el.get("name"), el.get("dimension", "none"), el.get("value")))
        elif t == "ConditionalDerivedVariable":
            cases = [(c.get("condition"), c.get("value")) for c in el if _ltag(c) == "Case"]
            derived.append((el.get("name"), el.get("dimension", "none"), cases))
        elif t == "TimeDerivative":
            ders[el.get("variable")] = el.get("value")
        elif t in ("OnStart", "OnCondition"):
            assigns, spike = [], False
            for a in el:
                if _ltag(a) == "StateAssignment":
                    assigns.append((a.get("variable"), a.get("value")))
                elif _ltag(a) == "EventOut":
                    if a.get("port") != "spike":
                        raise LemsError(f"{name}: EventOut port '{a.get('port')}' is not supported (spike)")
                    spike = True
            if t == "OnStart":
                on_start.extend(assigns)
            else:
                on_cond.append((el.get("test"), assigns, spike))
        elif t == "OnEvent":
            # a spike input port of the cell itself: no connection delivers spikes to it (events
            # reach cells through their synapses), so the block never runs
            continue
        elif t in ("Regime", "KineticScheme"):
            raise LemsError(f"{name}: <{t}> is not supported")
    return CellType(name=name, params=params, state=state, derived=_order_derived(name, derived), derivatives=ders,
                    on_start=on_start, on_conditions=on_cond, requirements=reqs, constants=consts,
                    derived_params=dparams, v_state=v_state, spikes=True)


@dataclass
class SpikeList:
    """A spikeArray: spikes at listed times (the reference runs it natively as a spike list,
    eden/Eden.cpp:6737-6790). Parameters: the times (us) in listed order, then +inf; state:
    the position of the next spike."""
    name: str = "spikeArray"
    params: list = field(default_factory=list)
    state: list = field(default_factory=lambda: [("pos", "none")])
    v_state: str = None

    @property
    def v_index(self):
        return -1


def emit_spike_list(fn_name, dialect):
    """Step of a spike list: on initialisation the position is 0; otherwise every listed time
    before the end of the step (t + dt) emits the spike and advances the position (the +inf
    after the last time stops the scan)."""
    q = "__device__ __forceinline__ " if dialect == "hip" else "static "
    return "\n".join([
        f"{q}void {fn_name}(const float *P, const float *S, float *Sn, long stride, float t, float dt, int init,"
        " float isyn, int rng_seed, long long rng_step, int *spike) {",
        "    /* spikeArray */",
        "    (void)stride; (void)isyn; (void)rng_seed; (void)rng_step;",
        "    int pos = (int)S[0];",
        "    int ev_spike = 0;",
        "    if (!init) {",
        "        while (P[pos] < t + dt) {",
        "            ev_spike = 1;",
        "            pos++;",
        "        }",
        "    } else {",
        "        pos = 0;",
        "    }",
        "    Sn[0] = (float)pos;",
        "    *spike = ev_spike;",
        "}"])


# ---------------------------------------------------------------- emitter

def emit_cell(ct, fn_name, dialect):
    """C / HIP text of one cell type's step function:

        void fn(const float *P, const float *S, float *Sn, long stride, float t, float dt,
                int init, float isyn, int rng_seed, long long rng_step, int *spike)

    P: the type's parameters in table order; S / Sn: the cell's present / next state, variable
    k at k * stride (they may alias); t: the step's time (time_f32); isyn: the external
    current (synapses then inputs); rng_seed, rng_step: the cell's RNG seed and the step, the
    stamp of random() (lems_randof, invocations numbered in emission order); *spike: set when
    an OnCondition with an EventOut fires."""
    q = "__device__ __forceinline__ " if dialect == "hip" else "static "
    L = [f"{q}void {fn_name}(const float *P, const float *S, float *Sn, long stride, float t, float dt, int init,"
         " float isyn, int rng_seed, long long rng_step, int *spike) {",
         f"    /* {ct.name} */",
         "    (void)S; (void)Sn; (void)stride; (void)t; (void)dt; (void)init; (void)isyn; (void)rng_seed;"
         " (void)rng_step;"]
    names, dims = {"t": "t", "__rng__": [0]}, {"t": dim("time")}
    for r, d in ct.requirements.items():
        if r in ("iSyn", "ISyn"):
            names[r], dims[r] = "isyn", dim(d)
    for k, (p, d) in enumerate(ct.params):
        L.append(f"    const float p_{p} = P[{k}];")
        names[p], dims[p] = f"p_{p}", dim(d or "none")
    for c, d, val in ct.constants:
        L.append(f"    const float c_{c} = {_lit(val)};")
        names[c], dims[c] = f"c_{c}", dim(d)
    for dp, expr in ct.derived_params:
        text, dd = infix_dim(parse(expr), names, dims)
        L.append(f"    const float dp_{dp} = {text};")
        names[dp], dims[dp] = f"dp_{dp}", dd
    for k, (s, d) in enumerate(ct.state):
        L.append(f"    float s_{s} = S[{k} * stride];")
        names[s], dims[s] = f"s_{s}", dim(d)
    for dv, d, expr in ct.derived:
        names[dv], dims[dv] = f"d_{dv}", dim(d)
        if isinstance(expr, str):
            L.append(f"    const float d_{dv} = {infix_dim(parse(expr), names, dims)[0]};")
        else:
            L.append(f"    float d_{dv} = 0.0f;")
            L.append("    if (0) { }")
            for cond, val in expr:
                if cond is not None:
                    L.append(f"    else if ({infix_dim(parse(cond), names, dims)[0]}) "
                             f"{{ d_{dv} = {infix_dim(parse(val), names, dims)[0]}; }}")
            default = [val for cond, val in expr if cond is None]
            if default:
                L.append(f"    else {{ d_{dv} = {infix_dim(parse(default[0]), names, dims)[0]}; }}")
    state_names = {s for s, _ in ct.state}

    def assign(items, indent):
        for var, expr in items:
            if var not in state_names:
                raise LemsError(f"{ct.name}: assignment to {var}, which is not a state variable")
            L.append(f"{indent}s_{var} = {infix_dim(parse(expr), names, dims)[0]}; n_{var} = s_{var};")

    for s, _ in ct.state:
        L.append(f"    float n_{s} = s_{s};")
    L.append("    int ev_spike = 0;")
    L.append("    if (init) {")
    assign(ct.on_start, "        ")
    L.append("    } else {")
    for k, (s, d) in enumerate(ct.state):
        if s in ct.derivatives:
            text, dd = infix_dim(parse(ct.derivatives[s]), names, dims)
            f = native(dd) + native(dim("time")) - native(dim(d))
            L.append(f"        const float der_{k} = {text};")
            L.append(f"        n_{s} = s_{s} + dt * der_{k}{f' * 1e{f}f' if f else ''};")
    for test, assigns, spike in ct.on_conditions:
        L.append(f"        if ({infix_dim(parse(test), names, dims)[0]}) {{")
        assign(assigns, "            ")
        if spike:
            L.append("            ev_spike = 1;")
        L.append("        }")
    L.append("    }")
    for k, (s, _) in enumerate(ct.state):
        L.append(f"    Sn[{k} * stride] = n_{s};")
    L.append("    *spike = ev_spike;")
    L.append("}")
    return "\n".join(L)


@dataclass
class LemsCells:
    """The point-neuron types of a model: per type its name, parameter vector, state size,
    membrane-voltage state index (-1 none) and the emitted source in both dialects."""
    type_names: list
    params: list
    n_state: list
    v_state: list
    state_names: list
    state_pow10: list     # per type, per state variable: power of ten of its native unit (SI)
    source_hip: str
    source_c: str


def build(types, custom=None):
    """types: list of (cell type name, {param: engine value}); custom: user CellTypes by name.
    Returns LemsCells whose functions are lems_cell_<k> and a dispatcher lems_cell(type, ...)."""
    import numpy as np
    hip, c, params, nst, vst, names, snames, spow = [], [], [], [], [], [], [], []
    custom = custom or {}
    for k, (tname, vals) in enumerate(types):
        if tname not in CELL_TYPES and tname not in custom:
            raise LemsError(f"cell type {tname} is not supported")
        ct = custom[tname] if tname in custom else CELL_TYPES[tname]
        if isinstance(ct, SpikeList):
            params.append(np.array(list(vals["times"]) + [np.inf], np.float32))
        else:
            missing = [p for p, _ in ct.params if p not in vals]
            if missing:
                raise LemsError(f"{tname}: parameters {missing} missing")
            params.append(np.array([vals[p] for p, _ in ct.params], np.float32))
        nst.append(len(ct.state))
        vst.append(ct.v_index)
        names.append(tname)
        snames.append([s for s, _ in ct.state])
        spow.append([native(dim(d)) for _, d in ct.state])
        if isinstance(ct, SpikeList):
            hip.append(emit_spike_list(f"lems_cell_{k}", "hip"))
            c.append(emit_spike_list(f"lems_cell_{k}", "c"))
        else:
            hip.append(emit_cell(ct, f"lems_cell_{k}", "hip"))
            c.append(emit_cell(ct, f"lems_cell_{k}", "c"))

    def dispatcher(q):
        L = [f"{q}void lems_cell(int type, const float *P, const float *S, float *Sn, long stride, float t, float dt,"
             " int init, float isyn, int rng_seed, long long rng_step, int *spike) {", "    switch (type) {"]
        for k in range(len(types)):
            L.append(f"    case {k}: lems_cell_{k}(P, S, Sn, stride, t, dt, init, isyn, rng_seed, rng_step, spike);"
                     " break;")
        L.append("    default: *spike = 0; break;")
        L.append("    }")
        L.append("}")
        return "\n".join(L)

    return LemsCells(names, params, nst, vst, snames, spow,
                     "\n\n".join(hip + [dispatcher("__device__ __forceinline__ ")]) + "\n",
                     "\n\n".join(c + [dispatcher("")]) + "\n")
