"""User-defined gate rate, time-course and steady-state ComponentTypes (LEMS) for ion channel
gates: a <ComponentType> extending baseVoltageDepRate (exposure r, per_time),
baseVoltageDepTime (t, time) or baseVoltageDepVariable (x, dimensionless), with Parameters,
Constants, DerivedParameters and (conditional) DerivedVariables of the membrane voltage v and
the network's temperature. The reference emits each such component inline in the channel's
gate code (DescribeLemsInline.SingleInstance, eden/Eden.cpp:5150-5160 for rates,
eden/NeuroML.cpp:4028-4108 for the parse); here each used instance becomes one function

    float lems_rate_<k>(float v)

with its parameter values and the temperature as literals, and a dispatcher
lems_rate(int k, float v), emitted in HIP (compiled with hiprtc into the library's
ext_channels_kernel) and C (the oracle). A gate rate of form EDENHIP_RATE_LEMS carries k in its
`rate` field. Expressions follow the dimension calculus of eden_amd.lems_cells (products and
quotients carry the power of ten between their operands' native units and the result's).
Concentration-dependent types (baseVoltageConcDep*) need ion pools, which this build does not
run, and are refused.
"""
from dataclasses import dataclass

from .lems import LemsError, _lit, _ltag, parse
from .lems_cells import dim, infix_dim

# base type -> (exposure, its LEMS dimension)
BASE_RATES = {"baseVoltageDepRate": ("r", "per_time"), "baseVoltageDepTime": ("t", "time"),
              "baseVoltageDepVariable": ("x", "none")}


@dataclass
class RateType:
    name: str
    base: str
    params: list          # (name, LEMS dimension)
    constants: list       # (name, LEMS dimension, engine value)
    derived_params: list  # (name, expression)
    derived: list         # (name, LEMS dimension, expression or [(condition or None, expression)])
    requirements: dict    # name -> LEMS dimension (v, temperature)


def base_rate_of(name, component_types):
    """The base rate type a ComponentType extends (through any chain), or None."""
    seen = set()
    while name in component_types and name not in seen:
        seen.add(name)
        name = component_types[name].get("extends", "")
    return name if name in BASE_RATES else None


def _symbols(tree, out):
    if tree[0] == "sym":
        out.add(tree[1])
    for x in tree[1:]:
        if isinstance(x, tuple):
            _symbols(x, out)


def _ordered(derived):
    """Derived variables in dependency order (each after those it reads)."""
    by_name = {d[0]: d for d in derived}
    deps = {}
    for n, _, expr in derived:
        s = set()
        texts = [expr] if isinstance(expr, str) else [x for case in expr for x in case if x is not None]
        for t in texts:
            _symbols(parse(t), s)
        deps[n] = {x for x in s if x in by_name and x != n}
    out, done = [], set()

    def visit(n, stack):
        if n in done:
            return
        if n in stack:
            raise LemsError(f"derived variable {n} depends on itself")
        for m in sorted(deps[n]):
            visit(m, stack | {n})
        done.add(n)
        out.append(by_name[n])

    for n, _, _ in derived:
        visit(n, frozenset())
    return out


def rate_type_from_lems(name, component_types, convert):
    """RateType of a user ComponentType extending a base rate type."""
    chain, n = [], name
    while n in component_types:
        chain.append(component_types[n])
        n = component_types[n].get("extends", "")
    if n not in BASE_RATES:
        raise LemsError(f"ComponentType {name} does not extend a voltage-dependent rate, time course or "
                        f"steady state ({', '.join(BASE_RATES)})")
    params, consts, reqs, dparams, dyn = [], [], {"v": "voltage"}, [], None
    for ct in reversed(chain):
        for el in ct:
            t = _ltag(el)
            if t == "Parameter":
                d = el.get("dimension", "none")
                dim(d)
                params = [p for p in params if p[0] != el.get("name")] + [(el.get("name"), d)]
            elif t == "DerivedParameter":
                dparams.append((el.get("name"), el.get("value")))
            elif t == "Constant":
                d = el.get("dimension", "none")
                dim(d)
                consts.append((el.get("name"), d, convert(el.get("value"), d)))
            elif t == "Requirement":
                if el.get("name") not in ("v", "temperature"):
                    raise LemsError(f"ComponentType {name}: requirement {el.get('name')} is not supported "
                                    "(v, temperature)")
                reqs[el.get("name")] = el.get("dimension", "none")
            elif t == "Dynamics":
                dyn = el
    if dyn is None:
        raise LemsError(f"ComponentType {name} has no Dynamics")
    derived = []
    for el in dyn:
        t = _ltag(el)
        if t == "DerivedVariable":
            if el.get("value") is None:
                raise LemsError(f"{name}: derived variable {el.get('name')} without a value is not supported")
            derived.append((el.get("name"), el.get("dimension", "none"), el.get("value")))
        elif t == "ConditionalDerivedVariable":
            cases = [(c.get("condition"), c.get("value")) for c in el if _ltag(c) == "Case"]
            derived.append((el.get("name"), el.get("dimension", "none"), cases))
        elif t in ("StateVariable", "TimeDerivative", "OnStart", "OnCondition", "OnEvent"):
            raise LemsError(f"{name}: a rate ComponentType with <{t}> is not supported")
    exposure = BASE_RATES[n][0]
    if exposure not in [d[0] for d in derived]:
        raise LemsError(f"{name}: no derived variable {exposure} to expose")
    return RateType(name, n, params, consts, dparams, _ordered(derived), reqs)


def emit_rate(rt, values, temperature, fn_name, dialect):
    """C / HIP text of one rate instance: float fn(float v), its parameters (engine values by
    name) and the temperature (K) as literals."""
    q = "__device__ __forceinline__ " if dialect == "hip" else "static "
    L = [f"{q}float {fn_name}(float v) {{", f"    /* {rt.name} */"]
    names, dims = {"v": "v"}, {"v": dim("voltage")}
    if "temperature" in rt.requirements:
        if temperature is None:
            raise LemsError(f"{rt.name}: needs the network's temperature")
        L.append(f"    const float temperature = {_lit(temperature)};")
        names["temperature"], dims["temperature"] = "temperature", dim("temperature")
    for p, d in rt.params:
        if p not in values:
            raise LemsError(f"{rt.name}: parameter {p} missing")
        L.append(f"    const float p_{p} = {_lit(values[p])};")
        names[p], dims[p] = f"p_{p}", dim(d)
    for c, d, val in rt.constants:
        L.append(f"    const float c_{c} = {_lit(val)};")
        names[c], dims[c] = f"c_{c}", dim(d)
    for dp, expr in rt.derived_params:
        text, dd = infix_dim(parse(expr), names, dims)
        L.append(f"    const float dp_{dp} = {text};")
        names[dp], dims[dp] = f"dp_{dp}", dd
    for dv, d, expr in rt.derived:
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
    L.append(f"    return d_{BASE_RATES[rt.base][0]};")
    L.append("}")
    return "\n".join(L)


@dataclass
class LemsRates:
    """The LEMS rate instances of a model (k -> type and parameter values) and their source."""
    entries: list         # (RateType, {param: engine value})
    source_hip: str
    source_c: str


def build(entries, temperature):
    hip, c = [], []
    for k, (rt, values) in enumerate(entries):
        hip.append(emit_rate(rt, values, temperature, f"lems_rate_{k}", "hip"))
        c.append(emit_rate(rt, values, temperature, f"lems_rate_{k}", "c"))

    def dispatcher(q):
        L = [f"{q}float lems_rate(int k, float v) {{", "    switch (k) {"]
        for k in range(len(entries)):
            L.append(f"    case {k}: return lems_rate_{k}(v);")
        L.append("    default: return 0.0f;")
        L.append("    }")
        L.append("}")
        return "\n".join(L)

    return LemsRates(list(entries), "\n\n".join(hip + [dispatcher("__device__ ")]) + "\n",
                     "\n\n".join(c + [dispatcher("")]) + "\n")
