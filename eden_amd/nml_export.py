"""The synthetic workloads written as NeuroML 2 / LEMS files, for EDEN's command line and for
eden_amd.neuroml (which reads them back into the same instantiated model).

`write_lems(model, path)` writes a LEMS file with the model's cells, channels, synapses,
inputs, network and simulation, and returns the path. It covers the models eden_amd.configs
builds: every cell a <cell> of its own morphology (one segment per compartment, explicit
proximal and distal points), HH squid channels by channelDensity per compartment, expTwo
synapse types, a pulseGenerator per pulse input, a poissonFiringSynapse per Poisson input,
one projection per synapse type with connectionWD elements, an OutputFile of the recorded
compartments' voltages and an EventOutputFile of the event columns. Float values are
written with 9 significant digits, so that reading them back gives the same float32.

The geometry of the configs' cells is given to the writer as the morphology the model was
built from (`morphologies`: per cell, parent, proximal, distal); the instantiated model
itself only keeps areas and resistances.
"""
import numpy as np

from . import _native as N

_FORMS = {N.RATE_EXP: "HHExpRate", N.RATE_EXPLINEAR: "HHExpLinearRate", N.RATE_SIGMOID: "HHSigmoidRate"}


def _g(x):
    return "%.9g" % float(np.float32(x))


def _channel_xml(i, ct):
    out = [f'  <ionChannelHH id="chan{i}" conductance="10pS">']
    for k in range(ct.n_gates):
        g = ct.gates[k]
        out.append(f'    <gateHHrates id="g{k}" instances="{g.instances}">')
        for tag, r in (("forwardRate", g.alpha), ("reverseRate", g.beta)):
            if r.form not in _FORMS:
                raise ValueError("fixed rates have no NeuroML rate element")
            # rate in per_ms: the engine's per-us value times 1e3, exact in the written digits
            # only when the per_ms value was the input; the configs build rates from per_ms values
            out.append(f'      <{tag} type="{_FORMS[r.form]}" rate="{_per_ms(r.rate)}per_ms" '
                       f'midpoint="{_g(r.midpoint)}mV" scale="{_g(r.scale)}mV"/>')
        if g.q10 != 1.0:
            out.append(f'      <q10Settings type="q10Fixed" fixedQ10="{_g(g.q10)}"/>')
        out.append('    </gateHHrates>')
    out.append('  </ionChannelHH>')
    return out


def _per_ms(rate_per_us):
    """The shortest per_ms decimal whose conversion (value * 10^-3, rounded to float32) gives
    back the stored per-us rate."""
    target = np.float32(rate_per_us)
    for digits in range(1, 18):
        s = "%.*g" % (digits, float(rate_per_us) * 1e3)
        if np.float32(float(s) * 1.0 * 10.0 ** -3 + 0.0) == target:
            return s
    raise ValueError(f"rate {rate_per_us} has no per_ms decimal")


def _same_cells(m, morphologies):
    """True when every cell has the same morphology and biophysics (one <cell> serves all)."""
    cb = m.cell_comp_begin
    if len(set(np.diff(cb).tolist())) != 1:
        return False
    n = int(cb[1] - cb[0])
    p0, x0, d0 = morphologies[0]
    for c in range(1, m.n_cells):
        p, x, d = morphologies[c]
        if not (np.array_equal(p, p0) and np.array_equal(x, x0) and np.array_equal(d, d0)):
            return False
    sl = lambda c: slice(m.comp_chan_begin[cb[c]], m.comp_chan_begin[cb[c] + n])
    for c in range(1, m.n_cells):
        if not (np.array_equal(np.diff(m.comp_chan_begin[cb[c]:cb[c] + n + 1]), np.diff(m.comp_chan_begin[:n + 1]))
                and np.array_equal(m.chan_type[sl(c)], m.chan_type[sl(0)])
                and np.array_equal(m.chan_gbase[sl(c)], m.chan_gbase[sl(0)])
                and np.array_equal(m.chan_erev[sl(c)], m.chan_erev[sl(0)])
                and np.array_equal(m.v0[cb[c]:cb[c] + n], m.v0[:n])):
            return False
    return True


def write_lems(model, path, morphologies, cm=1.0, ra=0.1, sim_seed=None, trajectory_file="trajectory.gen.txt",
               events_file="spikes.gen.txt"):
    """cm: specific capacitance (uF/cm^2) and ra: axial resistivity (kOhm cm) of every
    segment, as the model was built with. sim_seed: the simulation seed; by default the one
    the model's cell seeds come from (cell 0's seed is ReverseBits(seed), an involution)."""
    m = model
    if sim_seed is None:
        from .model import reverse_bits32
        sim_seed = reverse_bits32(int(m.cell_rng_seed[0]))
    nc = m.n_cells
    cb = m.cell_comp_begin
    shared = _same_cells(m, morphologies)
    n_types = 1 if shared else nc
    # cell type and population of each cell, and the cell's instance id in its population
    ctype = (lambda c: "cell0") if shared else (lambda c: f"cell{c}")
    pop = (lambda c: "p") if shared else (lambda c: f"p{c}")
    inst = (lambda c: c) if shared else (lambda c: 0)
    L = ['<Lems>', '  <Target component="sim1"/>', '<neuroml id="exported">']
    for i, ct in enumerate(m.channel_types):
        L += _channel_xml(i, ct)
    for i, st in enumerate(m.syn_types):
        L.append(f'  <expTwoSynapse id="syn{i}" gbase="{_g(st.gbase)}uS" erev="{_g(st.erev)}mV" '
                 f'tauRise="{_g(st.tau_rise)}us" tauDecay="{_g(st.tau_decay)}us"/>')
    for p in range(len(m.pulse_start)):
        L.append(f'  <pulseGenerator id="pulse{p}" delay="{_g(m.pulse_start[p])}us" '
                 f'duration="{_g(m.pulse_duration[p])}us" amplitude="{_g(m.pulse_amplitude[p])}nA"/>')
    for p in range(len(m.poisson_rate)):
        L.append(f'  <poissonFiringSynapse id="poisson{p}" averageRate="{_g(m.poisson_rate[p])}per_us" '
                 f'synapse="syn{int(m.poisson_syn_type[p])}" spikeTarget="./syn{int(m.poisson_syn_type[p])}"/>')
    for c in range(n_types):
        par, prox, dist = morphologies[c]
        n = int(cb[c + 1] - cb[c])
        L.append(f'  <cell id="cell{c}">')
        L.append('    <morphology id="morph%d">' % c)
        for k in range(n):
            L.append(f'      <segment id="{k}">')
            if par[k] >= 0:
                L.append(f'        <parent segment="{int(par[k])}"/>')
            L.append('        <proximal x="%s" y="%s" z="%s" diameter="%s"/>' % tuple(_g(v) for v in prox[k]))
            L.append('        <distal x="%s" y="%s" z="%s" diameter="%s"/>' % tuple(_g(v) for v in dist[k]))
            L.append('      </segment>')
        for k in range(n):
            L.append(f'      <segmentGroup id="s{k}"><member segment="{k}"/></segmentGroup>')
        L.append('    </morphology>')
        L.append('    <biophysicalProperties id="bio%d">' % c)
        L.append('      <membraneProperties>')
        for k in range(n):
            g = cb[c] + k
            for ch in range(m.comp_chan_begin[g], m.comp_chan_begin[g + 1]):
                L.append(f'        <channelDensity id="d{k}_{ch - m.comp_chan_begin[g]}" ionChannel="chan{int(m.chan_type[ch])}" '
                         f'condDensity="{_g(m.chan_gbase[ch])}mS_per_cm2" erev="{_g(m.chan_erev[ch])}mV" '
                         f'segmentGroup="s{k}"/>')
        vt = m.v_threshold[cb[c]]
        L.append(f'        <spikeThresh value="{_g(vt if np.isfinite(vt) else 0.0)}mV"/>')
        L.append(f'        <specificCapacitance value="{_g(cm)}uF_per_cm2"/>')
        L.append(f'        <initMembPotential value="{_g(m.v0[cb[c]])}mV"/>')
        L.append('      </membraneProperties>')
        L.append(f'      <intracellularProperties><resistivity value="{_g(ra)}kohm_cm"/></intracellularProperties>')
        L.append('    </biophysicalProperties>')
        L.append('  </cell>')
    L.append('  <network id="net">')
    if shared:
        L.append(f'    <population id="p" component="cell0" size="{nc}"/>')
    else:
        for c in range(nc):
            L.append(f'    <population id="p{c}" component="cell{c}" size="1"/>')
    comp_cell = np.repeat(np.arange(nc), np.diff(cb))
    for g in range(m.n_comps):
        c = int(comp_cell[g])
        k = int(g - cb[c])
        for p in range(m.comp_pulse_begin[g], m.comp_pulse_begin[g + 1]):
            L.append(f'    <inputList id="in{p}" component="pulse{p}" population="{pop(c)}">'
                     f'<inputW id="0" target="../{pop(c)}/{inst(c)}/{ctype(c)}" segmentId="{k}" '
                     f'weight="{_g(m.pulse_weight[p])}"/></inputList>')
        for p in range(m.comp_poisson_begin[g], m.comp_poisson_begin[g + 1]):
            L.append(f'    <inputList id="pin{p}" component="poisson{p}" population="{pop(c)}">'
                     f'<inputW id="0" target="../{pop(c)}/{inst(c)}/{ctype(c)}" segmentId="{k}" '
                     f'weight="{_g(m.poisson_weight[p])}"/></inputList>')
    # synapses: the connections of each synapse type in instance order, consecutive ones
    # between the same two populations in one projection; the reader numbers the instances of a
    # compartment and type in projection then connection order, which is then instance order
    if m.n_synapses:
        src = np.full(m.n_synapses, -1, np.int64)
        for g in range(m.n_comps):
            for k in range(m.comp_spike_begin[g], m.comp_spike_begin[g + 1]):
                src[m.spike_target[k]] = g
        if (src < 0).any():
            raise ValueError("every synapse needs a presynaptic compartment to be written as a connection")
        comp_of_syn = np.repeat(np.arange(m.n_comps), np.diff(m.comp_syn_begin))
        nproj = 0
        for t in range(len(m.syn_types)):
            sel = np.flatnonzero(m.syn_type == t)
            runs = []
            for s_ in sel:
                a, b = int(src[s_]), int(comp_of_syn[s_])
                key = (pop(int(comp_cell[a])), pop(int(comp_cell[b])))
                if not runs or runs[-1][0] != key:
                    runs.append((key, []))
                runs[-1][1].append((s_, a, b))
            for (pa, pb), lst in runs:
                L.append(f'    <projection id="proj{nproj}" presynapticPopulation="{pa}" '
                         f'postsynapticPopulation="{pb}" synapse="syn{t}">')
                nproj += 1
                for j, (s_, a, b) in enumerate(lst):
                    ca, cbb = int(comp_cell[a]), int(comp_cell[b])
                    L.append(f'      <connectionWD id="{j}" preCellId="../{pa}/{inst(ca)}/{ctype(ca)}" '
                             f'preSegmentId="{a - cb[ca]}" postCellId="../{pb}/{inst(cbb)}/{ctype(cbb)}" '
                             f'postSegmentId="{b - cb[cbb]}" weight="{_g(m.syn_weight[s_])}" '
                             f'delay="{_g(m.syn_delay[s_])}us"/>')
                L.append('    </projection>')
    L.append('  </network>')
    L.append('</neuroml>')
    L.append(f'  <Simulation id="sim1" length="{_g(m.t_final)}us" step="{_g(m.dt)}us" target="net" seed="{sim_seed}">')
    if len(m.record_comp):
        L.append(f'    <OutputFile id="traj" fileName="{trajectory_file}">')
        for i, g in enumerate(m.record_comp):
            c = int(comp_cell[g])
            L.append(f'      <OutputColumn id="v{i}" quantity="{pop(c)}/{inst(c)}/{ctype(c)}/{int(g - cb[c])}/v"/>')
        L.append('    </OutputFile>')
    if len(m.event_comp):
        L.append(f'    <EventOutputFile id="ev" fileName="{events_file}" format="TIME_ID">')
        for i, g in enumerate(m.event_comp):
            c = int(comp_cell[g])
            L.append(f'      <EventSelection id="{i}" select="{pop(c)}/{inst(c)}/{ctype(c)}/{int(g - cb[c])}" '
                     'eventPort="spike"/>')
        L.append('    </EventOutputFile>')
    L.append('  </Simulation>')
    L.append('</Lems>')
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
    return path
