"""User gate rate / time course / steady state ComponentTypes (eden_amd.lems_rates) on the CPU.

* EdenTest3's myHHExpRate and myHHSigmoidRate (tests/golden/neuroml/lems_rates.nml holds them
  as the reference's file does) in the HH squid channels give the very traces of the built-in
  HHExpRate and HHSigmoidRate: the generated expressions perform the built-in forms'
  operations, and the oracle sums the channels in distribution order either way.
* A temperature-dependent time course (EdenTest3's Bezaire_KvA_taun) and a voltage-dependent
  rate (Bezaire_Kdrfastngf_betaq), evaluated through the compiled C dialect, equal a float32
  evaluation of their LEMS expressions with the reference's unit factors.
* Concentration-dependent rate types and rate types with state are refused.
"""
import math
import os

import numpy as np
import pytest

from eden_amd import _native as N
from eden_amd import lems_rates, neuroml
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))

TYPES = '''
<ComponentType name="myHHExpRate" extends="baseVoltageDepRate">
  <Parameter name="rate" dimension="per_time"/><Parameter name="midpoint" dimension="voltage"/>
  <Parameter name="scale" dimension="voltage"/>
  <Dynamics><DerivedVariable name="r" exposure="r" value="rate * exp((v - midpoint)/scale)" dimension="per_time"/>
  </Dynamics></ComponentType>
<ComponentType name="myHHSigmoidRate" extends="baseVoltageDepRate">
  <Parameter name="rate" dimension="per_time"/><Parameter name="midpoint" dimension="voltage"/>
  <Parameter name="scale" dimension="voltage"/>
  <Dynamics><DerivedVariable name="r" exposure="r" value="rate / (1 + exp(0 - (v - midpoint)/scale))"
   dimension="per_time"/></Dynamics></ComponentType>
<ComponentType name="Bezaire_Kdrfastngf_betaq" extends="baseVoltageDepRate">
  <Constant name="TIME_SCALE" dimension="time" value="1 ms"/><Constant name="VOLT_SCALE" dimension="voltage" value="1 mV"/>
  <Dynamics><DerivedVariable name="V" dimension="none" value="v / VOLT_SCALE"/>
  <DerivedVariable name="r" dimension="per_time" value="((0.264) / (exp((V+34)/40))) / TIME_SCALE" exposure="r"/>
  </Dynamics></ComponentType>
<ComponentType name="Bezaire_KvA_taun" extends="baseVoltageDepTime">
  <Constant name="TIME_SCALE" dimension="time" value="1 ms"/><Constant name="VOLT_SCALE" dimension="voltage" value="1 mV"/>
  <Constant name="TEMP_SCALE" dimension="temperature" value="1 K"/>
  <Constant name="R" dimension="none" value="8.315"/><Constant name="F" dimension="none" value="96480"/>
  <Requirement name="temperature" dimension="temperature"/>
  <Dynamics><DerivedVariable name="V" dimension="none" value="v / VOLT_SCALE"/>
  <DerivedVariable name="T" dimension="none" value="temperature / TEMP_SCALE"/>
  <DerivedVariable name="alpha" dimension="none" value="exp((1e-3*(-3)*(V+33.6)*F) / (R*(T+0.01)))"/>
  <DerivedVariable name="beta" dimension="none" value="exp((1e-3*(-3)*0.6*(V+33.6)*F) / (R*(T+0.01)))"/>
  <DerivedVariable name="t" dimension="time" value="( beta / (0.02*(1 + alpha)) ) * TIME_SCALE" exposure="t"/>
  </Dynamics></ComponentType>
<ComponentType name="Bezaire_KvA_ssn" extends="baseVoltageDepVariable">
  <Constant name="VOLT_SCALE" dimension="voltage" value="1 mV"/>
  <Dynamics><DerivedVariable name="V" dimension="none" value="v / VOLT_SCALE"/>
  <DerivedVariable name="x" exposure="x" dimension="none" value="1 / (1 + exp(-(V + 40) / 8))"/></Dynamics>
  </ComponentType>
'''


def hh_model(tmp_path, user, extra_channel=""):
    """The HH squid cell (one compartment) with its rates built in or given by the user
    types; an optional extra channel density element."""
    ex, sg = ("myHHExpRate", "myHHSigmoidRate") if user else ("HHExpRate", "HHSigmoidRate")
    chans = (
        '<ionChannelHH id="leak" conductance="10pS"/>'
        '<ionChannelHH id="naChan" conductance="10pS" species="na">'
        '<gateHHrates id="m" instances="3"><forwardRate type="HHExpLinearRate" rate="1per_ms" midpoint="-40mV" '
        f'scale="10mV"/><reverseRate type="{ex}" rate="4per_ms" midpoint="-65mV" scale="-18mV"/></gateHHrates>'
        f'<gateHHrates id="h" instances="1"><forwardRate type="{ex}" rate="0.07per_ms" midpoint="-65mV" scale="-20mV"/>'
        f'<reverseRate type="{sg}" rate="1per_ms" midpoint="-35mV" scale="10mV"/></gateHHrates></ionChannelHH>'
        '<ionChannelHH id="kChan" conductance="10pS" species="k">'
        '<gateHHrates id="n" instances="4"><forwardRate type="HHExpLinearRate" rate="0.1per_ms" midpoint="-55mV" '
        f'scale="10mV"/><reverseRate type="{ex}" rate="0.125per_ms" midpoint="-65mV" scale="-80mV"/></gateHHrates>'
        '</ionChannelHH>'
        '<ionChannelHH id="kva" conductance="10pS" species="k"><gateHHtauInf id="n" instances="1">'
        '<timeCourse type="Bezaire_KvA_taun"/><steadyState type="Bezaire_KvA_ssn"/></gateHHtauInf></ionChannelHH>'
        '<ionChannelHH id="kdr" conductance="10pS" species="k"><gateHHrates id="n" instances="4">'
        '<forwardRate type="HHExpLinearRate" rate="0.42per_ms" midpoint="-9mV" scale="6mV"/>'
        '<reverseRate type="Bezaire_Kdrfastngf_betaq"/></gateHHrates></ionChannelHH>')
    p = tmp_path / ("hh_user.xml" if user else "hh_builtin.xml")
    p.write_text(
        '<Lems><Target component="s"/>' + TYPES + '<neuroml>' + chans +
        '<cell id="hh"><morphology id="m"><segment id="0"><proximal x="0" y="0" z="0" diameter="17.841242"/>'
        '<distal x="0" y="0" z="0" diameter="17.841242"/></segment></morphology><biophysicalProperties id="b">'
        '<membraneProperties><channelDensity id="leak" ionChannel="leak" condDensity="3.0 S_per_m2" erev="-54.3mV"/>'
        '<channelDensity id="na" ionChannel="naChan" condDensity="120.0 mS_per_cm2" erev="50.0 mV"/>'
        '<channelDensity id="k" ionChannel="kChan" condDensity="360 S_per_m2" erev="-77mV"/>' + extra_channel +
        '<spikeThresh value="-20mV"/><specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/>'
        '</membraneProperties><intracellularProperties><resistivity value="0.03 kohm_cm"/></intracellularProperties>'
        '</biophysicalProperties></cell><pulseGenerator id="pg" delay="10ms" duration="60ms" amplitude="0.08nA"/>'
        '<network id="n" temperature="6.3 degC"><population id="pop" component="hh" size="2"/>'
        '<explicitInput target="pop[0]" input="pg"/></network></neuroml>'
        '<Simulation id="s" length="100ms" step="0.025ms" target="n" seed="1"><OutputFile id="o" fileName="o.txt">'
        '<OutputColumn id="v" quantity="pop/0/hh/0/v"/><OutputColumn id="w" quantity="pop/1/hh/0/v"/></OutputFile>'
        '<EventOutputFile id="e" fileName="e.txt" format="TIME_ID"><EventSelection id="0" select="pop[0]" '
        'eventPort="spike"/></EventOutputFile></Simulation></Lems>')
    return neuroml.read(str(p))[0]


def test_user_types_equal_the_built_in_forms(tmp_path):
    mu, mb = hh_model(tmp_path, True), hh_model(tmp_path, False)
    # the declared (unused) kva and kdr channels give 3 instances either way; the user
    # version 4 more
    assert len(mb.lems_rates.entries) == 3 and len(mu.lems_rates.entries) == 3 + 4
    g = mu.channel_types[1].gates[1]
    assert (g.alpha.form, g.beta.form) == (N.RATE_LEMS, N.RATE_LEMS)
    tu, eu, ru, _ = oracle.run(mu)
    tb, eb, rb, _ = oracle.run(mb)
    assert np.array_equal(tu, tb) and np.array_equal(ru, rb) and len(rb) > 3


def test_temperature_and_voltage_rate_types_by_hand(tmp_path):
    m = hh_model(tmp_path, True)
    names = [rt.name for rt, _ in m.lems_rates.entries]
    lib = oracle.lems_library(m.lems_rates.source_c)
    lib.lems_rate.restype = __import__("ctypes").c_float
    lib.lems_rate.argtypes = [__import__("ctypes").c_int, __import__("ctypes").c_float]
    f = np.float32
    T = f(6.3 + 273.15)
    for v in (-80.0, -40.0, 0.0, 25.0):
        V = f(v) / f(1.0)
        # Bezaire_KvA_taun: t in ms, times TIME_SCALE (1000 us)
        a = f(math.exp(float((f(1e-3) * f(-3) * (V + f(33.6)) * f(96480)) / (f(8.315) * (T / f(1.0) + f(0.01))))))
        b = f(math.exp(float((f(1e-3) * f(-3) * f(0.6) * (V + f(33.6)) * f(96480)) /
                             (f(8.315) * (T / f(1.0) + f(0.01))))))
        t_us = (b / (f(0.02) * (f(1) + a))) * f(1000.0)
        got = lib.lems_rate(names.index("Bezaire_KvA_taun"), float(v))
        assert got == pytest.approx(float(t_us), rel=2e-6), v
        # Bezaire_Kdrfastngf_betaq: per ms, divided by TIME_SCALE (1000 us)
        r = (f(0.264) / f(math.exp(float((V + f(34)) / f(40))))) / f(1000.0)
        assert lib.lems_rate(names.index("Bezaire_Kdrfastngf_betaq"), float(v)) == pytest.approx(float(r), rel=2e-6)


def test_rate_types_with_concentration_or_state_refused():
    import xml.etree.ElementTree as ET
    conc = ET.fromstring('<ComponentType name="c" extends="baseVoltageConcDepRate"><Dynamics>'
                         '<DerivedVariable name="r" dimension="per_time" value="caConc"/></Dynamics></ComponentType>')
    st = ET.fromstring('<ComponentType name="s" extends="baseVoltageDepRate"><Dynamics>'
                       '<StateVariable name="q" dimension="none"/><DerivedVariable name="r" dimension="per_time" '
                       'value="q"/></Dynamics></ComponentType>')
    with pytest.raises(lems_rates.LemsError, match="does not extend"):
        lems_rates.rate_type_from_lems("c", {"c": conc}, lambda t, d: 0.0)
    with pytest.raises(lems_rates.LemsError, match="StateVariable"):
        lems_rates.rate_type_from_lems("s", {"s": st}, lambda t, d: 0.0)
