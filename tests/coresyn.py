"""The reference's EdenTest_CoreSynapses model (tests/golden/neuroml, input data) with the
constructs the generic synapse path does not cover removed: the blockingPlasticSynapse
projection, the continuous (graded) and electrical (gap junction) projections, and the
input list that only drives them. What remains is its chemical part: one HH cell driven by
a pulse, projecting onto passive cells through one synapse of each core type. A seed is
added to the simulation (the model draws no random numbers)."""
import os
import xml.etree.ElementTree as ET

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")
NS = "{http://www.neuroml.org/schema/neuroml2}"


def write_chemical_subset(out_dir, length_ms=None):
    ET.register_namespace("", NS[1:-1])
    t = ET.parse(os.path.join(GOLD, "EdenTest_CoreSynapses.nml"))
    net = t.getroot().find(NS + "network")
    for el in list(net):
        tag = el.tag.replace(NS, "")
        if tag in ("continuousProjection", "electricalProjection") or \
                (tag == "projection" and el.get("synapse") == "Syn_blockingPlasticSynapse") or \
                (tag == "inputList" and el.get("id") == "Stim_ContPre"):
            net.remove(el)
    t.write(os.path.join(out_dir, "EdenTest_CoreSynapses.nml"))
    text = open(os.path.join(GOLD, "LEMS_EdenTest_CoreSynapses.xml")).read()
    text = text.replace('<Simulation id="sim1"', '<Simulation id="sim1" seed="1"')
    if length_ms is not None:
        text = text.replace('length="50.000ms"', f'length="{length_ms}ms"')
    path = os.path.join(out_dir, "LEMS_EdenTest_CoreSynapses.xml")
    with open(path, "w") as f:
        f.write(text)
    return path
