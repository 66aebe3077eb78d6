"""The reference's EdenTest_CoreInputs model (tests/golden/neuroml, input data): passive cells
driven by one input of each core type (pulse, sine, ramp, voltage clamps, a weighted pulse),
spike sources (spikeArray, spikeGenerator) projecting onto two of them, and Izhikevich point
neurons driven by the dimensionless inputs. Removed: the timedSynapticInput list, an input
this build does not run. A seed is added to the simulation."""
import os
import xml.etree.ElementTree as ET

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")
NS = "{http://www.neuroml.org/schema/neuroml2}"


def write_model(out_dir):
    ET.register_namespace("", NS[1:-1])
    t = ET.parse(os.path.join(GOLD, "EdenTest_CoreInputs.nml"))
    net = t.getroot().find(NS + "network")
    for el in list(net):
        if el.tag.replace(NS, "") == "inputList" and el.get("component") == "Inp_timedSynapticInput":
            net.remove(el)
    t.write(os.path.join(out_dir, "EdenTest_CoreInputs.nml"))
    text = open(os.path.join(GOLD, "LEMS_EdenTest_CoreInputs.xml")).read()
    text = text.replace('<Simulation id="sim1"', '<Simulation id="sim1" seed="1"')
    path = os.path.join(out_dir, "LEMS_EdenTest_CoreInputs.xml")
    with open(path, "w") as f:
        f.write(text)
    return path
