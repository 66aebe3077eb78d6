"""The reference's EdenTest_ArtificialCells model (tests/golden/neuroml, input data): one cell
of each NeuroML core point-neuron type (integrate-and-fire, Izhikevich, adaptive exponential,
FitzHugh-Nagumo, Hindmarsh-Rose, the PyNN cells, a user-defined quadratic integrate-and-fire
ComponentType) beside passive physical cells, most driven by a current pulse, and a
spikeArray source population. Removed: the dimensionless pulse (pulseGeneratorDL) into the
izhikevichCell, an input this build does not run. A seed is added to the simulation."""
import os
import xml.etree.ElementTree as ET

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")
NS = "{http://www.neuroml.org/schema/neuroml2}"


def write_model(out_dir, length_ms=None):
    ET.register_namespace("", NS[1:-1])
    t = ET.parse(os.path.join(GOLD, "EdenTest_ArtificialCells.nml"))
    net = t.getroot().find(NS + "network")
    for el in list(net):
        tag = el.tag.replace(NS, "")
        if tag == "explicitInput" and el.get("input") == "pulseDL_1":
            net.remove(el)
    t.write(os.path.join(out_dir, "EdenTest_ArtificialCells.nml"))
    text = open(os.path.join(GOLD, "LEMS_EdenTest_ArtificialCells.xml")).read()
    text = text.replace('<Simulation id="sim1"', '<Simulation id="sim1" seed="1"')
    if length_ms is not None:
        text = text.replace('length="100.0ms"', f'length="{length_ms}ms"')
    path = os.path.join(out_dir, "LEMS_EdenTest_ArtificialCells.xml")
    with open(path, "w") as f:
        f.write(text)
    return path
