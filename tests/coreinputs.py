"""The reference's EdenTest_CoreInputs model (tests/golden/neuroml, input data): passive cells
driven by one input of each core type (pulse, sine, ramp, voltage clamps, a weighted pulse),
spike sources (spikeArray, spikeGenerator) projecting onto two of them, and Izhikevich point
neurons driven by the dimensionless inputs, and a timedSynapticInput (a spike list driving an
expOneSynapse). A seed is added to the simulation."""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")


def write_model(out_dir):
    shutil.copy(os.path.join(GOLD, "EdenTest_CoreInputs.nml"), os.path.join(out_dir, "EdenTest_CoreInputs.nml"))
    text = open(os.path.join(GOLD, "LEMS_EdenTest_CoreInputs.xml")).read()
    text = text.replace('<Simulation id="sim1"', '<Simulation id="sim1" seed="1"')
    path = os.path.join(out_dir, "LEMS_EdenTest_CoreInputs.xml")
    with open(path, "w") as f:
        f.write(text)
    return path
