"""The reference's EdenTest_CoreSynapses model (tests/golden/neuroml, input data): one HH
cell driven by a pulse, projecting onto passive cells through one synapse of each core type
(the blockingPlasticSynapse among them), and continuous (graded) and electrical projections
between two populations of passive cells. A seed is added to the simulation (the model
draws no random numbers), and the length can be cut."""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")


def write_model(out_dir, length_ms=None):
    shutil.copy(os.path.join(GOLD, "EdenTest_CoreSynapses.nml"), os.path.join(out_dir, "EdenTest_CoreSynapses.nml"))
    text = open(os.path.join(GOLD, "LEMS_EdenTest_CoreSynapses.xml")).read()
    text = text.replace('<Simulation id="sim1"', '<Simulation id="sim1" seed="1"')
    if length_ms is not None:
        text = text.replace('length="50.000ms"', f'length="{length_ms}ms"')
    path = os.path.join(out_dir, "LEMS_EdenTest_CoreSynapses.xml")
    with open(path, "w") as f:
        f.write(text)
    return path
