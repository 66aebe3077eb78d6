"""The reference's EdenTest_ArtificialCells model (tests/golden/neuroml, input data): one cell
of each NeuroML core point-neuron type (integrate-and-fire, Izhikevich, adaptive exponential,
FitzHugh-Nagumo, Hindmarsh-Rose, the PyNN cells, a user-defined quadratic integrate-and-fire
ComponentType) beside passive physical cells, most driven by a current pulse (the Izhikevich
cell by a dimensionless one), and a spikeArray source population. A seed is added to the
simulation; the model is otherwise as the reference holds it."""
import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")


def write_model(out_dir, length_ms=None):
    shutil.copy(os.path.join(GOLD, "EdenTest_ArtificialCells.nml"), os.path.join(out_dir, "EdenTest_ArtificialCells.nml"))
    text = open(os.path.join(GOLD, "LEMS_EdenTest_ArtificialCells.xml")).read()
    text = text.replace('<Simulation id="sim1"', '<Simulation id="sim1" seed="1"')
    if length_ms is not None:
        text = text.replace('length="100.0ms"', f'length="{length_ms}ms"')
    path = os.path.join(out_dir, "LEMS_EdenTest_ArtificialCells.xml")
    with open(path, "w") as f:
        f.write(text)
    return path
