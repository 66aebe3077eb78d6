"""The NeuroML / LEMS reader (eden_amd.neuroml), the exporter (eden_amd.nml_export) and the
command-line front end, on the CPU.

* The reference's own example model (examples/NML2_SingleCompHHCell.nml, kept under
  tests/golden/neuroml/ as input data, with a LEMS simulation file of this repository's)
  instantiates to the numbers its elements state, computed here independently: sphere area
  pi d^2, C = Cm A, densities in mS/cm^2, rates in 1/us, the pulse in us and nA
  (eden/NeuroML.cpp:305-425 fixes the engine units).
* Quantity conversion follows ParseQuantity's arithmetic (eden/NeuroML.cpp:3046-3077).
* Every synthetic workload written as LEMS by the exporter reads back into the same
  instantiated model, array for array (spike recipients per source as a set: their order
  does not change results, flags being idempotent).
* The reference's full EdenTest_CoreSynapses model, which holds constructs outside the
  per-step loop's mechanism set, is refused with an error naming the first of them.
"""
import dataclasses
import os

import numpy as np
import pytest

from eden_amd import configs, neuroml, nml_export

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")


def test_quantity_conversion():
    c = neuroml.convert
    assert c("-65mV", "voltage") == -65.0
    assert c("0.03 kohm_cm", "resistivity") == pytest.approx(0.03)
    assert np.float32(c("3.0 S_per_m2", "conductivity")) == np.float32(0.3)
    assert c("360 S_per_m2", "conductivity") == pytest.approx(36.0)
    assert c("1per_ms", "frequency") == 1e-3
    assert np.float32(c("5Hz", "frequency")) == np.float32(5e-6)   # 5 * 10^-6 in double, then float
    assert c("100ms", "time") == 100000.0
    assert c("0.08nA", "current") == 0.08
    assert c("20nS", "conductance") == pytest.approx(0.02)
    assert c("1.0 uF_per_cm2", "specific_capacitance") == 1.0
    assert c("0V", "voltage") == 0.0
    with pytest.raises(neuroml.NeuroMLError):
        c("5 furlongs", "time")
    with pytest.raises(neuroml.NeuroMLError):
        c("abc", "time")


def test_reference_hh_example_instantiates_to_its_stated_numbers():
    m, outs = neuroml.read(os.path.join(GOLD, "LEMS_HHCellExample.xml"))
    assert m.n_cells == 1 and m.n_comps == 1
    d = 17.841242
    area = np.float32(np.pi * np.float32(d) ** 2)
    assert m.area[0] == pytest.approx(area, rel=1e-6)
    assert m.capacitance[0] == pytest.approx(1.0 * area * 1e-2, rel=1e-6)     # uF/cm^2 * um^2 -> pF
    assert list(m.chan_type) == [0, 1, 2]                                     # passive, Na, K in file order
    np.testing.assert_allclose(m.chan_gbase, [0.3, 120.0, 36.0], rtol=1e-7)
    np.testing.assert_allclose(m.chan_erev, [-54.3, 50.0, -77.0], rtol=1e-7)
    na = m.channel_types[1]
    assert na.n_gates == 2 and na.gates[0].instances == 3 and na.gates[1].instances == 1
    assert na.gates[0].alpha.rate == np.float32(1e-3) and na.gates[0].alpha.midpoint == -40.0
    assert m.channel_types[2].gates[0].beta.rate == np.float32(0.125e-3)
    assert m.channel_types[0].n_gates == 0
    assert m.v0[0] == -65.0
    assert m.pulse_start[0] == 100000.0 and m.pulse_duration[0] == 100000.0
    assert m.pulse_amplitude[0] == np.float32(0.08)
    assert m.dt == 10.0 and m.t_final == 300000.0
    assert [c[1] for c in outs.trajectories[0].columns] == ["hhpop/0/hhcell/0/v"]
    assert outs.events[0].format == "TIME_ID"
    assert m.v_threshold[0] == -20.0


@pytest.mark.parametrize("which", ["config1", "config2", "config4", "config3", "mixed"])
def test_exported_workloads_read_back_to_the_same_model(tmp_path, which):
    m = {
        "config1": lambda: configs.hh_population(n_cells=30, t_final_ms=20.0),
        "config2": lambda: configs.random_trees(6, sizes=[20, 45, 80, 130, 260, 400], t_final_ms=20.0),
        "config4": lambda: configs.config4(n_cells=8, t_final_ms=10.0),
        "config3": lambda: configs.spiking_network(n_cells=400, in_degree=12, t_final_ms=10.0),
        "mixed": lambda: configs.mixed_network(),
    }[which]()
    path = nml_export.write_lems(m, str(tmp_path / f"{which}.xml"), m.morphologies)
    r, outs = neuroml.read(path)
    for f in dataclasses.fields(m):
        a, b = getattr(m, f.name), getattr(r, f.name)
        if f.name in ("name", "spike_target"):
            continue
        if f.name in ("channel_types", "syn_types"):
            assert len(a) == len(b)
            assert all(bytes(x) == bytes(y) for x, y in zip(a, b)), f.name
            continue
        if isinstance(a, np.ndarray):
            assert a.shape == b.shape, f.name
            assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f"), f.name
        else:
            assert a == b, f.name
    for g in range(m.n_comps):
        s0, s1 = m.comp_spike_begin[g], m.comp_spike_begin[g + 1]
        assert sorted(m.spike_target[s0:s1]) == sorted(r.spike_target[s0:s1])


def test_reference_core_synapses_model_is_refused_by_name():
    """The full model holds a blocking/plastic synapse, graded and gap-junction projections,
    which the per-step loop does not cover: refused, naming the first of them."""
    with pytest.raises(neuroml.NeuroMLError, match="blockingPlasticSynapse"):
        neuroml.read(os.path.join(GOLD, "LEMS_EdenTest_CoreSynapses.xml"))


def test_reader_errors_name_the_problem(tmp_path):
    p = tmp_path / "bad.xml"
    p.write_text('<Lems><Target component="s"/><neuroml><network id="n"><population id="a" component="nope" '
                 'size="2"/></network></neuroml><Simulation id="s" length="1ms" step="0.025ms" target="n" '
                 'seed="1"/></Lems>')
    with pytest.raises(neuroml.NeuroMLError, match="nope"):
        neuroml.read(str(p))


def test_cli_reports_missing_model_and_unknown_tokens(capsys):
    from eden_amd.__main__ import main
    assert main(["frobnicate"]) == 2
    out = capsys.readouterr().out
    assert 'skipping unknown token "frobnicate"' in out
    assert "NeuroML model not selected" in out
    assert main(["nml", "/nonexistent/model.xml"]) == 1
