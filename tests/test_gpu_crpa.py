"""CRPA on the GPU: a 2 x 2 array of crossed thin-wire dipole pairs (each element's X and Y probes
through a 90-degree hybrid: its RHCP channel) under an RHCP satellite overhead and two linearly
polarized jammers near the horizon. Ideal adaptation (one measurement pass per emitter, sources
switched and fields reset between passes) and blind adaptation (everything at once, jammers off
the carrier, I/Q snapshots, receiver noise): the MVDR weights null both jammers and lift the
SINR far above plain steering, and the adapted pass transmits with the weights."""
import numpy as np
import pytest

from em_lab_amd import crpa as CR
from em_lab_amd import emfdtd as E
from em_lab_amd.scene import crpa_crossed_dipoles

pytestmark = pytest.mark.gpu

LAM = 20.0


def _run(blind):
    sc = crpa_crossed_dipoles(LAM, blind)
    g = E.SimulationGrid(0)
    try:
        g.configure(sc.grid_config())
        s, e = sc.flat_materials()
        g.upload_materials(s, e, sc.subcell)
        mon = CR.CrpaMonitor(g)
        mon.configure(sc.crpa, sc.emitters, sc.sources, sc.ports, LAM, sc.cell_size)
        assert mon.phase == (CR.CrpaMonitor.LISTENING if blind else CR.CrpaMonitor.MEASURING)
        res = CR.run_crpa(g, mon, steps_per_frame=8, adapted_steps=200)
        assert mon.phase == CR.CrpaMonitor.ADAPTED and res.valid
        # the adapted pass transmits from the element feeds with the weights
        assert g.step_count >= 200 and np.abs(g.fields(E.LAYOUT_SOA)[0]).max() > 0
        for line in mon.report_lines():
            print(line, flush=True)
        return res, mon
    finally:
        g.close()


def test_crpa_ideal_nulls_jammers():
    print("case: crpa ideal, 2x2 crossed dipoles, 2 jammers", flush=True)
    res, mon = _run(False)
    sat = [e for e in res.emitters if e.satellite][0]
    jams = [e for e in res.emitters if not e.satellite]
    assert sat.az_deg == pytest.approx(mon.compass_angles((62.0, 58.0, 84.0))[0])
    for j in jams:
        assert j.output_db < j.input_db - 25.0          # nulled
        assert j.output_db < j.steering_out_db - 15.0   # far deeper than plain steering leaves it
        assert j.settle_db < -20.0                      # the pass reached its steady state
    assert res.sinr_mvdr_db > res.sinr_steering_db + 20.0
    assert res.sinr_mvdr_db > res.sinr_single_db + 20.0
    assert sat.cross_pol_db < -3.0                      # RHCP satellite arrives mostly in the RHCP channels


def test_crpa_blind_nulls_jammers():
    print("case: crpa blind, 2x2 crossed dipoles, 2 jammers", flush=True)
    res, mon = _run(True)
    assert res.blind and res.snapshots == 64
    jams = [e for e in res.emitters if not e.satellite]
    assert [round(e.offset_pct, 4) for e in jams] == [round(100 / 16, 4), round(-100 / 16, 4)]
    for j in jams:
        assert j.output_db < j.input_db - 15.0
    assert res.sinr_mvdr_db > res.sinr_steering_db + 15.0
