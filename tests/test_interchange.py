"""Objects exported from MATLAB as JSON (optifax_amd/matlab/optifax_export.m)
rebuilt by optifax_amd.interchange.  No MATLAB runs here: the JSON below is
written in the exporter's layout by hand, and the rebuilt objects are checked
against the same setup built directly (OPOConfig, NonlinearCrystal,
optic_response)."""
import json
import math
import os

import numpy as np
import pytest

from optifax_amd import interchange as ic
from optifax_amd import physics as ph

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def surface(coating, angle=0.0, gdd=0):
    return {"Coating": coating, "IncidentAngle": angle, "GDD": gdd}


def ring_export(tmp_path):
    xtal = {"class": "NonlinearCrystal", "Regime": "T", "S1": surface("AR"), "S2": surface("AR"),
            "Bulk": {"Material": "PPLN", "Length": 5e-3, "Temperature": 20},
            "GratingPeriod": 30.5e-6, "Uncertainty": [], "DutyCycleOffset": [], "StepSize": 1e-7,
            "Height": 1, "VerticalPosition": 1, "WaistPosition": 0, "Chi2": 2 * 27e-12,
            "Name": "PPLN"}
    oc = {"class": "Optic", "Regime": "R", "S1": surface(5), "S2": surface("AR"),
          "Bulk": {"Material": "FS", "Length": 1e-3, "Temperature": 20}, "Name": "OC"}
    sim = {"format": "optifax-export", "version": 1, "class": "OpticalSim",
           "SimWin": {"class": "SimWindow", "ReferenceWave": 1040e-9, "NumberOfPoints": 4096,
                      "TemporalRange": 12e-12, "TimeOffset": 0, "SpectralLimits": [300, 6000]},
           "Source": {"class": "Laser", "Wavelength": 1040e-9, "Waist": 50e-6,
                      "RepetitionRate": 1e9, "AveragePower": 2.0, "PulseDuration": 200e-15,
                      "SourceString": "sech", "PhaseString": None},
           "System": {"class": "Cavity", "Name": "ring", "OCPosition": 2,
                      "Optics": [xtal, oc]},
           "Delay": 0.3e-12, "RoundTrips": 10, "StepSize": 1e-7,
           "AdaptiveError": [2.5e-3, 5e-2], "ProgressPlots": 5, "Precision": "single",
           "PumpRadiusXtalIn": 50e-6, "DetectorPosition": 2}
    p = tmp_path / "sim.json"
    p.write_text(json.dumps(sim))
    return p


def test_exported_ring_rebuilds_the_same_device_instance(tmp_path):
    d = ic.load(ring_export(tmp_path))
    run = ic.opo_from(d, noise_seed=7)
    cfg = ph.OPOConfig(n=4096, window=12e-12, oc_band=None, delay=0.3e-12, noise_seed=7)
    b = cfg.build()
    inst = run["instance"]
    for k in ("cav", "ocm", "pump", "field0"):
        assert np.allclose(inst[k], b["ops"][k], rtol=1e-14, atol=0), k
    for k in ("beta_op", "alpha_w", "G33"):
        assert np.array_equal(np.asarray(inst["solver"][k]), np.asarray(b["solver"][k])), k
    for k in ("nSteps", "maxErr", "minErr", "sel", "n_cols", "w0", "dt"):
        assert inst["solver"][k] == b["solver"][k], k
    assert inst["delay"] == 0.3e-12


def test_components_and_refusals(tmp_path):
    d = ic.load(ring_export(tmp_path))
    sw = ic.simwindow_from(d["SimWin"])
    assert sw.n == 4096 and sw.spectral_limits == (300.0, 6000.0)
    xt = ic.crystal_from(d["System"]["Optics"][0])
    assert xt.local_period() == 30.5e-6 and xt.length == 5e-3
    # a fan-out grating at its vertical position
    fan = dict(d["System"]["Optics"][0], GratingPeriod=[29e-6, 31e-6], Height=1.0,
               VerticalPosition=0.5)
    assert ic.crystal_from(fan).local_period() == pytest.approx(30e-6)
    # MATLAB code cannot run here: an anonymous coating function is refused
    bad = dict(d["System"]["Optics"][1], S1=surface({"function": "@(lam) 0.5 + 0 * lam"}))
    with pytest.raises(ValueError):
        ic.optic_from(bad, sw)
    with pytest.raises(ValueError):           # a JSON file that is not an export
        ic.load(_plain(tmp_path))


def _plain(tmp_path):
    p = tmp_path / "plain.json"
    p.write_text(json.dumps({"class": "Laser"}))
    return p


def test_prism_optic_maps_its_path_function():
    """Optic("T", 'None', glass, @prismfn_PB, theta): the path length is the
    prism function of the index, as physics.path_length evaluates it."""
    sw = ph.SimWindow(1040e-9, 2 ** 12, 12e-12)
    gtab = ph.load_material_table(os.path.join(GOLDEN, "materials_transmission.csv"))
    prism = {"class": "Optic", "Regime": "T", "S1": surface("AR", 57.0), "S2": surface("AR", 57.0),
             "Bulk": {"Material": "H-ZLaF68C", "Length": {"function": "@prismfn_PB"},
                      "Temperature": 20}}
    regime, (t, r, phi) = ic.optic_from(prism, sw, glass_table=gtab)
    want = ph.bulk_dispersion(sw, "H-ZLaF68C", ph.prism_pb_path_length, gtab)
    ok = np.isfinite(want)
    assert regime == "T" and np.allclose(t, 1.0) and np.array_equal(phi[ok], want[ok])
