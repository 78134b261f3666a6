"""Objects exported from MATLAB (optifax_amd/matlab/optifax_export.m) as JSON,
rebuilt as this package's physics objects.

The toolbox stores its objects as MATLAB .mat files (Laser.m:133,
Optic.m:215, Cavity.m:153, NonlinearCrystal.m:324: obj.store), which only
MATLAB reads.  optifax_export.m writes the properties the engine uses as
JSON; this module reads that JSON:

    d = interchange.load("sim.json")                 # an exported OpticalSim
    run = interchange.opo_from(d, data_dir=..., glass_table=...)
    OPOBatch([run["instance"]], ...)

Component builders (each takes the exported dict of one object):
`simwindow_from`, `crystal_from`, `pump_from` (a sech^2 Laser), `optic_from`
(an Optic's transmission, reflection and phase through
physics.optic_response) and `chain_from` (a Cavity's optics around the
detector position, as OpticalSim.run propagates them: OpticalSim.m:274-283).

Function handles arrive as {"function": "<func2str>"}; the ones this package
has a restatement of are mapped (@prismfn_PB: a prism's path length), any
other raises, since MATLAB code cannot run here.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import physics as ph

KNOWN_FUNCTIONS = {
    "prismfn_PB": ph.prism_pb_path_length,
    "@prismfn_PB": ph.prism_pb_path_length,
}


def load(path):
    """The exported dict (checks the format tag)."""
    with open(path) as f:
        d = json.load(f)
    if d.get("format") != "optifax-export":
        raise ValueError(f"{path}: not an optifax_export file")
    return d


def _func(v, what):
    if isinstance(v, dict) and "function" in v:
        name = v["function"].strip()
        if name in KNOWN_FUNCTIONS:
            return KNOWN_FUNCTIONS[name]
        raise ValueError(f"{what}: MATLAB function {name!r} has no restatement here")
    return v


def _num(v, default=0.0):
    if v is None or (isinstance(v, list) and not v):
        return default
    return v


def simwindow_from(d):
    """SimWindow (SimWindow.m:3-10): reference wavelength, points, temporal
    range, spectral limits [nm].  A nonzero TimeOffset is not modelled."""
    if _num(d.get("TimeOffset"), 0.0) not in (0, 0.0):
        raise ValueError("SimWindow.TimeOffset other than 0 is not supported")
    return ph.SimWindow(float(d["ReferenceWave"]), int(d["NumberOfPoints"]),
                        float(d["TemporalRange"]),
                        spectral_limits_nm=tuple(float(x) for x in d["SpectralLimits"]))


def crystal_from(d):
    """NonlinearCrystal (NonlinearCrystal.m:3-60): the grating (one period,
    [P1 P2] fan-out, stepped list), its uncertainty and duty offset, height
    and vertical position, with the bulk's length, material, temperature and
    the crystal's Chi2."""
    if d.get("class") != "NonlinearCrystal":
        raise ValueError("not a NonlinearCrystal")
    period = _func(d["GratingPeriod"], "GratingPeriod")
    b = d["Bulk"]
    return ph.NonlinearCrystal(period=period, length=float(b["Length"]), chi2=float(d["Chi2"]),
                               temperature=float(b["Temperature"]),
                               duty_off=float(_num(d.get("DutyCycleOffset"), 0.0)),
                               material=str(b["Material"]),
                               uncertainty=float(_num(d.get("Uncertainty"), 0.0)),
                               height=float(_num(d.get("Height"), 1.0)),
                               vertical_position=float(_num(d.get("VerticalPosition"), 1.0)))


def pump_from(laser, simwin, radius):
    """The pump field of a sech^2 Laser (Laser.m:66-77, sechPulse.m): energy
    AveragePower / RepetitionRate, PulseDuration as the FWHM, in the area
    pi radius^2 (OpticalSim.PumpRadiusXtalIn)."""
    src = str(laser.get("SourceString") or "sech")
    if "sech" not in src.lower():
        raise ValueError(f"pump source {src!r}: only sech^2 pulses are built here")
    energy = float(laser["AveragePower"]) / float(laser["RepetitionRate"])
    return ph.sech_pump(simwin, float(laser["PulseDuration"]), energy, math.pi * radius ** 2)


def _coating(v):
    v = _func(v, "Coating")
    if isinstance(v, list):
        return np.asarray(v, dtype=np.float64)
    return v


def _gdd(v, data_dir):
    if v is None or (not isinstance(v, str) and float(np.sum(np.abs(np.atleast_1d(v)))) == 0.0):
        return None
    if isinstance(v, str):
        return ph.load_gdd(v if os.path.isabs(v) or data_dir is None else os.path.join(data_dir, v))
    raise ValueError("a GDD given as numbers is a placeholder in the reference (OpticalSurface.m:109)")


def optic_from(d, simwin, glass_table=None, transmission_table=None, data_dir=None):
    """(regime, (transmission, reflection, phase)) of an exported Optic
    through physics.optic_response (Optic.m:79-99)."""
    b = d["Bulk"]
    length = _func(b["Length"], "Bulk.Length")
    s1, s2 = d["S1"], d["S2"]
    return d["Regime"], ph.optic_response(
        simwin, d["Regime"], _coating(s1["Coating"]), str(b["Material"]), length,
        float(_num(s1.get("IncidentAngle"), 0.0)), _coating(s2["Coating"]), glass_table,
        float(b["Temperature"]), _gdd(s1.get("GDD"), data_dir), _gdd(s2.get("GDD"), data_dir),
        transmission_table=transmission_table, data_dir=data_dir)


def chain_from(cavity, simwin, detector, glass_table=None, transmission_table=None,
               data_dir=None):
    """A Cavity's optics as OpticalSim.run meets them after the crystal
    (OpticalSim.m:274-283): those before the detector position act on the
    output and the fed-back field, the detector optic couples out, those
    after it act on the fed-back field only.  The crystal's own entry is the
    solver and is skipped.  Returns (before, coupler, after): the chains as
    (power transmission, phase) and the coupler optic's response."""
    optics = cavity["Optics"]
    k = int(detector) - 1
    if not 0 <= k < len(optics):
        raise ValueError("detector position outside the cavity")

    def resp(o):
        return optic_from(o, simwin, glass_table, transmission_table, data_dir)

    before = [resp(o) for o in optics[:k] if o["class"] != "NonlinearCrystal"]
    after = [resp(o) for o in optics[k + 1:] if o["class"] != "NonlinearCrystal"]
    cpl = resp(optics[k])
    chain = lambda xs: ph.cavity_response(xs) if xs else None
    return chain(before), cpl, chain(after)


class _CouplerMasks:
    """A coupler optic as cavity_operators' output coupler: the fed-back
    power is the reflection of an "R" optic (the transmission of a "T"
    optic), the output the rest; its phase is the optic's."""

    def __init__(self, regime, response):
        t, r, phi = response
        self.fb, self.out = (r, t) if regime == "R" else (t, r)
        self.phi = np.nan_to_num(phi)

    def masks(self, simwin):
        return self.fb, self.out, self.phi


def opo_from(sim, glass_table=None, transmission_table=None, data_dir=None, noise_seed=0):
    """An exported OpticalSim as a device instance: the SimWindow, the
    crystal's solver inputs (StepSize, AdaptiveError as [min max] percent,
    ProgressPlots), the pump, the cavity operators from the chain around
    DetectorPosition, seeded one-photon noise, and the per-trip Delay.
    Returns dict(simwin, crystal, instance) with `instance` ready for
    solver.OPOBatch."""
    if sim.get("class") != "OpticalSim":
        raise ValueError("not an OpticalSim")
    sw = simwindow_from(sim["SimWin"])
    cav = sim["System"]
    xt = next((o for o in cav["Optics"] if o["class"] == "NonlinearCrystal"), None)
    if xt is None:
        raise ValueError("the cavity holds no NonlinearCrystal")
    crystal = crystal_from(xt)
    step = float(_num(sim.get("StepSize"), xt.get("StepSize", 1e-7)))
    si = crystal.solver_inputs(sw, step)
    err = np.atleast_1d(np.asarray(_num(sim.get("AdaptiveError"), [2.5e-3, 5e-2]), float))
    plots = int(_num(sim.get("ProgressPlots"), 5))
    si.update(minErr=float(err[0]), maxErr=float(err[-1]),
              sel=float(math.ceil(si["nSteps"] / max(plots - 1, 1))), n_cols=max(plots - 1, 0))
    radius = float(_num(sim.get("PumpRadiusXtalIn"), 50e-6))
    pump = pump_from(sim["Source"], sw, radius)
    noise = ph.quantum_noise(sw, math.pi * radius ** 2, noise_seed)
    before, cpl, after = chain_from(cav, sw, int(_num(sim.get("DetectorPosition"), 1)) or 1,
                                    glass_table, transmission_table, data_dir)
    ops = ph.cavity_operators(sw, si["n_xtal"], _CouplerMasks(*cpl), pump, noise,
                              chain_before_oc=before, chain_after_oc=after)
    delay = float(np.atleast_1d(sim.get("Delay", 0.0))[0])
    inst = dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"], ocm=ops["ocm"],
                delay=delay)
    return dict(simwin=sw, crystal=crystal, instance=inst, ops=ops, pump=pump, noise=noise)
