"""Warnings before a solve (bem/problems_checks.py:27-160): mesh too coarse
for the wavelength, a finite depth deep enough to be infinite, frequencies
above the body's first irregular frequency.  Host-side; they only log."""
import logging

import numpy as np

LOG = logging.getLogger(__name__)

WATER_DEPTH_THRESHOLD_FACTOR = 5  # water_depth > 5 wavelengths: infinite depth would do


def _check_wavelength_and_mesh_resolution(problems):
    """problems_checks.py:27-59: a panel radius above wavelength / 8."""
    risky = [pb for pb in problems if 0.0 < pb.wavelength < pb.body.minimal_computable_wavelength]
    if risky:
        radius = max(pb.body.mesh_including_lid.faces_radiuses.max() for pb in risky)
        wl = np.array([pb.wavelength for pb in risky])
        which = "all" if len(risky) == len(problems) else f"{len(risky)}"
        LOG.warning(f"Mesh resolution for {which} problem(s): the resolution of the mesh might be insufficient "
                    f"for wavelengths from {wl.min():.3f} to {wl.max():.3f} m (largest panel radius {radius:.3f} m > "
                    f"wavelength/8).")
    return risky


def _check_wavelength_and_water_depth(problems):
    """problems_checks.py:62-111: finite depth more than 5 wavelengths and
    twice the draft."""
    risky = [pb for pb in problems
             if pb.water_depth < np.inf and pb.water_depth > WATER_DEPTH_THRESHOLD_FACTOR * pb.wavelength
             and pb.water_depth > -2 * pb.body.mesh.vertices[:, 2].min()]
    if risky:
        which = "all" if len(risky) == len(problems) else f"{len(risky)}"
        LOG.warning(f"Water depth for {which} problem(s): deep finite water depth has been set for wavelengths "
                    f"shorter than water_depth/{WATER_DEPTH_THRESHOLD_FACTOR}. Computation might be faster by using "
                    "the infinite water depth instead: `water_depth=np.inf`.")
    return risky


def _check_wavelength_and_irregular_frequencies(problems):
    """problems_checks.py:114-146: omega above the first irregular
    frequency estimate of the body."""
    risky = [pb for pb in problems
             if pb.free_surface != np.inf and pb.body.first_irregular_frequency_estimate(g=pb.g) < pb.omega < np.inf]
    if risky:
        if any(pb.body.lid_mesh is None for pb in problems):
            recommendation = "Setting a lid for the floating body is recommended."
        else:
            recommendation = "The lid might need to be closer to the free surface."
        om = np.array([pb.omega for pb in risky])
        which = "all" if len(risky) == len(problems) else f"{len(risky)}"
        LOG.warning(f"Irregular frequencies for {which} problem(s): irregular frequencies might be encountered for "
                    f"omega from {om.min():.3f} to {om.max():.3f}.\n" + recommendation)
    return risky


def check_problems(problems):
    _check_wavelength_and_mesh_resolution(problems)
    _check_wavelength_and_water_depth(problems)
    _check_wavelength_and_irregular_frequencies(problems)
