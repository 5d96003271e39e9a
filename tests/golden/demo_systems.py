"""The input vectors of the reference's demo drivers, as systems for the golden fixtures.

The demo drivers (src/cauchy_estimator.cpp) hold fixed dynamics, noise scales and measurement
records; make_golden.py steps the reference build through each of them and stores its outputs
as fixtures.  These are test inputs, not part of the library."""
from typing import List

import numpy as np

from cauchyfriendly_amd.workloads import System, explicit_system


def reference_demo_systems() -> List[System]:
    """Input vectors held by the reference's demo drivers (src/cauchy_estimator.cpp:53-76, 93-119,
    160-196); the reference constructs these estimators without calling srand (glibc seed 1)."""
    out = []
    out.append(explicit_system(
        "ref_2state", [0.9, 0.1, -0.2, 1.1], [1, 0.3], [1.0, 1.0], [0.1], [0.2], np.eye(2), [0.10, 0.05],
        np.zeros(2), [0.0338, 0.2049, -2.3543, -0.6042, -0.2662, 0.1307, -0.2250, 0.1951, -0.2191, 0.0996],
        srand_seed=1))
    out.append(explicit_system(
        "ref_3state", [1.4, -0.6, -1.0, -0.2, 1.0, 0.5, 0.6, -0.6, -0.2], [.1, 0.3, -0.2], [1.0, 0.5, 0.2],
        [0.1], [0.2], np.eye(3), [0.10, 0.08, 0.05], np.zeros(3),
        [-1.2172011200334241, -0.35943271347277583, -0.52353301003957098, 0.5855389648301792,
         -0.8048243525901404, 0.34053610027255954, 1.0580483915838776, -0.55152999529515989,
         -0.72879029737003309, -0.82415138330170357, 0.0], srand_seed=1))
    out.append(explicit_system(
        "ref_3state_3msmts_horth", [1.4, -0.6, -1.0, -0.2, 1.0, 0.5, 0.6, -0.6, -0.2], [.1, 0.3, -0.2],
        [1, 0, 0, 0, 1, 0, 0, 0, 1], [0.1], [0.2, 0.15, 0.10], np.eye(3), [0.1, 0.08, 0.05], np.zeros(3),
        [0.10943250903225685, 0.32131358116921616, -0.39352816664526724,
         0.76258687662854907, -0.25344840215960657, 0.1578820974338809,
         0.52543601367678883, -0.67309502187832315, -0.37267005411252474,
         2.7335350536863903, -0.3754139600950176, 0.6986657326616188,
         0.52558307773279223, 0.82802377147093642, 0.98211422248186553], srand_seed=1))
    # 4 states (src/cauchy_estimator.cpp:121-145)
    Phi4 = [1.4, -0.6, -1.0, 0.0, -0.2, 1.0, 0.5, 0.0, 0.6, -0.6, -0.2, 0.0, 0.0, 0.0, 0.0, 0.5]
    out.append(explicit_system(
        "ref_4state", Phi4, [0.1, 0.3, -0.2, 0.4], [2.0, 0.5, 0.2, -0.1], [0.1], [0.2], np.eye(4),
        [0.1, 0.08, 0.05, 0.2], np.zeros(4),
        [-0.26300165310514712, -0.98289343232730964, -0.93317363235517392, -0.81311530427193779,
         -0.24140673945883995, 0.013971096637110103, -0.4842328985975715, -0.1607056967588112], srand_seed=1))
    # 5 states (src/cauchy_estimator.cpp:207-241)
    out.append(explicit_system(
        "ref_5state",
        [0.9, -0.6, -0.5, 0.5, 0.5, 0.05, 1.0, 0.25, -0.25, -0.25, 0.25, -0.6, 0.15, 0.55, 0.35,
         0.1, 0.0, -0.1, 0.4, -0.1, 0.05, 0.0, -0.05, 0.45, 0.75],
        [0.1, 0.3, -0.2, 0.4, -0.15], [2.0, 0.5, 0.2, -0.1, 0.4], [0.1], [0.2], np.eye(5),
        [0.1, 0.08, 0.05, 0.2, 0.3], np.zeros(5),
        [-0.23451108940494397, -1.6969544453434717, -1.2059957432220769, -2.7147437491778885,
         -1.3823660959009454, -0.29548720069792656, -1.4872516450002977, -1.5538990428366166,
         -1.7101257926043105], srand_seed=1))
    # 4 states, two process noises (src/cauchy_estimator.cpp:254-285; its 7-step array holds 5
    # measurements, the last two are zero)
    out.append(explicit_system(
        "ref_4state_2pnoise", [1.4, -0.6, -1.0, 0.0, -0.2, 1.0, 0.5, 0.0, 0.6, -0.6, -0.2, 0.0, 0.0, 0.0, 0.0, 1.0],
        [[0.1, 0.0], [0.3, 0.0], [0.2, 0.0], [0.0, -1.0]], [0.4165285461783826, -0.60, -1.0, 1.0], [0.1, 0.001],
        [0.2], np.eye(4), [0.4, 0.5, 0.6, 0.70], np.zeros(4),
        [-5.3335189550166655, -4.4110988021211845, -3.6610012492599329, -2.5741683288219699,
         -6.5109959475268671, 0.0, 0.0], srand_seed=1))
    # 4 states, two measurements per step (src/cauchy_estimator.cpp:299-335)
    out.append(explicit_system(
        "ref_4state_2msmts", Phi4, [0.1, 0.3, -0.2, 0.4], [[2.0, 0.5, 0.2, -0.1], [0.4, -0.7, 1.3, -1.5]], [0.1],
        [0.2, 0.15], np.eye(4), [0.1, 0.08, 0.05, 0.2], np.zeros(4),
        [[-0.2630016531051471, 1.5804720565253951], [-0.8123538549377811, 0.4001811238098553],
         [-0.8607383321320500, -1.1124356889621634], [-0.1026529815581601, -0.6794624240892977],
         [-3.9121237378676339, -1.3582285870633606], [-0.5498081141794045, 1.1925540511414112]], srand_seed=1))
    return out
