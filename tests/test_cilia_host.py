"""Host-side cilia helpers (no GPU): the intrinsic phase speeds of the reference's PAIR 1
mode and the beat names of the paired cilia."""
import math

import numpy as np

from cufcm_fil_amd import _lib as L
from cufcm_fil_amd.cilia import BEATS, _beat, pair_phase_speeds


def test_pair_mode_scales_the_second_half_of_the_filaments():
    """filament.cpp:240-244 with NPAIR = NFIL / 2: ids with id / NPAIR == 1 beat pair_dp
    times faster; with an odd count the last filament (id / NPAIR == 2) keeps omega0."""
    w = pair_phase_speeds(6, 2.0, 1.5)
    assert np.array_equal(w, [2.0, 2.0, 2.0, 3.0, 3.0, 3.0])
    w = pair_phase_speeds(7, 2.0 * math.pi, 0.5)
    assert np.array_equal(w, [2.0 * math.pi] * 3 + [math.pi] * 3 + [2.0 * math.pi])


def test_paired_beat_names_map_to_the_header_constants():
    assert _beat("bicilia") == L.CFM_BEAT_BICILIA == 6
    assert _beat("bicilia_longt") == L.CFM_BEAT_BICILIA_LONGT == 7
    assert len(BEATS) == 7
