"""The reference's on-disk formats (cufcm_fil_amd/simio.py): the reference's own
globals.ini, psi.dat round trips, .par lines in the split the reference's reader uses
(pyfile/visual/myIo.py:31-36: ' %% '), per-step lines of the layout its analysis scripts
index (pyfile/visual/visual.py:459-480: step first, 7 values per swimmer)."""
import os

import numpy as np
import pandas as pd

from cufcm_fil_amd import simio

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# the keys cilia_sim_main.cu:57-94 reads from [Parameters] and [Filenames]
PARAMS = ["nswim", "nseg", "nfil", "nblob", "ar", "seg_sep", "period", "sim_length",
          "spring_factor", "tilt_angle", "pair_dp", "wavnum", "wavnum_dia", "dimensionless_force",
          "fene_model", "force_noise_mag", "omega_spread", "index", "fil_x_dim", "fil_spacing",
          "blob_x_dim", "blob_spacing", "hex_num", "reverse_fil_direction_ratio", "twofil_angle",
          "force_mag"]
FILES = ["simulation_dir", "simulation_file", "filplacement_file_name", "blobplacement_file_name",
         "cufcm_config_file_name", "simulation_icstate_name", "simulation_bodystate_name"]


def test_reads_the_reference_globals(tmp_path):
    g = simio.read_globals(os.path.join(GOLD, "globals.ini"))
    for k in PARAMS:
        assert k in g["Parameters"], k
    for k in FILES:
        assert k in g["Filenames"], k
    assert int(g["Parameters"]["nfil"]) == 40 and g["Parameters"]["seg_sep"] == 2.6
    p = tmp_path / "globals.ini"
    simio.write_globals(p, g["Parameters"], g["Filenames"])
    assert simio.read_globals(p) == g


def test_psi_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    ph, an = rng.uniform(0, 6.28, 16), rng.normal(size=16)
    p = tmp_path / "psi.dat"
    simio.write_psi(p, 0.01, 1.05, ph, an)
    k, T, ph2, an2 = simio.read_psi(p, 16)
    assert (k, T) == (0.01, 1.05) and np.array_equal(ph2, ph) and np.array_equal(an2, an)
    assert len(open(p).read().strip().splitlines()) == 1
    simio.write_psi(p, 0.0, 1.0, ph)
    assert simio.read_psi(p, 16)[3] is None


def test_par_and_state_lines(tmp_path):
    w = simio.OutputWriter(str(tmp_path) + "/", "ciliate", output_digit=15)
    pars = {"NFIL": 2, "NSEG": 3, "NSWIM": 1, "NBLOB": 4, "DT": 0.01, "PRESCRIBED_CILIA": 1,
            "MU": 1.0}
    w.write_par(pars)
    df = pd.read_csv(w.path("par"), sep=" %% ", header=None, engine="python")
    assert {df.iloc[i, 1]: df.iloc[i, 0] for i in range(len(df))} == pars
    assert list(df.iloc[:3, 1]) == ["NFIL", "NSEG", "NSWIM"]
    rng = np.random.default_rng(1)
    for step in (1, 11):
        x = rng.normal(size=(6, 3))
        bx, bq = rng.normal(size=(1, 3)), rng.normal(size=(1, 4))
        w.states(step, x, bx, bq)
        w.true_states(step, 1.0, rng.uniform(0, 6, 2), rng.normal(size=2))
    steps, rows = simio.read_lines(w.path("seg_states"))
    assert list(steps) == [1, 11] and rows[1].size == 18
    np.testing.assert_allclose(rows[1], x.reshape(-1), rtol=1e-14)
    steps, rows = simio.read_lines(w.path("body_states"))
    np.testing.assert_allclose(rows[1][3:7], bq.reshape(-1), rtol=1e-14)
    line = open(w.path("true_states")).read().splitlines()[0].split()
    assert line[0] == "1" and float(line[1]) == 1.0 and len(line) == 2 + 4
    assert line[2] == f"{float(line[2]):.15e}"     # scientific, OUTPUT_DIGIT digits
