"""The .conf front end (include/hawaii_conf.h): parsing of the reference driver's format on
the CPU.  Runs on the GPU are in test_gpu_driver.py."""
import ctypes as C
import os

import pytest

import hawaiihybrid_amd as H

HERE = os.path.dirname(os.path.abspath(__file__))
REF_CONFIGS = "/root/reference/example_configs"


class Proc(C.Structure):
    _fields_ = [("op", C.c_int), ("arg", C.c_char * 512)]


class Conf(C.Structure):
    _fields_ = [("calculation_type", C.c_int), ("pair_state", C.c_int), ("temperature", C.c_double),
                ("reduced_mass", C.c_double), ("so_potential", C.c_char * 512), ("so_dipole", C.c_char * 512),
                ("niterations", C.c_long), ("total_trajectories", C.c_long), ("cvode_tolerance", C.c_double),
                ("sampling_time", C.c_double), ("max_trajectory_length", C.c_long), ("sampler_Rmin", C.c_double),
                ("sampler_Rmax", C.c_double), ("pesmin", C.c_double), ("initialM0_npoints", C.c_long),
                ("cf_filename", C.c_char * 512), ("sf_filename", C.c_char * 512), ("R0", C.c_double),
                ("Rcut", C.c_double), ("ppf", C.c_double), ("use_zimmermann_trick", C.c_int),
                ("n_monomers", C.c_int), ("monomer_type", C.c_int * 2), ("II", (C.c_double * 3) * 2),
                ("nproc", C.c_int), ("proc", Proc * 128),
                ("hep_ppf_niterations", C.c_long), ("hep_ppf_npoints", C.c_long),
                ("hep_m0_niterations", C.c_long), ("hep_m0_npoints", C.c_long),
                ("ntemperatures", C.c_int), ("temperatures", C.c_double * 64),
                ("nsatellite", C.c_int), ("satellite_temperatures", C.c_double * 64),
                ("nppf_ratios", C.c_int), ("ppf_ratios", C.c_double * 64),
                ("ncf_filenames", C.c_int), ("cf_filenames", (C.c_char * 512) * 64),
                ("initialM2_npoints", C.c_long), ("hep_m2_niterations", C.c_long), ("hep_m2_npoints", C.c_long),
                ("approx_frequency_max", C.c_double), ("average_time_between_collisions", C.c_double),
                ("spectrum_frequency_max", C.c_double)]


def parse_text(tmp_path, text):
    p = tmp_path / "x.conf"
    p.write_text(text)
    c = Conf()
    rc = H.lib().hh_conf_parse(str(p).encode(), C.byref(c))
    return rc, c


def parse(path):
    c = Conf()
    rc = H.lib().hh_conf_parse(path.encode(), C.byref(c))
    return rc, c


def test_parse_small_correlation_config():
    rc, c = parse(os.path.join(HERE, "conf", "cf-hear-small.conf"))
    assert rc == 0, H.lib().hh_last_error()
    assert c.calculation_type == 1 and c.pair_state == H.PAIR_FREE
    assert c.temperature == 300.0 and c.reduced_mass == 6.632039160324215e3
    assert c.niterations == 2 and c.total_trajectories == 512 and c.max_trajectory_length == 256
    assert c.cvode_tolerance == 1e-12 and c.sampling_time == 200.0 and c.Rcut == 40.0
    assert c.initialM0_npoints == 100 and c.ppf == 1.0
    assert c.cf_filename == b"CF-hear-small.txt" and c.n_monomers == 2 and list(c.monomer_type) == [1, 1]


@pytest.mark.parametrize("name", ["CF-F-CO2-Ar.conf", "CF-F-He-Ar.conf", "CF-B-CO2-Ar.conf"])
def test_parse_reference_example_configs(name):
    """The reference's own example configurations parse (read in place; skipped where the
    reference tree is absent)."""
    path = os.path.join(REF_CONFIGS, name)
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    rc, c = parse(path)
    assert rc == 0, H.lib().hh_last_error()
    assert c.calculation_type == 1 and c.n_monomers == 2
    assert c.total_trajectories > 0 and c.max_trajectory_length > 0 and c.Rcut > 0


def test_parse_reference_m2_settings():
    """CF-F-He-Ar.conf asks for the preliminary M2 over 10^6 points (src/driver.c:1193)."""
    path = os.path.join(REF_CONFIGS, "CF-F-He-Ar.conf")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    rc, c = parse(path)
    assert rc == 0, H.lib().hh_last_error()
    assert c.initialM0_npoints == 1000000 and c.initialM2_npoints == 1000000


def test_parse_phase_space_m2(tmp_path):
    """CALCULATE_PHASE_SPACE_M2 and the VEGAS M2 settings (src/driver.c:1106-1107, 1196-1197); the
    reference spells the phase-space types with CALCULATE_, a bare PHASE_SPACE_M0 is unknown."""
    rc, c = parse_text(tmp_path, """
&INPUT
  CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M2
  TEMPERATURE = 100.0
  HEP_M2_NITERATIONS = 8
  HEP_M2_NPOINTS = 500_000
  INITIALM2_NPOINTS = 1000
&END
""")
    assert rc == 0, H.lib().hh_last_error()
    assert c.calculation_type == 5 and c.hep_m2_niterations == 8 and c.hep_m2_npoints == 500000
    assert c.initialM2_npoints == 1000
    rc, _ = parse_text(tmp_path, "&INPUT\n  CALCULATION_TYPE = PHASE_SPACE_M0\n&END\n")
    assert rc < 0 and b"PHASE_SPACE_M0 is not part of this engine" in H.lib().hh_last_error()


def test_parse_reference_prmu_config():
    """The reference's PR_MU example (CO-Ar with a requantised rotor): the PR/mu settings parse;
    the rotor type is one this engine does not run."""
    path = os.path.join(REF_CONFIGS, "prmu_co_ar_req+poisson.conf")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    rc, c = parse(path)
    assert rc == 0, H.lib().hh_last_error()
    assert c.calculation_type == 6 and c.R0 == 40.0 and c.approx_frequency_max == 150.0
    assert c.average_time_between_collisions == 2.596e6 and c.max_trajectory_length == 65536
    assert c.sf_filename == b"SF-PRMU-CO-Ar-300.0.txt" and c.initialM2_npoints == 10000
    assert c.monomer_type[0] == 3 and c.monomer_type[1] == 1   # HH_MONOMER_OTHER + ATOM


def test_parse_reference_correlation_array_config():
    """The reference's CORRELATION_ARRAY example (six satellite temperatures of bound He-Ar)."""
    path = os.path.join(REF_CONFIGS, "CF-B-He-Ar.conf")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    rc, c = parse(path)
    assert rc == 0, H.lib().hh_last_error()
    assert c.calculation_type == 4 and c.pair_state == 2 and c.temperature == 100.0
    assert c.nsatellite == 6 and list(c.satellite_temperatures[:6]) == [50.0, 60.0, 70.0, 80.0, 90.0, 100.0]
    assert c.ncf_filenames == 6 and c.cf_filenames[0].value == b"CF-B-He-Ar-50.0.txt"
    assert c.nppf_ratios == 0


def test_parse_linear_molecule_and_processing(tmp_path):
    rc, c = parse_text(tmp_path, """
&INPUT
  CALCULATION_TYPE = correlation_single   ! keywords and values are case-insensitive
  PAIR_STATE = BOUND
  pair_reduced_mass = 3.817083069753919e+04
  TEMPERATURE = 100.0
  hep_m0_npoints = 1000                    ! VEGAS settings
&END
&MONOMER
  MONOMER_TYPE = LINEAR_MOLECULE
  II = {2.819827114123505e+05, 2.819827114123505e+05, 0.0}
&END
&MONOMER
  MONOMER_TYPE = ATOM
&END
&PROCESSING
  READ_CF("a.txt")
  DUP()
  CF_TO_SF()
  WRITE_SF("b.txt")
&END
""")
    assert rc == 0, H.lib().hh_last_error()
    assert c.pair_state == H.PAIR_BOUND and list(c.monomer_type) == [2, 1]
    assert c.II[0][0] == 2.819827114123505e5 and c.II[0][2] == 0.0
    assert c.nproc == 4 and [c.proc[i].op for i in range(4)] == [0, 4, 2, 9]
    assert c.proc[0].arg == b"a.txt" and c.proc[3].arg == b"b.txt"


def test_parse_processing_spectra(tmp_path):
    rc, c = parse_text(tmp_path, """
&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  SPECTRUM_FREQUENCY_MAX = 500.0
  READ_CF("a.txt")
  D4a()
  ALPHA()
  WRITE_SPECTRUM("b.txt")
  READ_SPECTRUM("b.txt")
  D1()
&END
""")
    assert rc == 0, H.lib().hh_last_error()
    assert c.spectrum_frequency_max == 500.0
    assert [c.proc[i].op for i in range(c.nproc)] == [0, 17, 12, 11, 10, 13]


def test_parse_baseline_smooth_int3(tmp_path):
    """FIT_BASELINE with no, a positional and a named index (the reference's
    tests/convert-cf-to-sf.conf), SMOOTH's named arguments (tests/smooth.conf; defaults
    window_size_min 21, window_size_step 0.1, window_size_delay 10, src/driver.c:208-210), INT3."""
    rc, c = parse_text(tmp_path, """
&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_CF("cf.txt")
  FIT_BASELINE()
  FIT_BASELINE(cf_extrapolation_begin_index=8000)
  FIT_BASELINE(8000)
  READ_SF("sf.txt")
  SMOOTH(grid_npoints=1200, grid_max=1200.0, window_size_delay=10, window_size_step=0.05)
  SMOOTH(grid_npoints=10, grid_max=5.0)
  INT3()
&END
""")
    assert rc == 0, H.lib().hh_last_error()
    ops = [c.proc[i].op for i in range(c.nproc)]
    assert ops == [0, 22, 22, 22, 1, 23, 23, 24]
    assert [c.proc[i].arg for i in (1, 2, 3)] == [b"-1", b"8000", b"8000"]
    f = [float(v) for v in c.proc[5].arg.split(b";")]
    assert f == [1200, 1200.0, 21, 0.05, 10]
    f = [float(v) for v in c.proc[6].arg.split(b";")]
    assert f == [10, 5.0, 21, 0.1, 10]


@pytest.mark.parametrize("text,needle", [
    ("&INPUT\n  FOO = 1\n&END\n", b"x.conf:2: unknown keyword FOO"),
    ("&INPUT\n  CALCULATION_TYPE = PR_MU_ARRAY\n&END\n", b"PR_MU_ARRAY is not part of this engine"),
    ("&INPUT\n  CALCULATION_TYPE = PROCESSING\n&END\n&PROCESSING\n  NORMALIZE()\n&END\n",
     b"NORMALIZE is not part of this engine"),
    # the reference's tests/fit-baseline-nargs.conf
    ("&INPUT\n  CALCULATION_TYPE = PROCESSING\n&END\n&PROCESSING\n  READ_CF(\"cf.txt\")\n  FIT_BASELINE(8000, 124)\n&END\n",
     b"FIT_BASELINE expects at most 1 arguments"),
    ("&INPUT\n  CALCULATION_TYPE = PROCESSING\n&END\n&PROCESSING\n  FIT_BASELINE(begin=8000)\n&END\n",
     b"expects a named argument cf_extrapolation_begin_index but got begin"),
    ("&INPUT\n  CALCULATION_TYPE = PROCESSING\n&END\n&PROCESSING\n  SMOOTH(1200)\n&END\n",
     b"SMOOTH accepts only named arguments"),
    ("&INPUT\n  TEMPERATURE = \"hot\"\n&END\n", b"TEMPERATURE expects a number"),
    ("&INPUT\n  TEMPERATURE = 300.0\n", b"not closed"),
])
def test_parse_errors_name_the_problem(tmp_path, text, needle):
    rc, _ = parse_text(tmp_path, text)
    assert rc < 0
    assert needle in H.lib().hh_last_error()


def test_parse_phase_space_m0(tmp_path):
    rc, c = parse_text(tmp_path, """
&INPUT
  CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M0
  PAIR_STATE = FREE_AND_METASTABLE
  pair_reduced_mass = 6.632039160324215e3
  TEMPERATURES = {100.0, 200.0, 300.0}
  sampler_Rmin = 3.0
  sampler_Rmax = 40.0
  HEP_M0_NITERATIONS = 10
  HEP_M0_NPOINTS = 200_000
&END
&MONOMER
  MONOMER_TYPE = ATOM
&END
&MONOMER
  MONOMER_TYPE = ATOM
&END
""")
    assert rc == 0, H.lib().hh_last_error()
    assert c.calculation_type == 3
    assert c.ntemperatures == 3 and list(c.temperatures[:3]) == [100.0, 200.0, 300.0]
    assert c.hep_m0_niterations == 10 and c.hep_m0_npoints == 200000


def test_parse_correlation_array(tmp_path):
    """CORRELATION_ARRAY's lists: satellite temperatures, one CF file name each and, optionally,
    the partition-function ratios (src/driver.c:1170-1247)."""
    rc, c = parse_text(tmp_path, """
&INPUT
  CALCULATION_TYPE = CORRELATION_ARRAY
  PAIR_STATE = BOUND
  TEMPERATURE = 100.0
  satellite_temperatures = {50.0, 60.0, 70.0}
  partial_partition_function_ratios = {1.5, 1.2, 1.0}
  cf_filenames = { "CF-50.txt",
                   "CF-60.txt", "CF-70.txt" }
&END
""")
    assert rc == 0, H.lib().hh_last_error()
    assert c.calculation_type == 4 and c.pair_state == 2
    assert c.nsatellite == 3 and list(c.satellite_temperatures[:3]) == [50.0, 60.0, 70.0]
    assert c.nppf_ratios == 3 and list(c.ppf_ratios[:3]) == [1.5, 1.2, 1.0]
    assert c.ncf_filenames == 3 and [c.cf_filenames[k].value for k in range(3)] == [b"CF-50.txt", b"CF-60.txt",
                                                                                     b"CF-70.txt"]
