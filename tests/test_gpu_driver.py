"""The .conf driver executable (hawaiihybrid_amd/hawaii_gpu) on the GPU: a CORRELATION_SINGLE
run reproduces the oracle's CF for the same seeded stream (the preliminary-M0 points consumed
first, as calculate_correlation_and_save does), written in the reference's file format, and a
PROCESSING run of the reference's CF -> SF known-answer test passes its CMP."""
import os
import subprocess

import numpy as np
import pytest

import hawaiihybrid_amd as H
import oracle_lib as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "hawaiihybrid_amd", "hawaii_gpu")


def run_driver(args, cwd):
    print("driver:", args, flush=True)
    return subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def read_columns(path):
    head, rows = [], []
    for line in open(path):
        if line.startswith("#"):
            head.append(line.strip())
        elif line.strip():
            rows.append([float(v) for v in line.split()])
    return head, np.array(rows)


def test_driver_correlation_single_matches_oracle(tmp_path):
    conf = os.path.join(HERE, "conf", "cf-hear-small.conf")
    r = run_driver([conf, "--seed", "42"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ITERATION 2/2: accumulated 512 trajectories" in r.stdout
    head, cf = read_columns(tmp_path / "CF-hear-small.txt")
    assert "# TEMPERATURE: 300.00" in head and "# AVERAGE OVER 512.00 TRAJECTORIES" in head
    assert "# MAXIMUM TRAJECTORY LENGTH: 256" in head
    assert cf.shape == (256, 2) and np.allclose(cf[:, 0], 200.0 * np.arange(256))
    # oracle: the first 100 accepted points went to the preliminary M0, then 2 x 256
    ic, _, _ = O.sample_ics(O.hear(), O.Sampler(O.IC_REJECTION, 1, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0), 42,
                            100 + 512)
    p = O.CFParams(1e-12, 1e-12, 200.0, 256, 40.0, O.HALVES_BOTH, 50000)
    ref = O.cf_rows(O.hear(), p, ic[100:], 8)
    cf_ref = ref["crln"].sum(0) / 512
    err = np.linalg.norm(cf[:, 1] - cf_ref) / np.linalg.norm(cf_ref)
    assert err < 1e-8, err
    _, sf = read_columns(tmp_path / "SF-hear-small.txt")
    assert sf.shape == (256, 2)


def test_driver_preliminary_m2_and_stream_order(tmp_path):
    """initialM2_npoints: the preliminary M2 (calculate_M2) averages |d mu/dt|^2 over the accepted
    points after the preliminary M0's, and the trajectories continue the stream after both
    (src/hawaii.c:2947-2965)."""
    conf = open(os.path.join(HERE, "conf", "cf-hear-small.conf")).read()
    conf = conf.replace("initialM2_npoints = 0", "initialM2_npoints = 200")
    conf = conf.replace("NITERATIONS         = 2", "NITERATIONS         = 1")
    conf = conf.replace("TOTAL_TRAJECTORIES  = 512", "TOTAL_TRAJECTORIES  = 128")
    conf = conf.replace("MaxTrajectoryLength = 256", "MaxTrajectoryLength = 64")
    conf = conf.replace('sf_filename = "SF-hear-small.txt"', "")
    (tmp_path / "m2.conf").write_text(conf)
    r = run_driver(["m2.conf", "--seed", "42"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    import re
    m = re.search(r"^M2 = ([0-9.e+-]+) \+/- ([0-9.e+-]+)", r.stdout, re.M)
    assert m, r.stdout[-2000:]
    ic, _, _ = O.sample_ics(O.hear(), O.Sampler(O.IC_REJECTION, 1, 300.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0), 42,
                            100 + 200 + 128)
    f = O.dipole_rate_sq(O.hear(), ic[100:300])
    assert float(m.group(1)) == pytest.approx(f.mean() * H.SECOND_COEFF, rel=1e-8)
    assert float(m.group(2)) == pytest.approx(f.std(ddof=1) / np.sqrt(200) * H.SECOND_COEFF, rel=1e-6)
    assert "M2 ESTIMATE FROM CF (9-point)" in r.stdout and "M2 ESTIMATE FROM CF (21-point)" in r.stdout
    _, cf = read_columns(tmp_path / "CF-hear-small.txt")
    p = O.CFParams(1e-12, 1e-12, 200.0, 64, 40.0, O.HALVES_BOTH, 50000)
    cf_ref = O.cf_rows(O.hear(), p, ic[300:], 8)["crln"].sum(0) / 128
    assert np.linalg.norm(cf[:, 1] - cf_ref) / np.linalg.norm(cf_ref) < 1e-8


def test_driver_correlation_array_matches_oracle(tmp_path):
    """CALCULATION_TYPE = CORRELATION_ARRAY: one file per satellite temperature, each with its
    weight sum as the trajectory count, against the oracle's rows of the same ensemble weighted
    with the oracle's restatement of src/hawaii.c:2639-2674."""
    conf = os.path.join(HERE, "conf", "cf-array-hear-bound-small.conf")
    r = run_driver([conf, "--seed", "42"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ITERATION 2/2: accumulated 256 trajectories" in r.stdout
    temps, ratios = [60.0, 100.0, 150.0], [1.25, 1.0, 0.75]
    ic, _, _ = O.sample_ics(O.hear(), O.Sampler(O.IC_REJECTION, 2, 100.0, 4.0, 40.0, O.PESMIN_HEAR, 0.0), 42, 256)
    ref = O.cf_rows(O.hear(), O.CFParams(1e-12, 1e-12, 200.0, 256, 40.0, O.HALVES_BOTH, 50000, 0), ic, 8)
    w, ns = O.cf_array_weights(O.hamiltonian(O.hear(), ic), 100.0, 2, O.PESMIN_HEAR, temps, ratios,
                               counted=O.array_counted(ref["status"]))
    want = O.ensemble_array_from_rows(ref["crln"], w, ns)
    for k, T in enumerate(temps):
        head, cf = read_columns(tmp_path / f"CF-B-hear-{T:.1f}.txt")
        assert f"# TEMPERATURE: {T:.2f}" in head
        assert f"# AVERAGE OVER {ns[k]:.2f} TRAJECTORIES" in head
        assert cf.shape == (256, 2)
        err = np.linalg.norm(cf[:, 1] - want[k]) / np.linalg.norm(want[k])
        print(f"  T={T}: nstar {ns[k]:.3f} rel L2 {err:.2e}", flush=True)
        assert err < 1e-8, (T, err)


def write_ref_format(path, x, y, T, ntraj, sf):
    with open(path, "w") as f:
        f.write(f"# HAWAII HYBRID v0.1\n# TEMPERATURE: {T:.2f}\n# AVERAGE OVER {ntraj:.2f} TRAJECTORIES\n")
        if sf:
            f.write(f"# SPECTRAL FUNCTION LENGTH: {len(x)}\n")
            f.writelines(f"{a:.10f}   {b:.10e}\n" for a, b in zip(x, y))
        else:
            f.write(f"# MAXIMUM TRAJECTORY LENGTH: {len(x)}\n")
            f.writelines(f"{a:.2f} {b:.10e}\n" for a, b in zip(x, y))


def test_driver_cf_to_sf_known_answer(tmp_path, golden):
    """The CF_TO_SF part of the reference's tests/convert-cf-to-sf.conf through the driver on
    the committed fixture.  The written SF matches the reference's SF file to 1e-9 of the peak
    everywhere and to 1e-4 relative wherever |SF| > 1e-8 of the peak.  (The reference's own CMP
    compares every value to 1e-4 relative; its SF tail falls to 1e-98, where two FFTs of
    different rounding agree to a few digits only, so CMP is exercised on its own below.)"""
    g = golden("cf_sf_hear_1000.npz")
    write_ref_format(tmp_path / "cf.txt", g["cf_t"], g["cf"], float(g["T"][()]), float(g["ntraj"][()]), False)
    (tmp_path / "kat.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_CF("cf.txt")
  CF_TO_SF()
  WRITE_SF("sf_out.txt")
&END
""")
    r = run_driver(["kat.conf"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    head, sf = read_columns(tmp_path / "sf_out.txt")
    assert "# SPECTRAL FUNCTION LENGTH: 65536" in head
    peak = np.abs(g["sf"]).max()
    assert np.max(np.abs(sf[:, 0] - g["sf_nu"])) < 1e-6
    assert np.max(np.abs(sf[:, 1] - g["sf"])) < 1e-9 * peak
    big = np.abs(g["sf"]) > 1e-8 * peak
    assert np.max(np.abs(sf[big, 1] / g["sf"][big] - 1.0)) < 1e-4


@pytest.mark.parametrize("perturb", [False, True])
def test_driver_cmp(tmp_path, golden, perturb):
    """CMP as the reference's execute_cmp: equal elements pass, a value moved by 1e-3 fails."""
    g = golden("cf_sf_hear_1000.npz")
    sf = g["sf"].copy()
    if perturb:
        sf[int(np.argmax(np.abs(sf)))] *= 1.001
    T, nt = float(g["T"][()]), float(g["ntraj"][()])
    write_ref_format(tmp_path / "a.txt", g["sf_nu"], g["sf"], T, nt, True)
    write_ref_format(tmp_path / "b.txt", g["sf_nu"], sf, T, nt, True)
    (tmp_path / "cmp.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_SF("a.txt")
  READ_SF("b.txt")
  CMP()
&END
""")
    r = run_driver(["cmp.conf"], str(tmp_path))
    if perturb:
        assert r.returncode != 0 and "CMP failed" in r.stderr
    else:
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "CMP: equal" in r.stdout


def test_driver_phase_space_m0_reproduces_reference_log(tmp_path):
    """CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M0 for CO2-Ar at 100 K (FREE, R in [4, 40]) against the
    reference's VEGAS M0 6.25651e-4 (+-0.070%, example_configs/CF-F-CO2-Ar.conf.out:186-189),
    within 4 combined standard errors."""
    (tmp_path / "m0.conf").write_text("""&INPUT
  CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M0
  PAIR_STATE = FREE_AND_METASTABLE
  TEMPERATURE = 100.0
  pair_reduced_mass = 3.817083069753919e+04
  sampler_Rmin = 4.0
  sampler_Rmax = 40.0
&END
&MONOMER
  MONOMER_TYPE = LINEAR_MOLECULE
  II = {2.819827114123505e+05, 2.819827114123505e+05, 0.0}
&END
&MONOMER
  MONOMER_TYPE = ATOM
&END
""")
    r = run_driver(["m0.conf", "--seed", "9"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    import re
    m = re.search(r"T = 1\.00e\+02 => M0: ([0-9.e+-]+) \+- ([0-9.e+-]+)", r.stdout)
    assert m, r.stdout[-2000:]
    val, err = float(m.group(1)), float(m.group(2))
    ref = 6.25651e-4
    assert abs(val - ref) < 4 * np.hypot(err, ref * 7.04919e-4), (val, err)


def test_driver_phase_space_m2_reproduces_reference_log(tmp_path):
    """CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M2 for CO2-Ar at 100 K (FREE, R in [4, 40]) against
    the reference's VEGAS M2 0.441296 (+-0.076%, example_configs/CF-F-CO2-Ar.conf.out:235-238),
    within 4 combined standard errors."""
    (tmp_path / "m2.conf").write_text("""&INPUT
  CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M2
  PAIR_STATE = FREE_AND_METASTABLE
  TEMPERATURE = 100.0
  pair_reduced_mass = 3.817083069753919e+04
  sampler_Rmin = 4.0
  sampler_Rmax = 40.0
&END
&MONOMER
  MONOMER_TYPE = LINEAR_MOLECULE
  II = {2.819827114123505e+05, 2.819827114123505e+05, 0.0}
&END
&MONOMER
  MONOMER_TYPE = ATOM
&END
""")
    r = run_driver(["m2.conf", "--seed", "9"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    import re
    m = re.search(r"T = 1\.00e\+02 => M2: ([0-9.e+-]+) \+- ([0-9.e+-]+)", r.stdout)
    assert m, r.stdout[-2000:]
    val, err = float(m.group(1)), float(m.group(2))
    ref = 0.441296
    assert abs(val - ref) < 4 * np.hypot(err, ref * 7.61738e-4), (val, err)


def test_driver_correlation_single_integrates_ppf_when_not_given(tmp_path):
    """Without PARTIAL_PARTITION_FUNCTION_RATIO the driver integrates it by VEGAS as the reference
    does; for He-Ar at 300 K the printed ratio is used for the CF and equals the PF integral over
    the analytic partition function (checked through the CF's M0 against the VEGAS M0)."""
    conf = open(os.path.join(HERE, "conf", "cf-hear-small.conf")).read()
    conf = "\n".join(l for l in conf.splitlines() if "PARTIAL_PARTITION_FUNCTION_RATIO" not in l.upper())
    (tmp_path / "noppf.conf").write_text(conf)
    r = run_driver(["noppf.conf", "--seed", "42"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    import re
    m = re.search(r"=> PPF ratio: ([0-9.e+-]+)", r.stdout)
    assert m and float(m.group(1)) > 0, r.stdout[-2000:]
    assert "=> M0:" in r.stdout and "=> M2:" in r.stdout


def test_driver_prmu_matches_oracle(tmp_path):
    """CALCULATION_TYPE = PR_MU (He-Ar, R0 = 12): the preliminary M0 and M2 over the stream's
    first 2 x 100 rejection-sampled points, then R0-injected trajectories from where the stream
    left off, those still inside R0 after L samples replaced by the next ones; the SF file
    against the oracle's series and spectra of the same trajectories."""
    L, dt, R0, n_iter = 1024, 200.0, 12.0, 32
    (tmp_path / "prmu.conf").write_text(f"""&INPUT
  CALCULATION_TYPE = PR_MU
  PAIR_STATE = FREE_AND_METASTABLE
  TEMPERATURE = 300.0
  pair_reduced_mass = 6.632039160324215e+03
  NITERATIONS = 2
  TOTAL_TRAJECTORIES = {2 * n_iter}
  CVODE_TOLERANCE = 1e-12
  sampling_time = {dt}
  MaxTrajectoryLength = {L}
  pesmin = -9.424412560638486e-05
  partial_partition_function_ratio = 1.0
  sampler_Rmin = 4.0
  sampler_Rmax = 40.0
  initialM0_npoints = 100
  initialM2_npoints = 100
  R0 = {R0}
  sf_filename = "SF-prmu.txt"
  ApproximateFrequencyMax = 300.0
&END
&MONOMER
  MONOMER_TYPE = ATOM
&END
&MONOMER
  MONOMER_TYPE = ATOM
&END
""")
    r = run_driver(["prmu.conf", "--seed", "11"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ITERATION 2/2: accumulated 64 trajectories" in r.stdout
    assert "M0 ESTIMATE FROM SF" in r.stdout and "M2 ESTIMATE FROM SF" in r.stdout
    sh, so = H.he_ar(), O.hear()
    rej = H.Sampler(H.IC_REJECTION, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_HEAR, 0.0)
    pts, _, draws = H.sample_ics_from(sh, rej, 11, 0, 200)
    import re
    m = re.search(r"^M0 = ([0-9.e+-]+)", r.stdout, re.M)
    d = np.array([np.dot(v, v) for v in (np.array([O.lib().orc_hear_dip(R)]) for R in pts[:100, 4])])
    assert float(m.group(1)) == pytest.approx(d.mean() * H.ZERO_COEFF, rel=1e-9)
    inj = H.Sampler(H.IC_R0_INJECT, H.PAIR_FREE, 300.0, 4.0, 40.0, H.PESMIN_HEAR, R0)
    ic, _, _ = H.sample_ics_from(sh, inj, 11, draws, 200)
    coeff = 4.3597447222071e-18 * 5.29177210903e-11 ** 6 * 2.4188843265857e-17 * R0 * R0 * dt * dt
    head, sf = read_columns(tmp_path / "SF-prmu.txt")
    nfreq = sf.shape[0]
    assert f"# SPECTRAL FUNCTION LENGTH: {nfreq}" in head and "# AVERAGE OVER 64.00 TRAJECTORIES" in head
    want, used = np.zeros(nfreq), 0
    for y0 in ic:
        status, n, vec = O.prmu_series(so, y0, dt, L + 2, R0)
        if status != 0 or n > L + 1:
            continue
        x = np.zeros((3, L))
        x[:, :n - 1] = vec[1:n].T
        want += coeff * (-y0[5] / O.MU_HEAR) * O.prmu_power(x, dt, nfreq)
        used += 1
        if used == 2 * n_iter:
            break
    assert used == 2 * n_iter
    want /= used
    step = 1.0 / (dt * 2.4188843265857e-17) / 2.99792458e10 / L
    assert np.allclose(sf[:, 0], step * np.arange(nfreq), rtol=1e-9, atol=1e-9)
    assert np.linalg.norm(sf[:, 1] - want) / np.linalg.norm(want) < 1e-8


def test_driver_desymmetrizations_and_alpha_known_answers(tmp_path, golden):
    """The reference's tests/desymmetrizations.conf and tests/compute-alpha.conf through the
    driver: ALPHA of the CF_TO_SF of CF-F-He-Ar-1000.0 without and with D1, D2, D3, and D4, D4a
    from the CF, written with WRITE_SPECTRUM and compared with the reference's expected spectra
    (every 4th point): 1e-9 of the peak everywhere, 1e-4 relative above 1e-6 of the peak (the
    SF's far tail, where two FFTs of different rounding agree to a few digits only, is raised by
    D2 and D3's factors growing with frequency).  SPECTRUM_FREQUENCY_MAX truncates the written
    spectrum."""
    g = golden("cf_sf_hear_1000.npz")
    s = golden("spectra_hear_1000.npz")
    write_ref_format(tmp_path / "cf.txt", g["cf_t"], g["cf"], float(g["T"][0]), float(g["ntraj"][0]), False)
    (tmp_path / "d.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_CF("cf.txt")
  DUP()
    CF_TO_SF()
    DUP()
      ALPHA()
      WRITE_SPECTRUM("SPCL.txt")
    DUP()
      D1()
      ALPHA()
      WRITE_SPECTRUM("SPD1.txt")
    DUP()
      D2()
      ALPHA()
      WRITE_SPECTRUM("SPD2.txt")
    D3()
    ALPHA()
    WRITE_SPECTRUM("SPD3.txt")
  DUP()
    D4()
    ALPHA()
    WRITE_SPECTRUM("SPD4.txt")
  D4a()
  ALPHA()
  DUP()
    WRITE_SPECTRUM("SPD4a.txt")
  READ_SPECTRUM("SPD4a.txt")
  CMP()
&END
""")
    r = run_driver(["d.conf"], str(tmp_path))
    assert r.returncode == 0 and "CMP: equal" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    for name in ("SPCL", "SPD1", "SPD2", "SPD3", "SPD4", "SPD4a"):
        head, sp = read_columns(tmp_path / f"{name}.txt")
        assert "# TEMPERATURE: 1000.00" in head and sp.shape == (65536, 2)
        got, want = sp[::4, 1], s[name]
        assert np.max(np.abs(sp[::4, 0] - s["nu"])) < 1e-6
        peak = np.abs(want).max()
        assert np.max(np.abs(got - want)) < 1e-9 * peak, name
        big = np.abs(want) > 1e-6 * peak
        assert np.max(np.abs(got[big] / want[big] - 1.0)) < 1e-4, name
    (tmp_path / "d4a.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  SPECTRUM_FREQUENCY_MAX = 1000.0
  READ_CF("cf.txt")
  D4a()
  ALPHA()
  WRITE_SPECTRUM("SPD4a-cut.txt")
&END
""")
    r = run_driver(["d4a.conf"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    _, sp = read_columns(tmp_path / "SPD4a-cut.txt")
    n = int(1000.0 / (s["nu"][1] / 4))
    assert sp.shape == (n, 2)
    want = s["SPD4a"][: (n + 3) // 4]
    got = sp[::4, 1]
    peak = np.abs(s["SPD4a"]).max()
    assert np.max(np.abs(got - want)) < 1e-9 * peak


CF4 = [(0.0, 1.0762846557e-33), (200.0, 1.0734463146e-33), (400.0, 1.0650019293e-33), (600.0, 1.0511598707e-33)]


def test_driver_average_cfs_known_answer(tmp_path):
    """The reference's tests/average-cfs.conf: two copies of tests/cf.txt (T 100 K, 3.4e6
    trajectories; its four rows are CF4) average to tests/cf-average.txt (the same rows over
    6.8e6 trajectories); a CF without a trajectory count is an error."""
    x, y = np.array([r[0] for r in CF4]), np.array([r[1] for r in CF4])
    write_ref_format(tmp_path / "cf.txt", x, y, 100.0, 3.4e6, False)
    write_ref_format(tmp_path / "cf-average.txt", x, y, 100.0, 6.8e6, False)
    (tmp_path / "nontraj.txt").write_text("# HAWAII HYBRID v0.1\n# TEMPERATURE: 100.00\n" +
                                          "".join(f"{a:.2f} {b:.10e}\n" for a, b in CF4))
    (tmp_path / "avg.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_CF("cf.txt")
  READ_CF("cf.txt")
  AVERAGE_CFS()
  DUP()
    WRITE_CF("out.txt")
  READ_CF("cf-average.txt")
  CMP()
  READ_CF("cf.txt")
  READ_CF("nontraj.txt")
  AVERAGE_CFS()
&END
""")
    r = run_driver(["avg.conf"], str(tmp_path))
    assert r.returncode != 0 and "CMP: equal" in r.stdout and "is invalid" in r.stderr, r.stdout + r.stderr
    head, cf = read_columns(tmp_path / "out.txt")
    assert "# AVERAGE OVER 6800000.00 TRAJECTORIES" in head and np.allclose(cf[:, 1], y, rtol=1e-12)


def test_driver_add_spectra_known_answer(tmp_path, golden):
    """The reference's tests/add-spectra.conf: the bound and free He-Ar spectra at 100 K summed by
    ADD_SPECTRA(n=2) and ADD_SPECTRA(wavenumber_step=0.05) (natural cubic spline resampling onto
    the common grid) against its expected SPCL-T files."""
    g = golden("add_spectra_hear_100.npz")
    for key, nt in (("B", 2808000.0), ("F", 500000000.0)):
        with open(tmp_path / f"{key}.txt", "w") as f:
            f.write(f"# HAWAII HYBRID v0.1\n# TEMPERATURE: 100.00\n# AVERAGE OVER {nt:.2f} TRAJECTORIES\n")
            f.writelines(f"{a:.3f} {b:.10e}\n" for a, b in zip(g[key + "_nu"], g[key]))
    (tmp_path / "add.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_SPECTRUM("B.txt")
  READ_SPECTRUM("F.txt")
  DUP2()
    ADD_SPECTRA(n=2)
    WRITE_SPECTRUM("T.txt")
  ADD_SPECTRA(wavenumber_step=0.05)
  WRITE_SPECTRUM("T005.txt")
&END
""")
    r = run_driver(["add.conf"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for out, key in (("T.txt", "T"), ("T005.txt", "T005")):
        head, sp = read_columns(tmp_path / out)
        assert "# AVERAGE OVER 502808000.00 TRAJECTORIES" in head
        assert sp.shape[0] == g[key].size
        assert np.max(np.abs(sp[:, 0] - g[key + "_nu"])) < 1e-9
        peak = np.abs(g[key]).max()
        assert np.max(np.abs(sp[:, 1] - g[key])) < 1e-9 * peak, out


def test_driver_spectral_moments(tmp_path, golden):
    """The reference's tests/compute-classical-moments.conf and compute-quantum-moments.conf
    sequence: M0 from the CF and from its SF agree (CF(0) ZeroCoeff / ALU^3 against the SF's
    Simpson integral), odd classical orders are 0, and the quantum M0 of the SF (int SF (1 +
    exp(-h nu / kT))) is printed; n as a named or positional argument."""
    g = golden("cf_sf_hear_1000.npz")
    write_ref_format(tmp_path / "cf.txt", g["cf_t"], g["cf"], float(g["T"][0]), float(g["ntraj"][0]), False)
    (tmp_path / "mn.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_CF("cf.txt")
  DUP()
    COMPUTE_Mn_CLASSICAL_DETAILED_BALANCE(n=0)
  CF_TO_SF()
  DUP()
    COMPUTE_Mn_CLASSICAL_DETAILED_BALANCE(0)
  DUP()
    COMPUTE_Mn_CLASSICAL_DETAILED_BALANCE(1)
  COMPUTE_Mn_QUANTUM_DETAILED_BALANCE(n=0)
&END
""")
    r = run_driver(["mn.conf"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    import re
    vals = re.findall(r"Spectral moment using (\w+) detailed balance \[based on (\w+)\] = ([0-9.e+-]+)", r.stdout)
    assert [v[:2] for v in vals] == [("classical", "CF"), ("classical", "SF"), ("classical", "SF"), ("quantum", "SF")]
    m0_cf, m0_sf, m1_sf, q0 = (float(v[2]) for v in vals)
    assert m0_cf == pytest.approx(float(g["cf"][0]) * H.ZERO_COEFF / 5.29177210903e-11 ** 3, rel=1e-5)
    assert m0_sf == pytest.approx(m0_cf, rel=1e-3) and m1_sf == 0.0 and q0 > 0


REF_HEAR_PLUGIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_hear.so")


def _m0_conf(monomers, plugin):
    """A short CALCULATE_PHASE_SPACE_M0 configuration with the given &MONOMER blocks and plugins."""
    mu = "6.632039160324215e+03" if "LINEAR" not in monomers else "3.817083069753919e+04"
    lines = ["&INPUT", "  CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M0", "  PAIR_STATE = FREE_AND_METASTABLE",
             "  TEMPERATURE = 300.0", f"  pair_reduced_mass = {mu}", "  sampler_Rmin = 4.0", "  sampler_Rmax = 40.0",
             "  HEP_M0_NITERATIONS = 3", "  HEP_M0_NPOINTS = 100000"]
    if plugin:
        lines += [f'  so_potential = "{plugin}"', f'  so_dipole = "{plugin}"']
    return "\n".join(lines + ["&END", monomers])


ATOM_ATOM = "&MONOMER\n  MONOMER_TYPE = ATOM\n&END\n&MONOMER\n  MONOMER_TYPE = ATOM\n&END\n"
LINEAR_ATOM = ("&MONOMER\n  MONOMER_TYPE = LINEAR_MOLECULE\n  II = {2.819827114123505e+05, 2.819827114123505e+05, 0.0}\n"
               "&END\n&MONOMER\n  MONOMER_TYPE = ATOM\n&END\n")


@pytest.mark.skipif(not os.path.exists(REF_HEAR_PLUGIN), reason="the reference-built He-Ar plugin is not built")
def test_driver_accepts_the_reference_hear_plugin(tmp_path):
    """The reference's own He-Ar plugin (PES-IDS/HeAr.h built into oracle/_ref) is the engine's
    built-in He-Ar surface: its pes_lab, dpes_lab and dipole_lab agree with the device surfaces at
    the check points, so the run goes ahead."""
    (tmp_path / "ok.conf").write_text(_m0_conf(ATOM_ATOM, REF_HEAR_PLUGIN))
    r = run_driver(["ok.conf", "--seed", "3"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "M0:" in r.stdout


@pytest.mark.skipif(not os.path.exists(REF_HEAR_PLUGIN), reason="the reference-built He-Ar plugin is not built")
def test_driver_refuses_a_plugin_that_is_not_the_builtin_system(tmp_path):
    """A LINEAR_MOLECULE + ATOM configuration whose plugin is a He-Ar surface (the situation of the
    reference's example_configs/CF-F-CO-Ar.conf: a linear molecule with another surface) is refused
    instead of being integrated on the CO2-Ar surfaces."""
    (tmp_path / "bad.conf").write_text(_m0_conf(LINEAR_ATOM, REF_HEAR_PLUGIN))
    r = run_driver(["bad.conf", "--seed", "3"], str(tmp_path))
    assert r.returncode != 0
    assert "not the engine's built-in CO2-Ar surface" in (r.stdout + r.stderr)


def test_driver_refuses_an_unloadable_plugin(tmp_path):
    """A plugin that cannot be loaded ends the run with an error (the reference's loader exits,
    src/driver.c:3083-3094), not a note and a fall-back to the built-in surface."""
    (tmp_path / "missing.conf").write_text(_m0_conf(ATOM_ATOM, str(tmp_path / "no-such-plugin.so")))
    r = run_driver(["missing.conf", "--seed", "3"], str(tmp_path))
    assert r.returncode != 0
    assert "cannot load the surface plugin" in (r.stdout + r.stderr)


def test_driver_fit_baseline_and_smooth(tmp_path, golden):
    """The reference's tests/convert-cf-to-sf.conf (FIT_BASELINE(), FIT_BASELINE(8000) and the
    named form, each followed by CF_TO_SF on the device) and tests/smooth.conf (SMOOTH, ALPHA)
    through the driver on the committed fixtures.  SFs: 1e-9 of the peak, 1e-4 relative above
    1e-8 of the peak; the baseline's own contribution (with minus without, 1e-5 of the SF) to
    2e-3 where it exceeds 1e-2 of its largest value (what the 11 printed digits of the two
    reference files resolve).  Smoothed spectrum: 1e-4 relative below 680 cm^-1, 1e-3 of the
    peak beyond, where the reference's ill-conditioned cubic fits carry rounding noise."""
    g, b = golden("cf_sf_hear_1000.npz"), golden("baseline_smooth_hear_1000.npz")
    T, nt = float(g["T"][()]), float(g["ntraj"][()])
    write_ref_format(tmp_path / "cf.txt", g["cf_t"], g["cf"], T, nt, False)
    write_ref_format(tmp_path / "sf.txt", g["sf_nu"], g["sf"], T, nt, True)
    (tmp_path / "p.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_CF("cf.txt")
  DUP()
    CF_TO_SF()
    WRITE_SF("plain.txt")
  DUP()
    FIT_BASELINE()
    CF_TO_SF()
    WRITE_SF("baseline.txt")
  DUP()
    FIT_BASELINE(cf_extrapolation_begin_index=8000)
    CF_TO_SF()
    WRITE_SF("custom.txt")
  FIT_BASELINE(8000)
  CF_TO_SF()
  WRITE_SF("custom2.txt")
  READ_SF("sf.txt")
  SMOOTH(grid_npoints=1200, grid_max=1200.0, window_size_delay=10, window_size_step=0.05)
  ALPHA()
  WRITE_SPECTRUM("sm.txt")
&END
""")
    r = run_driver(["p.conf"], str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    _, plain = read_columns(tmp_path / "plain.txt")
    for name, key in (("baseline", "sf_baseline"), ("custom", "sf_custom_baseline"), ("custom2", "sf_custom_baseline")):
        _, sf = read_columns(tmp_path / f"{name}.txt")
        want = b[key]
        peak = np.abs(want).max()
        assert np.max(np.abs(sf[:, 1] - want)) < 1e-9 * peak, name
        big = np.abs(want) > 1e-8 * peak
        assert np.max(np.abs(sf[big, 1] / want[big] - 1.0)) < 1e-4, name
        ours, theirs = sf[:, 1] - plain[:, 1], want - g["sf"]
        sel = np.abs(theirs) > 1e-2 * np.abs(theirs).max()
        err = np.max(np.abs(ours[sel] / theirs[sel] - 1.0))
        print(f"  {name}: baseline contribution rel err {err:.2e} over {sel.sum()} points", flush=True)
        assert err < 2e-3, name
    head, sm = read_columns(tmp_path / "sm.txt")
    assert sm.shape == (1200, 2)
    assert np.max(np.abs(sm[:, 0] - b["sm_nu"])) < 1e-6
    want = b["sm_alpha"]
    rel = np.abs(sm[1:, 1] / want[1:] - 1.0)
    assert rel[:679].max() < 1e-4
    assert np.max(np.abs(sm[:, 1] - want)) < 1e-3 * np.abs(want).max()


def test_driver_int3_stops_with_the_stack(tmp_path, golden):
    """INT3 (src/driver.c:2423-2446) prints the stack and ends the run with a failure."""
    g = golden("cf_sf_hear_1000.npz")
    write_ref_format(tmp_path / "cf.txt", g["cf_t"][:1024], g["cf"][:1024], 1000.0, 1.0, False)
    (tmp_path / "i.conf").write_text("""&INPUT
  CALCULATION_TYPE = PROCESSING
&END
&PROCESSING
  READ_CF("cf.txt")
  DUP()
  INT3()
  WRITE_CF("never.txt")
&END
""")
    r = run_driver(["i.conf"], str(tmp_path))
    assert r.returncode != 0
    assert "BREAKPOINT INTERRUPT ISSUED" in r.stdout and "1: CF of 1024 points" in r.stdout, r.stdout[-2000:]
    assert not (tmp_path / "never.txt").exists()
