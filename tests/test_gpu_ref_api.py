"""A program written like the reference's examples/correlation_co2_ar.cpp (tests/ref_api/), linked
against libhawaii.so, on the GPU:
* its CF (calculate_correlation_and_save: 1000 accepted points for the preliminary M0 first, then
  3 iterations over 600 trajectories from the same stream, FP64 at the reference's 1e-12) is the
  oracle's CF of those trajectories times the partition-function ratio, within 1e-6 relative L2
  (the bound of the FP64 engine tests on CO2-Ar), in the reference's file format;
* the returned CFnc carries the reference's fields (len, ntraj, Temperature, normalized);
* compute_Mn_from_cf_... gives hh_cf_moment's M0 and M2 of the written CF (M0 1e-9, M2 1e-6: the
  file keeps 11 significant digits and the M2 stencil cancels);
* idct_cf_to_sf + write_spectral_function equal the engine's CF -> SF of the same CF (1e-9);
* c_mpi_perform_integration(INTEGRAND_PF) is hh_vegas with the stream's seed (bit for bit)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import hawaiihybrid_amd as H
import oracle_lib as O
from test_ref_api import build_program

pytestmark = pytest.mark.gpu

ALU3 = 5.29177210903e-11 ** 3


def read_columns(path):
    head, rows = [], []
    for line in open(path):
        if line.startswith("#"):
            head.append(line.strip())
        elif line.strip():
            rows.append([float(v) for v in line.split()])
    return head, np.array(rows)


def test_examples_style_program_matches_oracle(tmp_path):
    exe = build_program(tmp_path)
    s = H.co2_ar()
    seed, L, total, niter, m0n = 42, 512, 600, 3, 1000
    cf_file, sf_file = tmp_path / "CF-F-300.0.txt", tmp_path / "SF-F-300.0.txt"
    print("ref api program: CO2-Ar, 600 trajectories, L=512", flush=True)
    r = subprocess.run([exe, repr(s.mu), repr(s.I), str(seed), str(L), str(total), str(niter), str(m0n),
                        str(cf_file), str(sf_file)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ITERATION 3/3: accumulated 600 trajectories" in r.stdout
    assert "CFNC len 512 ntraj 600.0 normalized 1 T 300.00" in r.stdout
    head, cf = read_columns(cf_file)
    assert "# TEMPERATURE: 300.00" in head and "# AVERAGE OVER 600.00 TRAJECTORIES" in head
    assert cf.shape == (L, 2) and np.allclose(cf[:, 0], 200.0 * np.arange(L))
    pesmin = -195.6337098547 / 2.1947463136320e5
    ic, _, _ = O.sample_ics(O.co2ar(), O.Sampler(O.IC_REJECTION, 1, 300.0, 4.0, 40.0, pesmin, 0.0), seed,
                            m0n + total)
    p = O.CFParams(1e-12, 1e-12, 200.0, L, 40.0, O.HALVES_BOTH, 50000)
    ref = O.cf_rows(O.co2ar(), p, ic[m0n:], 8)
    cf_ref = 2.68854e+05 * ref["crln"].sum(0) / total
    err = np.linalg.norm(cf[:, 1] - cf_ref) / np.linalg.norm(cf_ref)
    print(f"  CF rel L2 vs oracle {err:.3e}", flush=True)
    assert err < 1e-6
    # moments of the returned CF
    m = re.search(r"CF M0 ([0-9.e+-]+) M2 ([0-9.e+-]+)", r.stdout)
    m0 = H.cf_moment(cf[:, 1], 200.0, 0)
    m2 = H.cf_moment(cf[:, 1], 200.0, 2)
    assert float(m.group(1)) == pytest.approx(m0, rel=1e-9)
    assert float(m.group(2)) == pytest.approx(m2, rel=1e-6)
    # CF -> SF through the reference's name
    import torch
    eng = H.Engine(0)
    nu, sf, _ = eng.cf_to_sf(torch.from_numpy(np.ascontiguousarray(cf[:, 1])).to("cuda:0"), 200.0)
    _, sfw = read_columns(sf_file)
    sf = sf.cpu().numpy()
    assert np.max(np.abs(sfw[:, 1] - sf)) <= 1e-9 * np.max(np.abs(sf))
    assert np.allclose(sfw[:, 0], nu.cpu().numpy(), rtol=1e-9)
    # VEGAS through c_mpi_perform_integration: the engine's hh_vegas with the stream's seed
    m = re.search(r"PF ([0-9.e+-]+) \+- ([0-9.e+-]+) analytic ([0-9.e+-]+)", r.stdout)
    v = eng.vegas(s, H.INTEGRAND_PF, 300.0, 4.0, 40.0, pair_state=1, niterations=4, npoints=100000, seed=seed)
    assert float(m.group(1)) == pytest.approx(v.value, rel=1e-11)
    assert float(m.group(2)) == pytest.approx(v.error, rel=1e-11)
    assert float(m.group(3)) == pytest.approx(H.pf_analytic(s, 300.0), rel=1e-11)
    eng.close()
