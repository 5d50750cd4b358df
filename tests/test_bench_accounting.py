"""bench.py's own arithmetic, on the CPU: the algorithmic flop count behind
roofline.achieved, the traffic file's counter scaling, the lineshape rows the
spectrum errors are taken over, and the one-GPU refusal."""

import os
import sqlite3
import subprocess
import sys

import numpy as np
import pytest

import calcir_amd as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("n,nb", [(2048, 32), (8192, 32), (100, 32), (65, 32), (66, 32), (130, 8), (10, 4), (5, 2)])
def test_bulge_reflector_flops_closed_form(n, nb):
    # the reflectors of sweep s (s = 0 .. n-3) start at rows s+1, s+1+nb, ...
    # below n-1 and tile rows s+1 .. n-1, except that a block would start at
    # row n-1 (length 1) when nb divides n-2-s; 4 flops per element and
    # eigenvector column
    tot = n * (n - 1) // 2 - 1 - (n - 2) // nb
    assert bench.bulge_bt_flops(n, nb) == 4.0 * n * tot


def test_bulge_reflector_flops_at_the_bench_size():
    # the figure bench.py reports as roofline.flops_per_matrix (N = 2048):
    # 2 N^3 to within the O(N^2) terms
    f = bench.bulge_bt_flops(2048)
    assert f == 17170956288.0
    assert abs(f / (2.0 * 2048**3) - 1.0) < 2.0 / 2048


def _counter_db(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (dispatch_id integer, kernel_name text, "
                "counter_name text, value real)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", rows)
    con.commit()
    con.close()


def test_traffic_file_scales_each_counter_by_the_calibration_copy(tmp_path):
    q2 = "void cir::q2_apply_kernel<1024, true>(int, int, int, float const*, float*, int, int)"
    cp = "__amd_rocclr_copyBuffer"
    # FETCH: a timing dispatch of 8 matrices (7), the run's chunk (9, two
    # rows of one dispatch), the 1 GiB copy read as 2 GiB (scale 1/2)
    _counter_db(tmp_path / "f.db", [(3, cp, "FETCH_SIZE", 1048576.0), (3, cp, "FETCH_SIZE", 1048576.0),
                                    (7, q2, "FETCH_SIZE", 1000.0), (9, q2, "FETCH_SIZE", 3.0e6),
                                    (9, q2, "FETCH_SIZE", 1.0e6)])
    # WRITE: the copy read as exactly 1 GiB (scale 1)
    _counter_db(tmp_path / "w.db", [(3, cp, "WRITE_SIZE", 1048576.0), (7, q2, "WRITE_SIZE", 500.0),
                                    (9, q2, "WRITE_SIZE", 2.0e6)])
    t = bench.traffic_from(str(tmp_path / "f.db"), str(tmp_path / "w.db"), 600)
    # 600 frames: origins every 20 frames, 23 of them, 590 distinct frames
    assert t["matrices_per_dispatch"] == 590
    assert t["fetch_scale"] == 0.5 and t["write_scale"] == 1.0
    assert t["fetch_kib_raw"] == 4.0e6 and t["write_kib_raw"] == 2.0e6
    # (4e6 KiB / 2 + 2e6 KiB) over 590 matrices
    assert t["hbm_bytes_per_matrix"] == pytest.approx(4.0e6 * 1024 / 590, rel=1e-12)
    assert t["eigenvector_bytes_per_matrix"] == 2 * 4 * 2048 * 2048


def test_traffic_file_needs_the_calibration_copy(tmp_path):
    q2 = "cir::q2_apply_kernel<1024, true>"
    _counter_db(tmp_path / "f.db", [(9, q2, "FETCH_SIZE", 1.0e6)])
    _counter_db(tmp_path / "w.db", [(9, q2, "WRITE_SIZE", 1.0e6)])
    with pytest.raises(SystemExit, match="calibration copy"):
        bench.traffic_from(str(tmp_path / "f.db"), str(tmp_path / "w.db"), 600)


def test_bench_lineshape_rows_are_the_written_lineshape_file(tmp_path):
    # the spectrum errors are taken over the rows of <outf>_irls.dat: bench's
    # restatement gives the library writer's frequencies and values
    p = C.default_params(outf=str(tmp_path / "run"), ntcfpoints=150, nzeros=1898, avef=3400.0,
                         omegaStart=2800.0, omegaStop=4000.0, omegaStep=5, dt=0.010)
    nfft = p.ntcfpoints + p.nzeros
    nom = C.nomega(p)
    rng = np.random.default_rng(5)
    res = {k: rng.standard_normal(p.ntcfpoints) + 1j * rng.standard_normal(p.ntcfpoints)
           for k in ("tcf_ir", "tcf_vv", "tcf_vh")}
    res.update({k: rng.standard_normal(nfft) for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh")})
    res.update({k: rng.random(nom) for k in ("sw", "pw", "rw", "rmw")})
    C.write_outputs(p, res)
    for key, suffix in (("ftcf_ir", "irls"), ("ftcf_vv", "vvls"), ("ftcf_vh", "vhls")):
        f, v = bench.lineshape(p, res[key])
        ls = np.loadtxt(tmp_path / ("run_%s.dat" % suffix))
        assert ls.shape == (f.size, 2)
        # the file holds "%g" (6 significant digits) of the FP32 values
        np.testing.assert_allclose(ls[:, 0], f, rtol=1e-5, atol=0)
        np.testing.assert_allclose(ls[:, 1], v, rtol=1e-5, atol=1e-30)


def test_bench_lineshape_error_of_a_spectrum_against_itself_and_a_scaled_copy():
    p = C.default_params(ntcfpoints=150, nzeros=1898, avef=3400.0, omegaStart=2800.0, omegaStop=4000.0,
                         dt=0.010)
    nfft = p.ntcfpoints + p.nzeros
    x = np.exp(-0.5 * ((np.arange(nfft) - 40.0) / 9.0) ** 2) + 0.1
    res = {k: x for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh")}
    e = bench.spectrum_error(p, res, res)
    assert all(v["rel_l2"] == 0.0 and v["first_moment_shift_cm-1"] == 0.0 for v in e.values())
    # a uniform scale changes the norm by its factor and leaves the first
    # moment where it was
    e2 = bench.spectrum_error(p, {k: 1.001 * x for k in res}, res)
    for v in e2.values():
        assert v["rel_l2"] == pytest.approx(1e-3, rel=1e-9)
        assert abs(v["first_moment_shift_cm-1"]) < 1e-9


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_a_failing_extra_is_recorded_in_its_place(capsys):
    assert bench.guarded("ok", lambda: {"value": 1}) == {"value": 1}

    def boom():
        raise C.CalcirError(2, "out of memory")

    r = bench.guarded("8192-OH box", boom)
    assert set(r) == {"error"} and "out of memory" in r["error"]
    assert "8192-OH box failed" in capsys.readouterr().err


def test_cpu_baseline_reports_its_phases():
    # the CPU leg on a small box (32 waters, the bench's 8 tcf points): the
    # per-phase seconds of the oracle run sum to no more than its wall time
    xyz, box, times = C.synth_trajectory(32, 8, 77, box_nm=C.synth_box_nm(32))
    r = bench.cpu_baseline(xyz, box, times, 2)
    assert r["kind"] == "port" and r["cores"] == 2 and r["unit"] == "frames/s" and r["value"] > 0
    ph = r["phases_s"]
    assert set(ph) == {"field_hamiltonian", "eigh", "ipr_spectral_density", "propagate_correlate", "finalize"}
    assert all(v >= 0 for v in ph.values())
    wall = 8 / r["value"]
    assert sum(ph.values()) <= wall * 1.05 + 1e-3
    assert r["sample"].startswith("64-OH box")
