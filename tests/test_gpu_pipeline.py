"""End-to-end parity of the whole calcIR job on the GPU (cir_run through the
C ABI) against the CPU oracle's reference control flow."""

import subprocess

import numpy as np
import pytest

import calcir_amd as C
import oracle_lib as O
from cases import frames, rel, rel_l2
from conftest import announce

pytestmark = pytest.mark.gpu

# Small case with overlapping time origins (3 samples of 30 points, 10 frames
# apart), 128 waters, tip4p2005 model, H2O, Gruenbaum map.
KW = dict(species="H2O", model="tip4p2005", ntcfpoints=30, nsamples=3, sampleEvery=0.1, avef=3400.0,
          omegaStart=2800.0, omegaStop=4000.0, omegaStep=5, nzeros=2018, t1=0.26, dt=0.010)


def _run_pair(prec, mode, oracle_prec=O.F64, nwater=128, nframes=60, seed=1001, **over):
    kw = dict(KW)
    kw.update(over)
    p = C.default_params(precision=prec, mode=mode, **kw)
    xyz, box, times = frames(nwater, nframes, seed)
    announce("pipeline prec=%s mode=%s nwater=%d nframes=%d %s" % (prec, mode, nwater, nframes, over))
    g = C.run(p, xyz, box, times)
    o = O.run(oracle_prec, O.config_from_params(p), xyz, box, times)
    return p, g, o


def _check(g, o, tol_tcf, tol_spec, tol_dist, dlam=1e-3, n=256):
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(g[k], o[k]) < tol_tcf, (k, rel(g[k], o[k]))
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
        assert rel_l2(g[k], o[k]) < tol_spec, (k, rel_l2(g[k], o[k]))
    assert rel(g["sw"], o["sw"]) < tol_dist
    # Pw is a count: a mode whose eigenvalue lies within the precision's
    # eigenvalue error dlam (cm^-1) of a bin edge may change bin.  Expected
    # moves ~ mode-steps * 2 dlam / omegaStep; allow 4x that, plus 2.
    steps = g["nsteps"] * n  # mode-steps (n chromophores)
    unit = 1.0 / KW["omegaStep"] / steps
    moves = 0.5 * np.abs(g["pw"] - o["pw"]).sum() / unit
    allowed = 4 * steps * 2 * dlam / KW["omegaStep"] + 2
    assert moves <= allowed, (moves, allowed)
    assert abs(int(g["hist_skipped"]) - int(o["hist_skipped"])) <= allowed
    # Rw, Rmw (ratios) on bins whose counts agree; NaN where Pw = 0
    m = o["pw"] > 0
    assert np.all(np.isnan(g["rw"][~m]))
    same = m & (np.abs(g["pw"] - o["pw"]) < 0.5 * unit)
    assert rel(g["rw"][same], o["rw"][same]) < max(tol_dist, 1e-9) * 1e2
    assert rel(g["rmw"][same], o["rmw"][same]) < max(tol_dist, 1e-9) * 1e2


def test_pipeline_fp64_batched_matches_oracle_fp64():
    p, g, o = _run_pair("fp64", "batched")
    assert g["nsteps"] == 90 and g["nframes_used"] == 50
    _check(g, o, 1e-10, 1e-10, 1e-10, dlam=1e-9)


def test_pipeline_fp64_plain_matches_oracle_fp64():
    p, g, o = _run_pair("fp64", "plain")
    assert g["nsteps"] == 90 and g["nframes_used"] == 90
    _check(g, o, 1e-10, 1e-10, 1e-10, dlam=1e-9)


def test_pipeline_fp32_batched_within_1e4_of_oracle():
    # FP32 path against the reference algorithm: 1e-4 relative (TCFs, spectra)
    p, g, o = _run_pair("fp32", "batched")
    _check(g, o, 1e-4, 1e-4, 1e-4)
    _, _, o32 = _run_pair("fp32", "batched", oracle_prec=O.F32)
    _check(g, o32, 1e-4, 1e-4, 1e-4)


def test_pipeline_fp32_plain_within_1e4_of_oracle():
    p, g, o = _run_pair("fp32", "plain")
    _check(g, o, 1e-4, 1e-4, 1e-4)


def test_pipeline_bf16x3_within_1e3_of_fp64():
    p, g, o = _run_pair("bf16x3", "batched")
    _check(g, o, 1e-3, 1e-3, 1e-4)


@pytest.mark.parametrize("species,model,imap", [("HOD/D2O", "tip4p", 0), ("D2O", "e3b3", 1),
                                                ("HOD/H2O", "e3b2", 1)])
def test_pipeline_models_species_maps(species, model, imap):
    p, g, o = _run_pair("fp64", "batched", species=species, model=model, map=imap, omegaStart=2000.0,
                        omegaStop=4200.0)
    _check(g, o, 1e-10, 1e-10, 1e-10, dlam=1e-9)


def test_pipeline_uncoupled_closed_form():
    # HOD species: H is diagonal, F(t_k) = diag(exp(+i sum_{j=1..k} (w_i(t_j)-avef) dt/hbar))
    # (right endpoint), so C_IR(t_k) = sum_i mu_i(t_k).mu_i(0) exp(i phi_i(k)).
    p, g, _ = _run_pair("fp64", "batched", species="HOD/D2O", nsamples=1, ntcfpoints=12)
    xyz, box, _ = frames(128, 60, 1001)
    cfg = O.config_from_params(p)
    phase = 0.0
    ref = []
    mu0 = None
    for k in range(12):
        e = O.eproj(O.F64, cfg, xyz[k], box[k])
        s = O.sites(O.F64, cfg, xyz[k], box[k], e)
        mu = s[:, O.SITE_MU:O.SITE_MU + 3]
        if k == 0:
            mu0 = mu
        else:
            phase = phase + (s[:, O.SITE_W] - p.avef) * p.dt / 5.308837367
        ref.append(np.sum(np.sum(mu * mu0, axis=1) * np.exp(1j * phase)))
    ref = np.array(ref) * np.exp(-np.arange(12) * p.dt / (2 * p.t1))
    assert rel(g["tcf_ir"], ref) < 1e-10
    # C_IR(0) = sum |mu|^2 (real)
    assert abs(g["tcf_ir"][0].imag) < 1e-12


@pytest.mark.parametrize("prec,tol,dlam", [("fp64", 1e-10, 1e-9), ("fp32", 1e-4, 1e-3)])
def test_pipeline_plain_whole_gpu_eigensolver(prec, tol, dlam):
    # N = 512: the plain control flow diagonalises one Hamiltonian at a time,
    # which runs the tridiagonalisation across the whole GPU (tri1_* kernels)
    p, g, o = _run_pair(prec, "plain", nwater=256, nframes=20, seed=1003, nsamples=2, ntcfpoints=6,
                        sampleEvery=0.05)
    assert g["nsteps"] == 12
    _check(g, o, tol, tol, tol, dlam=dlam, n=512)


def test_pipeline_8192_oh_box_against_gpu_fp64():
    # 4096 waters (N = 8192, H 256 MiB per frame): the CPU oracle cannot
    # diagonalise this in test time, so the FP64 pipeline (pinned to the
    # oracle above) is the reference; FP32 within 1e-4, BF16x3 within 1e-3
    nwater, nframes = 4096, 3
    xyz, box, times = frames(nwater, nframes, 1005)
    kw = dict(KW, nsamples=1, ntcfpoints=3, sampleEvery=0.01)
    res = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        announce("pipeline N=8192 prec=%s" % prec)
        res[prec] = C.run(C.default_params(precision=prec, mode="batched", **kw), xyz, box, times)
    ref = res["fp64"]
    assert ref["nsteps"] == 3 and ref["nframes_used"] == 3
    for prec, tol in (("fp32", 1e-4), ("bf16x3", 1e-3)):
        g = res[prec]
        for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
            assert rel(g[k], ref[k]) < tol, (prec, k, rel(g[k], ref[k]))
        for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
            assert rel_l2(g[k], ref[k]) < tol, (prec, k, rel_l2(g[k], ref[k]))
        assert rel(g["sw"], ref["sw"]) < tol


@pytest.mark.parametrize("nwater", [50, 64, 333])
def test_pipeline_uneven_box_sizes(nwater):
    # boxes of uneven (and ragged: N = 100, 666) sizes, as in a batch of
    # independent boxes: batched FP32 within 1e-4 of the FP64 oracle
    p, g, o = _run_pair("fp32", "batched", nwater=nwater, nframes=24, seed=3000 + nwater, nsamples=2,
                        ntcfpoints=8, sampleEvery=0.05)
    assert g["nsteps"] == 16
    _check(g, o, 1e-4, 1e-4, 1e-4, n=2 * nwater)


def test_cli_reads_an_xtc_trajectory(tmp_path):
    # the drop-in executable on an XTC input (xdrfile's role, calcIR.cu:212-395)
    # gives the same files as the in-memory pipeline on the decoded frames
    xyz, box, times = frames(64, 30, 7)
    traj = tmp_path / "traj.xtc"
    C.xtc_write(traj, xyz, box, times)
    xq, bq, tq, _ = C.xtc_read(traj)
    inp = tmp_path / "input.inp"
    inp.write_text("xtcf %s\noutf %s\nmodel tip4p2005\nspecies H2O\nntcfpoints 10\nnsamples 2\n"
                   "sampleEvery 0.1\navef 3400\nomegaStart 2800\nomegaStop 4000\nnzeros 502\n"
                   % (traj, tmp_path / "cli"))
    announce("calcir CLI on an xtc file")
    r = subprocess.run([C.CLI_PATH, str(inp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "Found 256 atoms and 64 molecules." in r.stdout
    p = C.read_params(str(inp))
    p.outf = str(tmp_path / "mem").encode()
    C.write_outputs(p, C.run(p, xq, bq, tq))
    for suffix in ("irls", "vvls", "vhls", "irrtcf", "vhitcf", "Pw", "spdn"):
        a = np.loadtxt(tmp_path / ("cli_%s.dat" % suffix))
        b = np.loadtxt(tmp_path / ("mem_%s.dat" % suffix))
        assert a.shape == b.shape, suffix
        # both runs are FP32 with float atomics (sum order varies run to run)
        # and the files hold 6 significant digits
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-5 * np.abs(b).max(), err_msg=suffix)


THIRTEEN = ("irrtcf", "iritcf", "vvrtcf", "vvitcf", "vhrtcf", "vhitcf", "spdn", "Pw", "Rw", "Rmw", "irls",
            "vvls", "vhls")


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_cli_files_equal_the_oracle_files(tmp_path, prec):
    # one input file drives the drop-in executable (XTC read while the batched
    # run goes) and the CPU oracle on the decoded frames; each of the 13 files
    # of calcIR.cu:843-919 agrees.  FP64: to the 6 significant digits the
    # files hold.  BF16x3: TCFs and lineshapes within 1e-3 of the largest
    # entry; Pw within the bin moves of the eigenvalue error.
    nwater, nframes = 64, 40
    xyz, box, times = frames(nwater, nframes, 11)
    traj = tmp_path / "traj.xtc"
    C.xtc_write(traj, xyz, box, times)
    xq, bq, tq, _ = C.xtc_read(traj)
    inp = tmp_path / "input.inp"
    inp.write_text("xtcf %s\noutf %s\nmodel tip4p2005\nspecies H2O\nntcfpoints 12\nnsamples 3\n"
                   "sampleEvery 0.1\navef 3400\nomegaStart 2800\nomegaStop 4000\nomegaStep 5\nnzeros 500\n"
                   "precision %s\nmode batched\n" % (traj, tmp_path / "cli", prec))
    announce("calcir CLI prec=%s against the oracle's 13 files" % prec)
    r = subprocess.run([C.CLI_PATH, str(inp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    p = C.read_params(str(inp))
    p.outf = str(tmp_path / "orc").encode()
    o = O.run(O.F64, O.config_from_params(p), xq, bq, tq)
    C.write_outputs(p, o)
    steps = 3 * 12 * 2 * nwater  # mode-steps binned into Pw
    for s in THIRTEEN:
        a = np.loadtxt(tmp_path / ("cli_%s.dat" % s))
        b = np.loadtxt(tmp_path / ("orc_%s.dat" % s))
        assert a.shape == b.shape, s
        assert np.array_equal(a[:, 0], b[:, 0]), s  # time / frequency column
        y, z = a[:, 1], b[:, 1]
        if s in ("Pw", "Rw", "Rmw") and prec != "fp64":
            if s == "Pw":
                moves = 0.5 * np.abs(y - z).sum() * 5 * steps
                assert moves <= 4 * steps * 2 * 1e-3 / 5 + 2, moves
            continue
        assert np.array_equal(np.isnan(y), np.isnan(z)), s
        m = ~np.isnan(z)
        scale = np.abs(z[m]).max() if m.any() else 1.0
        if prec == "fp64":
            np.testing.assert_allclose(y[m], z[m], rtol=2e-5, atol=2e-6 * scale, err_msg=s)
        else:
            assert np.abs(y[m] - z[m]).max() <= 1e-3 * scale, s


def test_checkpoint_halt_and_resume_reproduce_the_uninterrupted_run(tmp_path):
    # max_frames stops a run at a chunk end with a checkpoint (CIR_INCOMPLETE);
    # resume continues from it.  Two halts and a final resume give the
    # uninterrupted result; periodic checkpoints leave the result unchanged.
    xyz, box, times = frames(128, 60, 1001)
    ck = str(tmp_path / "job")
    for prec, tol in (("fp64", 1e-12), ("fp32", 5e-5)):
        announce("checkpoint prec=%s uninterrupted" % prec)
        ref = C.run(C.default_params(precision=prec, mode="batched", chunk=8, **KW), xyz, box, times)
        announce("checkpoint prec=%s every 8 frames" % prec)
        per = C.run(C.default_params(precision=prec, mode="batched", chunk=8, cptf=ck, checkpoint_every=8, **KW),
                    xyz, box, times)
        assert (tmp_path / "job_ckpt.bin").exists()
        for k in ("tcf_ir", "tcf_vv", "tcf_vh", "sw", "pw"):
            assert rel(per[k], ref[k]) < tol, (prec, k)
        (tmp_path / "job_ckpt.bin").unlink()
        for step, (resume, maxf) in enumerate(((0, 16), (1, 16), (1, 0))):
            announce("checkpoint prec=%s leg %d resume=%d max_frames=%d" % (prec, step, resume, maxf))
            p = C.default_params(precision=prec, mode="batched", chunk=8, cptf=ck, resume=resume, max_frames=maxf,
                                 **KW)
            if maxf:
                with pytest.raises(C.CalcirError) as e:
                    C.run(p, xyz, box, times)
                assert e.value.status == C.INCOMPLETE
            else:
                g = C.run(p, xyz, box, times)
        assert g["nframes_used"] == ref["nframes_used"] == 50 and g["nsteps"] == ref["nsteps"]
        for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
            assert rel(g[k], ref[k]) < tol, (prec, k, rel(g[k], ref[k]))
        # FP32 runs differ by the order of their float atomics; the lineshapes
        # (sums with cancellation) move by ~1e-5 between runs
        for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
            assert rel_l2(g[k], ref[k]) < (tol if prec == "fp64" else 1e-4), (prec, k)
        assert rel(g["sw"], ref["sw"]) < tol
        # Pw sums weights with device atomics (order-dependent in the last bits)
        assert np.abs(g["pw"] - ref["pw"]).max() <= (1e-12 if prec == "fp64" else 1e-3) * ref["pw"].max()
        assert int(g["hist_skipped"]) == int(ref["hist_skipped"]) or prec != "fp64"


def test_checkpoint_guards(tmp_path):
    xyz, box, times = frames(64, 40, 77)
    ck = str(tmp_path / "g")
    kw = dict(KW, nsamples=2, ntcfpoints=10)
    announce("checkpoint guards: missing file")
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(precision="fp64", cptf=ck, resume=1, **kw), xyz, box, times)
    assert e.value.status == 2
    announce("checkpoint guards: halt")
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(precision="fp64", cptf=ck, chunk=4, max_frames=4, **kw), xyz, box, times)
    assert e.value.status == C.INCOMPLETE
    # another job (different avef, different frames) must not resume from it
    announce("checkpoint guards: other job")
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(precision="fp64", cptf=ck, resume=1, **dict(kw, avef=3450.0)), xyz, box, times)
    assert e.value.status == 1
    moved = xyz.copy()
    moved[0, 5, 0] += 0.01
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(precision="fp64", cptf=ck, resume=1, **kw), moved, box, times)
    assert e.value.status == 1
    x41, b41, t41 = frames(64, 41, 77)
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(precision="fp64", cptf=ck, resume=1, **kw), x41, b41, t41)
    assert e.value.status == 1
    # a truncated file is refused
    data = (tmp_path / "g_ckpt.bin").read_bytes()
    (tmp_path / "g_ckpt.bin").write_bytes(data[:len(data) // 2])
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(precision="fp64", cptf=ck, resume=1, **kw), xyz, box, times)
    assert e.value.status == 2
    # the reference control flow has no checkpoints
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(precision="fp64", mode="plain", cptf=ck, max_frames=4, **kw), xyz, box, times)
    assert e.value.status == 1


def test_cli_halts_and_resumes(tmp_path):
    # the drop-in executable: max_frames exits with status 3 and writes no
    # output files; resume 1 finishes and writes the same files as one run
    base = ("xtcf synth:64:40:7\nmodel tip4p2005\nspecies H2O\nntcfpoints 10\nnsamples 2\nsampleEvery 0.1\n"
            "avef 3400\nomegaStart 2800\nomegaStop 4000\nnzeros 502\nprecision fp64\nchunk 4\n")
    one = tmp_path / "one.inp"
    one.write_text(base + "outf %s\n" % (tmp_path / "one"))
    announce("calcir CLI uninterrupted")
    r = subprocess.run([C.CLI_PATH, str(one)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    leg = tmp_path / "leg.inp"
    leg.write_text(base + "outf %s\ncptf %s\nmax_frames 8\n" % (tmp_path / "two", tmp_path / "two"))
    announce("calcir CLI halted")
    r = subprocess.run([C.CLI_PATH, str(leg)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 3, r.stdout[-2000:]
    assert "resume 1" in r.stdout and not (tmp_path / "two_irls.dat").exists()
    leg.write_text(base + "outf %s\ncptf %s\nresume 1\n" % (tmp_path / "two", tmp_path / "two"))
    announce("calcir CLI resumed")
    r = subprocess.run([C.CLI_PATH, str(leg)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    for suffix in ("irls", "vvls", "vhls", "irrtcf", "vhitcf", "Pw", "spdn"):
        a = np.loadtxt(tmp_path / ("two_%s.dat" % suffix))
        b = np.loadtxt(tmp_path / ("one_%s.dat" % suffix))
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(b).max(), err_msg=suffix)


def test_run_many_uneven_boxes_equal_single_runs():
    # independent boxes of uneven size (ragged N = 100, 128, 256, 512, 666),
    # mixed precisions, at once on one GPU: each job's result is its own run's
    sizes = [(64, "fp64"), (50, "fp32"), (256, "fp32"), (128, "fp64"), (333, "bf16x3")]
    jobs = []
    for j, (nw, prec) in enumerate(sizes):
        xyz, box, times = frames(nw, 24, 4000 + j)
        p = C.default_params(precision=prec, mode="batched", **dict(KW, nsamples=2, ntcfpoints=8, sampleEvery=0.05))
        jobs.append((p, xyz, box, times))
    announce("run_many %d jobs" % len(jobs))
    outs = C.run_many(jobs, nthreads=4)
    for (p, xyz, box, times), g, (nw, prec) in zip(jobs, outs, sizes):
        announce("run_many single check nwater=%d %s" % (nw, prec))
        ref = C.run(p, xyz, box, times)
        tol = 1e-12 if prec == "fp64" else 5e-5  # FP32-class: float atomics reorder sums
        for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
            assert rel(g[k], ref[k]) < tol, (nw, prec, k, rel(g[k], ref[k]))
        assert g["nsteps"] == ref["nsteps"] == 16


def test_run_many_gives_back_its_pooled_memory_and_names_a_failed_job():
    # cir_run_many's jobs take their buffers from a stream-ordered memory pool
    # (the 1024-water job alone holds several hundred MB at once): after the
    # call the device memory is free again, also when a job fails, and the
    # other jobs' results equal their single runs
    import torch

    sizes = [(1024, "bf16x3", "batched"), (256, "fp32", "plain"), (128, "fp64", "batched")]
    jobs = []
    for j, (nw, prec, mode) in enumerate(sizes):
        xyz, box, times = frames(nw, 16, 4100 + j)
        p = C.default_params(precision=prec, mode=mode, **dict(KW, nsamples=2, ntcfpoints=6, sampleEvery=0.05))
        jobs.append((p, xyz, box, times))
    announce("run_many pooled: warm-up")
    C.run_many(jobs, nthreads=3)  # first calls time launch shapes and load code
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    announce("run_many pooled: memory")
    outs = C.run_many(jobs, nthreads=3)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 < (64 << 20), (free0, free1)
    for (p, xyz, box, times), g, (nw, prec, mode) in zip(jobs, outs, sizes):
        announce("run_many pooled: single check nwater=%d %s %s" % (nw, prec, mode))
        ref = C.run(p, xyz, box, times)
        tol = 1e-12 if prec == "fp64" else 5e-5
        for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
            assert rel(g[k], ref[k]) < tol, (nw, prec, k, rel(g[k], ref[k]))
    bad = jobs[2][1].copy()
    bad[3, 5, 1] = np.nan
    announce("run_many pooled: a failing job")
    with pytest.raises(C.CalcirError, match="job 1"):
        C.run_many([jobs[0], (jobs[2][0], bad, jobs[2][2], jobs[2][3])], nthreads=2)
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    assert free0 - free2 < (64 << 20), (free0, free2)


def test_reunitarisation_in_the_pipeline():
    # FP64: F is unitary to rounding, so restoring the Gram matrices changes
    # nothing beyond 1e-12; BF16x3 with it stays within 1e-3 of FP64 and close
    # to BF16x3 without it; a halted and resumed run keeps the recorded G0
    xyz, box, times = frames(128, 60, 1001)
    announce("reunitarise fp64")
    a = C.run(C.default_params(precision="fp64", **KW), xyz, box, times)
    b = C.run(C.default_params(precision="fp64", reunitarise_every=5, **KW), xyz, box, times)
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(b[k], a[k]) < 1e-12, (k, rel(b[k], a[k]))
    announce("reunitarise bf16x3")
    c = C.run(C.default_params(precision="bf16x3", reunitarise_every=3, **KW), xyz, box, times)
    d = C.run(C.default_params(precision="bf16x3", **KW), xyz, box, times)
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
        assert rel_l2(c[k], a[k]) < 1e-3 and rel_l2(c[k], d[k]) < 1e-4, k
    with pytest.raises(C.CalcirError):
        C.run(C.default_params(precision="fp64", mode="plain", reunitarise_every=5, **KW), xyz, box, times)


def test_reunitarisation_survives_checkpoint_resume(tmp_path):
    xyz, box, times = frames(128, 60, 1001)
    ck = str(tmp_path / "ru")
    kw = dict(KW, precision="fp64", chunk=8, reunitarise_every=4)
    announce("reunitarise + checkpoint: uninterrupted")
    ref = C.run(C.default_params(**kw), xyz, box, times)
    announce("reunitarise + checkpoint: halt")
    with pytest.raises(C.CalcirError) as e:
        C.run(C.default_params(cptf=ck, max_frames=24, **kw), xyz, box, times)
    assert e.value.status == C.INCOMPLETE
    announce("reunitarise + checkpoint: resume")
    g = C.run(C.default_params(cptf=ck, resume=1, **kw), xyz, box, times)
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(g[k], ref[k]) < 1e-12, (k, rel(g[k], ref[k]))


def _blocks(xyz, box, sizes):
    first = 0
    for k in sizes:
        yield xyz[first:first + k], box[first:first + k]
        first += k
    if first < xyz.shape[0]:
        yield xyz[first:], box[first:]


@pytest.mark.parametrize("prec,tol,chunk,sizes", [
    ("fp64", 1e-12, 8, (1, 7, 13, 1, 2)),  # ragged blocks across chunk ends
    ("fp64", 1e-12, 0, ()),                 # the whole trajectory in one push
    ("bf16x3", 5e-5, 16, (1,) * 60),        # frame by frame
])
def test_streamed_run_equals_cir_run(prec, tol, chunk, sizes):
    # cir_stream_begin / _push / _end give cir_run's results: the same
    # schedule, the same chunks, frames copied as they arrive
    xyz, box, times = frames(128, 60, 1001)
    p = C.default_params(precision=prec, mode="batched", chunk=chunk, **KW)
    announce("stream prec=%s chunk=%d blocks=%d" % (prec, chunk, len(sizes)))
    ref = C.run(p, xyz, box, times)
    g = C.run_streamed(p, times, _blocks(xyz, box, sizes), xyz.shape[1])
    assert g["nframes_used"] == ref["nframes_used"] == 50 and g["nsteps"] == ref["nsteps"]
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(g[k], ref[k]) < tol, (k, rel(g[k], ref[k]))
    # two runs of an FP32-class job differ by the order of their float atomics;
    # the lineshapes (sums with cancellation) move by ~1e-5 between runs
    tol_spec = tol if prec == "fp64" else 1e-4
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
        assert rel_l2(g[k], ref[k]) < tol_spec, (k, rel_l2(g[k], ref[k]))
    assert rel(g["sw"], ref["sw"]) < tol
    assert np.abs(g["pw"] - ref["pw"]).max() <= (1e-12 if prec == "fp64" else 1e-3) * ref["pw"].max()


def test_streamed_triclinic_run_equals_cir_run():
    from cases import triclinic_frames
    xyz, box9, times = triclinic_frames(64, 40, 31)
    box9 = np.ascontiguousarray(box9.reshape(len(times), 9))
    p = C.default_params(precision="fp64", mode="batched", chunk=8, triclinic=1,
                         **dict(KW, nsamples=2, ntcfpoints=12))
    announce("stream triclinic fp64")
    ref = C.run(p, xyz, box9, times)
    g = C.run_streamed(p, times, _blocks(xyz, box9, (5, 5, 11)), xyz.shape[1])
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(g[k], ref[k]) < 1e-12, k


def test_stream_guards():
    import ctypes
    xyz, box, times = frames(32, 24, 5)
    L = C.lib()
    kw = dict(KW, nsamples=2, ntcfpoints=8, sampleEvery=0.05)
    natoms = xyz.shape[1]

    def ctx_for(**over):
        p = C.default_params(precision="fp64", mode="batched", **dict(kw, **over))
        ctx = ctypes.c_void_p()
        C.check(L.cir_create(ctypes.byref(p), natoms, None, ctypes.byref(ctx)))
        return ctx

    def push(ctx, first, count):
        x = np.ascontiguousarray(xyz[first:first + count])
        b = np.ascontiguousarray(box[first:first + count])
        return L.cir_stream_push(ctx, first, count, x.ctypes.data, b.ctypes.data)

    announce("stream guards")
    ctx = ctx_for()
    try:
        # push without begin; frames out of order close the stream
        assert push(ctx, 0, 2) == 1
        C.check(L.cir_stream_begin(ctx, len(times), times.ctypes.data))
        C.check(push(ctx, 0, 4))
        assert push(ctx, 5, 2) == 1
        assert "expected frame 4" in L.cir_last_error().decode()
        assert push(ctx, 4, 2) == 1  # the failed stream is closed
        # a used frame never pushed: CIR_ERR_FRAME at the end
        C.check(L.cir_stream_begin(ctx, len(times), times.ctypes.data))
        C.check(push(ctx, 0, 10))
        assert L.cir_stream_end(ctx) == 6
        # frames past the end, non-finite coordinates
        C.check(L.cir_stream_begin(ctx, len(times), times.ctypes.data))
        C.check(push(ctx, 0, 20))
        x = np.ascontiguousarray(np.concatenate([xyz[20:], xyz[:1]]))
        b = np.ascontiguousarray(np.concatenate([box[20:], box[:1]]))
        assert L.cir_stream_push(ctx, 20, 5, x.ctypes.data, b.ctypes.data) == 1
        C.check(L.cir_stream_begin(ctx, len(times), times.ctypes.data))
        bad = np.array(xyz[:1])
        bad[0, 3, 1] = np.nan
        assert L.cir_stream_push(ctx, 0, 1, bad.ctypes.data, np.ascontiguousarray(box[:1]).ctypes.data) == 10
        # a complete stream after the failures gives cir_run's result
        C.check(L.cir_stream_begin(ctx, len(times), times.ctypes.data))
        C.check(push(ctx, 0, len(times)))
        C.check(L.cir_stream_end(ctx))
        r, bufs, _ = C._collect(L, ctx, C.default_params(precision="fp64", mode="batched", **kw), False)
        g = C._results_dict(r, bufs, None)
    finally:
        L.cir_destroy(ctx)
    ref = C.run(C.default_params(precision="fp64", mode="batched", **kw), xyz, box, times)
    assert rel(g["tcf_ir"], ref["tcf_ir"]) < 1e-12
    # plain mode and the checkpoint keys are refused at begin
    for over in (dict(mode="plain"), dict(checkpoint_every=4), dict(max_frames=4)):
        ctx = ctx_for(**over) if "mode" not in over else None
        if ctx is None:
            p = C.default_params(precision="fp64", **dict(kw, **over))
            ctx = ctypes.c_void_p()
            C.check(L.cir_create(ctypes.byref(p), natoms, None, ctypes.byref(ctx)))
        try:
            assert L.cir_stream_begin(ctx, len(times), times.ctypes.data) == 1, over
        finally:
            L.cir_destroy(ctx)


def test_origin_stride_frames_equals_sample_every():
    # origins every 10 frames of 10 fs select the same frames as sampleEvery
    # 0.1 ps, so the run is the same
    xyz, box, times = frames(64, 40, 12)
    kw = dict(KW, nsamples=2, ntcfpoints=12)
    announce("origin_stride_frames 10 vs sampleEvery 0.1")
    a = C.run(C.default_params(precision="fp64", **kw), xyz, box, times)
    b = C.run(C.default_params(precision="fp64", origin_stride_frames=10, **dict(kw, sampleEvery=7.0)), xyz, box,
              times)
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(b[k], a[k]) < 1e-12, k


def test_cli_rounds_the_frame_spacing_at_the_file_precision(tmp_path):
    # calcIR.cu:342-345: frame_dt = round(dt * prec) / prec with the XTC file's
    # own precision.  Frames 0.0125 ps apart in a precision-10000 file give
    # frame_dt 0.0125 (at precision 1000 it would be 0.013 and the time check
    # would refuse the run after a few points).
    xyz, box, _ = frames(32, 40, 3)
    times = (np.arange(40) * 0.0125).astype(np.float32)
    traj = tmp_path / "fine.xtc"
    C.xtc_write(traj, xyz, box, times, precision=10000.0)
    inp = tmp_path / "input.inp"
    inp.write_text("xtcf %s\noutf %s\nmodel tip4p2005\nspecies H2O\nntcfpoints 20\nnsamples 2\n"
                   "sampleEvery 0.1\ndt 0.0125\navef 3400\nomegaStart 2000\nomegaStop 4500\nnzeros 100\n"
                   "precision fp64\n" % (traj, tmp_path / "cli"))
    announce("calcir CLI, frame spacing at precision 10000")
    r = subprocess.run([C.CLI_PATH, str(inp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "Frame time offset is: 0.012500" in r.stdout
    xq, bq, tq, _ = C.xtc_read(traj)
    p = C.read_params(str(inp))
    p.time_prec = 10000.0
    o = O.run(O.F64, O.config_from_params(p), xq, bq, tq)
    a = np.loadtxt(tmp_path / "cli_irrtcf.dat")
    np.testing.assert_allclose(a[:, 1], o["tcf_ir"].real, rtol=2e-5, atol=2e-6 * np.abs(o["tcf_ir"].real).max())
    # the same frames at the default precision 1000 are refused by the time check
    p.time_prec = 1000.0
    with pytest.raises(C.CalcirError):
        C.run(p, xq, bq, tq)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("bf16x3", 1e-4)])
def test_graph_frame_loops_equal_direct_launches(prec, tol, monkeypatch):
    # each chunk's frame loop (propagation, origin bookkeeping,
    # re-unitarisation, correlations) runs as one captured graph;
    # CIR_NO_GRAPHS=1 launches the same kernels one by one.  Several chunks
    # (chunk 8), re-unitarisation inside the graphs; the graph runs alternate
    # with direct runs and each is checked against the direct result, to the
    # run-to-run rounding of the float atomics (2e-14 seen in FP64 spectra,
    # 1.2e-5 in BF16x3 ones; a wrong frame or buffer in a graph is O(1)).
    xyz, box, times = frames(128, 60, 1001)
    p = C.default_params(precision=prec, mode="batched", chunk=8, reunitarise_every=4, **KW)
    monkeypatch.setenv("CIR_NO_GRAPHS", "1")
    announce("graph vs direct: direct %s" % prec)
    ref = C.run(p, xyz, box, times)
    for rep in range(2):
        monkeypatch.delenv("CIR_NO_GRAPHS")
        announce("graph vs direct: graphs %s, run %d" % (prec, rep))
        g = C.run(p, xyz, box, times)
        for k in ("tcf_ir", "tcf_vv", "tcf_vh", "ftcf_ir", "ftcf_vv", "ftcf_vh", "sw", "pw"):
            assert rel(g[k], ref[k]) <= tol, (k, rel(g[k], ref[k]))
        monkeypatch.setenv("CIR_NO_GRAPHS", "1")
        d = C.run(p, xyz, box, times)
        for k in ("tcf_ir", "ftcf_ir", "sw"):
            assert rel(d[k], ref[k]) <= tol, (k, rel(d[k], ref[k]))


@pytest.mark.parametrize("prec,mode", [("bf16x3", "batched"), ("fp32", "plain")])
def test_deterministic_runs_are_bitwise_equal(prec, mode):
    # deterministic 1: two-stage eigensolver (no float atomics), one K segment
    # in the propagation, histograms/density summed in a fixed order, default
    # launch shapes; two runs give the same bits, and agree with the default
    # run to the FP32 bar
    xyz, box, times = frames(128, 60, 1001)
    kw = dict(KW, chunk=16, reunitarise_every=4) if mode == "batched" else dict(KW, nsamples=1, ntcfpoints=8)
    p = C.default_params(precision=prec, mode=mode, deterministic=1, **kw)
    announce("deterministic %s %s: two runs" % (prec, mode))
    a = C.run(p, xyz, box, times)
    b = C.run(p, xyz, box, times)
    for k in ("tcf_ir", "tcf_vv", "tcf_vh", "ftcf_ir", "ftcf_vv", "ftcf_vh", "sw", "pw", "rw", "rmw"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    d = C.run(C.default_params(precision=prec, mode=mode, **kw), xyz, box, times)
    for k in ("tcf_ir", "ftcf_ir", "sw"):
        assert rel(a[k], d[k]) < 1e-4, (k, rel(a[k], d[k]))


def test_deterministic_needs_fp32_and_the_two_stage_sizes():
    xyz, box, times = frames(16, 4, 1)  # N = 32 < 66
    for prec in ("fp64", "fp32"):
        p = C.default_params(precision=prec, deterministic=1, **dict(KW, nsamples=1, ntcfpoints=2))
        with pytest.raises(C.CalcirError, match="deterministic needs"):
            C.run(p, xyz, box, times)
