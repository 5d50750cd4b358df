"""The C ABI boundary without a GPU: libcalcir loads and exports every symbol
include/calcir.h declares; defaults, the input-file parser, model/species
resolution, frame selection, the 13 output writers, the synthetic frame
source and the calcir command line (calcIR.cu:9-87, 323-352, 843-919,
1539-1643)."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import calcir_amd as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "calcir.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cir_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.lib()
    names = _declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
        assert name in C.SIGNATURES, name
    # and nm agrees that they are exported, C-linkage functions
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r" T (cir_[a-z0-9_]+)$", out, flags=re.M))
    assert set(names) <= exported


def test_defaults_are_the_reference_defaults():
    p = C.default_params()
    assert p.xtcf == b"traj.xtc" and p.outf == b"spec" and p.cptf == b"spec" and p.model == b"e3b3"
    assert p.species == b" "
    assert (p.ntcfpoints, p.nsamples, p.omegaStep, p.natom_mol, p.nchrom_mol, p.nzeros, p.imap) == \
        (150, 1, 5, 4, 2, 25600, 0)
    assert (p.sampleEvery, p.omegaStart, p.omegaStop, p.beginTime) == (10.0, 2000.0, 5000.0, 0.0)
    assert p.dt == pytest.approx(0.010) and p.t1 == pytest.approx(0.260) and p.avef == pytest.approx(3415.2)
    assert p.cutoff == pytest.approx(0.7831)
    assert C.nomega(p) == 601
    # the reference's H choice and charged sites; GROMACS' default precision
    assert (p.oh_per_molecule, p.msite_bisector, p.time_prec) == (0, 0, 1000.0)


def test_parser_keys_comments_and_quirks(tmp_path):
    f = tmp_path / "input.inp"
    f.write_text(
        "xtcf            nvt.xtc                     // the xtc file\n"
        "outf            test                        // prefix\n"
        "dt              0.020\n"
        "ntcfpoints      400\n"
        "nsamples        2\n"
        "sampleEvery     4.0\n"
        "t1              0.300\n"
        "avef            3400.0\n"
        "omegaStart      2000\n"
        "omegaStop       5000\n"
        "#avef            2415.2                      // commented keys are unknown\n"
        "omegaStep       5\n"
        "natom_mol       4\n"
        "nchrom_mol      2\n"
        "nzeros          25600\n"
        "model           tip4p\n"
        "beginTime       1.5\n"
        "map             1\n"
        "species         HOD/D2O\n"
        "cutoff          0.75\n"
        "precision       bf16x3\n"
        "mode            plain\n"
        "checkpoint_every 40\n"
        "resume          1\n"
        "max_frames      200\n"
        "adjoint_store   2\n"
        "triclinic       1\n"
        "reunitarise_every 25\n"
        "origin_stride_frames 20\n")
    p = C.read_params(str(f))
    assert p.xtcf == b"nvt.xtc" and p.outf == b"test" and p.model == b"tip4p" and p.species == b"HOD/D2O"
    assert p.ntcfpoints == 400 and p.nsamples == 2 and p.imap == 1
    assert p.dt == pytest.approx(0.02) and p.t1 == pytest.approx(0.3) and p.avef == pytest.approx(3400.0)
    assert p.beginTime == pytest.approx(1.5) and p.sampleEvery == pytest.approx(4.0)
    assert p.cutoff == pytest.approx(0.75) and p.precision == C.PREC_BF16X3 and p.mode == C.MODE_PLAIN
    assert (p.ckpt_every, p.resume, p.max_frames) == (40, 1, 200)
    assert (p.adj_store, p.triclinic, p.reunit_every) == (2, 1, 25)
    assert C.default_params(reunitarise_every=7).reunit_every == 7
    assert p.origin_stride == 20 and C.default_params().origin_stride == 0
    # origins every 20 frames of 0.01 ps: sample 3, point 5 is frame 65, as
    # with sampleEvery = 0.2 ps
    q = C.default_params(origin_stride_frames=20, sampleEvery=99.0, dt=0.01)
    assert C.frame_index(q, 3, 5, 0.01)[0] == 65
    assert C.frame_index(C.default_params(sampleEvery=0.2, dt=0.01), 3, 5, 0.01)[0] == 65
    assert (C.default_params().adj_store, C.default_params().triclinic) == (0, 0)
    assert (C.default_params().ckpt_every, C.default_params().resume, C.default_params().max_frames) == (0, 0, 0)
    assert C.default_params(checkpoint_every=5).ckpt_every == 5
    # a line holding one token takes the next line's first token as its value
    g = tmp_path / "quirk.inp"
    g.write_text("outf\nspec2 extra words\nntcfpoints 7\n")
    q = C.read_params(str(g))
    assert q.outf == b"spec2" and q.ntcfpoints == 7


def test_parser_warns_on_unknown_key(tmp_path):
    f = tmp_path / "a.inp"
    f.write_text("#avef 2415.2\nbogus 3\nntcfpoints 12\n")
    code = ("import calcir_amd as C; p = C.read_params(%r, verbose=True); print('N', p.ntcfpoints)" % str(f))
    out = subprocess.run(["python", "-c", code], cwd=ROOT, stdout=subprocess.PIPE, text=True).stdout
    assert "WARNING: Parameter #avef in input file" in out
    assert "WARNING: Parameter bogus in input file" in out
    assert "N 12" in out


def test_parser_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(C.CalcirError) as e:
        C.read_params(str(tmp_path / "nope.inp"))
    assert e.value.status == 2


@pytest.mark.parametrize("model,val", [("tip4p", 0), ("e3b2", 0), ("tip4p2005", 1), ("e3b3", 1)])
def test_resolve_model(model, val):
    assert C.resolve_model(model) == val


@pytest.mark.parametrize("species,val", [("H2O", 0), ("HOD/D2O", 1), ("HOD/H2O", 2), ("D2O", 3)])
def test_resolve_species(species, val):
    assert C.resolve_species(species) == val


def test_resolve_errors():
    with pytest.raises(C.CalcirError) as e:
        C.resolve_model("spce")
    assert e.value.status == 3
    with pytest.raises(C.CalcirError) as e:
        C.resolve_species(" ")
    assert e.value.status == 4


def _fake_results(p):
    nt, nom, nfft = p.ntcfpoints, C.nomega(p), p.ntcfpoints + p.nzeros
    rng = np.random.default_rng(0)
    res = {k: rng.standard_normal(nt) + 1j * rng.standard_normal(nt) for k in ("tcf_ir", "tcf_vv", "tcf_vh")}
    res.update({k: rng.standard_normal(nfft) for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh")})
    res.update({k: rng.random(nom) for k in ("sw", "pw", "rw", "rmw")})
    return res


def test_writers_produce_the_thirteen_files(tmp_path):
    p = C.default_params(outf=str(tmp_path / "run"), ntcfpoints=10, nzeros=118, avef=3400.0, omegaStart=2800.0,
                         omegaStop=4000.0, dt=0.01)
    res = _fake_results(p)
    C.write_outputs(p, res)
    names = sorted(os.listdir(tmp_path))
    expect = sorted("run_" + s for s in ["irrtcf.dat", "iritcf.dat", "vvrtcf.dat", "vvitcf.dat", "vhrtcf.dat",
                                         "vhitcf.dat", "spdn.dat", "Pw.dat", "Rw.dat", "Rmw.dat", "irls.dat",
                                         "vvls.dat", "vhls.dat"])
    assert names == expect
    # TCF lines: "%g %g \n" (trailing space), time = i*dt
    lines = open(tmp_path / "run_irrtcf.dat").read().split("\n")
    assert lines[0] == "0 %g " % np.float32(res["tcf_ir"][0].real)
    assert lines[3] == "%g %g " % (np.float32(3 * np.float32(0.01)), np.float32(res["tcf_ir"][3].real))
    iml = open(tmp_path / "run_vhitcf.dat").read().split("\n")
    assert iml[2] == "%g %g " % (np.float32(2 * np.float32(0.01)), np.float32(res["tcf_vh"][2].imag))
    # distributions on the omega grid, "%g %g\n"
    pw = open(tmp_path / "run_Pw.dat").read().split("\n")
    assert len(pw) == C.nomega(p) + 1 and pw[-1] == ""
    assert pw[1] == "%g %g" % (np.float32(2805.0), np.float32(res["pw"][1]))
    # lineshape: descending frequency inside [omegaStart, omegaStop]
    ls = np.loadtxt(tmp_path / "run_irls.dat")
    assert np.all(np.diff(ls[:, 0]) < 0)
    assert ls[:, 0].max() <= 4000.0 and ls[:, 0].min() >= 2800.0
    n = p.ntcfpoints + p.nzeros
    factor = np.float32(2 * 3.14159265359 * 5.308837367 / (np.float32(0.01) * np.float32(n)))
    # first row of the 'positive' half: i = 0 -> freq = avef, value y0/(factor n)
    row = [r for r in ls if abs(r[0] - 3400.0) < 1e-3][0]
    assert row[1] == pytest.approx(np.float32(res["ftcf_ir"][0]) / (factor * n), rel=1e-5)


def test_lineshape_rows_keep_the_reference_loop_bounds(tmp_path):
    # calcIR.cu:899-919: the first loop runs i = n/2 .. n-1, the second
    # i = 0 .. ntcfpoints + nzeros/2 - 1 (operator precedence: not n/2), so
    # with a window holding every frequency the file has n - n/2 +
    # ntcfpoints + nzeros/2 rows, the second loop's last ones at frequencies
    # below avef - (n/2) factor taken from bins the first loop also wrote
    nt, nz = 10, 21
    p = C.default_params(outf=str(tmp_path / "w"), ntcfpoints=nt, nzeros=nz, avef=3400.0, omegaStart=-1e7,
                         omegaStop=1e7, omegaStep=1000000, dt=0.01)  # (a coarse grid for the distributions)
    n = nt + nz
    res = _fake_results(p)
    res["ftcf_ir"] = np.arange(n, dtype=np.float64) + 1.0  # bin i holds i + 1
    C.write_outputs(p, res)
    ls = np.loadtxt(tmp_path / "w_irls.dat")
    assert ls.shape[0] == (n - n // 2) + nt + nz // 2 == 36
    factor = np.float32(2 * 3.14159265359 * 5.308837367 / (np.float32(0.01) * np.float32(n)))
    scale = factor * np.float32(n)
    bins = np.rint(ls[:, 1] * scale).astype(int) - 1
    assert list(bins) == list(range(n // 2, n)) + list(range(0, nt + nz // 2))
    # frequencies: (n - i) factor + avef in the first loop, avef - i factor
    # in the second, so the second loop's bins past n/2 lie below avef - (n/2) factor
    assert np.all(ls[n - n // 2:, 0] <= 3400.0 + 1e-3)
    assert ls[-1, 0] == pytest.approx(3400.0 - (nt + nz // 2 - 1) * factor, rel=1e-5)  # "%g"
    assert ls[-1, 0] < 3400.0 - (n // 2) * factor


def test_synthetic_trajectory_is_seeded_and_rigid():
    a = C.synth_trajectory(64, 6, 42)
    b = C.synth_trajectory(64, 6, 42)
    c = C.synth_trajectory(64, 6, 43)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    xyz, box, times = a
    assert xyz.shape == (6, 256, 3) and np.allclose(times, np.arange(6) * 0.01)
    assert box[0, 0] == pytest.approx(1.564 * (64 / 128) ** (1 / 3), rel=1e-6)
    o, h1, h2, m = xyz[:, 0::4], xyz[:, 1::4], xyz[:, 2::4], xyz[:, 3::4]
    assert np.allclose(np.linalg.norm(h1 - o, axis=2), 0.09572, atol=2e-6)
    assert np.allclose(np.linalg.norm(h2 - o, axis=2), 0.09572, atol=2e-6)
    assert np.allclose(np.linalg.norm(m - o, axis=2), 0.01546, atol=2e-6)
    cosang = np.sum((h1 - o) * (h2 - o), axis=2) / 0.09572 ** 2
    assert np.allclose(np.degrees(np.arccos(np.clip(cosang, -1, 1))), 104.52, atol=2e-3)
    assert np.all((o >= 0) & (o < box[0, 0]))
    # no hard-core contacts: H...H >= 0.18 nm, H...M >= 0.21 nm between molecules
    L = box[0, 0]
    for f in range(6):
        hs = np.concatenate([h1[f], h2[f]])
        mol = np.concatenate([np.arange(64), np.arange(64)])
        d = hs[:, None, :] - hs[None, :, :]
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=2)
        r[mol[:, None] == mol[None, :]] = 9
        assert r.min() >= 0.18 - 1e-6
        d = hs[:, None, :] - m[f][None, :, :]
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=2)
        r[mol[:, None] == np.arange(64)[None, :]] = 9
        assert r.min() >= 0.21 - 1e-6


def _cli(args, cwd):
    return subprocess.run([C.CLI_PATH] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def test_cli_usage_and_input_errors(tmp_path):
    r = _cli([], tmp_path)
    assert r.returncode == 1 and "Include as the first argument the name of an input file" in r.stdout
    r = _cli(["a", "b"], tmp_path)
    assert r.returncode == 1
    r = _cli([str(tmp_path / "missing.inp")], tmp_path)
    assert r.returncode == 1 and "ERROR: Could not open" in r.stdout
    f = tmp_path / "bad.inp"
    f.write_text("model spce\nspecies H2O\n")
    r = _cli([str(f)], tmp_path)
    assert r.returncode == 1 and "WARNING: model: spce is not recognized" in r.stdout
    f.write_text("model tip4p\n")
    r = _cli([str(f)], tmp_path)
    assert r.returncode == 1 and "WARNING: species:   is not recognized" in r.stdout
    f.write_text("model tip4p\nspecies H2O\nxtcf nothere.xtc\n")
    r = _cli([str(f)], tmp_path)
    assert r.returncode == 1 and "could not be opened" in r.stdout
    # the echo of calcIR.cu:70-87
    assert "\tSetting model to tip4p\n" in r.stdout and "\tSetting species to H2O\n" in r.stdout


def test_rule_and_precision_keys_parse(tmp_path):
    f = tmp_path / "input.inp"
    f.write_text("nchrom_mol 1\nnatom_mol 3\noh_per_molecule 1\nmsite_bisector 1\ntime_precision 10000\n"
                 "deterministic 1\n")
    p = C.read_params(str(f))
    assert (p.nchrom_mol, p.natom_mol, p.oh_per_molecule, p.msite_bisector) == (1, 3, 1, 1)
    assert p.time_prec == 10000.0
    assert p.deterministic == 1
    assert C.default_params().deterministic == 0


def test_trajectory_shapes_are_checked_before_the_library_reads_them():
    xyz, box, times = C.synth_trajectory(4, 3, 1)
    p = C.default_params(species="H2O")
    C.check_trajectory(p, xyz, box, times)
    bad = [(xyz.reshape(3, -1), box, times),                  # [nframes, natoms*3]
           (xyz[:, :15], box, times),                         # natoms not a multiple of natom_mol
           (xyz, box[:2], times),                             # a box per frame
           (xyz, box, times[:2]),                             # a time per frame
           (xyz, np.zeros((3, 9), np.float32), times)]        # 9 floats without the triclinic key
    for args in bad:
        with pytest.raises(ValueError):
            C.check_trajectory(p, *args)
        with pytest.raises(ValueError):
            C.run(p, *args)
    pt = C.default_params(species="H2O", triclinic=1)
    C.check_trajectory(pt, xyz, np.zeros((3, 3, 3), np.float32), times)
    with pytest.raises(ValueError):
        C.check_trajectory(pt, xyz, box, times)
    with pytest.raises(ValueError):
        C.run_many([(p, xyz, box, times[:1])])


@pytest.mark.parametrize("over,natoms,status", [
    (dict(map=2), 16, 5),                    # calcIR.cu:1207-1242 leaves the map arrays unset: refused
    (dict(map=-1), 16, 5),
    (dict(model="spce"), 16, 3),             # calcIR.cu:327-330
    (dict(species="H2S"), 16, 4),            # calcIR.cu:336-339
    (dict(), 15, 1),                         # natoms not a multiple of natom_mol
    (dict(natom_mol=2), 16, 1),
    (dict(ntcfpoints=0), 16, 1),
    (dict(omegaStart=4000.0, omegaStop=3000.0), 16, 1),
    (dict(reunitarise_every=-1), 16, 1),
    (dict(triclinic=2), 16, 1),
])
def test_create_refuses_bad_parameters_before_looking_for_a_device(over, natoms, status):
    kw = dict(species="H2O")
    kw.update(over)
    p = C.default_params(**kw)
    L = C.lib()
    c = ctypes.c_void_p()
    assert L.cir_create(ctypes.byref(p), natoms, None, ctypes.byref(c)) == status
    assert not c.value
    assert L.cir_last_error().decode()


def test_create_with_good_parameters_needs_a_device():
    p = C.default_params(species="H2O")
    L = C.lib()
    c = ctypes.c_void_p()
    s = L.cir_create(ctypes.byref(p), 16, None, ctypes.byref(c))
    if s == 0:  # a machine with a GPU
        L.cir_destroy(c)
    else:
        assert s == 11 and not c.value  # CIR_ERR_NODEVICE
