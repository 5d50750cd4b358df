"""GPU parity against the oracle for the molecule layouts the reference
accepts (three-site water, one OH chromophore per molecule, with the
reference's chrom % 2 and atom-index semantics) and for the generalisations
beyond it (the oh_per_molecule and msite_bisector rules, triclinic boxes);
the CLI on XTC files of these layouts and the lineshape adjoint in a
triclinic box.  The oracle is pinned by closed forms, brute-force sums,
lattice invariance and reduction to the two-OH, four-site case."""

import subprocess

import numpy as np
import pytest

import calcir_amd as C
import oracle_lib as O
from calcir_amd.adjoint import LineshapeSession, lineshape_map
from cases import frames, rel, rel_l2, triclinic_frames
from conftest import announce

pytestmark = pytest.mark.gpu

KW = dict(species="H2O", model="tip4p2005", ntcfpoints=12, nsamples=3, sampleEvery=0.1, avef=3400.0,
          omegaStart=2800.0, omegaStop=4000.0, omegaStep=5, nzeros=500, t1=0.26, dt=0.010)


def _cmp(g, o, tol):
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(g[k], o[k]) < tol, (k, rel(g[k], o[k]))
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
        assert rel_l2(g[k], o[k]) < tol, (k, rel_l2(g[k], o[k]))
    assert rel(g["sw"], o["sw"]) < tol


@pytest.mark.parametrize("prec,mode,tol", [("fp64", "batched", 1e-10), ("fp64", "plain", 1e-10),
                                           ("fp32", "batched", 1e-4)])
def test_triclinic_pipeline_matches_oracle(prec, mode, tol):
    xyz, b9, times = triclinic_frames(128, 40, 1001)
    p = C.default_params(precision=prec, mode=mode, triclinic=1, **KW)
    announce("triclinic pipeline prec=%s mode=%s" % (prec, mode))
    g = C.run(p, xyz, b9, times)
    o = O.run(O.F64, O.config_from_params(p), xyz, b9, times)
    _cmp(g, o, tol)


def test_zero_shear_triclinic_equals_rectangular():
    xyz, box, times = frames(128, 40, 1001)
    b9 = np.zeros((box.shape[0], 3, 3), np.float32)
    for k in range(3):
        b9[:, k, k] = box[:, k]
    announce("triclinic zero shear vs rectangular")
    a = C.run(C.default_params(precision="fp64", **KW), xyz, box, times)
    b = C.run(C.default_params(precision="fp64", triclinic=1, **KW), xyz, b9, times)
    _cmp(b, a, 1e-12)
    # an unreduced or malformed box is refused
    bad = b9.copy()
    bad[:, 1, 0] = 0.8 * bad[:, 0, 0]
    with pytest.raises(C.CalcirError):
        C.run(C.default_params(precision="fp64", triclinic=1, **KW), xyz, bad, times)
    bad = b9.copy()
    bad[:, 0, 1] = 0.1
    with pytest.raises(C.CalcirError):
        C.run(C.default_params(precision="fp64", triclinic=1, **KW), xyz, bad, times)


def _layout(xyz, natom_mol):
    if natom_mol == 3:
        return np.ascontiguousarray(xyz.reshape(xyz.shape[0], -1, 4, 3)[:, :, :3, :].reshape(xyz.shape[0], -1, 3))
    return xyz


@pytest.mark.parametrize("natom_mol,nchrom_mol,rules", [
    (3, 2, {}), (4, 1, {}), (3, 1, {}),  # the reference's own semantics
    (3, 2, dict(msite_bisector=1)), (4, 1, dict(oh_per_molecule=1)),
    (3, 1, dict(msite_bisector=1, oh_per_molecule=1))])
def test_three_site_and_one_oh_pipelines_match_oracle(natom_mol, nchrom_mol, rules):
    xyz, box, times = frames(128, 40, 1001)
    xyz = _layout(xyz, natom_mol)
    p = C.default_params(precision="fp64", natom_mol=natom_mol, nchrom_mol=nchrom_mol, **rules, **KW)
    announce("pipeline natom_mol=%d nchrom_mol=%d rules=%s" % (natom_mol, nchrom_mol, rules))
    g = C.run(p, xyz, box, times)
    o = O.run(O.F64, O.config_from_params(p), xyz, box, times)
    _cmp(g, o, 1e-10)
    assert np.allclose(g["pw"], o["pw"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("natom_mol,nchrom_mol,extra", [(3, 2, ""), (4, 1, ""), (3, 1, "msite_bisector 1\n"),
                                                        (4, 1, "oh_per_molecule 1\n")])
def test_cli_layouts_match_the_oracle(tmp_path, natom_mol, nchrom_mol, extra):
    # the drop-in executable on an XTC file of three-site or four-site water,
    # with the reference's natom_mol / nchrom_mol semantics (or a rule key),
    # against the oracle's 13 files written from the same decoded frames
    xyz, box, times = frames(64, 30, 17)
    xyz = _layout(xyz, natom_mol)
    traj = tmp_path / "w.xtc"
    C.xtc_write(traj, xyz, box, times)
    xq, bq, tq, _ = C.xtc_read(traj)
    inp = tmp_path / "input.inp"
    inp.write_text("xtcf %s\noutf %s\nmodel tip4p2005\nspecies H2O\nntcfpoints 10\nnsamples 2\n"
                   "sampleEvery 0.1\navef 3400\nomegaStart 2800\nomegaStop 4000\nnzeros 502\nprecision fp64\n"
                   "natom_mol %d\nnchrom_mol %d\n%s" % (traj, tmp_path / "cli", natom_mol, nchrom_mol, extra))
    announce("calcir CLI natom_mol=%d nchrom_mol=%d %s" % (natom_mol, nchrom_mol, extra.strip()))
    r = subprocess.run([C.CLI_PATH, str(inp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    p = C.read_params(str(inp))
    p.outf = str(tmp_path / "orc").encode()
    o = O.run(O.F64, O.config_from_params(p), xq, bq, tq)
    C.write_outputs(p, o)
    for s in ("irrtcf", "iritcf", "vvrtcf", "vhrtcf", "spdn", "Pw", "Rw", "Rmw", "irls", "vvls", "vhls"):
        a = np.loadtxt(tmp_path / ("cli_%s.dat" % s))
        b = np.loadtxt(tmp_path / ("orc_%s.dat" % s))
        assert a.shape == b.shape, s
        y, z = a[:, 1], b[:, 1]
        assert np.array_equal(np.isnan(y), np.isnan(z)), s
        m = ~np.isnan(z)
        scale = np.abs(z[m]).max() if m.any() else 1.0
        np.testing.assert_allclose(y[m], z[m], rtol=2e-5, atol=2e-6 * scale, err_msg=s)


def test_cli_reads_a_triclinic_xtc(tmp_path):
    xyz, b9, times = triclinic_frames(64, 30, 7)
    traj = tmp_path / "tri.xtc"
    C.xtc_write(traj, xyz, b9, times)
    xq, bq, tq, _ = C.xtc_read(traj, full_box=True)
    assert np.array_equal(bq, b9)
    inp = tmp_path / "input.inp"
    inp.write_text("xtcf %s\noutf %s\nmodel tip4p2005\nspecies H2O\nntcfpoints 10\nnsamples 2\n"
                   "sampleEvery 0.1\navef 3400\nomegaStart 2800\nomegaStop 4000\nnzeros 502\nprecision fp64\n"
                   % (traj, tmp_path / "cli"))
    announce("calcir CLI on a triclinic xtc")
    r = subprocess.run([C.CLI_PATH, str(inp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "Triclinic box" in r.stdout
    p = C.read_params(str(inp))
    p.triclinic = 1
    p.outf = str(tmp_path / "mem").encode()
    C.write_outputs(p, C.run(p, xq, bq, tq))
    for suffix in ("irls", "vvls", "vhls", "irrtcf", "Pw", "spdn"):
        a = np.loadtxt(tmp_path / ("cli_%s.dat" % suffix))
        b = np.loadtxt(tmp_path / ("mem_%s.dat" % suffix))
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(b).max(), err_msg=suffix)


def test_adjoint_in_a_triclinic_box():
    xyz, b9, times = triclinic_frames(64, 40, 2024)
    p = C.default_params(precision="fp64", triclinic=1, **KW)
    coef = C.map_coefs(0, 0)
    idx, _, scale = lineshape_map(p)
    announce("adjoint triclinic fp64")
    sess = LineshapeSession(p, xyz, b9, times)
    ft = sess.forward(coef)
    o = O.run(O.F64, O.config_from_params(p), xyz, b9, times)
    assert np.abs(ft - o["ftcf_ir"]).max() <= 1e-10 * np.abs(o["ftcf_ir"]).max()
    target = ft[idx] * scale * 1.05
    r = ft[idx] * scale - target
    dL = np.zeros_like(ft)
    np.add.at(dL, idx, 2 * r * scale)
    g = sess.backward(dL)
    sess.close()
    for name in ("w0", "w1", "mu1", "x1", "wi1"):
        c = C.COEF_NAMES.index(name)
        d = 1e-4 * abs(coef[c])
        vals = []
        for sgn in (1.0, -1.0):
            cf = coef.copy()
            cf[c] += sgn * d
            ftp = O.run(O.F64, O.config_from_params(p, coef=cf), xyz, b9, times)["ftcf_ir"]
            rr = ftp[idx] * scale - target
            vals.append(float(rr @ rr))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g[c] - fd) <= 1e-4 * abs(fd), (name, g[c], fd)
