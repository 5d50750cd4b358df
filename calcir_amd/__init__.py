"""calcir_amd -- MI355X (gfx950) spectroscopy hot path of calcIR.

Python mirror of calcIR's driver (calcIR.cu:9-919) over the C ABI of
``libcalcir.so`` (``include/calcir.h``): the input-file keys and defaults,
model/species strings, the sample x frame pipeline and the 13 output files.
Device arrays are torch tensors on the HIP device; per-stage calls run on
torch's current stream.

The HIP library is required: importing works without it, but every call
that needs it raises ``CalcirError`` naming the missing file (there is no CPU
fallback in this package).
"""

import ctypes
import os

import numpy as np

__all__ = [
    "CalcirError", "Params", "Results", "lib", "default_params", "read_params", "write_outputs",
    "synth_box_nm", "synth_trajectory", "run", "run_device", "nomega", "resolve_model", "resolve_species",
    "frame_index", "PREC_FP32", "PREC_FP64", "PREC_BF16X3", "MODE_PLAIN", "MODE_BATCHED", "NSITE",
    "SITE_E", "SITE_W", "SITE_X", "SITE_P", "SITE_MUP", "SITE_U", "SITE_R", "SITE_MU", "SITE_A",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcalcir.so")
CLI_PATH = os.path.join(HERE, "bin", "calcir")

PREC_FP32, PREC_FP64, PREC_BF16X3 = 0, 1, 2
PHASES = ("sites", "hamiltonian", "tridiag", "bisect", "invit", "backtransform", "ipr", "sw", "propagate",
          "correlate", "finalize", "bulge", "bulge_backtransform")
MODE_PLAIN, MODE_BATCHED = 0, 1
NSITE = 20
SITE_E, SITE_W, SITE_X, SITE_P, SITE_MUP, SITE_U, SITE_R, SITE_MU, SITE_A = 0, 1, 2, 3, 4, 5, 8, 11, 14
STRLEN = 256
INCOMPLETE = 12  # CIR_INCOMPLETE: stopped at max_frames after writing a checkpoint


class CalcirError(RuntimeError):
    """A libcalcir call returned a non-zero status (or the library is missing)."""

    def __init__(self, status, msg):
        super().__init__("%s (status %d)" % (msg, status))
        self.status = status


class Params(ctypes.Structure):
    """cir_params: the input keys of calcIR.cu:1556-1638 plus extensions."""

    _fields_ = [
        ("xtcf", ctypes.c_char * STRLEN), ("outf", ctypes.c_char * STRLEN),
        ("cptf", ctypes.c_char * STRLEN), ("model", ctypes.c_char * STRLEN),
        ("species", ctypes.c_char * STRLEN),
        ("ntcfpoints", ctypes.c_int), ("nsamples", ctypes.c_int), ("sampleEvery", ctypes.c_float),
        ("omegaStep", ctypes.c_int), ("natom_mol", ctypes.c_int), ("nchrom_mol", ctypes.c_int),
        ("nzeros", ctypes.c_int), ("imap", ctypes.c_int),
        ("dt", ctypes.c_float), ("t1", ctypes.c_float), ("avef", ctypes.c_float),
        ("beginTime", ctypes.c_float), ("omegaStart", ctypes.c_float), ("omegaStop", ctypes.c_float),
        ("cutoff", ctypes.c_float), ("precision", ctypes.c_int), ("mode", ctypes.c_int),
        ("chunk", ctypes.c_int), ("ckpt_every", ctypes.c_int), ("resume", ctypes.c_int),
        ("max_frames", ctypes.c_int), ("adj_store", ctypes.c_int), ("triclinic", ctypes.c_int),
        ("reunit_every", ctypes.c_int), ("origin_stride", ctypes.c_int),
        ("oh_per_molecule", ctypes.c_int), ("msite_bisector", ctypes.c_int), ("time_prec", ctypes.c_float),
        ("deterministic", ctypes.c_int),
    ]


_D = ctypes.POINTER(ctypes.c_double)


class Results(ctypes.Structure):
    """cir_results."""

    _fields_ = [
        ("tcf_ir", _D), ("tcf_vv", _D), ("tcf_vh", _D),
        ("ftcf_ir", _D), ("ftcf_vv", _D), ("ftcf_vh", _D),
        ("sw", _D), ("pw", _D), ("rw", _D), ("rmw", _D),
        ("nsteps", ctypes.c_int64), ("nframes_used", ctypes.c_int64),
        ("hist_skipped", ctypes.c_int64), ("seconds", ctypes.c_double),
    ]


_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_ll = ctypes.c_longlong

# name -> (restype, argtypes) for every symbol of include/calcir.h
SIGNATURES = {
    "cir_strerror": (ctypes.c_char_p, [_i]),
    "cir_last_error": (ctypes.c_char_p, []),
    "cir_params_default": (None, [ctypes.POINTER(Params)]),
    "cir_params_read": (_i, [ctypes.c_char_p, ctypes.POINTER(Params), _i]),
    "cir_params_echo": (None, [ctypes.POINTER(Params)]),
    "cir_resolve_model": (_i, [ctypes.c_char_p, ctypes.POINTER(_i)]),
    "cir_resolve_species": (_i, [ctypes.c_char_p, ctypes.POINTER(_i)]),
    "cir_nomega": (_i, [ctypes.POINTER(Params)]),
    "cir_minimage": (_f, [_f, _f]),
    "cir_frame_index": (_i, [ctypes.POINTER(Params), _i, _i, _f, ctypes.POINTER(_f)]),
    "cir_write_outputs": (_i, [ctypes.POINTER(Params), ctypes.POINTER(Results)]),
    "cir_synth_box_nm": (_f, [_i]),
    "cir_synth_trajectory": (_i, [_i, _i, ctypes.c_uint64, _f, _vp, _vp]),
    "cir_xtc_scan": (_i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "cir_xtc_read": (_i, [ctypes.c_char_p, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "cir_xtc_write": (_i, [ctypes.c_char_p, _i, _i, _vp, _vp, _vp, _f, _i]),
    "cir_xtc_read_full": (_i, [ctypes.c_char_p, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "cir_xtc_times": (_i, [ctypes.c_char_p, _i, _vp, _vp]),
    "cir_xtc_open": (_i, [ctypes.c_char_p, _i, ctypes.POINTER(_vp)]),
    "cir_xtc_next": (_i, [_vp, _i, _vp, _vp, _vp]),
    "cir_xtc_close": (None, [_vp]),
    "cir_xtc_offsets": (_i, [ctypes.c_char_p, _i, _vp]),
    "cir_xtc_seek": (_i, [_vp, _i]),
    "cir_xtc_precision": (_i, [ctypes.c_char_p, _i, ctypes.POINTER(_f)]),
    "cir_xtc_write_full": (_i, [ctypes.c_char_p, _i, _i, _vp, _vp, _vp, _f, _i]),
    "cir_main": (_i, [_i, ctypes.POINTER(ctypes.c_char_p)]),
    "cir_device_count": (_i, [ctypes.POINTER(_i)]),
    "cir_create": (_i, [ctypes.POINTER(Params), _i, _vp, ctypes.POINTER(_vp)]),
    "cir_run": (_i, [_vp, _i, _vp, _vp, _vp]),
    "cir_run_device": (_i, [_vp, _i, _vp, _vp, _vp]),
    "cir_stream_begin": (_i, [_vp, _i, _vp]),
    "cir_stream_push": (_i, [_vp, _i, _i, _vp, _vp]),
    "cir_stream_end": (_i, [_vp]),
    "cir_get_results": (_i, [_vp, ctypes.POINTER(Results)]),
    "cir_run_many": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _i]),
    "cir_reunit_workspace": (ctypes.c_size_t, [_i]),
    "cir_reunitarise": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "cir_set_profiling": (_i, [_vp, _i]),
    "cir_phase_times": (_i, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_ll), _i]),
    "cir_tuning_report": (_i, [ctypes.c_char_p, _i]),
    "cir_destroy": (None, [_vp]),
    "cir_map_coefs": (_i, [_i, _i, _vp]),
    "cir_get_coefs": (_i, [_vp, _vp]),
    "cir_set_coefs": (_i, [_vp, _vp]),
    "cir_forward_ls": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "cir_backward_ls": (_i, [_vp, _vp, _vp]),
    "cir_eproj_sites": (_i, [_i, _i, _i, _i, _i, _i, _i, _i, _f, _vp, _vp, _vp, _vp]),
    "cir_hamiltonian": (_i, [_i, _i, _i, _i, _i, _i, _i, _f, _vp, _vp, _vp, _vp]),
    "cir_eigh_workspace": (ctypes.c_size_t, [_i, _i, _i, _i]),
    "cir_eigh_batched": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "cir_ipr_hist": (_i, [_i, _i, _i, _i, _i, _f, _f, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cir_spectral_density": (_i, [_i, _i, _i, _i, _f, _f, _f, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cir_propagate": (_i, [_i, _i, _i, _i, _i, _f, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cir_propagate_workspace": (ctypes.c_size_t, [_i, _i, _i]),
    "cir_propagate_noalloc": (_i, [_i, _i, _i, _i, _i, _f, _vp, _vp, _vp, _vp, _vp]),
    "cir_propagate_noalloc_max_bytes": (ctypes.c_size_t, []),
    "cir_correlate": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cir_spectrum_c2r": (_i, [_i, _i, _i, _vp, _vp, _vp]),
    "cir_gemm": (_i, [_i, _i, _i, _i, _i, _vp, _ll, _ll, _ll, _vp, _ll, _ll, _ll, _vp, _ll, _ll, _ll,
                      ctypes.c_double, ctypes.c_double, _vp]),
}


def lib():
    """The loaded libcalcir.so; raises CalcirError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CalcirError(-1, "libcalcir.so not found at %s: build it with "
                                  "`python -m calcir_amd.build` (no CPU fallback exists)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        L = lib()
        raise CalcirError(status, "%s: %s" % (L.cir_strerror(status).decode(), L.cir_last_error().decode()))


def _setstr(p, field, value):
    b = value.encode() if isinstance(value, str) else value
    if len(b) >= STRLEN:
        raise ValueError("%s longer than %d bytes" % (field, STRLEN - 1))
    setattr(p, field, b)


def default_params(**kw):
    """Defaults of calcIR.cu:36-59, then keyword overrides (input-file key names;
    ``map`` may be given for ``imap``, ``checkpoint_every`` for ``ckpt_every``;
    precision/mode accept their strings)."""
    p = Params()
    lib().cir_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k == "map":
            k = "imap"
        if k == "checkpoint_every":
            k = "ckpt_every"
        if k == "adjoint_store":
            k = "adj_store"
        if k == "reunitarise_every":
            k = "reunit_every"
        if k == "origin_stride_frames":
            k = "origin_stride"
        if k == "time_precision":
            k = "time_prec"
        if k == "precision" and isinstance(v, str):
            v = {"fp32": PREC_FP32, "fp64": PREC_FP64, "bf16x3": PREC_BF16X3}[v]
        if k == "mode" and isinstance(v, str):
            v = {"plain": MODE_PLAIN, "batched": MODE_BATCHED}[v]
        if k in ("xtcf", "outf", "cptf", "model", "species"):
            _setstr(p, k, v)
        elif any(k == f[0] for f in Params._fields_):
            setattr(p, k, v)
        else:
            raise KeyError("unknown parameter %r" % k)
    return p


def read_params(path, params=None, verbose=False):
    """ir_init (calcIR.cu:1539-1643) on an input file, starting from params
    (defaults if None)."""
    p = params if params is not None else default_params()
    check(lib().cir_params_read(path.encode(), ctypes.byref(p), 1 if verbose else 0))
    return p


def nomega(p):
    return lib().cir_nomega(ctypes.byref(p))


def resolve_model(model):
    v = ctypes.c_int()
    check(lib().cir_resolve_model(model.encode(), ctypes.byref(v)))
    return v.value


def resolve_species(species):
    v = ctypes.c_int()
    check(lib().cir_resolve_species(species.encode(), ctypes.byref(v)))
    return v.value


def frame_index(p, sample, point, frame_dt):
    t = ctypes.c_float()
    f = lib().cir_frame_index(ctypes.byref(p), sample, point, frame_dt, ctypes.byref(t))
    return f, t.value


def synth_box_nm(nwater):
    return lib().cir_synth_box_nm(nwater)


def synth_trajectory(nwater, nframes, seed, box_nm=None):
    """Seeded synthetic rigid-water trajectory: (xyz [nframes, 4*nwater, 3]
    float32 nm, box [nframes, 3] float32, times [nframes] float32 ps)."""
    L = synth_box_nm(nwater) if box_nm is None else float(box_nm)
    xyz = np.zeros((nframes, 4 * nwater, 3), np.float32)
    times = np.zeros(nframes, np.float32)
    check(lib().cir_synth_trajectory(nwater, nframes, seed, L, xyz.ctypes.data, times.ctypes.data))
    box = np.full((nframes, 3), L, np.float32)
    return xyz, box, times


def xtc_scan(path):
    """(natoms, nframes) of a GROMACS XTC file."""
    na, nf = ctypes.c_int(), ctypes.c_int()
    check(lib().cir_xtc_scan(os.fsencode(path), ctypes.byref(na), ctypes.byref(nf)))
    return na.value, nf.value


def xtc_read(path, first=0, count=None, full_box=False):
    """Frames of an XTC file: (xyz [count, natoms, 3] nm, box diagonal
    [count, 3] -- or with full_box the matrices [count, 3, 3], rows a, b, c --
    times [count] ps, steps [count])."""
    natoms, nframes = xtc_scan(path)
    count = nframes - first if count is None else count
    xyz = np.zeros((count, natoms, 3), np.float32)
    box = np.zeros((count, 3, 3) if full_box else (count, 3), np.float32)
    times = np.zeros(count, np.float32)
    steps = np.zeros(count, np.int32)
    fn = lib().cir_xtc_read_full if full_box else lib().cir_xtc_read
    check(fn(os.fsencode(path), first, count, natoms, xyz.ctypes.data, box.ctypes.data, times.ctypes.data,
             steps.ctypes.data))
    return xyz, box, times, steps


def xtc_times(path):
    """(times [nframes] ps, box matrices [nframes, 3, 3]) of an XTC file from
    the frame headers alone (cir_xtc_times skips the coordinate blocks)."""
    _, nframes = xtc_scan(path)
    times = np.zeros(nframes, np.float32)
    box = np.zeros((nframes, 3, 3), np.float32)
    check(lib().cir_xtc_times(os.fsencode(path), nframes, times.ctypes.data, box.ctypes.data))
    return times, box


def xtc_precision(path, frame=1):
    """Coordinate precision stored in a frame of an XTC file (-1 for frames of
    at most 9 atoms, which carry none)."""
    v = ctypes.c_float()
    check(lib().cir_xtc_precision(os.fsencode(path), int(frame), ctypes.byref(v)))
    return v.value


def xtc_offsets(path):
    """Byte offset of every frame of an XTC file (cir_xtc_offsets)."""
    _, nframes = xtc_scan(path)
    off = np.zeros(nframes, np.int64)
    check(lib().cir_xtc_offsets(os.fsencode(path), nframes, off.ctypes.data))
    return off


def xtc_blocks(path, block, first=0):
    """Consecutive frames of an XTC file from frame `first` (cir_xtc_open,
    cir_xtc_seek, cir_xtc_next), yielded as (xyz [k, natoms, 3], box
    matrices [k, 3, 3], times [k]) with k <= block."""
    natoms, nframes = xtc_scan(path)
    L = lib()
    rd = ctypes.c_void_p()
    check(L.cir_xtc_open(os.fsencode(path), natoms, ctypes.byref(rd)))
    try:
        if first:
            check(L.cir_xtc_seek(rd, int(first)))
        for f0 in range(int(first), nframes, block):
            k = min(block, nframes - f0)
            xyz = np.zeros((k, natoms, 3), np.float32)
            box = np.zeros((k, 3, 3), np.float32)
            times = np.zeros(k, np.float32)
            check(L.cir_xtc_next(rd, k, xyz.ctypes.data, box.ctypes.data, times.ctypes.data))
            yield xyz, box, times
    finally:
        L.cir_xtc_close(rd)


def xtc_write(path, xyz, box, times, precision=1000.0, append=False):
    """Writes frames (xyz [nframes, natoms, 3] nm, box diagonal [nframes, 3]
    or matrices [nframes, 3, 3], times [nframes] ps) as XTC with the given
    precision."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    box = np.ascontiguousarray(box, np.float32)
    times = np.ascontiguousarray(times, np.float32)
    nframes, natoms = xyz.shape[0], xyz.shape[1]
    fn = lib().cir_xtc_write_full if box.ndim == 3 else lib().cir_xtc_write
    check(fn(os.fsencode(path), natoms, nframes, xyz.ctypes.data, box.ctypes.data, times.ctypes.data,
             float(precision), int(bool(append))))


def _results_buffers(p):
    nt, nom, nfft = p.ntcfpoints, nomega(p), p.ntcfpoints + p.nzeros
    bufs = {k: np.zeros(2 * nt) for k in ("tcf_ir", "tcf_vv", "tcf_vh")}
    bufs.update({k: np.zeros(nfft) for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh")})
    bufs.update({k: np.zeros(nom) for k in ("sw", "pw", "rw", "rmw")})
    r = Results()
    for k, v in bufs.items():
        setattr(r, k, v.ctypes.data_as(_D))
    return r, bufs


NCOEF = 13
COEF_NAMES = ("w0", "w1", "w2", "x0", "x1", "p0", "p1", "mu0", "mu1", "mu2", "wi0", "wi1", "wi2")


def map_coefs(imap, ispecies):
    """The 13 map coefficients of a (map, species) row (calcIR.cu:1203-1242),
    in COEF_NAMES order."""
    c = np.zeros(NCOEF)
    check(lib().cir_map_coefs(int(imap), int(ispecies), c.ctypes.data))
    return c


def check_trajectory(p, xyz, box, times):
    """Host arrays of a trajectory as float32 contiguous numpy arrays, checked
    against what the C ABI reads through their pointers: xyz [nframes, natoms,
    3] with natoms a multiple of natom_mol; box [nframes, 3] (the diagonal),
    or with the triclinic key [nframes, 9] / [nframes, 3, 3]; times
    [nframes].  Raises ValueError otherwise."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    box = np.ascontiguousarray(box, np.float32)
    times = np.ascontiguousarray(times, np.float32)
    if xyz.ndim != 3 or xyz.shape[2] != 3 or xyz.shape[0] < 1 or xyz.shape[1] < 1:
        raise ValueError("xyz must be [nframes, natoms, 3], got %s" % (xyz.shape,))
    nframes, natoms = xyz.shape[0], xyz.shape[1]
    if p.natom_mol <= 0 or natoms % p.natom_mol:
        raise ValueError("natoms %d is not a multiple of natom_mol %d" % (natoms, p.natom_mol))
    width = 9 if p.triclinic else 3
    if box.shape[0] != nframes or box.reshape(nframes, -1).shape[1] != width:
        raise ValueError("box must hold %d floats per frame for %d frames (triclinic %d), got %s"
                         % (width, nframes, p.triclinic, box.shape))
    if times.ndim != 1 or times.size != nframes:
        raise ValueError("times must hold one value per frame (%d), got %s" % (nframes, times.shape))
    return xyz, box, times


def run(p, xyz, box, times, stream=None, profile=False, coefs=None):
    """The whole calcIR job (every sample, calcIR.cu:360-919) on one GPU.

    xyz: [nframes, natoms, 3] float32 nm (host numpy); box [nframes, 3];
    times [nframes] ps.  Returns a dict with complex TCFs (decayed and
    1/nsamples scaled), the C2R outputs, Sw, Pw, Rw, Rmw and counters.
    coefs (13 doubles, COEF_NAMES order) replaces the map table row."""
    xyz, box, times = check_trajectory(p, xyz, box, times)
    nframes, natoms = xyz.shape[0], xyz.shape[1]
    L = lib()
    ctx = ctypes.c_void_p()
    check(L.cir_create(ctypes.byref(p), natoms, stream, ctypes.byref(ctx)))
    phases = None
    try:
        if coefs is not None:
            cf = np.ascontiguousarray(coefs, np.float64)
            if cf.shape != (NCOEF,):
                raise ValueError("coefs must hold %d values" % NCOEF)
            check(L.cir_set_coefs(ctx, cf.ctypes.data))
        if profile:
            check(L.cir_set_profiling(ctx, 1))
        check(L.cir_run(ctx, nframes, xyz.ctypes.data, box.ctypes.data, times.ctypes.data))
        r, bufs, phases = _collect(L, ctx, p, profile)
    finally:
        L.cir_destroy(ctx)
    return _results_dict(r, bufs, phases)


def _collect(L, ctx, p, profile):
    r, bufs = _results_buffers(p)
    check(L.cir_get_results(ctx, ctypes.byref(r)))
    phases = None
    if profile:
        secs = (ctypes.c_double * len(PHASES))()
        cnts = (ctypes.c_longlong * len(PHASES))()
        check(L.cir_phase_times(ctx, secs, cnts, len(PHASES)))
        phases = {name: (secs[i], cnts[i]) for i, name in enumerate(PHASES)}
    return r, bufs, phases


def _results_dict(r, bufs, phases):
    out = {}
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        out[k] = bufs[k][0::2] + 1j * bufs[k][1::2]
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh", "sw", "pw", "rw", "rmw"):
        out[k] = bufs[k]
    out.update(nsteps=r.nsteps, nframes_used=r.nframes_used, hist_skipped=r.hist_skipped,
               seconds=r.seconds, phases=phases)
    return out


def run_streamed(p, times, blocks, natoms, profile=False, coefs=None):
    """cir_run fed block by block (cir_stream_begin / _push / _end): times
    [nframes] ps of the whole trajectory, blocks an iterable of (xyz, box)
    arrays of consecutive frames in trajectory order ([k, natoms, 3] nm and
    [k, 3], or [k, 9] with the triclinic key).  Only the frames the samples
    use are copied to the device; the trajectory is never resident.  Returns
    what run() returns."""
    times = np.ascontiguousarray(times, np.float32)
    if times.ndim != 1 or times.size < 2:
        raise ValueError("times must hold one value per trajectory frame")
    if int(natoms) <= 0 or p.natom_mol <= 0 or int(natoms) % p.natom_mol:
        raise ValueError("natoms must be a positive multiple of natom_mol")
    width = 9 if p.triclinic else 3
    L = lib()
    ctx = ctypes.c_void_p()
    check(L.cir_create(ctypes.byref(p), int(natoms), None, ctypes.byref(ctx)))
    try:
        if coefs is not None:
            cf = np.ascontiguousarray(coefs, np.float64)
            if cf.shape != (NCOEF,):
                raise ValueError("coefs must hold %d values" % NCOEF)
            check(L.cir_set_coefs(ctx, cf.ctypes.data))
        if profile:
            check(L.cir_set_profiling(ctx, 1))
        check(L.cir_stream_begin(ctx, times.size, times.ctypes.data))
        first = 0
        for xyz, box in blocks:
            xyz = np.ascontiguousarray(xyz, np.float32)
            box = np.ascontiguousarray(box, np.float32)
            k = xyz.shape[0]
            if xyz.shape != (k, int(natoms), 3) or box.shape[0] != k or box.reshape(k, -1).shape[1] != width:
                raise ValueError("block shapes %s / %s do not match natoms %d and %d box floats per frame"
                                 % (xyz.shape, box.shape, int(natoms), width))
            if first + k > times.size:
                raise ValueError("more frames pushed than times given")
            check(L.cir_stream_push(ctx, first, k, xyz.ctypes.data, box.ctypes.data))
            first += k
        check(L.cir_stream_end(ctx))
        r, bufs, phases = _collect(L, ctx, p, profile)
    finally:
        L.cir_destroy(ctx)
    return _results_dict(r, bufs, phases)


def tuning_report():
    """The launch configurations this process has chosen by timing
    (cir_tuning_report): {(family, bucket...): (choice, [ms per candidate])}."""
    L = lib()
    n = L.cir_tuning_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.cir_tuning_report(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        head, _, rest = line.partition(" -> ")
        fam, *shape = head.split()
        choice, _, times = rest.partition(" (")
        out[(fam,) + tuple(int(x) for x in shape)] = (int(choice), [float(x) for x in times.rstrip(" ms)").split()])
    return out


def run_many(jobs, nthreads=8):
    """Independent jobs at once on one GPU (cir_run_many): jobs is a list of
    (params, xyz, box, times) as for run(); returns their result dicts."""
    L = lib()
    n = len(jobs)
    ctxs = (ctypes.c_void_p * n)()
    keep = []
    try:
        for j, (p, xyz, box, times) in enumerate(jobs):
            xyz, box, times = check_trajectory(p, xyz, box, times)
            keep.append((p, xyz, box, times))
            c = ctypes.c_void_p()
            check(L.cir_create(ctypes.byref(p), xyz.shape[1], None, ctypes.byref(c)))
            ctxs[j] = c.value
        nfr = (ctypes.c_int * n)(*[k[1].shape[0] for k in keep])
        px = (ctypes.c_void_p * n)(*[k[1].ctypes.data for k in keep])
        pb = (ctypes.c_void_p * n)(*[k[2].ctypes.data for k in keep])
        pt = (ctypes.c_void_p * n)(*[k[3].ctypes.data for k in keep])
        check(L.cir_run_many(n, ctxs, nfr, px, pb, pt, int(nthreads)))
        outs = []
        for j in range(n):
            r, bufs = _results_buffers(keep[j][0])
            check(L.cir_get_results(ctypes.c_void_p(ctxs[j]), ctypes.byref(r)))
            out = {}
            for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
                out[k] = bufs[k][0::2] + 1j * bufs[k][1::2]
            for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh", "sw", "pw", "rw", "rmw"):
                out[k] = bufs[k]
            out.update(nsteps=r.nsteps, nframes_used=r.nframes_used, hist_skipped=r.hist_skipped, seconds=r.seconds)
            outs.append(out)
        return outs
    finally:
        for j in range(n):
            if ctxs[j]:
                L.cir_destroy(ctypes.c_void_p(ctxs[j]))


def run_device(p, xyz, box, times, profile=False):
    """run() with the trajectory already resident on the device: xyz, box are
    cuda float32 torch tensors [nframes, natoms, 3] / [nframes, 3]; times a
    host array.  Runs on torch's current stream."""
    import torch

    if not (xyz.is_cuda and box.is_cuda and xyz.dtype == torch.float32 and box.dtype == torch.float32):
        raise ValueError("xyz and box must be cuda float32 tensors")
    if xyz.dim() != 3 or xyz.shape[2] != 3 or tuple(box.shape) != (xyz.shape[0], 3):
        raise ValueError("xyz must be [nframes, natoms, 3] and box [nframes, 3]")
    xyz, box = xyz.contiguous(), box.contiguous()
    times = np.ascontiguousarray(times, np.float32)
    if times.size != xyz.shape[0]:
        raise ValueError("times must have one entry per frame")
    L = lib()
    ctx = ctypes.c_void_p()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.cir_create(ctypes.byref(p), xyz.shape[1], stream, ctypes.byref(ctx)))
    phases = None
    try:
        if profile:
            check(L.cir_set_profiling(ctx, 1))
        check(L.cir_run_device(ctx, xyz.shape[0], ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(box.data_ptr()),
                               times.ctypes.data))
        r, bufs = _results_buffers(p)
        check(L.cir_get_results(ctx, ctypes.byref(r)))
        if profile:
            secs = (ctypes.c_double * len(PHASES))()
            cnts = (ctypes.c_longlong * len(PHASES))()
            check(L.cir_phase_times(ctx, secs, cnts, len(PHASES)))
            phases = {name: (secs[i], cnts[i]) for i, name in enumerate(PHASES)}
    finally:
        L.cir_destroy(ctx)
    out = {}
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        out[k] = bufs[k][0::2] + 1j * bufs[k][1::2]
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh", "sw", "pw", "rw", "rmw"):
        out[k] = bufs[k]
    out.update(nsteps=r.nsteps, nframes_used=r.nframes_used, hist_skipped=r.hist_skipped,
               seconds=r.seconds, phases=phases)
    return out


def write_outputs(p, res):
    """The 13 output files of calcIR.cu:843-919 for a result dict of run()."""
    bufs = {}
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        z = np.asarray(res[k], np.complex128)
        a = np.empty(2 * z.size)
        a[0::2], a[1::2] = z.real, z.imag
        bufs[k] = a
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh", "sw", "pw", "rw", "rmw"):
        bufs[k] = np.ascontiguousarray(res[k], np.float64)
    r = Results()
    for k, v in bufs.items():
        setattr(r, k, v.ctypes.data_as(_D))
    check(lib().cir_write_outputs(ctypes.byref(p), ctypes.byref(r)))
