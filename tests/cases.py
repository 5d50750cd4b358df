"""Shared small parity cases (seeded synthetic frames)."""
import functools

import numpy as np

import calcir_amd as C


@functools.lru_cache(maxsize=None)
def frames(nwater, nframes, seed):
    xyz, box, times = C.synth_trajectory(nwater, nframes, seed)
    xyz.setflags(write=False)
    box.setflags(write=False)
    times.setflags(write=False)
    return xyz, box, times


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@functools.lru_cache(maxsize=None)
def triclinic_frames(nwater, nframes, seed, shear=(0.15, 0.1, 0.12)):
    """Seeded rigid-water frames in a reduced triclinic box: the cubic
    synthetic frames with every molecule moved rigidly so that its oxygen's
    fractional coordinates are kept in the box a = (L,0,0), b = (s0 L,L,0),
    c = (s1 L,s2 L,L).  Returns (xyz, box [nframes, 3, 3], times)."""
    xyz, box, times = frames(nwater, nframes, seed)
    L = float(box[0, 0])
    M = np.array([[L, 0, 0], [shear[0] * L, L, 0], [shear[1] * L, shear[2] * L, L]], np.float64)
    x = xyz.astype(np.float64).reshape(nframes, nwater, 4, 3)
    frac = x[:, :, 0, :] / L
    shift = frac @ M - x[:, :, 0, :]
    out = (x + shift[:, :, None, :]).reshape(nframes, 4 * nwater, 3).astype(np.float32)
    b9 = np.repeat(M[None].astype(np.float32), nframes, axis=0)
    out.setflags(write=False)
    b9.setflags(write=False)
    return out, b9, times
