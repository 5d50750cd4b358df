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
