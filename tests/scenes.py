"""Shared small scenes for the parity tests (seeded, built by the product's own
host voxelization, which is itself pinned to the reference's voxelization)."""
from __future__ import annotations

import functools

import numpy as np

import solargpu_amd as sg
import oracle_lib as orc

SITE = (15.65, 46.56, 1)


@functools.lru_cache(maxsize=None)
def sky(site=SITE, seed=20260101):
    az_alt = sg.sun_positions(*site)
    dni, dhi = sg.synth_weather(seed, az_alt)
    return az_alt, dni, dhi, orc.sky_bins(az_alt)


@functools.lru_cache(maxsize=None)
def synth_scene(seed, w, h, cell=1.0, density=0.35, trees=4):
    heights = sg.synth_dsm(seed, w, h, cell, density, trees)
    return sg.assemble_scene(heights, cell)


def scene_from_heights(heights, cell=1.0):
    return sg.assemble_scene(np.asarray(heights, np.float32), cell)


@functools.lru_cache(maxsize=None)
def oracle_lists(seed, w, h, cell=1.0, density=0.35, trees=4):
    sc = synth_scene(seed, w, h, cell, density, trees)
    return orc.viewshed(sc, sky()[3])
