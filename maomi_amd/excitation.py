"""Photon order and axial envelope (call surface of maomi/excitation.py:5-49)."""

import numpy as np

ORDERS = {"one-photon": 1, "two-photon": 2, "three-photon": 3, "temporal-focusing": 2}


def excitation_order(config):
    """Photon order of config.scaling (excitation.py:8-11)."""
    if config.scaling not in ORDERS:
        raise ValueError(f"Unknown excitation scaling: {config.scaling}")
    return ORDERS[config.scaling]


def temporal_profile(size, length_um, dz_um, *, legacy_offset=0.0):
    """1/sqrt(1 + (i/z_R)^2) with z_R = L/(dz 2 sqrt 3), i = ceil(size/2) - (1..size) + offset dz.

    The offset keeps the legacy applyTemporalFocusing convention (excitation.py:14-29).
    """
    if (int(size) != size or size < 1
            or not np.isfinite([length_um, dz_um, legacy_offset]).all()
            or min(length_um, dz_um) <= 0):
        raise ValueError("Positive finite axial length and sampling required")
    zr = length_um / (dz_um * 2 * np.sqrt(3))
    idx = np.ceil(size / 2) - np.arange(1, size + 1) + legacy_offset * dz_um
    return 1 / np.sqrt(1 + (idx / zr) ** 2)


def axial_profile(config, size, part="focus"):
    """Per-plane weights: ones unless temporal focusing (excitation.py:32-49)."""
    if config.scaling != "temporal-focusing":
        return np.ones(size, np.float32)
    if part == "focus":
        dz, offset = 1 / config.vres, 0.0
    else:
        pp = config.prop_um * config.vres
        depth = config.psf_um[2] * config.vres
        dz = pp / config.vres
        offset = ((depth * config.vres + 2 * pp) / (pp * pp) if part == "top"
                  else -depth * config.vres / (pp * pp))
    return temporal_profile(size, config.temporal_length_um, dz,
                            legacy_offset=offset).astype(np.float32)
