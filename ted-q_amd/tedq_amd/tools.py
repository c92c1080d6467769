"""Bit helpers used by ``probs(..., after_state=True)`` (tedq/tools/helpers.py:22-52)."""

import math

import torch


def dec_to_bin(x, size):
    """Binary digits of ``x`` as a list of ``log2(size)`` ints, most significant first."""
    width = int(math.log2(size))
    return [int(c) for c in bin(x)[2:].zfill(width)]


def rescale_state(state):
    """List of the amplitudes of ``state`` divided by its 2-norm."""
    scale = torch.sqrt(torch.sum(torch.abs(state) ** 2))
    return [s / scale for s in state]
