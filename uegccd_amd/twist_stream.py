"""Twist-averaging / r_s-scan request streams for the solve server.

The reference averages over twists one after another (source/Main.f90:112-168,
twists drawn at random in units of 2 pi/L); here a stream is a seeded list of
independent requests with Poisson arrival times, submitted to the server as
they arrive.
"""
from __future__ import annotations

import numpy as np

from .inputfile import UEGInfo, parse_input, standard_input

ELECTRONS = (14, 38, 54)
RS_VALUES = (0.5, 1.0, 2.0, 5.0)
CUTOFFS = (9, 16, 25)


def stream_requests(n=256, seed=1234, rate=None, electrons=ELECTRONS, rs_values=RS_VALUES,
                    cutoffs=CUTOFFS):
    """n requests (nel, cutoff, r_s, twist): per request a twist uniform in
    [-0.5, 0.5)^3, then N, r_s and the basis cutoff (MaxKPoint) drawn from one
    generator seeded with `seed`; arrival times in seconds are Poisson at
    `rate` per second (all zero when rate is None)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        tw = tuple(float(x) for x in rng.uniform(-0.5, 0.5, 3))
        nel = int(rng.choice(electrons))
        rs = float(rng.choice(rs_values))
        ec = int(rng.choice(cutoffs))
        reqs.append((nel, ec, rs, tw))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n)) if rate else np.zeros(n)
    return reqs, arrivals


def request_info(req, cutoff_twisted=False) -> UEGInfo:
    """The Input of one request (CCD only, the reference's default keys).  The
    basis cutoff is the reference's |n|^2 <= E_c; with cutoff_twisted the
    basis is the twisted sphere |n + twist|^2 <= E_c instead (an opt-in of
    the library's basis builders)."""
    nel, ec, rs, tw = req
    info = parse_input(standard_input(nel, ec, rs, tw))
    info.CutoffTwisted = bool(cutoff_twisted)
    return info
