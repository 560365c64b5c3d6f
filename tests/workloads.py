"""Generated designs for the parity tests, chosen for the light they transmit.

Random draws of the benchmark distributions mostly land no ray in a bin (their
information is exactly 0), which makes a comparison of information trivial.
`transmitting` keeps the designs whose FP64 probe (cpd_hit_fraction over
4 wavelengths x 256 rays) lands at least `min_hit` of the rays in a bin;
configurations 5..7 are already drawn that way by the generator."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import compound_prism_designer as cpd


def transmitting(config: int, n: int, first: int = 0, min_hit: float = 0.1, limit: int = 20000):
    """The first n designs of a configuration from index `first` on whose hit
    fraction is >= min_hit: returns (params, designs, samplings, hit fractions)."""
    P, D, S, H = [], [], [], []
    i = first
    while len(D) < n:
        if i - first >= limit:
            raise RuntimeError(f"config {config}: fewer than {n} transmitting designs in {limit} draws")
        m = max(16, 4 * (n - len(D)))
        p, d, s, _ = cpd.generate_designs(config, i, m)
        with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as pool:  # ctypes releases the GIL
            hs = list(pool.map(cpd.hit_fraction, d))
        for pp, dd, ss, h in zip(p, d, s, hs):
            if h >= min_hit and len(D) < n:
                P.append(pp)
                D.append(dd)
                S.append(ss)
                H.append(h)
        i += m
    return P, D, S, np.array(H)
