"""Shared test helpers: oracle construction and field comparison."""
import numpy as np

import oracle


def build_oracle(sc, twin=False):
    """The scene on the CPU oracle, in the reference's init order."""
    o = oracle.Oracle(sc.settings, twin=twin)
    o.set_rect_limits(*sc.container)
    for v in sc.volumes:
        o.add_volume(v)
    for b in sc.bodies:
        o.add_boundary(b.samples, b.can_move, b.mass, b.R, b.com)
    return o


def max_rel(a, b):
    """max |a - b| / max |b|: the field-wise max-relative error."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())
