import numpy as np


def rel_fro_lower(x, ref):
    """||tril(x - ref)||_F / ||tril(ref)||_F over the lower triangle incl. diagonal."""
    d = np.tril(x - ref)
    return float(np.linalg.norm(d) / np.linalg.norm(np.tril(ref)))


def rel_fro(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))
