"""Host logic of the drop-ins' resident-graph cache: graphs are found by
content, so an in-place edge insertion (the reference's updetaGraph flow
changes ``indices`` and ``degree`` of the same arrays) gives a new key, and
the same CSR in new arrays finds the resident graph."""
import numpy as np

from flsg import dropin
from flsg.graph import powerlaw_graph_small


def test_fingerprint_follows_content():
    g = powerlaw_graph_small()
    ip, ind, deg = g.indptr.copy(), g.indices.copy(), g.degree.copy()
    k0 = dropin.fingerprint(ip, ind, deg)
    assert dropin.fingerprint(ip.copy(), ind.copy(), deg.copy()) == k0
    deg[5] += 1                       # an insertion raises a degree in place
    assert dropin.fingerprint(ip, ind, deg) != k0
    deg[5] -= 1
    assert dropin.fingerprint(ip, ind, deg) == k0
    ind[3] = (ind[3] + 1) % g.n       # a rewired entry with the same degrees
    assert dropin.fingerprint(ip, ind, deg) != k0
    # dtype and shape are part of the key
    assert dropin.fingerprint(ip.astype(np.int32), g.indices, g.degree) != k0
    # non-contiguous views hash their content
    a = np.arange(20, dtype=np.int64)
    assert dropin.fingerprint(a[::2]) == dropin.fingerprint(np.ascontiguousarray(a[::2]))
