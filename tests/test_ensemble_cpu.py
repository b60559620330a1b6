"""Union basins of ensembles (host side, no GPU)."""
import numpy as np

from shyhip import ensemble, mesh


def _members():
    return [mesh.lagoon(n, seed=s, lx=lx, ly=lx / 2, nlv=4, tile=16)
            for n, s, lx in ((600, 1, 3000.0), (900, 2, 4000.0), (400, 3, 2500.0))]


def test_union_offsets_and_ranges():
    bs = _members()
    u = ensemble.union_basin(bs)
    assert u.basin.nkn == sum(b.nkn for b in bs) and u.basin.nel == sum(b.nel for b in bs)
    for m, b in enumerate(bs):
        k0, k1 = u.node_start[m], u.node_start[m + 1]
        e0, e1 = u.elem_start[m], u.elem_start[m + 1]
        assert (k1 - k0, e1 - e0) == (b.nkn, b.nel)
        # the member's elements use only its own nodes, in its own numbering
        assert np.array_equal(u.basin.nen3v[e0:e1] - k0, b.nen3v)
        assert np.array_equal(u.basin.xgv[k0:k1], b.xgv) and np.array_equal(u.basin.hm3v[e0:e1], b.hm3v)
        assert np.array_equal(u.basin.ilhkv[k0:k1], b.ilhkv)
        # boundary ids shifted by the boundaries of the members before it
        ib = u.basin.ibnd[k0:k1]
        assert np.array_equal(ib > 0, b.ibnd > 0)
        assert np.array_equal(ib[b.ibnd > 0], b.ibnd[b.ibnd > 0] + u.bnd_start[m])
        assert np.array_equal(u.basin.ibtyp[u.bnd_start[m] + 1:u.bnd_start[m + 1] + 1], b.ibtyp[1:])
        assert np.array_equal(u.basin.bnd_nodes[u.basin.ibnd[u.basin.bnd_nodes] > u.bnd_start[m]][:0], [])
    assert len(u.basin.ibtyp) == u.bnd_start[-1] + 1
    assert np.array_equal(np.sort(u.basin.bnd_nodes), np.nonzero(u.basin.ibnd)[0])


def test_union_rejects_mixed_vertical_grids():
    a = mesh.lagoon(400, seed=1, lx=2000.0, ly=1000.0, nlv=4, tile=8)
    b = mesh.lagoon(400, seed=2, lx=2000.0, ly=1000.0, nlv=6, tile=8)
    try:
        ensemble.union_basin([a, b])
    except ValueError:
        return
    raise AssertionError("mixed vertical grids accepted")
