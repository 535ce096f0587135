"""The library's own topology (panel ids, Ar, port rhs) equals the oracle's
restatement of lse_generate_panelIDs / lse_compute_Ae_matrix / executer :165-192."""
import numpy as np
import pytest

from oracle import topology
from tests import problems
from voxhenry_amd import structures

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('name', ['wire', 'hollow', 'bus', 'spiral', 'loop', 'strip'])
def test_topology_matches_oracle(ctx, name):
    import voxhenry_amd as vh
    from voxhenry_amd import solver
    print('topology', name, flush=True)
    if name == 'wire':
        geo = problems.wire_vhr()
        geo['well'] = geo['well_cond']
    elif name == 'hollow':
        geo = structures.bar(9, 6, 5)
        geo['sigma'][3:6, 2:4, 1:4] = 0.0
    elif name == 'bus':
        geo = structures.bus(nbars=3, length=20)
    elif name == 'spiral':
        geo = structures.spiral(turns=2, outer=30, width=3, height=2, spacing=2)
    elif name == 'loop':
        geo = structures.loop(8, section=3)
    else:
        geo = structures.strip(12, 4)
    s = topology.build_system(dict(geo, well_cond=geo['well']))
    idx = solver.nonempty(geo)
    assert np.array_equal(idx, s['idxS'])
    G = vh.Geometry(ctx, geo['sigma'].shape, geo['dx'], idx)
    rhs = G.build_topology(geo['ports_P'], geo['ports_N'], geo['well'])
    assert G.num_node == s['num_node']
    assert np.array_equal(rhs.T, s['rhs'])
    # Ar through the matvec's node rows: y_node = Ar J
    G.generate_green()
    G.set_impedance(solver.impedance(geo, idx, 1e9))
    rng = np.random.default_rng(0)
    x = np.zeros(G.n, complex)
    x[:G.num_curr] = rng.standard_normal(G.num_curr)
    y = G.matvec(x, 1j)
    assert np.allclose(y[G.num_curr:], s['Ar'] @ x[:G.num_curr], rtol=0, atol=1e-12)
