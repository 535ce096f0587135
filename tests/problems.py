"""Test problems built with the oracle (the checker): geometry dicts in the
format of oracle.vhr.read_vhr, plus helpers to load them into the device
geometry."""
import os

import numpy as np

from oracle import vhr, topology

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def bar(L, M, N, dx=1e-6, sigma=5.8e7, freq=(1e9,), lam=None):
    """Fully filled bar; port P on the +x face of x=L, N on the -x face of x=1."""
    s = np.full((L, M, N), float(sigma))
    P = np.array([[L, j + 1, k + 1, 2] for k in range(N) for j in range(M)], dtype=np.int64)
    Nn = np.array([[1, j + 1, k + 1, 1] for k in range(N) for j in range(M)], dtype=np.int64)
    return dict(sigma=s, lambdaL=lam, freq=np.array(freq, float), dx=dx, ports_P=[P], ports_N=[Nn],
                well_cond=np.zeros((0, 4), np.int64))


def wire_vhr():
    return vhr.read_vhr(os.path.join(GOLDEN, 'wire_len2.0u_dia4.0u.vhr'))


def numex1():
    """numex1 straight bar 60x20x20 @0.5um (reduced from its .vhr by make_fixtures.py)."""
    d = np.load(os.path.join(GOLDEN, 'numex1_geometry.npz'))
    L, M, N = (int(v) for v in d['dims'])
    g = bar(L, M, N, dx=float(d['dx']), sigma=float(d['sigma']), freq=d['freq'])
    g['ports_P'] = [d['P']]
    g['ports_N'] = [d['N']]
    return g


def voxel_z(geo, sysd, f):
    zr, zi = topology.impedance(geo, sysd['idxS'], f)
    z = zr.ravel(order='F')[sysd['idxS']].astype(complex)
    if zi is not None:
        z = z + zi.ravel(order='F')[sysd['idxS']]
    return z


def device_geometry(ctx, geo, sysd, pad_mode=1, f=1e9, tensor='gpu'):
    import voxhenry_amd as vh
    G = vh.Geometry(ctx, sysd['dims'], geo['dx'], sysd['idxS'], pad_mode=pad_mode)
    if tensor == 'gpu':
        G.generate_green()
    G.set_incidence(sysd['Ar'])
    G.set_impedance(voxel_z(geo, sysd, f))
    return G


def numex3():
    """numex3 circular coil (circular_coil_loop150.0u_rad5.0u.vhr, 310x310x10
    @1um, 74 680 voxels, one port), reduced by make_fixtures.py; freq holds the
    49 frequencies of its committed R/L table."""
    d = np.load(os.path.join(GOLDEN, 'numex3_geometry.npz'))
    dims = tuple(int(v) for v in d['dims'])
    s = np.zeros(int(np.prod(dims)))
    s[d['idx']] = float(d['sigma'])
    rl = np.load(os.path.join(GOLDEN, 'numex_RL.npz'))
    return dict(sigma=s.reshape(dims, order='F'), lambdaL=None, freq=rl['freq3'], dx=float(d['dx']),
                ports_P=[d['P']], ports_N=[d['N']], well_cond=d['well'])


def structure(geo):
    """Oracle geometry dict (well_cond) -> product structure dict (well)."""
    out = dict(geo)
    out['well'] = out.pop('well_cond')
    return out


def oracle_geometry(st):
    """Product structure dict (well) -> oracle geometry dict (well_cond)."""
    out = dict(st)
    out['well_cond'] = out.pop('well')
    return out
