"""Index sets, panel IDs, incidence matrix Ae, port nodes and right-hand sides.

Restated from:
* src_pre_process/pre_define_structure_params.m:2-16   (idxS, idxS5)
* src_lse_formation/lse_generate_panelIDs.m:1-176      (panel numbering)
* src_lse_formation/lse_compute_Ae_matrix.m:21-88      (Ae, 16 nnz per voxel)
* src_lse_formation/lse_assign_exc_grnd_nodes.m:1-94   (port nodes)
* VoxHenry_executer.m:159-193                          (rhs, Ar)

All indices here are 0-based; the numbering ORDER is the reference's.
Pinned by: exact structural checks (Ae column sums / node degrees, the numex1
golden solution's node layout) and the closed-form DC resistance of a bar,
tests/test_oracle_topology.py and tests/test_oracle_solve.py.  The two-fluid
impedance (VoxHenry_executer.m:243-254) is pinned by the parallel-channel
identity z = 1/(sigma + 1/(j w mu lam^2)) and its sigma = 0 limit
j w mu lam^2 (tests/test_oracle_matvec.py::test_two_fluid_impedance_identity).
"""
import numpy as np
import scipy.sparse as sp


def structure_params(sigma, lambdaL=None):
    """idxS: 0-based column-major linear indices of non-empty voxels
    (pre_define_structure_params.m:5-13)."""
    s = np.abs(sigma.ravel(order='F')) > 1e-12
    if lambdaL is not None:
        s = s | (np.abs(lambdaL.ravel(order='F')) > 1e-12)
    idxS = np.flatnonzero(s)
    nD = sigma.size
    idxS5 = np.concatenate([idxS + c * nD for c in range(5)])   # :16
    return idxS, idxS5


def _panel_axis(idxS, dims, axis):
    """One axis of lse_generate_panelIDs.m: voxels visited with ``axis`` the
    fastest loop index (x: nn,mm,ll; y: ll,nn,mm; z: mm,ll,nn -- :35-37, :78-80,
    :122-124).  A voxel gets (left, right) = (ind, ind+1); ind advances by one
    more when the +axis neighbour is empty or outside (:53-63)."""
    L, M, N = dims
    nv = len(idxS)
    sub = np.array(np.unravel_index(idxS, dims, order='F'))          # (3, nv)
    occ = np.zeros(dims, dtype=bool)
    occ[tuple(sub)] = True
    # loop orders, slowest..fastest
    order = {0: (2, 1, 0), 1: (0, 2, 1), 2: (1, 0, 2)}[axis]
    key = np.zeros(nv, dtype=np.int64)
    for a in order:
        key = key * dims[a] + sub[a]
    perm = np.argsort(key, kind='stable')
    nxt = sub.copy()
    nxt[axis] += 1
    inside = nxt[axis] < dims[axis]
    has_next = np.zeros(nv, dtype=bool)
    has_next[inside] = occ[tuple(nxt[:, inside])]
    no_next_in_order = (~has_next[perm]).astype(np.int64)
    c = np.concatenate([[0], np.cumsum(no_next_in_order)[:-1]])     # jumps before k
    left_in_order = 1 + np.arange(nv) + c                              # 1-based
    out = np.zeros((nv, 2), dtype=np.int64)
    out[perm, 0] = left_in_order
    out[perm, 1] = left_in_order + 1
    return out, int(out.max()) if nv else 0


def panel_ids(idxS, dims):
    """all_panels_ids (nv x 6, 0-based node ids, columns -x,+x,-y,+y,-z,+z) and
    num_nodes, lse_generate_panelIDs.m:31-172."""
    px, nx = _panel_axis(idxS, dims, 0)
    py, ny = _panel_axis(idxS, dims, 1)
    pz, nz = _panel_axis(idxS, dims, 2)
    ids = np.concatenate([px, py + nx, pz + nx + ny], axis=1) - 1     # :164-171
    return ids, nx + ny + nz


def ae_matrix(ids, num_nodes):
    """Ae (num_nodes x 5nv), lse_compute_Ae_matrix.m:21-88."""
    nv = ids.shape[0]
    k = np.arange(nv)
    h = 0.5                                                           # const_lin, :21
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r); cols.append(c); vals.append(np.full(nv, v))
    add(ids[:, 0], k, -1.0); add(ids[:, 1], k, 1.0)                   # Jx  :30-31
    add(ids[:, 2], k + nv, -1.0); add(ids[:, 3], k + nv, 1.0)         # Jy  :32-33
    add(ids[:, 4], k + 2 * nv, -1.0); add(ids[:, 5], k + 2 * nv, 1.0) # Jz  :34-35
    add(ids[:, 0], k + 3 * nv, h); add(ids[:, 1], k + 3 * nv, h)      # J2d (x) :44-45
    add(ids[:, 2], k + 3 * nv, -h); add(ids[:, 3], k + 3 * nv, -h)    # J2d (y) :53-54
    add(ids[:, 0], k + 4 * nv, h); add(ids[:, 1], k + 4 * nv, h)      # J3d (x) :63-64
    add(ids[:, 2], k + 4 * nv, h); add(ids[:, 3], k + 4 * nv, h)      # J3d (y) :72-73
    add(ids[:, 4], k + 4 * nv, -2 * h); add(ids[:, 5], k + 4 * nv, -2 * h)  # J3d (z) :81-82
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(num_nodes, 5 * nv))


def _nodes_of(idxS, dims, ids, rows):
    """Node ids for (x,y,z,side) rows, lse_assign_exc_grnd_nodes.m:14-31."""
    out = []
    pos = {int(v): i for i, v in enumerate(idxS)} if len(rows) else {}
    for x, y, z, side in rows:
        lin = (x - 1) + dims[0] * ((y - 1) + dims[1] * (z - 1))
        if lin not in pos:
            raise ValueError('port node on empty voxel (%d,%d,%d)' % (x, y, z))
        out.append(ids[pos[lin], side - 1])
    return np.array(out, dtype=np.int64)


def build_system(geo):
    """Everything the per-frequency loop needs, VoxHenry_executer.m:110-193.

    Returns dict with idxS, idxS5, ids, Ae (full), Ar (port rows removed),
    rhs (n x nports, real), grnd, inject node lists, dims, dx.
    """
    sigma = geo['sigma']
    dims = sigma.shape
    idxS, idxS5 = structure_params(sigma, geo['lambdaL'])
    ids, num_nodes = panel_ids(idxS, dims)
    Ae = ae_matrix(ids, num_nodes)
    nv = len(idxS)
    num_curr = 5 * nv
    nports = len(geo['ports_P'])
    grnd, inj = [], []
    rhs = np.zeros((num_curr + num_nodes, nports))
    Ae_csr = Ae.tocsr()
    for p in range(nports):
        pi = _nodes_of(idxS, dims, ids, geo['ports_P'][p])
        pg = _nodes_of(idxS, dims, ids, geo['ports_N'][p])
        if p == 0 and len(geo['well_cond']):
            pg = np.concatenate([pg, _nodes_of(idxS, dims, ids, geo['well_cond'])])
        inj.append(pi); grnd.append(pg)
        # exc_vect = sum(-Ae(port_nodeid_4_injectcurr,:),1)   :175-177
        rhs[:num_curr, p] = -np.asarray(Ae_csr[pi, :].sum(axis=0)).ravel()
    grnd = np.concatenate(grnd) if grnd else np.zeros(0, np.int64)
    inj = np.concatenate(inj) if inj else np.zeros(0, np.int64)
    keep = np.ones(num_nodes, dtype=bool)
    keep[grnd] = False
    keep[inj] = False
    Ar = Ae_csr[keep, :].tocsc()                                       # :181-192
    rhs = rhs[:num_curr + Ar.shape[0], :]                              # :190
    return dict(idxS=idxS, idxS5=idxS5, ids=ids, Ae=Ae, Ar=Ar, keep=keep, rhs=rhs,
                grnd=grnd, inject=inj, dims=dims, dx=geo['dx'], nv=nv,
                num_curr=num_curr, num_node=Ar.shape[0])


def impedance(geo, idxS, freq):
    """z_real (L,M,N) and z_imag (L,M,N complex, or None), VoxHenry_executer.m:231-259."""
    mu = 4 * np.pi * 1e-7
    omega = 2 * np.pi * freq
    sigma = geo['sigma']
    lam = geo['lambdaL']
    with np.errstate(divide='ignore', invalid='ignore'):
        if lam is None:
            return 1.0 / sigma, None
        den = (omega * mu * sigma * lam ** 2) ** 2 + 1
        z_real = (sigma * ((omega * mu * lam ** 2) ** 2)) / den
        z_imag = (1j * omega * mu * lam ** 2) / den
        lam_nz = np.abs(lam) > 1e-12
        sig_nz = np.abs(sigma) > 1e-12
        sel = (~lam_nz) & sig_nz
        z_real[sel] = 1.0 / sigma[sel]
        z_imag[sel] = 0.0
    return z_real, z_imag
