"""``.vhr`` reader, restated from src_pre_process/pre_input_file.m:1-133.

Pinned by: parsing the reference's own input files (voxel counts, port-node
counts, frequency lists) in tests/test_oracle_topology.py.
"""
import re
import numpy as np

# pre_input_file.m:10 -- side codes map to panel columns 1..6 (-x,+x,-y,+y,-z,+z)
SIDES = {'-x': 1, '+x': 2, '-y': 3, '+y': 4, '-z': 5, '+z': 6}


def read_vhr(path):
    """Return dict(sigma, lambdaL, freq, dx, ports_P, ports_N, well_cond).

    ``sigma``/``lambdaL`` are (L,M,N) float arrays (lambdaL None when the file
    has no ``superconductor`` keyword, pre_input_file.m:112-114).  ``ports_P`` /
    ``ports_N`` are lists (one per port, in order of first appearance,
    pre_input_file.m:80-89) of int arrays (k,4) holding 1-based x,y,z and the
    side code 1..6.  ``well_cond`` collects nodes whose type is neither P nor N
    (pre_input_file.m:100-102).
    """
    freq = []
    dx = 1e-6                                   # pre_input_file.m:22
    sigma = lam = None
    supercond = False
    names, P, N, well = [], [], [], []
    with open(path, 'r') as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        tok = [t for t in re.split(r"[ \f\n\r\t\v=,]+", lines[i].strip()) if t != '']
        i += 1
        if not tok:
            continue
        key = tok[0].lower()
        if key == 'freq':
            freq += [float(t) for t in tok[1:]]
        elif key == 'dx' and len(tok) >= 2:
            dx = float(tok[1])
        elif key == 'lmn' and len(tok) >= 4:
            L, M, Nn = int(tok[1]), int(tok[2]), int(tok[3])
            sigma = np.zeros((L, M, Nn))
            lam = np.zeros((L, M, Nn))
        elif key == 'superconductor':
            supercond = True
        elif key == 'startvoxellist':
            # textscan 'V %d %d %d %f [%f]' stops at the first non-matching line
            while i < len(lines):
                t = lines[i].split()
                if not t or t[0] != 'V':
                    break
                x, y, z = int(t[1]) - 1, int(t[2]) - 1, int(t[3]) - 1
                sigma[x, y, z] = float(t[4])
                if supercond:
                    lam[x, y, z] = float(t[5])
                i += 1
        elif key == 'n' and len(tok) >= 7:
            pname = tok[1].lower()
            if pname not in names:
                names.append(pname)
                P.append([])
                N.append([])
            pn = names.index(pname)
            row = [int(tok[3]), int(tok[4]), int(tok[5]), SIDES[tok[6]]]
            kind = tok[2].upper()
            if kind == 'P':
                P[pn].append(row)
            elif kind == 'N':
                N[pn].append(row)
            else:
                well.append(row)
    as_arr = lambda l: np.array(l, dtype=np.int64).reshape(-1, 4)
    return dict(sigma=sigma, lambdaL=lam if supercond else None,
                freq=np.array(freq if freq else [1.0]), dx=dx,
                ports_P=[as_arr(p) for p in P], ports_N=[as_arr(n) for n in N],
                well_cond=as_arr(well))
