"""The product's .vhr reader and writer (structures.read_vhr / write_vhr,
pre_input_file.m:36-114 / pre_output_file.m:44-130) against the oracle's
reader: the reference's own input decks parse to the same structure, and
write -> read returns what was written (normal and two-fluid conductors,
several ports, grounded faces)."""
import glob
import os

import numpy as np
import pytest

from oracle import vhr
from tests import problems
from voxhenry_amd import structures

REF_INPUTS = '/root/reference/Input_files'


def same(a, b):
    """A product structure (well) equals an oracle one (well_cond)."""
    assert a['sigma'].shape == b['sigma'].shape
    assert np.array_equal(a['sigma'], b['sigma'])
    assert (a['lambdaL'] is None) == (b['lambdaL'] is None)
    if a['lambdaL'] is not None:
        assert np.array_equal(a['lambdaL'], b['lambdaL'])
    assert np.array_equal(a['freq'], b['freq']) and a['dx'] == b['dx']
    assert len(a['ports_P']) == len(b['ports_P']) == len(a['ports_N']) == len(b['ports_N'])
    for x, y in zip(a['ports_P'] + a['ports_N'], b['ports_P'] + b['ports_N']):
        assert np.array_equal(x, y)
    assert np.array_equal(a['well'], b['well_cond'])


def decks():
    out = [os.path.join(problems.GOLDEN, 'wire_len2.0u_dia4.0u.vhr')]
    if os.path.isdir(REF_INPUTS):
        out += sorted(glob.glob(os.path.join(REF_INPUTS, '*.vhr')))
    return out


@pytest.mark.parametrize('path', decks(), ids=os.path.basename)
def test_read_vhr_matches_oracle_reader(path):
    same(structures.read_vhr(path), vhr.read_vhr(path))


@pytest.mark.parametrize('path', decks(), ids=os.path.basename)
def test_reference_deck_round_trip(path, tmp_path):
    """read -> write -> read of every reference input deck is the identity."""
    a = structures.read_vhr(path)
    out = tmp_path / 'copy.vhr'
    structures.write_vhr(str(out), a, title=os.path.basename(path))
    b = structures.read_vhr(str(out))
    same(b, vhr.read_vhr(str(out)))
    for k in ('sigma', 'freq'):
        assert np.array_equal(a[k], b[k])
    assert a['dx'] == b['dx']
    assert all(np.array_equal(x, y) for x, y in zip(a['ports_P'] + a['ports_N'], b['ports_P'] + b['ports_N']))


def test_write_vhr_synthetic(tmp_path):
    """Multi-port two-fluid bundle with a grounded face and non-round numbers."""
    st = structures.parallel('strip', 3, 12, 4, freq=np.logspace(6, 10, 13))
    st['sigma'] = st['sigma'] * (1 + 1e-9 * np.pi)
    st['well'] = np.array([[1, 1, 1, 5], [2, 1, 1, 5]], np.int64)
    p = tmp_path / 's.vhr'
    structures.write_vhr(str(p), st)
    txt = p.read_text()
    assert txt.count('\nfreq=') == 2 and 'Superconductor' in txt and 'N portGND G 2 1 1 -z' in txt
    b = structures.read_vhr(str(p))
    assert np.array_equal(b['sigma'], st['sigma']) and np.array_equal(b['lambdaL'], st['lambdaL'])
    assert np.array_equal(b['freq'], st['freq'])
    # the reader opens a port for every port name, the grounded faces'
    # portGND too (pre_input_file.m:80-89): it comes back as a fourth, empty port
    assert len(b['ports_P']) == 4 and len(b['ports_P'][3]) == 0 and len(b['ports_N'][3]) == 0
    for x, y in zip(b['ports_P'][:3] + b['ports_N'][:3], st['ports_P'] + st['ports_N']):
        assert np.array_equal(x, y)
    assert np.array_equal(b['well'], st['well'])
    same(b, vhr.read_vhr(str(p)))
