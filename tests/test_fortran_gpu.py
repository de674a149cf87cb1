"""The flang-built Fortran host running vcc.F's call sequence through the library."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import rccsd, vcc_loop

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, 'aces-ii_amd', 'fortran', 'xvcc_amd')


def write_input(path, blk, o, v):
    from aces2 import lists
    L = lists.c1_integral_lists(blk, o)
    with open(path, 'wb') as f:
        np.array([o, v, len(L)], dtype=np.int64).tofile(f)
        for lst, (tl, tr, data) in L.items():
            nr = {13: o * o, 16: v * v, 10: o * o, 30: v * v, 23: o * v, 233: v * (v + 1) // 2}[lst]
            nc = data.size // nr
            np.array([lst, tl, tr, nr, nc], dtype=np.int64).tofile(f)
            np.ascontiguousarray(data, dtype=np.float64).ravel().tofile(f)
        np.diag(blk['fock']).astype(np.float64).tofile(f)


@pytest.mark.gpu
def test_fortran_host_matches_oracle(tmp_path):
    from aces2 import synth
    assert os.path.exists(EXE), 'Fortran driver not built (run __graft_entry__.build())'
    print('case fortran xvcc', flush=True)
    o, v = 5, 19
    eps, B = synth.draw(o, v, 41, sigma=synth.scaled_sigma(o, v))
    blk, eri = synth.blocks_numpy(eps, B, o)
    ref = vcc_loop.run(blk['fock'], rccsd.blocks(eri, o), o, maxcyc=60, conv=9, order=6)
    inp = tmp_path / 'lists.bin'
    write_input(str(inp), blk, o, v)
    out = subprocess.run([EXE, str(inp), '9', '60', '6'], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    eng = [float(x) for x in re.findall(r'The total correlation energy is\s+(\S+) a\.u\.', out.stdout)]
    res = [float(x.rstrip('.')) for x in re.findall(r'Largest element of residual vector :\s+(\S+)', out.stdout)]
    # printed: E(MP1), then E(T_new) per cycle, then the final E on convergence
    assert len(res) == ref['cycles']
    np.testing.assert_allclose(eng[:-1], ref['energy'], rtol=0, atol=2e-12 + 1e-11)
    np.testing.assert_allclose(res, ref['resmiter'], rtol=1e-6, atol=1e-13)
    assert abs(eng[-1] - ref['final_energy']) < 1e-10
    assert 'Amplitude equations converged' in out.stdout
