"""Write tests/golden/o2_sto3g.json from the reference's regression test zmat.015a.

Run in the development container only (reads /root/reference).  Everything
stored is data: the O STO-3G basis of test/KEEP/GENBAS, the geometry implied
by the printed nuclear repulsion energy, the SCF occupation by D2h irrep, and
the per-cycle CC energies (total and the AA/AB spin-case lines of CMPENG) and
largest residual elements printed in
test/OUTPUT/out.zmat.015a (first geometry, MP1 start, CC_EXPORDER = 5,
CC_CONV = 7), the xlambda trace that follows (largest residual and DIIS
residual elements per Lambda cycle, cycle count) and the xdens natural orbital
occupation numbers of that geometry, plus the TEST.DAT values of test/zmat.015a.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import molint  # noqa: E402

REF = '/root/reference/test'


def main():
    shells = molint.read_genbas_entry(os.path.join(REF, 'KEEP', 'GENBAS'), 'O:STO-3G')
    with open(os.path.join(REF, 'OUTPUT', 'out.zmat.015a')) as f:
        lines = f.read().split('\n')
    # first geometry: lines from the first SCF to the first "Amplitude equations converged"
    nucrep = float(re.search(r'Nuclear repulsion energy :\s+(\S+)', '\n'.join(lines)).group(1))
    i0 = next(k for k, s in enumerate(lines) if 'E(SCF)=' in s)
    escf = float(lines[i0].split()[1])
    i1 = next(k for k, s in enumerate(lines) if 'Amplitude equations converged' in s)
    energies, resid, eaa, eab = [], [], [], []
    for s in lines[i0:i1 + 6]:
        m = re.search(r'The total correlation energy is\s+(\S+) a\.u\.', s)
        if m:
            energies.append(float(m.group(1)))
        m = re.search(r'The (AA|AB) contribution to the correlation energy is:\s+(\S+) a\.u\.', s)
        if m:
            (eaa if m.group(1) == 'AA' else eab).append(float(m.group(2)))
        m = re.search(r'Largest element of residual vector :\s+(\S+)\.$', s)
        if m:
            resid.append(float(m.group(1)))
    cycles = int(re.search(r'converged in\s+(\d+)', lines[i1]).group(1))
    # xlambda and xdens of the first geometry
    j0 = next(k for k, s in enumerate(lines) if 'Entering xlambda' in s)
    j1 = next(k for k, s in enumerate(lines) if 'Lambda equations converged' in s)
    lres, ldiis = [], []
    for s in lines[j0:j1]:
        m = re.search(r'Largest element of residual vector :\s+(\S+)\.$', s)
        if m:
            lres.append(float(m.group(1)))
        m = re.search(r'Largest element of DIIS residual\s+:\s+(\S+)\.$', s)
        if m:
            ldiis.append(float(m.group(1)))
    lcycles = int(re.search(r'converged in\s+(\d+)', lines[j1]).group(1))
    k0 = next(k for k, s in enumerate(lines) if 'Natural orbital occupation numbers' in s)
    nocc_vals = []
    for s in lines[k0 + 2:k0 + 10]:
        if 'Trace' in s or s.strip().startswith('---'):
            break
        nocc_vals += [float(x) for x in s.split()]
    trace = float(re.search(r'Trace of density matrix :\s+(\S+)\.', '\n'.join(lines[k0:k0 + 10])).group(1))
    with open(os.path.join(REF, 'zmat.015a')) as f:
        zm = f.read()
    totenerg = float(re.search(r'd TOTENERG\s+(\S+)', zm).group(1))
    scfeneg = float(re.search(r'd SCFENEG\s+(\S+)', zm).group(1))
    coords = [float(x) for x in re.findall(r'^\s+(\S+)\s+0\.1E-08', zm.split('d COORD')[1].split('d NUCREP')[0],
                                           re.M)]
    fixture = {
        'source': 'test/zmat.015a, test/OUTPUT/out.zmat.015a (first geometry), test/KEEP/GENBAS O:STO-3G',
        'basis_O': [{'l': l, 'exps': exps, 'coefs': cm.tolist()} for (l, exps, cm) in shells],
        'first_geometry': {
            'z_bohr': 32.0 / nucrep, 'nucrep': nucrep, 'escf': escf,
            'cc_exporder': 5, 'cc_conv': 7,
            'mp2_and_cycle_energies': energies[:cycles + 1],
            'final_energy': energies[cycles + 1],
            'aa_contributions': eaa[:cycles + 2],
            'ab_contributions': eab[:cycles + 2],
            'resmiter': resid[:cycles],
            'cycles': cycles,
            'lambda_resmiter': lres[:lcycles],
            'lambda_resmdiis': ldiis[:lcycles],
            'lambda_cycles': lcycles,
            'natural_occupations': nocc_vals,
            'density_trace': trace,
        },
        'final_geometry': {'z_bohr': coords[2], 'totenerg': totenerg, 'scfeneg': scfeneg},
        'occupation_d2h': {'(1,1,1)': 3, '(-1,1,1)': 1, '(1,-1,1)': 1, '(1,1,-1)': 2, '(1,-1,-1)': 1},
        'nocc': 8,
    }
    with open(os.path.join(HERE, 'o2_sto3g.json'), 'w') as f:
        json.dump(fixture, f, indent=1)


if __name__ == '__main__':
    main()
