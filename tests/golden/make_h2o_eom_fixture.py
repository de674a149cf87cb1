"""Write tests/golden/h2o_eom.json from the reference's regression test zmat.054a.

Run in the development container only (reads /root/reference).  Everything
stored is data: the Z-matrix of test/zmat.054a (H; O 1 R; H 2 R 1 A with
R = 0.957 A, A = 104.5 degrees; BASIS=DZP, SPHERICAL=OFF, the same
test/KEEP/GENBAS entries as tests/golden/h2o_dzp.json holds) and the numbers
test/OUTPUT/out.zmat.054a printed for it:

  nuclear repulsion, E(SCF), the CCSD total energy, the number of basis
  functions and of occupied orbitals per irrep of C2v in the reference's
  order, and per symmetry block (ESTATE_SYM=1/1/1/1, one root each) the
  EOM-CCSD excitation energy of the right-hand and of the left-hand Davidson
  solve, from its "Total EOM-CCSD electronic energy" lines (total minus the
  CCSD energy) and its "Converged eigenvalue" lines.
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = '/root/reference/test/OUTPUT/out.zmat.054a'
FLOAT = r'([-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?)'


def main():
    with open(OUT) as f:
        lines = f.read().split('\n')
    txt = '\n'.join(lines)
    nucrep = float(re.search(r'Nuclear repulsion energy :\s+' + FLOAT, txt).group(1))
    escf = float(re.search(r'E\(SCF\)=\s+' + FLOAT, txt).group(1))
    eccsd = float(re.search(r'CCSD\s+energy is\s+' + FLOAT, txt).group(1))
    k = next(i for i, s in enumerate(lines) if re.search(r'Irrep\s+# of functions', s))
    nfun = []
    for s in lines[k + 1:k + 10]:
        m = re.match(r'\s+(\d+)\s+(\d+)\s*$', s)
        if not m:
            break
        nfun.append(int(m.group(2)))
    pop = [int(x) for x in re.search(r'Alpha population by irrep:\s+([\d ]+)', txt).group(1).split()]
    blocks = []
    starts = [i for i, s in enumerate(lines) if 'Beginning symmetry block' in s]
    for n, i0 in enumerate(starts):
        i1 = starts[n + 1] if n + 1 < len(starts) else len(lines)
        seg = lines[i0:i1]
        sym = int(re.search(r'block\s+(\d+)\.', seg[0]).group(1))
        ev = [float(re.search(FLOAT, s.split(':')[1]).group(1)) for s in seg if 'Converged eigenvalue' in s]
        tot = [float(re.search(FLOAT, s.split('energy')[1]).group(1)) for s in seg
               if s.startswith('   Total EOM-CCSD electronic energy')]
        blocks.append({'symmetry_block': sym, 'right_eigenvalue': ev[0], 'left_eigenvalue': ev[1],
                       'right_from_total': tot[0] - eccsd, 'left_from_total': tot[1] - eccsd})
    fixture = {
        'source': 'test/zmat.054a and test/OUTPUT/out.zmat.054a (EXCITE=EOMEE, ESTATE_SYM=1/1/1/1)',
        'zmatrix': {'R_angstrom': 0.957, 'A_degree': 104.5, 'bohr_angstrom': 0.529177249},
        'basis': 'DZP (Cartesian d), as in h2o_dzp.json',
        'nocc': sum(pop),
        'nucrep': nucrep,
        'escf': escf,
        'ccsd_total_energy': eccsd,
        'irrep_functions': nfun,
        'irrep_occupied': pop,
        'eom_blocks': blocks,
    }
    with open(os.path.join(HERE, 'h2o_eom.json'), 'w') as f:
        json.dump(fixture, f, indent=1)


if __name__ == '__main__':
    main()
