"""Write tests/golden/h2o_dzp.json from the reference's regression tests zmat.013a / zmat.014a.

Run in the development container only (reads /root/reference).  Everything
stored is data: the H and O DZP basis sets of test/KEEP/GENBAS, the first
geometry of the two runs (the same Z-matrix, R = 0.943056 A, A = 105.96879
degrees, Cartesian d functions), and the numbers the reference printed for it:

  test/OUTPUT/out.zmat.013a (RHF-CCSD, CC_EXPORDER 5, CC_CONV 7)
    nuclear repulsion, E(SCF), the MP2 energies of xintprc (E2(AA), E2(AB), total),
    every CMPENG line of the CC run (AA, AB, total), RESMITER / RESMDIIS per cycle,
    the final CCSD energy, the xlambda RESMITER / RESMDIIS trace and the xdens
    natural orbital occupation numbers;
  test/OUTPUT/out.zmat.014a (RHF-CCSD(T), same first geometry)
    the CCSD energy (equal to the run above) and the triples pieces of xvcc's
    TRPS: E4T(BAA), E4T(AAA), E5ST(A), E5ST(B), E[4]T, E[5]ST and CCSD(T).
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import molint  # noqa: E402

REF = '/root/reference/test'
FLOAT = r'([-+]?\d*\.\d+(?:[EeDd][-+]?\d+)?)'


def _f(x):
    return float(x.replace('D', 'E').replace('d', 'e'))


def first_block(lines, start_pat, end_pat):
    i0 = next(k for k, s in enumerate(lines) if re.search(start_pat, s))
    i1 = next(k for k, s in enumerate(lines[i0:], i0) if re.search(end_pat, s))
    return i0, i1


def cc_trace(lines):
    i0, i1 = first_block(lines, r'Entering xvcc', r'Amplitude equations converged')
    tot, aa, ab, res, dres = [], [], [], [], []
    for s in lines[i0:i1 + 4]:
        m = re.search(r'The total correlation energy is\s+' + FLOAT, s)
        if m:
            tot.append(_f(m.group(1)))
        m = re.search(r'The (AA|AB) contribution to the correlation energy is:\s+' + FLOAT, s)
        if m:
            (aa if m.group(1) == 'AA' else ab).append(_f(m.group(2)))
        m = re.search(r'Largest element of residual vector :\s+' + FLOAT, s)
        if m:
            res.append(_f(m.group(1)))
        m = re.search(r'Largest element of DIIS residual\s+:\s+' + FLOAT, s)
        if m:
            dres.append(_f(m.group(1)))
    cycles = int(re.search(r'converged in\s+(\d+)', lines[i1]).group(1))
    return {'cycles': cycles, 'mp2_and_cycle_energies': tot[:cycles + 1], 'final_energy': tot[cycles + 1],
            'aa_contributions': aa[:cycles + 2], 'ab_contributions': ab[:cycles + 2],
            'resmiter': res[:cycles], 'resmdiis': dres[:cycles]}


def main():
    basis = {el: molint.read_genbas_entry(os.path.join(REF, 'KEEP', 'GENBAS'), el + ':DZP') for el in ('H', 'O')}
    with open(os.path.join(REF, 'OUTPUT', 'out.zmat.013a')) as f:
        a = f.read().split('\n')
    with open(os.path.join(REF, 'OUTPUT', 'out.zmat.014a')) as f:
        b = f.read().split('\n')
    txt = '\n'.join(a)
    nucrep = _f(re.search(r'Nuclear repulsion energy :\s+' + FLOAT, txt).group(1))
    escf = _f(re.search(r'E\(SCF\)=\s+' + FLOAT, txt).group(1))
    mp2 = {k: _f(re.search(r'%s\s+=\s+' % re.escape(k) + FLOAT, txt).group(1))
           for k in ('E2(AA)', 'E2(AB)', 'E2(TOT)')}
    cc = cc_trace(a)
    ccsd_total = _f(re.search(r'CCSD\s+energy is\s+' + FLOAT, txt).group(1))
    # xlambda of the first geometry
    j0, j1 = first_block(a, r'Entering xlambda', r'Lambda equations converged')
    lres, ldiis = [], []
    for s in a[j0:j1]:
        m = re.search(r'Largest element of residual vector :\s+' + FLOAT, s)
        if m:
            lres.append(_f(m.group(1)))
        m = re.search(r'Largest element of DIIS residual\s+:\s+' + FLOAT, s)
        if m:
            ldiis.append(_f(m.group(1)))
    lcycles = int(re.search(r'converged in\s+(\d+)', a[j1]).group(1))
    k0 = next(k for k, s in enumerate(a) if 'Natural orbital occupation numbers' in s)
    occ = []
    for s in a[k0 + 2:k0 + 12]:
        if 'Trace' in s or s.strip().startswith('---') or not s.strip():
            break
        occ += [float(x) for x in s.split()]
    trace = _f(re.search(r'Trace of density matrix :\s+' + FLOAT, '\n'.join(a[k0:k0 + 12])).group(1))
    # triples of zmat.014a, first geometry
    t0 = next(k for k, s in enumerate(b) if 'VCC Triple Excitation Contributions' in s)
    tb = '\n'.join(b[t0:t0 + 60])
    b_txt = '\n'.join(b[:t0])
    ccsd_014 = _f(re.findall(r'CCSD\s+energy is\s+' + FLOAT, b_txt)[0])
    cc_014 = cc_trace(b)
    trip = {
        'e4t_baa': _f(re.search(r'E4TBAA\s+' + FLOAT, tb).group(1)),
        'e4t_aaa': _f(re.search(r'E4TAAA\s+' + FLOAT, tb).group(1)),
        'e5st_a': _f(re.search(r'E5ST A\s+' + FLOAT, tb).group(1)),
        'e5st_b': _f(re.search(r'E5ST B\s+' + FLOAT, tb).group(1)),
        'e4t': _f(re.search(r'E4T \(T\(CCSD\)\) =\s+' + FLOAT, tb).group(1)),
        'e5st': _f(re.search(r'E5ST\s+=\s+' + FLOAT, tb).group(1)),
        'ccsd': _f(re.search(r'@TRPS-I, CCSD\s+=\s+' + FLOAT, tb).group(1)),
        'ccsd_plus_t_ccsd': _f(re.search(r'CCSD \+ T\(CCSD\) =\s+' + FLOAT, tb).group(1)),
        'ccsd_t': _f(re.search(r'CCSD\(T\)\s+=\s+' + FLOAT, tb).group(1)),
    }
    fixture = {
        'source': 'test/zmat.013a, test/zmat.014a, test/OUTPUT/out.zmat.013a and out.zmat.014a (first geometry), '
                  'test/KEEP/GENBAS H:DZP and O:DZP',
        'basis': {el: [{'l': l, 'exps': exps, 'coefs': cm.tolist()} for (l, exps, cm) in sh]
                  for el, sh in basis.items()},
        'zmatrix': {'R_angstrom': 0.943056, 'A_degree': 105.96879, 'bohr_angstrom': 0.529177249},
        'nocc': 5,
        'nucrep': nucrep,
        'escf': escf,
        'mp2': mp2,
        'cc_exporder': 5,
        'cc_conv': 7,
        'cc': cc,
        'ccsd_total_energy': ccsd_total,
        'lambda_resmiter': lres[:lcycles],
        'lambda_resmdiis': ldiis[:lcycles],
        'lambda_cycles': lcycles,
        'natural_occupations': occ,
        'density_trace': trace,
        'ccsd_t_run': {'ccsd_total_energy': ccsd_014, 'cc_final_energy': cc_014['final_energy'],
                       'cc_cycles': cc_014['cycles'], 'triples': trip},
    }
    with open(os.path.join(HERE, 'h2o_dzp.json'), 'w') as f:
        json.dump(fixture, f, indent=1)


if __name__ == '__main__':
    main()
