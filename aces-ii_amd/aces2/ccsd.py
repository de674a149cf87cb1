"""Host driver of the closed-shell CCSD amplitude path (the xvcc CC loop).

Mirrors the reference's loop ``vcc/vcc.F:798-1437`` on top of the C-ABI
engine (``include/aces2_amd.h``): the MP1 start written by xintprc, one
``aces_ccsd_cycle`` per iteration (increments, NEWT2, CMPENG, DODIIS0,
DRMOVE), convergence on |RESMITER| < 10**-CC_CONV, non-convergence raising
(the reference exits with status 1, ``vcc/vcc.F:1439-1441``).

Device memory is torch-allocated plumbing; no torch op runs on the path.
"""
import ctypes

from . import _lib

# 'refine': fp64 cycle whose eight large products run as bf16x3 products of
# the operands' differences to an fp64 anchor evaluation
PREC = {'fp64': 0, 'fp32': 1, 'bf16x3': 2, 'refine': 3}


class NotConverged(RuntimeError):
    """The Coupled-Cluster equations did not converge (``vcc/vcc.F:1439``)."""


class LambdaNotConverged(RuntimeError):
    """The Lambda equations did not converge (``lambda/vlamcc.F``)."""


class CCSD:
    _p = 'aces_ccsd_'

    def __init__(self, o, v, blocks, stream=None):
        import torch
        self.o, self.v = o, v
        self.torch = torch
        self.stream = stream
        h = ctypes.c_void_p()
        _lib.call(self._p + 'create', o, v, stream, ctypes.byref(h))
        self.h = h
        dev = {k: torch.as_tensor(x, device='cuda', dtype=torch.float64).contiguous() for k, x in blocks.items()}
        _lib.call(self._p + 'setup', self.h, dev['fock'].data_ptr(), dev['ovov'].data_ptr(), dev['oovv'].data_ptr(),
                  dev['ovoo'].data_ptr(), dev['oooo'].data_ptr(), dev['ovvv'].data_ptr(), dev['list233'].data_ptr())
        del dev

    def close(self):
        if self.h:
            getattr(_lib.lib(), self._p + 'destroy')(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return getattr(_lib.lib(), self._p + 'device_bytes')(self.h)

    def set_amplitudes(self, t1, t2):
        torch = self.torch
        t1 = torch.as_tensor(t1, device='cuda', dtype=torch.float64).contiguous()
        t2 = torch.as_tensor(t2, device='cuda', dtype=torch.float64).contiguous()
        _lib.call(self._p + 'set_amplitudes', self.h, t1.data_ptr(), t2.data_ptr())

    def get_amplitudes(self):
        torch = self.torch
        o, v = self.o, self.v
        t1 = torch.empty((o, v), device='cuda', dtype=torch.float64)
        t2 = torch.empty((o, o, v, v), device='cuda', dtype=torch.float64)
        _lib.call(self._p + 'get_amplitudes', self.h, t1.data_ptr(), t2.data_ptr())
        return t1, t2

    def start_mp1(self):
        _lib.call(self._p + 'start_mp1', self.h)

    def increments(self, prec='fp64'):
        torch = self.torch
        o, v = self.o, self.v
        i1 = torch.empty((o, v), device='cuda', dtype=torch.float64)
        i2 = torch.empty((o, o, v, v), device='cuda', dtype=torch.float64)
        _lib.call(self._p + 'increments', self.h, PREC[prec], i1.data_ptr(), i2.data_ptr())
        return i1, i2

    def energy(self, t1, t2):
        torch = self.torch
        t1 = torch.as_tensor(t1, device='cuda', dtype=torch.float64).contiguous()
        t2 = torch.as_tensor(t2, device='cuda', dtype=torch.float64).contiguous()
        e = ctypes.c_double()
        _lib.call(self._p + 'energy', self.h, t1.data_ptr(), t2.data_ptr(), ctypes.byref(e))
        return e.value

    def set_profiling(self, on=True):
        _lib.call(self._p + 'set_profiling', self.h, 1 if on else 0)

    def profile(self, reset=True):
        out = (ctypes.c_double * 21)()
        _lib.call(self._p + 'profile', self.h, out, 1 if reset else 0)
        names = ('ladder', 'wbuild', 'ring', 'other', 'adjoint', 'adjoint_lp', 'triples')
        return {nm: {'ms': out[3 * c], 'flops': out[3 * c + 1], 'launches': int(out[3 * c + 2])}
                for c, nm in enumerate(names)}

    def triples(self, per_triple=False, prec='fp64'):
        """(T) of CCSD(T) at the current amplitudes (the reference's TRPS):
        dict with e4t, e5st (E(T) = e4t + e5st), e4t_aaa, e4t_baa, ms, flops,
        ntriples; with ``per_triple`` also the (ntriples, 3) device array of each
        occupied triple's unweighted sums.  ``prec``: 'fp64' or 'bf16x3' (the
        W products as split bf16 products)."""
        torch = self.torch
        out = (ctypes.c_double * 8)()
        o = self.o
        nt = o * (o + 1) * (o + 2) // 6
        per = torch.empty((nt, 3), device='cuda', dtype=torch.float64) if per_triple else None
        _lib.call(self._p + 'triples', self.h, PREC[prec], out, per.data_ptr() if per is not None else None)
        r = {'e4t': out[0], 'e5st': out[1], 'et': out[0] + out[1], 'e4t_aaa': out[2], 'e4t_baa': out[3],
             'ms': out[4], 'flops': out[5], 'ntriples': int(out[6])}
        if per_triple:
            r['per_triple'] = per
        return r

    def eom(self, nroots=1, irrep=0, orb_irrep=None, tol=1e-7, maxit=100, vectors=False):
        """EOM-CCSD singlet excitation energies at the current (converged)
        amplitudes: dict with omega (hartree), resid, iters and, with
        ``vectors``, the left eigenvectors (device, (nroots, o v + o^2 v^2)).
        ``orb_irrep``: orbital irreps (occupied then virtual) to restrict to the
        excitations of ``irrep``; None: no symmetry."""
        torch = self.torch
        om = (ctypes.c_double * nroots)()
        rs = (ctypes.c_double * nroots)()
        it = ctypes.c_int()
        irr = None
        if orb_irrep is not None:
            irr = (ctypes.c_int * len(orb_irrep))(*[int(x) for x in orb_irrep])
        vec = None
        if vectors:
            nt = self.o * self.v + (self.o * self.v) ** 2
            vec = torch.empty((nroots, nt), device='cuda', dtype=torch.float64)
        _lib.call(self._p + 'eom', self.h, irr, int(irrep), int(nroots), float(tol), int(maxit), om, rs,
                  ctypes.byref(it), vec.data_ptr() if vec is not None else None)
        out = {'omega': list(om), 'resid': list(rs), 'iters': it.value}
        if vectors:
            out['vectors'] = vec
        return out

    def set_fock(self, fock):
        torch = self.torch
        f = torch.as_tensor(fock, device='cuda', dtype=torch.float64).contiguous()
        _lib.call(self._p + 'set_fock', self.h, f.data_ptr())

    def set_diis_order(self, order):
        _lib.call(self._p + 'set_diis_order', self.h, order)

    def reset_diis(self):
        _lib.call(self._p + 'reset_diis', self.h)

    def reset_anchor(self):
        """The next 'refine' cycle records a new fp64 anchor."""
        _lib.call(self._p + 'reset_anchor', self.h)

    def cycle(self, icycle, prec='fp64', tol=1e-7):
        st = (ctypes.c_double * 8)()
        _lib.call(self._p + 'cycle', self.h, icycle, PREC[prec], tol, st)
        return {'energy': st[0], 'resmiter': st[1], 'resmdiis': st[2], 'converged': bool(st[3]),
                'final_energy': st[4], 'ms_increments': st[5], 'ms_cycle': st[6]}

    def run(self, maxcyc=50, conv=7, order=6, start='mp1', prec='fp64', log=None, raise_on_fail=True,
            polish_conv=None, polish='refine'):
        """The reference's CC loop; returns the per-cycle history.  With a
        low-precision ``prec`` and ``polish_conv``, the loop switches to
        ``polish`` cycles once |RESMITER| < 10^-conv (DIIS restarted, as its
        low-precision error vectors would hold the extrapolation at their
        accuracy) and runs to 10^-polish_conv; ``polish`` is 'fp64' or 'refine'
        (fp64 accuracy from bf16x3 products of the differences to the amplitudes
        of the switch, anchored by the first polish cycle's fp64 evaluation).
        ``hist['precision']`` records each cycle's form."""
        torch = self.torch
        self.set_diis_order(order)
        self.reset_diis()
        if start == 'mp1':
            self.start_mp1()
        elif start == 'current':
            pass
        elif start == 'zero':
            self.set_amplitudes(torch.zeros(self.o, self.v), torch.zeros(self.o, self.o, self.v, self.v))
        t1, t2 = self.get_amplitudes()
        hist = {'energy': [self.energy(t1, t2)], 'resmiter': [], 'resmdiis': [], 'ms_cycle': [],
                'converged': False}
        tol = 10.0 ** (-conv)
        phase, ic0 = prec, 0
        hist['precision'] = []
        for icycle in range(1, maxcyc + 1):
            st = self.cycle(icycle - ic0, phase, tol)
            hist['energy'].append(st['energy'])
            hist['resmiter'].append(st['resmiter'])
            hist['resmdiis'].append(st['resmdiis'])
            hist['ms_cycle'].append(st['ms_cycle'])
            hist['precision'].append(phase)
            if log:
                log(icycle, st)
            if st['converged'] and phase == prec and prec in ('fp32', 'bf16x3') and polish_conv is not None:
                phase, tol, ic0 = polish, 10.0 ** (-polish_conv), icycle
                self.reset_diis()
                if polish == 'refine':
                    self.reset_anchor()
                continue
            if st['converged']:
                hist['converged'] = True
                hist['final_energy'] = st['final_energy']
                hist['cycles'] = icycle
                break
        if not hist['converged'] and raise_on_fail:
            raise NotConverged('@VCC: The Coupled-Cluster equations did not converge!')
        return hist

    # ---- differentiation form ----
    def adjoint(self, l1, l2, density=False):
        """r = dL/dT for spatial multipliers (l1, l2) at the engine's amplitudes;
        with ``density`` also dL/dfock (the response one-particle density)."""
        torch = self.torch
        o, v = self.o, self.v
        n = o + v
        l1 = torch.as_tensor(l1, device='cuda', dtype=torch.float64).contiguous()
        l2 = torch.as_tensor(l2, device='cuda', dtype=torch.float64).contiguous()
        r1 = torch.empty((o, v), device='cuda', dtype=torch.float64)
        r2 = torch.empty((o, o, v, v), device='cuda', dtype=torch.float64)
        dens = torch.empty((n, n), device='cuda', dtype=torch.float64) if density else None
        _lib.call(self._p + 'adjoint', self.h, l1.data_ptr(), l2.data_ptr(), r1.data_ptr(), r2.data_ptr(),
                  dens.data_ptr() if density else None)
        return (r1, r2, dens) if density else (r1, r2)

    def lambda_start(self):
        _lib.call(self._p + 'lambda_start', self.h)

    def set_lambda_precision(self, prec):
        """'fp64' or 'refine' (mixed) Lambda cycles."""
        _lib.call(self._p + 'set_lambda_precision', self.h, PREC[prec])

    def set_lambda_diis_order(self, order):
        _lib.call(self._p + 'set_lambda_diis_order', self.h, order)

    def lambda_cycle(self, icycle, tol=1e-7):
        st = (ctypes.c_double * 8)()
        _lib.call(self._p + 'lambda_cycle', self.h, icycle, tol, st)
        return {'resmiter': st[1], 'resmdiis': st[2], 'converged': bool(st[3]), 'ms_cycle': st[6]}

    def get_lambda(self, aces_norm=False):
        torch = self.torch
        o, v = self.o, self.v
        l1 = torch.empty((o, v), device='cuda', dtype=torch.float64)
        l2 = torch.empty((o, o, v, v), device='cuda', dtype=torch.float64)
        _lib.call(self._p + 'get_lambda', self.h, l1.data_ptr(), l2.data_ptr(), 1 if aces_norm else 0)
        return l1, l2

    def set_lambda(self, l1, l2, aces_norm=False):
        torch = self.torch
        l1 = torch.as_tensor(l1, device='cuda', dtype=torch.float64).contiguous()
        l2 = torch.as_tensor(l2, device='cuda', dtype=torch.float64).contiguous()
        _lib.call(self._p + 'set_lambda', self.h, l1.data_ptr(), l2.data_ptr(), 1 if aces_norm else 0)

    def density(self):
        """Response one-particle density dE_corr/dfock of the current multipliers."""
        torch = self.torch
        n = self.o + self.v
        d = torch.empty((n, n), device='cuda', dtype=torch.float64)
        _lib.call(self._p + 'density', self.h, d.data_ptr())
        return d

    def solve_lambda(self, maxcyc=50, conv=7, order=6, log=None, raise_on_fail=True, prec='fp64'):
        """xlambda's loop (``lambda/vlamcc.F:599-830``) at the engine's amplitudes.
        ``prec='refine'``: residuals from an fp64 anchor plus the linear part
        on the difference with its large products in bf16x3."""
        self.lambda_start()
        self.set_lambda_diis_order(order)
        self.set_lambda_precision(prec)
        self.lambda_start()
        hist = {'resmiter': [], 'resmdiis': [], 'ms_cycle': [], 'converged': False}
        tol = 10.0 ** (-conv)
        for icycle in range(1, maxcyc + 1):
            st = self.lambda_cycle(icycle, tol)
            hist['resmiter'].append(st['resmiter'])
            hist['resmdiis'].append(st['resmdiis'])
            hist['ms_cycle'].append(st['ms_cycle'])
            if log:
                log(icycle, st)
            if st['converged']:
                hist['converged'] = True
                hist['cycles'] = icycle
                break
        if not hist['converged'] and raise_on_fail:
            raise LambdaNotConverged('The Lambda equations did not converge')
        return hist


class SymCCSD(CCSD):
    """The same driver over the irrep-blocked engine (``aces_sym_*``).

    ``occ``/``vir`` give the orbitals per irrep (nirr = 1, 2, 4 or 8, irrep
    product = XOR of 0-based labels); ``blocks`` are dense C1 arrays with the
    orbitals irrep-sorted inside the occupied and the virtual space.
    """
    _p = 'aces_sym_'

    def __init__(self, occ, vir, blocks, stream=None):
        import torch
        self.occ, self.vir = list(occ), list(vir)
        self.o, self.v = sum(occ), sum(vir)
        self.torch = torch
        self.stream = stream
        nirr = len(occ)
        h = ctypes.c_void_p()
        O = (ctypes.c_int * nirr)(*occ)
        V = (ctypes.c_int * nirr)(*vir)
        _lib.call('aces_sym_create', nirr, O, V, stream, ctypes.byref(h))
        self.h = h
        dev = {k: torch.as_tensor(x, device='cuda', dtype=torch.float64).contiguous() for k, x in blocks.items()
               if k in ('fock', 'ovov', 'oovv', 'ovoo', 'oooo', 'ovvv', 'list233')}
        _lib.call('aces_sym_setup', self.h, dev['fock'].data_ptr(), dev['ovov'].data_ptr(), dev['oovv'].data_ptr(),
                  dev['ovoo'].data_ptr(), dev['oooo'].data_ptr(), dev['ovvv'].data_ptr(), dev['list233'].data_ptr())
        del dev

    def energy(self, t1=None, t2=None):
        """Energy of the engine's current amplitudes (after set_amplitudes when given)."""
        if t1 is not None:
            self.set_amplitudes(t1, t2)
        e = ctypes.c_double()
        _lib.call('aces_sym_energy', self.h, ctypes.byref(e))
        return e.value

    @property
    def packed_t2(self):
        return _lib.lib().aces_sym_packed_t2(self.h)

    def run(self, maxcyc=50, conv=7, order=6, start='mp1', prec='fp64', log=None, raise_on_fail=True,
            polish_conv=None, polish='refine'):
        torch = self.torch
        self.set_diis_order(order)
        self.reset_diis()
        if start == 'mp1':
            self.start_mp1()
        elif start == 'zero':
            self.set_amplitudes(torch.zeros(self.o, self.v), torch.zeros(self.o, self.o, self.v, self.v))
        hist = {'energy': [self.energy()], 'resmiter': [], 'resmdiis': [], 'ms_cycle': [], 'converged': False}
        tol = 10.0 ** (-conv)
        phase, ic0 = prec, 0
        hist['precision'] = []
        for icycle in range(1, maxcyc + 1):
            st = self.cycle(icycle - ic0, phase, tol)
            hist['energy'].append(st['energy'])
            hist['resmiter'].append(st['resmiter'])
            hist['resmdiis'].append(st['resmdiis'])
            hist['ms_cycle'].append(st['ms_cycle'])
            hist['precision'].append(phase)
            if log:
                log(icycle, st)
            if st['converged'] and phase == prec and prec in ('fp32', 'bf16x3') and polish_conv is not None:
                phase, tol, ic0 = polish, 10.0 ** (-polish_conv), icycle
                self.reset_diis()
                if polish == 'refine':
                    self.reset_anchor()
                continue
            if st['converged']:
                hist['converged'] = True
                hist['final_energy'] = st['final_energy']
                hist['cycles'] = icycle
                break
        if not hist['converged'] and raise_on_fail:
            raise NotConverged('@VCC: The Coupled-Cluster equations did not converge!')
        return hist
