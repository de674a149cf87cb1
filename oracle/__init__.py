"""CPU oracle for the closed-shell CCSD amplitude path of ACES II ``xvcc``.

TEST INFRASTRUCTURE ONLY.  Nothing in the product library (``aces-ii_amd/``)
may import, call or link anything in this package; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and there only as the checker.

Modules
-------
spinorb   Stanton-Gauss spin-orbital CCSD residual (the general form of the
          equations ``vcc/genint.F`` and ``vcc/vcc.F:798-1437`` evaluate).
brute     Second-quantised brute force: <mu| exp(-T) H exp(T) |0> in the full
          determinant space.  Pins ``spinorb`` for arbitrary (T1, T2).
rccsd     Closed-shell (RHF, spatial-orbital) restatement used for the larger
          parity cases and the CPU baseline; pinned against ``spinorb``.
vcc_loop  The reference's Jacobi/DIIS iteration (NEWT2, CMPENG, DODIIS0,
          DRMOVE, SYMETRIZE_T2, ABTOAA) driving ``rccsd``.
fci       Determinant-space FCI for two-electron checks (CCSD == FCI).
lambda_   Lambda (adjoint) equations and the response one-particle density,
          restated from the Lagrangian; pinned by finite differences.
"""
