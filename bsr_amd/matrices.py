"""Host-side containers for the matrices of one partial wave (module
bsr_matrix of the reference, src/bsr/BSR_MAT3/mod_intmat.f90:2-29)."""
from __future__ import annotations

import numpy as np


class WaveMatrices:
    """hcc(ns,ns,nch(nch+1)/2), hcb(ns,nch,ncp), hbb(ncp(ncp+1)/2),
    acf(nch,nch,0:mk), htarg, otarg; all Fortran-ordered."""

    def __init__(self, ns, nch, ncp, mk):
        nij = nch * (nch + 1) // 2
        self.ns, self.nch, self.ncp, self.mk = ns, nch, ncp, mk
        self.hcc = np.zeros((ns, ns, nij), order="F")
        self.hcb = np.zeros((ns, nch, max(ncp, 1)), order="F")
        self.hbb = np.zeros(max(ncp * (ncp + 1) // 2, 1), order="F")
        self.acf = np.zeros((nch, nch, mk + 1), order="F")
        self.htarg = np.zeros(nij, order="F")
        self.otarg = np.zeros(nij, order="F")

    def dense(self):
        """Full symmetric nhm x nhm matrix, nhm = nch*ns + ncp."""
        ns, nch, ncp = self.ns, self.nch, self.ncp
        n = nch * ns + ncp
        A = np.zeros((n, n))
        for i in range(1, nch + 1):
            for j in range(1, i + 1):
                b = self.hcc[:, :, i * (i - 1) // 2 + j - 1]
                A[(i - 1) * ns: i * ns, (j - 1) * ns: j * ns] = b
                if i != j:
                    A[(j - 1) * ns: j * ns, (i - 1) * ns: i * ns] = b.T
        if ncp:
            V = self.hcb.reshape(ns * nch, ncp, order="F")
            A[: nch * ns, nch * ns:] = V
            A[nch * ns:, : nch * ns] = V.T
            P = np.zeros((ncp, ncp))
            q = 0
            for ip in range(ncp):
                for jp in range(ip + 1):
                    P[ip, jp] = P[jp, ip] = self.hbb[q]
                    q += 1
            A[nch * ns:, nch * ns:] = P
        return A
