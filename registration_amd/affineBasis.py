"""Affine bases of the global affine part of a registration, drop-in for py-lddmm's
affineBasis module (py-lddmm/affineBasis.py:3-94).

AffineBasis(dim, affine) holds a (dim^2 + dim, affineDim) matrix whose columns span the
allowed transforms, stacked as [vec(A); b]: 'affine' (symmetric trace-free, scaling, rotation,
translation; affineDim = 2 (dim(dim-1)/2 + dim)), 'similitude' (scaling, rotation,
translation), 'euclidean' (rotation, translation), 'translation', anything else = none
(affineDim 0).  rotComp / simComp / affComp / transComp index the columns of each kind.  The
reference indexes one entry with a float ((i+1.), affineBasis.py:41), which NumPy no longer
accepts; the index is the same integer here.
"""
from __future__ import annotations

import numpy as np


class AffineBasis:
    def __init__(self, dim=3, affine='affine'):
        u = 1.0 / np.sqrt(2.0)
        dimSym = (dim * (dim - 1)) // 2
        codes = {'affine': 1, 'similitude': 2, 'euclidean': 3, 'translation': 4}
        affCode = codes.get(affine, 5)
        self.dim = dim
        self.rotComp, self.simComp, self.affComp, self.transComp = [], [], [], []
        self.affineDim = {1: 2 * (dimSym + dim), 2: dimSym + 1 + dim, 3: dimSym + dim,
                          4: dim, 5: 0}[affCode]
        if affCode == 5:
            self.basis = []
            return
        self.basis = np.zeros([2 * (dimSym + dim), self.affineDim])
        k = 0
        if affCode <= 1:
            for i in range(dimSym):
                for j in range(i + 1, dim):
                    self.basis[i * dim + j, k] = u
                    self.basis[j * dim + i, k] = u
                    k += 1
            for i in range(dim - 1):
                uu = np.sqrt((1 - 1.0 / (i + 2))) / (i + 1.)
                self.basis[(i + 1) * dim + (i + 1), k] = np.sqrt(1 - 1.0 / (i + 2))
                for j in range(i + 1):
                    self.basis[j * dim + j, k] = -uu
                k += 1
            self.affComp = list(range(k))
        if affCode <= 2:
            k0 = k
            for i in range(dim):
                self.basis[i * dim + i, k] = 1. / np.sqrt(dim)
            k += 1
            self.simComp = list(range(k0, k))
        if affCode <= 3:
            k0 = k
            for i in range(dim):
                for j in range(i + 1, dim):
                    self.basis[i * dim + j, k] = u
                    self.basis[j * dim + i, k] = -u
                    k += 1
            self.rotComp = list(range(k0, k))
        k0 = k
        for i in range(dim):
            self.basis[i + dim ** 2, k] = 1
            k += 1
        self.transComp = list(range(k0, k))

    def getTransforms(self, Afft):
        """[A (T, dim, dim), b (T, dim)] of the coefficients Afft (T, affineDim)
        (affineBasis.py:67-79)."""
        Afft = np.asarray(Afft)
        Tsize = Afft.shape[0]
        dim2 = self.dim ** 2
        A = [np.zeros([Tsize, self.dim, self.dim]), np.zeros([Tsize, self.dim])]
        if self.affineDim > 0:
            AB = (self.basis[np.newaxis, :, :] * Afft[:, np.newaxis, :]).sum(axis=2)
            A[0] = AB[:, 0:dim2].reshape([Tsize, self.dim, self.dim])
            A[1] = AB[:, dim2:dim2 + self.dim]
        return A

    def integrateFlow(self, Afft):
        """The composed affine map [X (T+1, dim, dim), x (T+1, dim)] of the Euler steps
        I + dt A_t (affineBasis.py:81-94)."""
        Afft = np.asarray(Afft)
        Tsize = Afft.shape[0]
        X = [np.zeros([Tsize + 1, self.dim, self.dim]), np.zeros([Tsize + 1, self.dim])]
        A = self.getTransforms(Afft)
        dt = 1.0 / Tsize
        eye = np.eye(self.dim)
        X[0][0] = eye
        for t in range(Tsize):
            B = eye + dt * A[0][t]
            X[0][t + 1] = np.dot(B, X[0][t])
            X[1][t + 1] = np.dot(B, X[1][t]) + dt * A[1][t]
        return X
