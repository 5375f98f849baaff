"""Learned separable bases for the transform-coded snapshots.

The snapshot coder (rtmhip_snap_dct_*) transforms 8 x 8 x 32 bricks of each
snapshot with one orthonormal matrix per axis and keeps the snapshot's largest
coefficients.  The DCT-II is its default basis.  A learned basis replaces it
with the principal directions of the snapshot lines themselves: the
eigenvectors of the second-moment matrix of every 32-cell z line, 8-cell x line
and 8-cell y line of the training snapshots' bricks (the factor matrices of a
higher-order SVD of the brick population), strongest first.  It is trained on
other shots' snapshots of the same model and passed to the forward call, which
keeps it in the store for the backward pass.

    learner = BasisLearner()
    for region in training_snapshots:          # [nxe][nye][nze - 2 blen]
        learner.add(region, blen)
    basis = learner.basis()                    # float32 [32*32 + 8*8 + 8*8]
"""
import numpy as np

NZ, NXY = 32, 8
BASIS_SIZE = NZ * NZ + 2 * NXY * NXY


class BasisLearner:
    """Second moments of the brick lines of the snapshots added so far."""

    def __init__(self):
        self.cz = np.zeros((NZ, NZ))
        self.cx = np.zeros((NXY, NXY))
        self.cy = np.zeros((NXY, NXY))
        self.lines = 0

    def add(self, region, blen):
        """Adds one snapshot region ([nxe][nye][nze - 2 blen], the layout of a
        snapshot buffer slot, whose first z plane is z = blen of the grid).
        Bricks are placed as the coder places them: x, y from 0 and z at
        multiples of 32 of the grid, cells outside the region zero."""
        r = np.asarray(region)
        nx, ny, lz = r.shape
        zoff = blen % NZ
        pz = -(-(zoff + lz) // NZ) * NZ
        py = -(-ny // NXY) * NXY
        for x0 in range(0, nx, NXY):
            slab = np.zeros((NXY, py, pz))
            slab[:min(NXY, nx - x0), :ny, zoff:zoff + lz] = r[x0:x0 + NXY]
            # [8 x][py/8][8 y][pz/32][32 z] -> bricks [nb][8][8][32]
            b = slab.reshape(NXY, py // NXY, NXY, pz // NZ, NZ).transpose(1, 3, 0, 2, 4)
            b = b.reshape(-1, NXY, NXY, NZ)
            zl = b.reshape(-1, NZ)
            self.cz += zl.T @ zl
            xl = b.transpose(0, 2, 3, 1).reshape(-1, NXY)
            self.cx += xl.T @ xl
            yl = b.transpose(0, 1, 3, 2).reshape(-1, NXY)
            self.cy += yl.T @ yl
            self.lines += zl.shape[0]

    @staticmethod
    def _rows(c):
        w, v = np.linalg.eigh(c)
        rows = v[:, ::-1].T.copy()
        # deterministic signs: each vector's largest component positive
        big = np.abs(rows).argmax(axis=1)
        rows *= np.sign(rows[np.arange(len(rows)), big])[:, None]
        return rows

    def basis(self):
        """The learned basis packed for the coder: z rows (32 x 32), then x
        rows (8 x 8), then y rows (8 x 8), row-major float32, each row a basis
        vector, strongest direction first."""
        if self.lines == 0 or not np.trace(self.cz) > 0:
            raise ValueError("no non-zero training lines")
        return np.concatenate([self._rows(self.cz).ravel(), self._rows(self.cx).ravel(),
                               self._rows(self.cy).ravel()]).astype(np.float32)


def dct_basis():
    """The coder's default basis (orthonormal DCT-II), packed like basis()."""
    def d(n):
        i = np.arange(n)
        m = np.cos(np.pi * (2 * i[None, :] + 1) * i[:, None] / (2 * n)) * np.sqrt(2.0 / n)
        m[0] /= np.sqrt(2.0)
        return m
    return np.concatenate([d(NZ).ravel(), d(NXY).ravel(), d(NXY).ravel()]).astype(np.float32)


def unpack(basis):
    """(z [32][32], x [8][8], y [8][8]) of a packed basis."""
    b = np.asarray(basis, np.float64).ravel()
    assert b.size == BASIS_SIZE
    return (b[:NZ * NZ].reshape(NZ, NZ), b[NZ * NZ:NZ * NZ + 64].reshape(NXY, NXY),
            b[NZ * NZ + 64:].reshape(NXY, NXY))
