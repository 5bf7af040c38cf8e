// osb_eigzero.hpp -- the structural zeros of the Euler eigenvector matrices
// of direction DIR: entries that are zero for every state (never set, or a
// product with a zero direction selector k_d).  The face kernels skip their
// terms in the characteristic projections.  Plain constexpr integer code,
// shared with a host check against the oracle's eigensystem.
#pragma once

namespace osb {

// Entries of the left eigenvectors that are zero for every state: never set,
// or a product with a zero direction selector k_d.  2-D: the shear row's last
// entry and its entry 1 + DIR.  3-D: in rows r = 0..2, entry 4 unless DIR = r,
// and entry 1 + j unless DIR = r or (r != j and DIR = 3 - r - j).
template <int NDIM, int DIR>
__host__ __device__ constexpr bool l_zero(int r, int m) {
    if (NDIM == 2) return r == 1 && (m == 3 || m == 1 + DIR);
    if (NDIM == 3 && r <= 2 && DIR != r) {
        if (m == 4) return true;
        if (m >= 1) {
            const int j = m - 1;
            return r == j || DIR != 3 - r - j;
        }
    }
    return false;
}

// The same for the right eigenvectors R[m][k].  2-D: in column 1, row 0 and
// row 1 + DIR.  3-D, in columns k = 0..2: row 0 where k != DIR; in row
// m = 1 + c of momentum component c, every k != DIR when c = DIR, and k = c
// alone when c != DIR (six entries).
template <int NDIM, int DIR>
__host__ __device__ constexpr bool r_zero(int m, int k) {
    if (NDIM == 2) return k == 1 && (m == 0 || m == 1 + DIR);
    if (NDIM == 3 && k <= 2) {
        if (m == 0) return k != DIR;
        if (m <= 3) return m - 1 == DIR ? k != DIR : k == m - 1;
    }
    return false;
}

}  // namespace osb
