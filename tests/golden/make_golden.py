"""Writes tests/golden/transforms.npz: scipy's DST-II/III, DCT-II/III and
numpy's FFT of seeded complex vectors and 2D fields.

scipy.fft's unnormalised dst/dct types 2 and 3 are FFTW's RODFT10/RODFT01 and
REDFT10/REDFT01 (the kinds the reference plans, transformer.cpp:112-151).
Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20120131)
    out = {}
    for n in (8, 16, 40, 64, 192, 256):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x{n}"] = x
        out[f"fft{n}"] = np.fft.fft(x)
        out[f"ifft{n}"] = np.fft.ifft(x) * n
        for name, f, t in (("RODFT10", sf.dst, 2), ("RODFT01", sf.dst, 3), ("REDFT10", sf.dct, 2),
                           ("REDFT01", sf.dct, 3)):
            out[f"{name}{n}"] = f(x.real, type=t) + 1j * f(x.imag, type=t)
    # 2D fields: 16x16 (the reference's transformer test grid) and 40x24
    for (sy, sx) in ((16, 16), (24, 40)):
        f = rng.standard_normal((sy, sx)) + 1j * rng.standard_normal((sy, sx))
        key = f"{sy}x{sx}"
        out[f"field{key}"] = f

        def both(fn, t, axes):
            return fn(f.real, type=t, axes=axes) + 1j * fn(f.imag, type=t, axes=axes)

        out[f"FFT{key}"] = np.fft.fft2(f)
        out[f"iFFT{key}"] = np.fft.ifft2(f) * f.size
        out[f"FFTx{key}"] = np.fft.fft(f, axis=1)
        out[f"FFTy{key}"] = np.fft.fft(f, axis=0)
        out[f"DST{key}"] = both(sf.dstn, 2, (0, 1))
        out[f"iDST{key}"] = both(sf.dstn, 3, (0, 1))
        out[f"DSTx{key}"] = both(sf.dstn, 2, (1,))
        out[f"iDSTy{key}"] = both(sf.dstn, 3, (0,))
        out[f"DCT{key}"] = both(sf.dctn, 2, (0, 1))
        out[f"iDCTx{key}"] = both(sf.dctn, 3, (1,))
    # Hermitian matrices and their eigenvalues (LAPACK via scipy.linalg.eigh)
    import scipy.linalg as sl
    for n in (2, 8, 33):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        A = A + A.conj().T
        out[f"herm{n}"] = A
        out[f"eig{n}"] = sl.eigh(A, eigvals_only=True)
    np.savez_compressed(os.path.join(HERE, "transforms.npz"), **out)


if __name__ == "__main__":
    main()
