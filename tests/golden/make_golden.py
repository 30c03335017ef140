"""Writes the small golden fixtures under tests/golden/ from the FP64 oracle.

The reference repository ships no expected outputs for this path, so the
fixtures are the oracle's own vectors, each cross-checked here against the
dense closed-form operator before it is written.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
from oracle import ddscat_oracle as O  # noqa: E402


def main():
    rng = np.random.default_rng(20261202)
    dims = (4, 3, 5)
    nat = int(np.prod(dims))
    akd = 0.41
    iocc = (rng.random(nat) < 0.7).astype(np.int16)
    iocc[0] = 1
    cx = lambda s: s * (rng.normal(size=3 * nat) + 1j * rng.normal(size=3 * nat))
    x = cx(1.0) * np.tile(iocc != 0, 3)
    adia, aoff = cx(1.0), cx(0.3)
    go = O.green_octant(*dims, akd)
    yN = O.matvec_fft("N", x, adia, aoff, go, *dims, iocc)
    yC = O.matvec_fft("C", x, adia, aoff, go, *dims, iocc)
    A = O.dense_operator(*dims, akd, adia, aoff, iocc)
    assert np.allclose(A @ x, yN, rtol=1e-13, atol=1e-13)
    np.savez_compressed(HERE / "matvec_small.npz", dims=np.array(dims), akd=np.array(akd), iocc=iocc, x=x,
                        adia=adia, aoff=aoff, green_octant=go, y_N=yN, y_C=yC)
    print("wrote", HERE / "matvec_small.npz")


if __name__ == "__main__":
    main()
