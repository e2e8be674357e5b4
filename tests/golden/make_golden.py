"""Write the small golden fixtures of the FP64 oracle (inputs and outputs).

The reference ships no golden vectors for this path and cannot run here
(Julia + FFTW), so these fixtures are the oracle's own results, pinned by the
closed-form tests in tests/test_oracle.py; they guard against drift of the
oracle and give the GPU tests fixed expected values.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gbg_oracle as O  # noqa: E402


def main():
    # 8x7x6: transforms and operators of a seeded field (odd and even axes)
    g = O.Grid3D((8, 7, 6))
    m = O.Multipliers.build(g)
    f = np.random.default_rng(6006).standard_normal((3,) + g.physical_shape)
    F = O.to_spectral(f)
    np.savez(os.path.join(HERE, "ops_8x7x6.npz"), phys=f, spec=F,
             curl=O.curl(F, m), div=O.divergence(F, m), lap=O.laplacian(F, m),
             ilap=O.inverse_laplacian(F, m), proj=O.helmholtz_projection(F, m),
             grad0=O.gradient(F[0], m), icurl=O.inverse_curl(F, m),
             inner=np.array(O.inner_spectral(F, F, g)))
    # 16^3: the u x omega term of the seeded solenoidal field
    g = O.Grid3D((16, 16, 16))
    m = O.Multipliers.build(g)
    u = O.synthetic_field(g, 2001, vector=True, m=m)
    nh, E, Z = O.uxw(u, g, O.dealias_cutoff(16), m)
    np.savez(os.path.join(HERE, "uxw_16.npz"), u_hat=u, n_hat=nh, energy=np.array(E),
             enstrophy=np.array(Z))
    # 64^3: the 2/3 mask, bit-packed, and checksums of the config-1 pipeline
    g = O.Grid3D((64, 64, 64))
    m = O.Multipliers.build(g)
    kc = O.dealias_cutoff(64)
    mask = O.band_mask(m, 0.0, kc).astype(bool)
    a = O.synthetic_field(g, 1001, m=m)
    b = O.synthetic_field(g, 1002, m=m)
    p = O.dealiased_product(a, b, g, kc, m)
    np.savez(os.path.join(HERE, "config1_64.npz"), mask_bits=np.packbits(mask.ravel()),
             mask_shape=np.array(mask.shape), kept=np.array(int(mask.sum())),
             a_norm=np.array(np.linalg.norm(a)), b_norm=np.array(np.linalg.norm(b)),
             p_norm=np.array(np.linalg.norm(p)), p_first=p[:2, :3, :4],
             grad_norm=np.array(np.linalg.norm(O.gradient(a, m))))

    # the 2/3 mask at the headline sizes and the batch sizes: kept-mode count
    # and SHA-256 of the bit-packed mask (storage order)
    import json
    shapes = [(64, 64, 64), (512, 512, 512), (1024, 1024, 1024), (32, 32, 32), (48, 48, 48),
              (96, 96, 96), (128, 128, 128), (128, 128, 256), (256, 64, 64), (256, 256, 256)]
    out = {}
    for n in shapes:
        g = O.Grid3D(n)
        m = O.Multipliers.build(g)
        kc = O.dealias_cutoff(min(n))
        kept, digest = O.mask_digest(m, kc)
        out["x".join(map(str, n))] = {"kc": kc, "kept": kept, "sha256": digest}
    with open(os.path.join(HERE, "mask_digests.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
