"""The oracle against its committed golden fixtures (CPU)."""

import os

import numpy as np

from oracle import gbg_oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_ops_8x7x6():
    d = load("ops_8x7x6.npz")
    g = O.Grid3D((8, 7, 6))
    m = O.Multipliers.build(g)
    F = O.to_spectral(d["phys"])
    np.testing.assert_allclose(F, d["spec"], rtol=0, atol=1e-12)
    for name, val in (("curl", O.curl(F, m)), ("div", O.divergence(F, m)), ("lap", O.laplacian(F, m)),
                      ("ilap", O.inverse_laplacian(F, m)), ("proj", O.helmholtz_projection(F, m)),
                      ("grad0", O.gradient(F[0], m)), ("icurl", O.inverse_curl(F, m))):
        np.testing.assert_allclose(val, d[name], rtol=0, atol=1e-11, err_msg=name)
    assert abs(O.inner_spectral(F, F, g) - float(d["inner"])) < 1e-12 * float(d["inner"])


def test_uxw_16():
    d = load("uxw_16.npz")
    g = O.Grid3D((16, 16, 16))
    m = O.Multipliers.build(g)
    u = O.synthetic_field(g, 2001, vector=True, m=m)
    np.testing.assert_array_equal(u, d["u_hat"])
    nh, E, Z = O.uxw(u, g, O.dealias_cutoff(16), m)
    assert O.rel_l2(nh, d["n_hat"]) < 1e-13
    assert abs(E - float(d["energy"])) < 1e-13 and abs(Z - float(d["enstrophy"])) < 1e-12


def test_config1_64():
    d = load("config1_64.npz")
    g = O.Grid3D((64, 64, 64))
    m = O.Multipliers.build(g)
    kc = O.dealias_cutoff(64)
    mask = O.band_mask(m, 0.0, kc).astype(bool)
    bits = np.unpackbits(d["mask_bits"])[: mask.size].astype(bool).reshape(tuple(d["mask_shape"]))
    assert np.array_equal(mask, bits)
    assert int(mask.sum()) == int(d["kept"])
    a = O.synthetic_field(g, 1001, m=m)
    b = O.synthetic_field(g, 1002, m=m)
    assert abs(np.linalg.norm(a) - float(d["a_norm"])) < 1e-12 * float(d["a_norm"])
    p = O.dealiased_product(a, b, g, kc, m)
    assert abs(np.linalg.norm(p) - float(d["p_norm"])) < 1e-12 * float(d["p_norm"])
    np.testing.assert_allclose(p[:2, :3, :4], d["p_first"], rtol=0, atol=1e-9)


def test_mask_digests():
    """The 2/3 mask at the headline and batch sizes: the oracle's bit-packed
    mask digest equals the committed one, and the kept-mode count equals an
    integer lattice count (n1 >= 0 on the folded axis, |n|^2 <= floor(kc^2)
    on 2 pi boxes) done without the oracle."""
    import json
    import math
    with open(os.path.join(G, "mask_digests.json")) as fh:
        digests = json.load(fh)
    assert len(digests) >= 10
    for key, want in digests.items():
        n = tuple(int(v) for v in key.split("x"))
        g = O.Grid3D(n)
        m = O.Multipliers.build(g)
        kc = O.dealias_cutoff(min(n))
        assert kc == want["kc"]
        kept, digest = O.mask_digest(m, kc)
        assert (kept, digest) == (want["kept"], want["sha256"]), key
        R = math.floor(kc * kc + 1e-9)
        count = 0
        for n3 in range(-(n[2] // 2), (n[2] + 1) // 2):
            for n2 in range(-(n[1] // 2), (n[1] + 1) // 2):
                r = R - n2 * n2 - n3 * n3
                if r >= 0:
                    count += min(n[0] // 2, math.isqrt(r)) + 1
        assert count == kept, key
