"""The byte model behind bench.py's roofline fractions, on the CPU.

`kept_ranges` (the index ranges the fused u x omega schedule keeps) is pinned
to the oracle's 2/3 mask (`band_mask`, the restatement of the reference's
band predicate): the kept ball, the (kx, ky) pencils holding a kept mode, the
kept kx and ky counts.  `uxw_stage_bytes` then moves five fields from the z
pass to the y pass where Ny <= 512 and six above, which is the schedule the
library runs (the y pass forms the kx and ky terms of w where Ny <= 512)."""

import numpy as np
import pytest

import bench
from oracle import gbg_oracle as O


@pytest.mark.parametrize("n", [(32, 32, 32), (48, 32, 40), (16, 64, 24), (64, 16, 16)])
def test_kept_ranges_match_the_oracle_mask(n):
    g = O.Grid3D(n)
    m = O.Multipliers.build(g)
    kc = O.dealias_cutoff(n[0])
    mask = O.band_mask(m, 0.0, kc) > 0            # (Nz, Ny, Nh)
    r = bench.kept_ranges(n, kc)
    assert r["ball"] == int(mask.sum())
    pencils = np.any(mask, axis=0)                 # (Ny, Nh): a kept mode on the pencil
    assert r["disk"] == int(pencils.sum())
    kx_kept = np.any(pencils, axis=0)
    assert r["nhk"] == int(kx_kept.sum()) and kx_kept[:r["nhk"]].all()
    assert r["nyc"] == int(np.any(pencils, axis=1).sum())


def test_stage_bytes_five_fields_up_to_ny_512():
    for n, nf in (((64, 512, 64), 5), ((64, 1024, 64), 6)):
        kc = n[0] / 3
        r = bench.kept_ranges(n, kc)
        sb = bench.uxw_stage_bytes(n, kc, 8)
        disk = n[2] * r["disk"] * 16
        assert sb["K1_z_inverse_curl"] - 3 * r["ball"] * 16 == nf * disk
        assert sb["K2_y_inverse"] - 6 * n[2] * n[1] * r["nhk"] * 16 == nf * disk
        # FP32 moves half the bytes of FP64
        assert bench.uxw_stage_bytes(n, kc, 4) == {k: v // 2 for k, v in sb.items()}
