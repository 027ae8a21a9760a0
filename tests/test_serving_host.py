"""Host-side pieces of the serving layer (no GPU): the request stream and the
autotuner's key."""
import numpy as np

from mtjgpu import serving as S


def test_request_stream_is_deterministic_and_bounded():
    a = S.request_stream(64, seed=5)
    b = S.request_stream(64, seed=5)
    assert a == b
    arr = [r["arrival"] for r in a]
    assert arr == sorted(arr) and arr[0] == 0.0
    for r in a:
        nx, ny, nz = r["shape"]
        assert r["dim"] in (2, 3)
        assert (nz == 1) == (r["dim"] == 2)
        assert 500 <= nx * ny * nz <= 2 * 262144


def test_shape_key_separates_stencils_and_scales():
    def grid(dim, m):
        from mtjgpu import problems as P
        n, _, rp, ci, va = P.poisson2d(m) if dim == 2 else P.poisson3d(m)
        return n, rp, ci, va, np.ones(n)
    k2 = S.shape_key([grid(2, 64)])
    k3 = S.shape_key([grid(3, 16)])
    # 2-D 5-point rows hold at most 5 entries, 3-D 7-point rows up to 7
    assert k2[2] == 0.0 and k3[2] > 0.5
    assert abs(sum(k2[:5]) - 1.0) < 0.11 and abs(sum(k3[:5]) - 1.0) < 0.11
    assert S.shape_key([grid(2, 64)]) == k2  # deterministic
    assert S.shape_key([grid(2, 256)])[5] > k2[5]  # the scale term grows with n
