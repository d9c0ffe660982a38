"""The mixed form decides pairs near the cutoff as the FP64 form does.

Quantized positions move each coordinate by up to half a step (L / 2^33),
so a pair within that distance of the cutoff could be taken by one form and
dropped by the other; the force of sep_lj_shift jumps by F(rc) = -0.039 at
rc = 2.5, which would put the whole jump into the force error.  Pairs that
close are decided on the FP64 records.  Two atoms placed just inside and
just outside the cutoff (by 1e-11, far below the quantization step of
1.2e-9 at L = 20): the mixed form takes exactly the pairs the FP64 form
takes, and its force is the FP64 one to FP32 precision.
"""
import numpy as np
import pytest

from seplib_amd import sep as S
from seplib_amd.sim import SepSystem

pytestmark = pytest.mark.gpu

RC = 2.5


def _pair_force(product, r, precision, resident):
    L = 20.0
    # an irrational-looking offset, so the quantized positions are not exact
    x0 = np.array([5.1234567891, 7.3456789012, 9.5678901234])
    d = np.array([0.6, -0.48, 0.64])  # unit vector
    w = dict(x=np.array([x0, x0 + r * d]), v=np.zeros((2, 3)), m=np.ones(2), q=np.zeros(2),
             type=np.array([b"A", b"A"]), L=np.array([L, L, L]), mol=np.full(2, -1))
    s = SepSystem.from_workload(product, w, cf=RC, dt=0.005, skin=0.3)
    try:
        s.gpu_precision(precision)
        if resident:
            s.gpu_resident(True)
        s.reset_retval()
        s.reset_force()
        s.force_pairs("AA", RC, "sep_lj_shift")
        if resident:
            s.gpu_sync()
        return np.array(s.view["f"]), s.ret.epot
    finally:
        s.close()


@pytest.mark.parametrize("resident", [False, True], ids=["sync", "resident"])
def test_mixed_takes_the_pairs_fp64_takes(product, resident):
    for r in (RC * (1.0 - 4e-12), RC * (1.0 + 4e-12), RC - 1e-11, RC + 1e-11):
        f64, e64 = _pair_force(product, r, S.SEP_GPU_FP64, resident)
        f32, e32 = _pair_force(product, r, S.SEP_GPU_MIXED, resident)
        assert (np.abs(f64).max() > 0) == (np.abs(f32).max() > 0), r
        if np.abs(f64).max() > 0:
            # F(rc) = 48 rc^-13 - 24 rc^-7 in magnitude, about 0.039
            assert abs(np.linalg.norm(f64[0]) - 0.0389994) < 1e-4
            assert np.abs(f32 - f64).max() <= 1e-6 * np.abs(f64).max(), r
