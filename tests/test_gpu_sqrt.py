"""The engine's FP32 square root (the reference's sqrtf in length(),
src/vector_arithmetic.h) against the correctly rounded root, on the device,
for every FP32 bit pattern: positive and negative values, zeros,
subnormals, infinities and NaNs."""

import pytest

from sphere_amd import engine


@pytest.mark.gpu
def test_fp32_sqrt_matches_correct_rounding_for_every_value():
    assert engine.selftest_sqrt32(0, 1 << 32) == 0
