"""The fast MVS kernel's divide, reciprocal, square root and constant divide
(csrc/rn_ops.hpp) against the compiler's IEEE expansions, bit for bit, over
2^31 random operand sets drawn across the ranges the kernel relies on.  A
single differing bit would break the kernel's bitwise agreement with the
CPU reference."""
import pytest

pytestmark = pytest.mark.gpu


def test_range_ops_equal_ieee(hip):
    for seed in (1, 0x5eed):
        bad = hip.check_arith(1 << 31, seed)
        print(f"seed {seed:#x}: mismatches {bad}", flush=True)
        assert bad == {"div": 0, "rcp": 0, "sqrt": 0, "div_const": 0}
