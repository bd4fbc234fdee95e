"""java.util.Random replica of the oracle and of the library's host generator, pinned by the
JDK's published outputs (Starfish's shared generator is `new Random(0)`,
core/common/Starfish.java:238)."""
import ctypes as C

import oracle as O
from starfish_amd.plugin import Context


def test_known_answers_of_java_util_random():
    # values printed by any JDK for these seeds
    assert O.java_ints(0, 1)[0] == -1155484576          # new Random(0).nextInt()
    assert O.java_ints(42, 1)[0] == -1170105035         # new Random(42).nextInt()
    d = O.java_doubles(0, 2)
    assert d[0] == 0.730967787376657                    # new Random(0).nextDouble()
    assert d[1] == 0.24053641567148587


def test_next_double_uses_53_bits_from_two_draws():
    # nextDouble = ((next(26) << 27) + next(27)) * 2^-53, i.e. two consecutive nextInt() calls
    # give the same high bits: next(32) >> 6 == next(26) and next(32) >> 5 == next(27)
    ints = O.java_ints(7, 2).astype("int64") & 0xFFFFFFFF
    d = O.java_doubles(7, 1)[0]
    hi = ints[0] >> 6
    lo = ints[1] >> 5
    assert d == ((hi << 27) + lo) * 2.0 ** -53


def test_bounded_ints_are_in_range_and_power_of_two_path():
    v = O.java_ints(3, 1000, bound=10)
    assert v.min() >= 0 and v.max() < 10
    p = O.java_ints(3, 1000, bound=16)
    assert p.min() >= 0 and p.max() < 16


def test_library_generator_scrambles_the_seed_like_the_jdk():
    ctx = Context(-1)
    try:
        st = C.c_int64()
        ctx.call("sf_ctx_set_rng", 0, 0)
        ctx.call("sf_ctx_get_rng_state", C.byref(st))
        assert st.value == 0x5DEECE66D                  # (seed ^ 0x5DEECE66D) & (2^48 - 1)
        ctx.call("sf_ctx_set_rng", 42, 0)
        ctx.call("sf_ctx_get_rng_state", C.byref(st))
        assert st.value == (42 ^ 0x5DEECE66D)
    finally:
        ctx.close()
