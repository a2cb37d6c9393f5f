"""Integer identities that HIP kernels use in place of the reference's
arithmetic, checked exhaustively on the CPU.  Each names the kernel that
relies on it.  CPU only."""

import numpy as np


def byte_lerp(x, y, r):
    """v_lerp_u8 on one byte lane: (x + y + r) >> 1, r in {0, 1}."""
    return (x + y + r) >> 1


def test_three_to_one_lerp_from_byte_means():
    # resize_body.h up2x4_group: (3a + b + 2) >> 2, the generic lerp
    # (src/resize/resize_linear_generic_u8_neon.h) at fraction 64 of 256, as
    # (a + m + r) >> 1 with m = (a + b + 1) >> 1 and r = 1 when a + b is
    # even.
    a, b = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    m = byte_lerp(a, b, 1)
    r = 1 - ((a ^ b) & 1)
    assert np.array_equal(byte_lerp(a, m, r), (3 * a + b + 2) >> 2)
    # the mean and the rounding bit are symmetric in a and b, so one pair
    # serves L(a, b, 64) and L(a, b, 192) = L(b, a, 64)
    assert np.array_equal(byte_lerp(b, m, r), (a + 3 * b + 2) >> 2)


def test_three_to_one_lerp_rounding_bit_is_needed():
    # the rounding bit is not constant: with r = 0 or r = 1 throughout,
    # some byte pairs come out wrong
    a, b = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    m = byte_lerp(a, b, 1)
    want = (3 * a + b + 2) >> 2
    assert not np.array_equal(byte_lerp(a, m, 0), want)
    assert not np.array_equal(byte_lerp(a, m, 1), want)


def test_median5_network_header_is_the_checked_generator_output(tmp_path):
    # median5_net.h must be exactly what gen_median5_net.py writes, and the
    # generator's network must equal numpy's median on every window of
    # random and low-entropy 6x8 footprints (its own check, run here)
    import importlib.util
    import os
    csrc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "yocto-mirrors_gitlab_kleidi_kleidicv_amd", "csrc")
    spec = importlib.util.spec_from_file_location("gen_median5_net",
                                                  os.path.join(csrc, "gen_median5_net.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    net, outputs = gen.build()
    live = gen.live_ops(net, outputs)
    gen.check(net, live, outputs)
    out = tmp_path / "median5_net.h"
    gen.emit(net, live, outputs, str(out))
    with open(os.path.join(csrc, "median5_net.h")) as fh:
        assert out.read_text() == fh.read()
    assert len(live) == 558 and len(outputs) == 8
