"""Extract known-answer data from the reference's own tests into JSON fixtures.

Run in the build container, where the reference tree is mounted:

    python tests/golden/extract_fixtures.py /root/reference

It reads only the literal numbers of the tests (inputs and expected outputs)
and writes them to tests/golden/reference_kats.json.  No reference source
text is stored.

Sources (paths relative to the reference root):
  test/api/test_gaussian_blur.cpp:717-881        ValidImageSize{3,5,7,9}
  test/api/test_separable_filter_2d.cpp:129-262  5x5 u8 overflow KATs
  test/api/test_warp_perspective.cpp:700-742      WarpPerspectiveLinear.SimpleResize
  test/api/test_yuv420sp_to_rgb.cpp:42-53         fixed NV12 inputs
  test/api/test_resize_linear.cpp:1016-1187       ResizeLinearU8Params tables
  test/unit_neon/test_resize_linear_generic_u8.cpp:251-389  generator sizes
  test/api/test_blur_and_downsample.cpp:179-206   BlurAndDownsample.Minimal_u8
  kleidicv/src/conversions/yuv42x_coefficients.h:95-105  YUV -> RGB weights
  test/api/test_yuv420sp_to_rgb.cpp:112-160       the test's integer formula
"""

import json
import os
import re
import sys


def _ints(text):
    return [int(v) for v in re.findall(r"-?\d+", text)]


def _block(text, start_marker):
    """Text of the brace-balanced block that follows start_marker."""
    i = text.index(start_marker)
    j = text.index("{", i)
    depth = 0
    for k in range(j, len(text)):
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return text[j:k + 1]
    raise ValueError(start_marker)


def _sets(block, name):
    """Rows assigned by `name.set(row, col, {...})`, in order of appearance."""
    rows = []
    for m in re.finditer(name + r"\.set\((\d+),\s*(\d+),\s*\{([^}]*)\}\)", block):
        rows.append((int(m.group(1)), int(m.group(2)), _ints(m.group(3))))
    return rows


def gaussian(ref):
    text = open(os.path.join(ref, "test/api/test_gaussian_blur.cpp")).read()
    out = []
    for k in (3, 5, 7, 9):
        block = _block(text, "TYPED_TEST(GaussianBlur, ValidImageSize%dx%d)" % (k, k))
        src = _sets(block, "src")
        exp = _sets(block, "expected")
        n = k - 1
        # the first n expected rows belong to sigma 0, the next n to sigma 2.25
        out.append({
            "kernel": k, "size": n, "border": "REVERSE",
            "src": [r[2] for r in src],
            "expected_sigma0": [r[2] for r in exp[:n]],
            "expected_sigma2_25": [r[2] for r in exp[n:2 * n]],
            "sigma2_25_tolerance": 1,
        })
    return out


def separable(ref):
    text = open(os.path.join(ref, "test/api/test_separable_filter_2d.cpp")).read()
    out = []
    for name in ("5x5_U8OverflowSequence", "5x5_U8OverflowMax",
                 "5x5_U8OverflowVectorNEON", "5x5_U8OverflowVectorSC"):
        block = _block(text, "TEST(SeparableFilter2D, %s)" % name)
        out.append({
            "name": name, "border": "REPLICATE",
            "src": [r[2] for r in _sets(block, "src")],
            "kernel_x": _sets(block, "kernel_x")[0][2],
            "kernel_y": _sets(block, "kernel_y")[0][2],
            "expected": [r[2] for r in _sets(block, "dst_expected")],
        })
    return out


def warp(ref):
    text = open(os.path.join(ref, "test/api/test_warp_perspective.cpp")).read()
    block = _block(text, "TYPED_TEST(WarpPerspectiveLinear, SimpleResize)")
    dims = {m.group(1): int(m.group(2)) for m in
            re.finditer(r"const size_t (\w+) = (\d+);", block)}
    src_rows = _sets(block, "source")
    exp_rows = _sets(block, "expected")
    mat = re.search(r"little_scale\[\]\s*=\s*\{([^}]*)\}", block).group(1)
    src = [[0] * dims["src_w"] for _ in range(dims["src_h"])]
    for r, c, vals in src_rows:
        src[r][c:c + len(vals)] = vals
    exp = [[0] * dims["dst_w"] for _ in range(dims["dst_h"])]
    for r, c, vals in exp_rows:
        exp[r][c:c + len(vals)] = vals
    return {
        "src": src, "expected": exp,
        "transform": [float(v) for v in re.findall(r"-?\d+(?:\.\d+)?", mat)],
        "interpolation": "LINEAR", "border": "REPLICATE",
    }


def nv12(ref):
    text = open(os.path.join(ref, "test/api/test_yuv420sp_to_rgb.cpp")).read()
    block = _block(text, "void execute_test(")
    y = [[0] * 4 for _ in range(5)]
    for r, c, vals in _sets(block, "input_y"):
        y[r][c:c + len(vals)] = vals
    uv = [[0] * 4 for _ in range(3)]
    for r, c, vals in _sets(block, "input_uv"):
        uv[r][c:c + len(vals)] = vals
    return {"y": y, "uv": uv}


def nv12_coefficients(ref):
    """The 20-bit YUV -> RGB weights, read twice: from the library's table
    (kYWeight, kUVWeights) and from the integer formula of the NV12 test's
    expected-value function; the two must agree."""
    text = open(os.path.join(ref, "kleidicv/src/conversions/yuv42x_coefficients.h")).read()
    y = int(re.search(r"kYWeight\s*=\s*/\*[^*]*\*/\s*(-?\d+)", text).group(1))
    uv = _ints(re.sub(r"/\*.*?\*/", "", _block(text, "kUVWeights[4] =")))
    lib = {"y": y, "rv": uv[0], "gu": uv[1], "gv": uv[2], "bu": uv[3]}
    t = open(os.path.join(ref, "test/api/test_yuv420sp_to_rgb.cpp")).read()
    body = _block(t, "void calculate_expected(")
    r = re.search(r"int32_t r = \(\((\d+) \* y\) \+ \((\d+) \* v\) \+ \(1 << (\d+)\)\) >> (\d+)",
                  body)
    g = re.search(r"int32_t g =\s*\(\((\d+) \* y\) - \((\d+) \* u\) - \((\d+) \* v\) \+ "
                  r"\(1 << (\d+)\)\) >> (\d+)", body)
    b = re.search(r"int32_t b = \(\((\d+) \* y\) \+ \((\d+) \* u\) \+ \(1 << (\d+)\)\) >> (\d+)",
                  body)
    test = {"y": int(r.group(1)), "rv": int(r.group(2)), "gu": -int(g.group(2)),
            "gv": -int(g.group(3)), "bu": int(b.group(2)), "shift": int(r.group(4))}
    assert int(g.group(1)) == test["y"] == int(b.group(1))
    assert int(r.group(3)) + 1 == test["shift"] == int(g.group(5)) == int(b.group(4))
    for k in lib:
        assert lib[k] == test[k], (k, lib[k], test[k])
    # the test's offsets: Y' = max(Y - 16, 0), U' = U - 128, V' = V - 128
    assert "std::max(0, y - 16)" in body and "u -= 128" in body and "v -= 128" in body
    test.update({"y_offset": 16, "uv_offset": 128})
    return test


def resize_params(ref):
    text = open(os.path.join(ref, "test/api/test_resize_linear.cpp")).read()
    i = text.index("INSTANTIATE_TEST_SUITE_P(\n    ResizeLinear, ResizeLinearU8Params")
    j = text.index("testing::Values(", i) + len("testing::Values(")
    k = text.index("}}}));", j) + 3
    body = re.sub(r"//[^\n]*", "", text[j:k])
    body = body.replace("Pu8", "").replace("{", "[").replace("}", "]")
    cases = json.loads("[" + body + "]")
    return [{"src": c[0], "expected": c[1]} for c in cases]


def generator_sizes(ref):
    text = open(os.path.join(ref, "test/unit_neon/test_resize_linear_generic_u8.cpp")).read()
    out = []
    for m in re.finditer(r"generator_test<(\d+),\s*(\d+),\s*(\d+)>\((\d*)\)", text):
        out.append({"channels": int(m.group(1)), "src_width": int(m.group(2)),
                    "dst_width": int(m.group(3)),
                    "xfrac_tolerance": int(m.group(4) or 0)})
    return out


def blur_and_downsample(ref):
    text = open(os.path.join(ref, "test/api/test_blur_and_downsample.cpp")).read()
    block = _block(text, "TEST(BlurAndDownsample, Minimal_u8)")
    src = [r[2] for r in sorted(_sets(block, "input"))]
    exp = [r[2] for r in sorted(_sets(block, "expected"))]
    return [{"border": "REPLICATE", "channels": 1, "src": src, "expected": exp}]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = {
        "gaussian_valid_image_size": gaussian(ref),
        "separable_overflow": separable(ref),
        "warp_simple_resize": warp(ref),
        "nv12_fixed_inputs": nv12(ref),
        "resize_u8_params": resize_params(ref),
        "resize_generator_sizes": generator_sizes(ref),
        "blur_and_downsample_minimal": blur_and_downsample(ref),
        "nv12_coefficients": nv12_coefficients(ref),
    }
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_kats.json"), "w") as f:
        json.dump(data, f, indent=None, separators=(",", ":"))
        f.write("\n")
    print("wrote", sum(len(v) if isinstance(v, list) else 1 for v in data.values()),
          "fixture groups")


if __name__ == "__main__":
    main()
