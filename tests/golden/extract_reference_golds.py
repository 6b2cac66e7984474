"""Write reference_golds.json: the golden values of the reference's own unit
tests for the edge-assembly path, transcribed as data.

Run in a container where the reference tree is mounted (it reads the C++ test
sources as text and parses the numeric arrays out of them):

    python tests/golden/extract_reference_golds.py /root/reference

Sources (reference paths):
  unit_tests/edge_kernels/UnitTestContinuityAdvEdge.C:16-106
  unit_tests/edge_kernels/UnitTestScalarAdvDiffEdge.C:20-90 (serial golds)
  unit_tests/edge_kernels/UnitTestMomentumAdvDiffEdge.C:16-230
  unit_tests/UnitTestPecletFunction.C:36-110
  unit_tests/ngp_algorithms/UnitTestGeometryAlg.C:20-100
  unit_tests/ngp_algorithms/UnitTestMdotAlg.C:21-78 (MdotEdgeAlg on one
    unit-cube Hex8: the field fills and the expected edge value)

Only numbers are kept.  The inputs that generate them (one unit-cube Hex8,
the trigonometric test fields, scheme options) are restated in
tests/fixtures.py from the same tests.
"""
import json
import math
import os
import re
import sys


def _numbers(text):
    text = re.sub(r"//[^\n]*", "", text)
    return [float(x) for x in re.findall(
        r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?", text)]


def _array(src, name):
    """Numbers of `name = {...};` or `name[..] = {...};` (nested braces ok)."""
    m = re.search(r"\b%s\b(?:\s*\[[^\]]*\])*\s*=\s*\{" % re.escape(name), src)
    if not m:
        raise KeyError(name)
    i = m.end() - 1
    depth = 0
    for j in range(i, len(src)):
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
            if depth == 0:
                return _numbers(src[i:j + 1])
    raise ValueError(name)


def main(ref):
    rd = lambda p: open(os.path.join(ref, p)).read()
    cont = rd("unit_tests/edge_kernels/UnitTestContinuityAdvEdge.C")
    scal = rd("unit_tests/edge_kernels/UnitTestScalarAdvDiffEdge.C")
    mom = rd("unit_tests/edge_kernels/UnitTestMomentumAdvDiffEdge.C")
    out = {
        "source": "Exawind/nalu-wind unit tests (numbers only)",
        "continuity_adv_edge": {
            "rhs": _array(cont, "rhs"),
            "lhs": _array(cont, "lhs"),
            "rhs_tol": 1.0e-12,
            "lhs_tol": 1.0e-15,
        },
        "scalar_adv_diff_edge_tpetra": {
            name: _array(scal, name)
            for name in ["rowOffsets_serial", "cols_serial", "vals_serial",
                         "fixed_vals_serial", "dirichlet_vals_serial",
                         "rhs_serial", "fixed_rhs_serial",
                         "dirichlet_rhs_serial"]
        },
        "momentum_adv_diff_edge": {
            "rhs": _array(mom, "rhs"),
            "lhs": _array(mom, "lhs"),
            "tol": 1.0e-12,
            "num_sum_into_calls": 12,
        },
        "peclet_classic": {"A": 5.0, "hybrid": 1.0,
                           "pe": [0.0, 1.0, math.sqrt(5.0), 1e5],
                           "fac": [0.0, 1.0 / 6.0, 0.5, 1.0], "tol": 1e-6},
        "peclet_tanh": {"c1": 5000.0, "c2": 200.0,
                        "x": [-5000.0 - 2000.0, 5000.0, 5000.0 + 2000.0],
                        "fac": [0.0, 0.5, 1.0], "tol": 1e-6},
        "geometry_unit_hex": {"dual_volume": 0.125, "element_volume": 1.0,
                              "area_mag_sq": 0.0625, "num_edges": 12,
                              "tol": 1e-16},
    }
    mdot = rd("unit_tests/ngp_algorithms/UnitTestMdotAlg.C")
    body = mdot[mdot.index("NGP_mdot_calc_edge"):]
    body = body[:body.index("TEST_F", 10)]
    fill = lambda field: float(re.search(
        r"field_fill\(\s*([-+0-9.eE]+)\s*,\s*\*%s\s*\)" % field, body).group(1))
    near = re.search(r"EXPECT_NEAR\(\s*mdot\[0\]\s*,\s*([-+0-9.eE]+)", body)
    out["mdot_edge_unit_hex"] = {
        "density": fill("density_"), "velocity": fill("velocity_"),
        "pressure": fill("pressure_"), "dpdx": fill("dpdx_"),
        "interp_together": 1.0, "mdot": float(near.group(1)),
        "tol": float(re.search(r"tol\s*=\s*([-+0-9.eE]+)", body).group(1)),
        "num_edges": 12}
    out["scalar_adv_diff_edge_tpetra"]["tol"] = 1.0e-14
    out["scalar_adv_diff_edge_tpetra"]["fix_node_id"] = 1
    out["scalar_adv_diff_edge_tpetra"]["fix_ref_pressure"] = 1.0
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "reference_golds.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
