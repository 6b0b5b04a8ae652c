"""Writes fdm_matrix_N4.json: the 16x16 expected matrix held by the
reference's own test (src/tests/fdm_matrix.cpp:47-64, the periodic 5-point
operator for a 2D N=4 lattice).  Run from the repo root with the reference
checked out at /root/reference; the JSON it writes is committed, so the test
suite never needs the reference."""
import json
import re
import sys

src_path = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/tests/fdm_matrix.cpp"
src = open(src_path).read()
blk = src[src.index("correct_matrix[] = {") + len("correct_matrix[] = {"):]
blk = blk[: blk.index("};")]
vals = [int(v) for v in re.findall(r"-?\d+", blk)]
assert len(vals) == 256
rows = ",\n".join("  " + json.dumps(vals[r * 16:(r + 1) * 16]) for r in range(16))
with open("tests/golden/fdm_matrix_N4.json", "w") as f:
    f.write('{"source": "reference src/tests/fdm_matrix.cpp:47-64 (correct_matrix, '
            '2D N=4, stencil_setting 0)",\n "N": 4, "ndims": 2,\n "matrix": [\n'
            + rows + "\n ]}\n")
