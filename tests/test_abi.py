"""The C ABI without a GPU: the library loads, exports every entry point of include/crvhip.h, the ctypes structs
match the header's layout, and the generated math constants equal the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

from crvhip import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "crvhip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^int (crv_\w+)\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libcrvhip.so not built")
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.EXPORTS)


def test_struct_field_counts_match_header():
    text = open(HEADER).read()
    body = re.search(r"typedef struct crv_params \{(.*?)\} crv_params;", text, re.S).group(1)
    fields = []
    for line in body.splitlines():
        line = line.split("/*")[0].strip()
        if not line or line.startswith("/*"):
            continue
        m = re.match(r"(u?int\d+_t|double)\s+(.*);", line)
        if m:
            fields += [f.strip() for f in m.group(2).split(",")]
    assert [f for f, _ in L.Params._fields_] == fields
    assert ctypes.sizeof(L.Params) % 8 == 0


def test_missing_library_fails_loudly(tmp_path):
    L._lib_saved = L._lib
    try:
        L._lib = None
        with pytest.raises(RuntimeError, match="not built"):
            L.load(str(tmp_path / "nope.so"))
    finally:
        L._lib = L._lib_saved


def test_generated_constants_equal_the_oracle():
    from oracle import fmath as F, rand as R
    text = open(os.path.join(ROOT, "mu2e-fastsim_amd", "csrc", "fmath_consts.h")).read()
    vals = {}
    for name, body in re.findall(r"#define (FM_\w+) (.+)", text):
        body = body.strip()
        if body.startswith("{"):
            vals[name] = [float.fromhex(x.strip()) for x in body.strip("{}").split(",")]
        elif not name.endswith("_N"):
            vals[name] = float.fromhex(body)
    c = F.constants()
    for k in ("LN2_HI", "LN2_LO", "INV_LN2", "PIO2_1", "PIO2_2", "PIO2_3", "TWO_OVER_PI", "PI_2", "TWO_PI",
              "SQRT_HALF", "EXP_HI", "EXP_LO"):
        assert vals["FM_" + k] == c[k], k
    for k in ("EXP_C", "LOG_C", "SIN_C", "COS_C", "ASIN_C"):
        assert vals["FM_" + k + "_INIT"] == c[k], k
    assert vals["FM_STIRLING_INIT"] == R.STIRLING and vals["FM_FACT_INIT"] == R.FACT
    assert vals["FM_LN_SQRT_2PI"] == R.LN_SQRT_2PI and vals["FM_E"] == np.e


def _header_fields(struct):
    text = open(HEADER).read()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        m = re.match(r"(?:const )?(?:u?int\d+_t|double)\s*(.*)", decl)
        if m:
            fields += [f.strip().lstrip("*").strip() for f in m.group(1).split(",")]
    return fields


@pytest.mark.parametrize("struct,cls", [("crv_batch_opts", "BatchOpts"), ("crv_timing", "Timing"),
                                        ("crv_pulses", "Pulses"), ("crv_clusters", "Clusters")])
def test_ctypes_structs_follow_the_header(struct, cls):
    assert [f for f, _ in getattr(L, cls)._fields_] == _header_fields(struct)
