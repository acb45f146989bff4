"""Writes tests/golden/ref_table_hashes.json: SHA-256 digests of the lookup
tables hard-coded in the reference (Jure-Z/ASTC_implementation), parsed from
its sources under /root/reference.  Only the digests are committed; the
tests regenerate each table from the ASTC specification formulas and compare
digests.

Tables (reference file:line):
  quant_to_unquant   shaders/pass07_weights_and_error_for_bm.wgsl:12-37
  weight_scramble    code/physical_compression.cpp:249-274
  color_u2q          shaders/pass14_pack_color_endpoints.wgsl:29-319
  color_pquant       code/physical_compression.cpp:284-591
  quant_mode_table   shaders/pass11_best_color_combination_for_mode_1part.wgsl:12-63
  prev_next          shaders/pass16_realign_weights.wgsl:13-190
  integer_of_trits   code/physical_compression.cpp:44-?
  integer_of_quints  code/physical_compression.cpp:?-243
"""
import hashlib
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_table_hashes.json")


def wgsl_array(path, name):
    txt = open(path).read()
    i = txt.index("const " + name)
    j = txt.index("(", txt.index("array", i) + 5)
    k = txt.index(");", j)
    body = re.sub(r"//[^\n]*", "", txt[j + 1:k])
    return [int(x, 0) for x in re.findall(r"-?0x[0-9a-fA-F]+|-?\d+", body)]


def c_array(path, name):
    txt = open(path).read()
    i = txt.index(name)
    j = txt.index("{", i)
    depth, k = 0, j
    while True:
        if txt[k] == "{":
            depth += 1
        elif txt[k] == "}":
            depth -= 1
            if depth == 0:
                break
        k += 1
    body = re.sub(r"//[^\n]*", "", txt[j:k + 1])
    return [int(x, 0) for x in re.findall(r"0x[0-9a-fA-F]+|\d+", body)]


def digest(values):
    return hashlib.sha256(",".join(str(int(v)) for v in values).encode()).hexdigest()


def main():
    sh = os.path.join(REF, "shaders")
    pc = os.path.join(REF, "code", "physical_compression.cpp")
    tables = {
        "quant_to_unquant": wgsl_array(os.path.join(sh, "pass07_weights_and_error_for_bm.wgsl"), "QUANT_TO_UNQUANT"),
        "weight_scramble": c_array(pc, "transfet_tables[12]"),
        "color_u2q": wgsl_array(os.path.join(sh, "pass14_pack_color_endpoints.wgsl"), "COLOR_UNQUANT_TO_QUANT_TABLE"),
        "color_pquant": c_array(pc, "color_uquant_to_scrambled_pquant_tables[17][256]"),
        "quant_mode_table": wgsl_array(os.path.join(sh, "pass11_best_color_combination_for_mode_1part.wgsl"),
                                       "QUANT_MODE_TABLE"),
        "prev_next": wgsl_array(os.path.join(sh, "pass16_realign_weights.wgsl"), "ALL_PREV_NEXT_VALUES"),
        "integer_of_trits": c_array(pc, "integer_of_trits[3][3][3][3][3]"),
        "integer_of_quints": c_array(pc, "integer_of_quints[5][5][5]"),
    }
    out = {k: {"count": len(v), "sha256": digest(v)} for k, v in tables.items()}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
