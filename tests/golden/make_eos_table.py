"""Writes eos_wb_coefficients.json: the 13 + 13 coefficients of the
Wuppertal-Budapest rational fit of calculatePressure (EOS_TYPE 2,
src/EquationOfState.cpp:20-46 of derekeverett/freestream-milne), read from the
reference's source text as data.  Run in a checkout of the reference:

    python tests/golden/make_eos_table.py /root/reference/src/EquationOfState.cpp
"""
import json
import os
import re
import sys


def main(src):
    text = open(src).read()
    coef = {}
    for name, val in re.findall(r"float\s+([ab]\d+)\s*=\s*([-+0-9.eE]+)\s*;", text):
        coef[name] = float(val)
    a = [coef["a%d" % k] for k in range(13)]
    b = [coef["b%d" % k] for k in range(13)]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eos_wb_coefficients.json")
    with open(out, "w") as f:
        json.dump({"source": "src/EquationOfState.cpp:20-46", "a": a, "b": b}, f, indent=1)
        f.write("\n")
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/EquationOfState.cpp")
