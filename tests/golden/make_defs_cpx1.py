"""Extract the reference's complex 1D known-answer vectors as JSON data.

Reads test/unit/suites/defs_cpx1.h of the reference checkout (text only; it is
a table of inputs and expected outputs for N = 2, 3, 5, 7, 330, printed to 4-5
decimals, used by test/unit/suites/api_cpx1.c:157-225 at tolerance 1e-3) and
writes tests/golden/defs_cpx1.json.  Run in the build container only:

    python tests/golden/make_defs_cpx1.py /root/reference
"""
import json
import os
import re
import sys


def parse(text: str) -> dict:
    out = {}
    for m in re.finditer(r"fft(\d+)_(in|out)\s*\[[^\]]*\]\s*=\s*\{(.*?)\};", text, re.S):
        n, kind, body = int(m.group(1)), m.group(2), m.group(3)
        pairs = re.findall(r"\{\s*([-0-9.eE+]+)\s*,\s*([-0-9.eE+]+)\s*\}", body)
        vals = [[float(a), float(b)] for a, b in pairs]
        assert len(vals) == n, (n, kind, len(vals))
        out.setdefault(str(n), {})[kind] = vals
    return out


def main() -> None:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "test", "unit", "suites", "defs_cpx1.h")
    with open(src) as f:
        data = parse(f.read())
    data["_source"] = "test/unit/suites/defs_cpx1.h:5-227 (tolerance 1e-3, api_cpx1.c:10)"
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "defs_cpx1.json")
    with open(dst, "w") as f:
        json.dump(data, f, indent=0)
    print("wrote", dst, sorted(k for k in data if not k.startswith("_")))


if __name__ == "__main__":
    main()
