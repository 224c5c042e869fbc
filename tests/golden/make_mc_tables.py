"""Write tests/golden/mc_tables.json: the three marching-cubes tables as data.

Reads the reference's table text (cu_marching_cubes_v1.cu:471-505 edgeTable,
:509-767 triTable[256][16], :770-1028 numVertsTable) from /root/reference and
stores the integers as a JSON fixture.  The fixture is data (the classic
Lorensen-Cline / Bourke case tables the reference uploads at :406-411); it is
checked by tests/test_tables.py against sha256 prefixes of the parsed lists and
against the case invariants the tables must satisfy.

Run only in a container that has the reference mounted:
    python tests/golden/make_mc_tables.py
"""
import hashlib
import json
import os
import re
import sys

REF = "/root/reference/cu_marching_cubes_v1.cu"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mc_tables.json")


def _grab(src, name):
    m = re.search(r"const int " + name + r"\s*\[[^\]]*\](\[[^\]]*\])?\s*=", src)
    if m is None:
        raise SystemExit(f"table {name} not found in {REF}")
    j = src.index("{", m.end())
    depth = 0
    for k in range(j, len(src)):
        if src[k] == "{":
            depth += 1
        elif src[k] == "}":
            depth -= 1
            if depth == 0:
                break
    body = re.sub(r"//[^\n]*", "", src[j:k + 1])
    return [int(t, 0) for t in re.findall(r"-?0x[0-9a-fA-F]+|-?\d+", body)]


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} is not present; the committed fixture is authoritative")
    src = open(REF).read()
    edge = _grab(src, "edgeTable")
    tri = _grab(src, "triTable")
    num = _grab(src, "numVertsTable")
    assert len(edge) == 256 and len(tri) == 4096 and len(num) == 256
    tri2 = [tri[i * 16:(i + 1) * 16] for i in range(256)]
    doc = {
        "source": "cu_marching_cubes_v1.cu:471-1028",
        "edgeTable": edge,
        "triTable": tri2,
        "numVertsTable": num,
        "sha256_16": {
            "edgeTable": hashlib.sha256(str(edge).encode()).hexdigest()[:16],
            "triTable": hashlib.sha256(str(tri2).encode()).hexdigest()[:16],
            "numVertsTable": hashlib.sha256(str(num).encode()).hexdigest()[:16],
        },
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, separators=(",", ":"))
        f.write("\n")
    print("wrote", OUT, doc["sha256_16"])


if __name__ == "__main__":
    main()
