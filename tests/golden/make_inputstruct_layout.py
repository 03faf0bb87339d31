"""Writes inputstruct_layout.json: the member list (name, ctypes type) of the parameter block the reference's
Python front end packs for omegaMain (class ``parameters``, _pack_ = 1,
source/Python/omegatomo/projector/proj.py:1959-2282), read as text from the reference sources.

The fixture is the binary interface omegaMain must accept; tests/test_omega_main_layout.py checks
include/omega_main.h against it. Run from the repository root:  python tests/golden/make_inputstruct_layout.py
"""
import json
import os
import re

SRC = "/root/reference/source/Python/omegatomo/projector/proj.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "inputstruct_layout.json")


def main():
    text = open(SRC).read()
    start = text.index("class parameters(ctypes.Structure):")
    block = text[start:text.index("]", text.index("_fields_", start))]
    assert re.search(r"_pack_\s*=\s*1", block), "the block is expected to be packed to one byte"
    members = []
    for m in re.finditer(r"\(\s*'(\w+)'\s*,\s*ctypes\.(POINTER\(\s*ctypes\.(\w+)\s*\)|\w+)\s*\)", block):
        name, kind, pointee = m.group(1), m.group(2), m.group(3)
        members.append([name, f"*{pointee}" if pointee else kind])
    with open(OUT, "w") as f:
        json.dump({"source": "proj.py class parameters", "pack": 1, "members": members}, f, indent=0)
    print(f"{len(members)} members -> {OUT}")


if __name__ == "__main__":
    main()
