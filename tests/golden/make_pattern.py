"""Extract the 512-point rBRIEF sampling pattern (OpenCV's ``bit_pattern_31_``)
from the reference's ``orbd.cu:97-353`` as plain data.

The table is data that must be identical for descriptor bits to agree; it is
written as a bare list of integers (no source text) to the two ``.inc`` files
that consume it: the product's ``cuda-orb_amd/csrc/orb_pattern31.inc`` and the
oracle's own copy ``oracle/pattern31.inc``.  Run from the repo root while the
reference tree is mounted at /root/reference.
"""
import hashlib
import pathlib
import re
import sys

REF = pathlib.Path("/root/reference/orbd.cu")
ROOT = pathlib.Path(__file__).resolve().parents[2]
EXPECTED_SHA256 = "2164181aea6ff9ac426ca512d5130d15e1f6e3cd47b1cbdd568bbe1e55d49023"


def parse(text):
    i = text.index("bit_pattern_31_[256 * 4] = {")
    j = text.index("};", i)
    body = re.sub(r"/\*.*?\*/", "", text[i:j], flags=re.S).split("{", 1)[1]
    return [int(x) for x in re.findall(r"-?\d+", body)]


def digest(nums):
    return hashlib.sha256(bytes(n & 255 for n in nums)).hexdigest()


def render(nums, header):
    lines = [header]
    for k in range(0, len(nums), 16):
        lines.append(", ".join(str(v) for v in nums[k:k + 16]) + ",")
    return "\n".join(lines) + "\n"


def main():
    nums = parse(REF.read_text())
    assert len(nums) == 1024 and digest(nums) == EXPECTED_SHA256
    hdr = ("/* 512 (x, y) sampling points of the 31x31 rBRIEF pattern (OpenCV bit_pattern_31_),\n"
           "   as data; extracted by tests/golden/make_pattern.py. */")
    for out in (ROOT / "cuda-orb_amd/csrc/orb_pattern31.inc", ROOT / "oracle/pattern31.inc"):
        out.write_text(render(nums, hdr))
        print("wrote", out)


if __name__ == "__main__":
    sys.exit(main())
