"""Write tests/golden/synth_sha256.json: SHA-256 of the seeded synthetic images
(orb_synth_image / orb_synth_warp) for the workload sizes, so a change of the
generator's bytes on any host is caught.  Run from the repo root after build."""
import hashlib
import json
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "cuda-orb_amd"))
import orb_amd  # noqa: E402

CASES = [(0, 640, 480, 40), (1, 1920, 1080, 80), (3, 1280, 720, 120), (7, 333, 257, 8)]


def main():
    out = []
    for seed, w, h, n in CASES:
        a = orb_amd.synth_image(seed, w, h, n)
        b, H = orb_amd.synth_warp(seed, a, w, h)
        out.append({"seed": seed, "w": w, "h": h, "nshapes": n,
                    "a": hashlib.sha256(a.tobytes()).hexdigest(),
                    "b": hashlib.sha256(b.tobytes()).hexdigest(),
                    "H": [float(v) for v in H.ravel()]})
    (ROOT / "tests/golden/synth_sha256.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
