"""Writes tests/golden/wire/*.bin: the cereal binary serialisation of fixed
OnlineSpikesPayload / SorterParameters values, produced by the reference's own
headers (src/Networking/onlineSpikesPayload.h, sorterParameters.h) and its
vendored cereal, compiled by ``make -C oracle ref`` (LP64 build: `long` is
8 bytes).  Run in a container that holds the reference sources:

    python tests/golden/make_wire_vectors.py
"""
import shutil
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]


def main() -> None:
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], check=True)
    out = ROOT / "tests" / "golden" / "wire"
    if out.exists():
        shutil.rmtree(out)
    out.mkdir(parents=True)
    subprocess.run([str(ROOT / "oracle" / "_ref" / "ref_wire"), str(out)], check=True)


if __name__ == "__main__":
    main()
