"""Copies the reference's end-to-end example data into fixtures (run in the container that holds
the reference; the fixtures travel, the reference does not).

* grids: examples/grids/naca0004_129x129.plot3d and naca0012_129x129.plot3d, gzipped as they are;
* expected outputs: examples/response/Ma0p05.out, Ma0p5.out and examples/gain/Ma0p*.out, the
  texts the reference's CI compares byte for byte (.github/workflows/tests.yml:47-61).
"""

import gzip
import shutil
from pathlib import Path

REF = Path("/root/reference/examples")
OUT = Path(__file__).resolve().parent / "examples"


def main() -> None:
    OUT.mkdir(exist_ok=True)
    for name in ("naca0004_129x129.plot3d", "naca0012_129x129.plot3d"):
        with open(REF / "grids" / name, "rb") as src, \
                gzip.GzipFile(OUT / (name + ".gz"), "wb", mtime=0) as dst:
            shutil.copyfileobj(src, dst)
    for case in ("Ma0p05", "Ma0p5"):
        shutil.copyfile(REF / "response" / f"{case}.out", OUT / f"response_{case}.out")
    for src in sorted((REF / "gain").glob("Ma0p*.out")):
        shutil.copyfile(src, OUT / f"gain_{src.name}")


if __name__ == "__main__":
    main()
