"""Writes the golden fixtures of tests/golden/ from the reference tree.

* philox_ref.json: draws of the reference's own openrand::Philox header
  (src/PatchOperation/Random/philox.h), compiled by `make -C oracle ref` into
  oracle/_ref/philox_ref from the sources where they lie.
* bond-19.190-3.800.npz: the data of the reference's bond-table fixture
  Unit_test/Compute/bond-19.190-3.800.dat (two columns, X and Y).

Run in a container that has /root/reference: python tests/golden/make_golden.py
"""
import json
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = Path("/root/reference")


def main():
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], check=True)
    out = subprocess.run([str(ROOT / "oracle" / "_ref" / "philox_ref")], check=True, capture_output=True, text=True)
    data = json.loads(out.stdout)
    (HERE / "philox_ref.json").write_text(json.dumps(data, indent=1) + "\n")
    tab = np.loadtxt(REF / "Unit_test" / "Compute" / "bond-19.190-3.800.dat")
    np.savez_compressed(HERE / "bond-19.190-3.800.npz", X=tab[:, 0], Y=tab[:, 1])
    print(f"wrote {len(data)} philox records and a {len(tab)}-row bond table")


if __name__ == "__main__":
    main()
