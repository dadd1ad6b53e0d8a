"""Copies the input data of Macaw's Butadiene example (Examples/
Butadiene_5modes_CMF: the CMF2 input deck, the Hamiltonian and expectation
operator files and the two tabulated potentials) into butadiene_5modes/,
and the Dormand-Prince deck of Examples/Butadiene_5modes as
but_5D_dopri.inp.  Data only: the reference's source is not copied.

    python tests/golden/copy_butadiene.py [REFERENCE_ROOT]
"""
import os
import shutil
import sys

FILES = ("but_5D.inp", "ham_interp.op", "e1.op", "e12.op", "q1.op", "q2.interp", "qcoup.interp")


def main(root="/root/reference"):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "butadiene_5modes")
    os.makedirs(out, exist_ok=True)
    src = os.path.join(root, "Examples", "Butadiene_5modes_CMF")
    for f in FILES:
        shutil.copyfile(os.path.join(src, f), os.path.join(out, f))
    shutil.copyfile(os.path.join(root, "Examples", "Butadiene_5modes", "but_5D.inp"),
                    os.path.join(out, "but_5D_dopri.inp"))


if __name__ == "__main__":
    main(*sys.argv[1:])
