"""Copy the reference's own data fixtures used by these tests into tests/golden/.

cubesphere_vectors.bin is the reference's golden vector file
(bake/tests/data/cubesphere_vectors.bin, format in bake/tests/gen_vectors.py:
magic b"CSV2", int32 K, K records '<idddddidd' = face, u, v, x, y, z, face2,
u2, v2 from the reference's scalar to_sphere / from_sphere under glibc).  It
is data, copied byte for byte, so the tests run without a reference checkout.
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/bake/tests/data/cubesphere_vectors.bin"

if __name__ == "__main__":
    if not os.path.exists(REF):
        sys.exit(f"{REF} not found")
    shutil.copyfile(REF, os.path.join(HERE, "cubesphere_vectors.bin"))
    print("copied", REF)
