"""Copy the reference's element-test INP meshes (data files of its test suite,
tests/element_test/C3D4.inp, C3D8.inp, C3D10.inp) into tests/golden/inp/ as
gzip fixtures, so INP ingestion is tested without the reference present.

Run once where the reference is mounted:  python tests/golden/copy_inp_fixtures.py
"""
import gzip
import os
import shutil

SRC = '/root/reference/tests/element_test'
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'inp')

if __name__ == '__main__':
    os.makedirs(DST, exist_ok=True)
    for name in ('C3D4', 'C3D8', 'C3D10'):
        with open(os.path.join(SRC, name + '.inp'), 'rb') as fi, \
                gzip.GzipFile(os.path.join(DST, name + '.inp.gz'), 'wb', mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        print('wrote', os.path.join(DST, name + '.inp.gz'))
